#include "pinned_pool.h"

#include <stdexcept>
#include <string>

#include "slot_builder.h"

namespace oc {

static size_t sizeClass(size_t bytes)
{
    size_t c = size_t(1) << 20;
    while (c < bytes)
        c <<= 1;
    return c;
}

PinnedPool::~PinnedPool()
{
    for (auto& kv : itsFree)
        pinnedRelease(kv.second);
}

std::shared_ptr<uint8_t> PinnedPool::acquire(size_t bytes)
{
    const size_t capacity = sizeClass(bytes);
    void* p               = nullptr;
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        auto it = itsFree.find(capacity);
        if (it != itsFree.end())
        {
            p = it->second;
            itsFree.erase(it);
            itsCached -= capacity;
        }
    }
    if (!p)
        p = pinnedAllocate(capacity);  // throws on failure
    std::shared_ptr<PinnedPool> self = shared_from_this();
    return std::shared_ptr<uint8_t>(static_cast<uint8_t*>(p),
                                    [self, capacity](uint8_t* q) { self->recycle(q, capacity); });
}

void PinnedPool::recycle(void* p, size_t capacity)
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        if (itsCached + capacity <= itsMaxCached)
        {
            itsFree.emplace(capacity, p);
            itsCached += capacity;
            return;
        }
    }
    pinnedRelease(p);
}

}  // namespace oc
