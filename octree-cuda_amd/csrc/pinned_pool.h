// Page-locked host buffers for device-to-host result copies (the copy engines reach full PCIe
// speed only into pinned memory), recycled by size class. Buffers hold a reference to the pool,
// so a result can outlive the service that produced it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>

namespace oc {

class PinnedPool : public std::enable_shared_from_this<PinnedPool>
{
public:
    explicit PinnedPool(size_t maxCachedBytes) : itsMaxCached(maxCachedBytes) {}
    ~PinnedPool();
    PinnedPool(const PinnedPool&) = delete;
    PinnedPool& operator=(const PinnedPool&) = delete;

    // A buffer of at least `bytes` bytes; returned to the pool when the last reference goes.
    std::shared_ptr<uint8_t> acquire(size_t bytes);

private:
    void recycle(void* p, size_t capacity);

    size_t itsMaxCached;
    size_t itsCached = 0;
    std::mutex itsMutex;
    std::multimap<size_t, void*> itsFree;  // capacity -> buffer
};

}  // namespace oc
