#include "host_copy.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace oc {

namespace {

constexpr size_t PARALLEL_MIN = 8u << 20;  // below this one thread copies
constexpr size_t PART_BYTES   = 2u << 20;  // parts handed to the threads

// A fixed pool of threads that take parts of one copy at a time (copies from several callers
// queue on a mutex).
class CopyPool
{
public:
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n  = std::max(1u, std::min(8u, hw > 1 ? hw - 1 : 1u));
        for (unsigned i = 0; i < n; ++i)
            itsThreads.emplace_back([this] { work(); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(itsMutex);
            itsStop = true;
        }
        itsWake.notify_all();
        for (auto& t : itsThreads)
            t.join();
    }

    void copy(void* dst, const void* src, size_t bytes)
    {
        std::lock_guard<std::mutex> job(itsJobMutex);
        {
            std::lock_guard<std::mutex> lk(itsMutex);
            itsDst   = static_cast<uint8_t*>(dst);
            itsSrc   = static_cast<const uint8_t*>(src);
            itsBytes = bytes;
            itsParts = (bytes + PART_BYTES - 1) / PART_BYTES;
            itsNext.store(0);
            itsDone.store(0);
            itsGeneration++;
        }
        itsWake.notify_all();
        run();
        std::unique_lock<std::mutex> lk(itsMutex);
        itsFinished.wait(lk, [&] { return itsDone.load() == itsParts; });
    }

private:
    // takes parts until none is left
    void run()
    {
        for (;;)
        {
            const size_t p = itsNext.fetch_add(1);
            if (p >= itsParts)
                return;
            const size_t off = p * PART_BYTES;
            std::memcpy(itsDst + off, itsSrc + off, std::min(PART_BYTES, itsBytes - off));
            if (itsDone.fetch_add(1) + 1 == itsParts)
            {
                std::lock_guard<std::mutex> lk(itsMutex);
                itsFinished.notify_all();
            }
        }
    }

    void work()
    {
        uint64_t seen = 0;
        for (;;)
        {
            {
                std::unique_lock<std::mutex> lk(itsMutex);
                itsWake.wait(lk, [&] { return itsStop || itsGeneration != seen; });
                if (itsStop)
                    return;
                seen = itsGeneration;
            }
            run();
        }
    }

    std::vector<std::thread> itsThreads;
    std::mutex itsJobMutex;
    std::mutex itsMutex;
    std::condition_variable itsWake, itsFinished;
    bool itsStop           = false;
    uint64_t itsGeneration = 0;
    uint8_t* itsDst        = nullptr;
    const uint8_t* itsSrc  = nullptr;
    size_t itsBytes        = 0;
    size_t itsParts        = 0;
    std::atomic<size_t> itsNext{0}, itsDone{0};
};

CopyPool& pool()
{
    static CopyPool p;
    return p;
}

}  // namespace

void parallelMemcpy(void* dst, const void* src, size_t bytes)
{
    if (bytes < PARALLEL_MIN)
    {
        std::memcpy(dst, src, bytes);
        return;
    }
    pool().copy(dst, src, bytes);
}

}  // namespace oc
