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
// queue on a mutex). A copy's parts are set up only when no thread is still inside the previous
// copy's loop, and a thread enters a copy's loop only under the lock that published it.
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

    // the pieces cut into parts of at most PART_BYTES, copied by the pool and the caller
    void copy(const CopyPiece* pieces, size_t count)
    {
        std::lock_guard<std::mutex> job(itsJobMutex);
        {
            std::unique_lock<std::mutex> lk(itsMutex);
            itsIdle.wait(lk, [&] { return itsBusy == 0; });
            itsParts.clear();
            for (size_t i = 0; i < count; ++i)
                for (size_t off = 0; off < pieces[i].bytes; off += PART_BYTES)
                    itsParts.push_back({static_cast<uint8_t*>(pieces[i].dst) + off,
                                        static_cast<const uint8_t*>(pieces[i].src) + off,
                                        std::min(PART_BYTES, pieces[i].bytes - off)});
            itsNext.store(0);
            itsDone.store(0);
            itsGeneration++;
        }
        itsWake.notify_all();
        run();
        std::unique_lock<std::mutex> lk(itsMutex);
        itsFinished.wait(lk, [&] { return itsDone.load() == itsParts.size(); });
    }

private:
    // takes parts until none is left
    void run()
    {
        const size_t count = itsParts.size();
        for (;;)
        {
            const size_t p = itsNext.fetch_add(1);
            if (p >= count)
                return;
            const CopyPiece& part = itsParts[p];
            std::memcpy(part.dst, part.src, part.bytes);
            if (itsDone.fetch_add(1) + 1 == count)
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
                itsBusy++;
            }
            run();
            {
                std::lock_guard<std::mutex> lk(itsMutex);
                if (--itsBusy == 0)
                    itsIdle.notify_all();
            }
        }
    }

    std::vector<std::thread> itsThreads;
    std::mutex itsJobMutex;
    std::mutex itsMutex;
    std::condition_variable itsWake, itsFinished, itsIdle;
    bool itsStop           = false;
    uint64_t itsGeneration = 0;
    uint32_t itsBusy       = 0;  // pool threads inside run()
    std::vector<CopyPiece> itsParts;
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
    const CopyPiece piece{dst, src, bytes};
    pool().copy(&piece, 1);
}

void parallelMemcpy(const CopyPiece* pieces, size_t count)
{
    size_t bytes = 0;
    for (size_t i = 0; i < count; ++i)
        bytes += pieces[i].bytes;
    if (bytes < PARALLEL_MIN)
    {
        for (size_t i = 0; i < count; ++i)
            std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
        return;
    }
    pool().copy(pieces, count);
}

}  // namespace oc
