// BuildService: continuous batching of octree builds on one GPU.
//
// A fixed pool of slots, each a worker thread that owns an OctreeBuilder (its own HIP stream and
// a grow-only device arena that stays resident between requests). Requests (configured sessions)
// wait in one FIFO queue; a free slot takes the head when the projected device footprint of all
// slots stays within the memory budget (the head's footprint is estimated from its point count
// and output factor, E = N * factor). Builds on different slots overlap on the device; each
// finished build is copied to host memory and handed to its session, and the slot takes the next
// request. An idle slot gives its arena back when the head of the queue does not fit.
//
// The reference has no counterpart: it builds one cloud per process (src/main.cpp:55-109).
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "builder.h"
#include "pinned_pool.h"
#include "session.h"

namespace oc {

class BuildService
{
public:
    BuildService(int device, uint32_t slots, uint64_t memoryBudget);
    ~BuildService();
    BuildService(const BuildService&) = delete;
    BuildService& operator=(const BuildService&) = delete;

    static BuildService* fromHandle(void* handle);

    // Validates the session's configuration and queues it. The session (and its cloud buffer)
    // must stay untouched until wait() has returned for it.
    void submit(Session* session);
    // Waits until the session's build has finished; false on timeout. Throws with the build's
    // error message when it failed.
    bool wait(Session* session, uint32_t timeoutMs);

    // Removes a queued session, or waits for its running build (a session being destroyed).
    void forget(Session* session);

    // Slots that take requests (1..slots); the rest stay idle. Used by the slot-count tuner.
    void setActiveSlots(uint32_t active);

    // [0] submitted [1] completed [2] failed [3] queued now [4] running now [5] max queue depth
    // [6] p50 latency ms [7] p99 latency ms [8] mean latency ms [9] reserved device bytes
    // [10] memory budget [11] active slots [12] slots [13] mean queue wait ms [14] max latency ms
    // [15] peak reserved device bytes
    std::vector<double> stats();
    void resetStats();

private:
    struct Slot
    {
        std::unique_ptr<OctreeBuilder> builder;
        uint64_t reserved = 0;  // device bytes the slot holds or will hold for its current request
        bool busy         = false;
        std::thread worker;
    };

    void run(uint32_t index);
    uint64_t reservedExcept(uint32_t index) const;

    int itsDevice;
    uint64_t itsBudget;
    uint32_t itsActive;
    std::vector<std::unique_ptr<Slot>> itsSlots;
    std::shared_ptr<PinnedPool> itsPinned;  // results are copied into page-locked buffers
    std::deque<Session*> itsQueue;
    std::mutex itsMutex;
    std::condition_variable itsWork;  // slots wait for requests or memory
    std::condition_variable itsDone;  // callers wait for their requests
    bool itsStop = false;

    uint64_t itsSubmitted = 0, itsCompleted = 0, itsFailed = 0, itsMaxQueue = 0, itsPeakReserved = 0;
    double itsPhaseMs[3] = {0, 0, 0};  // summed host time of completed requests: setup, generate, capture
    std::vector<double> itsLatencies, itsWaits;
};

}  // namespace oc
