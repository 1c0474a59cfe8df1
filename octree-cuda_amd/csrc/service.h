// Build services: continuous batching of octree builds on one GPU. The reference has no
// counterpart: it builds one cloud per process (src/main.cpp:55-109).
//
// Requests (configured sessions, BuildRequest) wait in one FIFO queue and are admitted against a
// device-memory budget by their estimated footprint (E = N * output factor drives it). Two kinds:
//
// BuildService: a fixed pool of slots, each a worker thread that owns a SlotBuilder (on the GPU an
// OctreeBuilder: its own HIP stream and a grow-only device arena that stays resident between
// requests). A free slot takes the head of the queue when the projected footprint of all slots
// stays within the budget; builds on different slots overlap on the device. A slot's finished
// build is copied to host memory while the slot goes on with its next request, and is handed to
// its session as soon as the copies have landed. An idle slot gives its arena back when the head
// of the queue does not fit.
//
// BatchService: a few workers, each with a BatchBuilder. A free worker takes every queued request
// that fits the budget (up to a batch size) and builds them together, one launch per stage
// for all of them; requests that arrive meanwhile wait for the next batch, which another worker
// may already be uploading or building.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "build_request.h"
#include "pinned_pool.h"
#include "slot_builder.h"

namespace oc {

// What the C ABI's service entry points use.
class ServiceBase
{
public:
    virtual ~ServiceBase() = default;
    static ServiceBase* fromHandle(void* handle);

    // Validates the request's configuration and queues it. The request (and its cloud buffer)
    // must stay untouched until wait() has returned for it.
    virtual void submit(BuildRequest* request) = 0;
    // Waits until the request's build has finished; false on timeout. Throws with the build's
    // error message when it failed.
    virtual bool wait(BuildRequest* request, uint32_t timeoutMs) = 0;
    // Removes a queued request, or waits for its running build (a request being destroyed).
    virtual void forget(BuildRequest* request) = 0;
    // Workers that take requests (1..workers); the rest stay idle.
    virtual void setActiveSlots(uint32_t active) = 0;
    // [0] submitted [1] completed [2] failed [3] queued now [4] running now [5] max queue depth
    // [6] p50 latency ms [7] p99 latency ms [8] mean latency ms [9] reserved device bytes
    // [10] memory budget [11] active workers [12] workers [13] mean queue wait ms [14] max latency ms
    // [15] peak reserved device bytes [16] [17] [18] mean host ms per completed request in setup,
    // generate and result capture (slot service) [19] batches built [20] mean requests per batch
    virtual std::vector<double> stats() = 0;
    virtual void resetStats() = 0;
};

// Bookkeeping both services share: queue, admission counters, latencies.
class ServiceCore : public ServiceBase
{
public:
    bool wait(BuildRequest* request, uint32_t timeoutMs) override;
    void forget(BuildRequest* request) override;
    std::vector<double> stats() override;
    void resetStats() override;

protected:
    ServiceCore(uint64_t memoryBudget, uint32_t workers);
    void enqueue(BuildRequest* request, uint64_t footprint);
    // under the lock: marks a request finished (latency recorded) or failed
    void finishLocked(BuildRequest* request, const std::string& error, const double phase[3]);
    virtual uint64_t reservedLocked() const = 0;
    virtual uint32_t runningLocked() const  = 0;

    uint64_t itsBudget;
    uint32_t itsWorkers;
    uint32_t itsActive;
    std::deque<BuildRequest*> itsQueue;
    std::mutex itsMutex;
    std::condition_variable itsWork;  // workers wait for requests or memory
    std::condition_variable itsDone;  // callers wait for their requests
    bool itsStop = false;

    uint64_t itsSubmitted = 0, itsCompleted = 0, itsFailed = 0, itsMaxQueue = 0, itsPeakReserved = 0;
    uint64_t itsBatches = 0, itsBatchedRequests = 0;
    double itsPhaseMs[3] = {0, 0, 0};  // summed host time of completed requests: setup, generate, capture
    std::vector<double> itsLatencies, itsWaits;
};

class BuildService final : public ServiceCore
{
public:
    BuildService(int device, uint32_t slots, uint64_t memoryBudget);
    ~BuildService() override;
    BuildService(const BuildService&) = delete;
    BuildService& operator=(const BuildService&) = delete;

    void submit(BuildRequest* request) override;
    void setActiveSlots(uint32_t active) override;

private:
    struct Slot
    {
        std::unique_ptr<SlotBuilder> builder;
        uint64_t reserved = 0;  // device bytes the slot holds or will hold for its current request
        bool busy         = false;
        std::thread worker;
    };

    void run(uint32_t index);
    uint64_t reservedExcept(uint32_t index) const;
    uint64_t reservedLocked() const override { return reservedExcept(~0u); }
    uint32_t runningLocked() const override;

    std::vector<std::unique_ptr<Slot>> itsSlots;
    std::shared_ptr<PinnedPool> itsPinned;  // results are copied into page-locked buffers
};

class BatchService final : public ServiceCore
{
public:
    BatchService(int device, uint32_t workers, uint32_t maxBatch, uint64_t memoryBudget);
    ~BatchService() override;
    BatchService(const BatchService&) = delete;
    BatchService& operator=(const BatchService&) = delete;

    void submit(BuildRequest* request) override;
    void setActiveSlots(uint32_t active) override;

private:
    struct Worker
    {
        std::unique_ptr<BatchSlotBuilder> builder;
        uint64_t reserved = 0;
        bool busy         = false;
        std::thread thread;
    };

    void run(uint32_t index);
    uint64_t reservedExcept(uint32_t index) const;
    uint64_t reservedLocked() const override { return reservedExcept(~0u); }
    uint32_t runningLocked() const override;

    uint32_t itsMaxBatch;
    std::vector<std::unique_ptr<Worker>> itsWorkersList;
    std::shared_ptr<PinnedPool> itsPinned;
};

}  // namespace oc
