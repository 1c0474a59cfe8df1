#include "service.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

#include "log.h"

namespace oc {

using Clock = std::chrono::steady_clock;

static double milliseconds(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

ServiceBase* ServiceBase::fromHandle(void* handle)
{
    auto* s = static_cast<ServiceBase*>(handle);
    if (!s)
        throw std::runtime_error("No BuildService is currently initialized!");
    return s;
}

// ---- shared bookkeeping -------------------------------------------------------------------------

ServiceCore::ServiceCore(uint64_t memoryBudget, uint32_t workers)
    : itsBudget(memoryBudget), itsWorkers(workers), itsActive(workers)
{
}

void ServiceCore::enqueue(BuildRequest* request, uint64_t footprint)
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        auto& r = request->request();
        if (r.state == BuildRequest::Request::QUEUED || r.state == BuildRequest::Request::RUNNING)
            throw std::runtime_error("the session is already submitted");
        r.error.clear();
        r.owner     = this;
        r.state     = BuildRequest::Request::QUEUED;
        r.footprint = footprint;
        r.submitted = r.started = r.finished = Clock::now();
        itsQueue.push_back(request);
        itsSubmitted++;
        itsMaxQueue = std::max<uint64_t>(itsMaxQueue, itsQueue.size());
    }
    itsWork.notify_all();
}

void ServiceCore::finishLocked(BuildRequest* request, const std::string& error, const double phase[3])
{
    auto& r    = request->request();
    r.finished = Clock::now();
    if (error.empty())
    {
        r.state = BuildRequest::Request::DONE;
        itsCompleted++;
        itsLatencies.push_back(milliseconds(r.submitted, r.finished));
        itsWaits.push_back(milliseconds(r.submitted, r.started));
        for (int p = 0; p < 3; ++p)
            itsPhaseMs[p] += phase[p];
    }
    else
    {
        r.state = BuildRequest::Request::FAILED;
        r.error = error;
        itsFailed++;
        log(LOG_ERROR, "service build failed: " + error);
    }
}

bool ServiceCore::wait(BuildRequest* request, uint32_t timeoutMs)
{
    std::unique_lock<std::mutex> lk(itsMutex);
    auto& r = request->request();
    if (r.state == BuildRequest::Request::NONE)
        throw std::runtime_error("the session was not submitted to this service");
    const bool finished = itsDone.wait_for(lk, std::chrono::milliseconds(timeoutMs), [&] {
        return r.state == BuildRequest::Request::DONE || r.state == BuildRequest::Request::FAILED;
    });
    if (!finished)
        return false;
    if (r.state == BuildRequest::Request::FAILED)
        throw std::runtime_error(r.error);
    return true;
}

void ServiceCore::forget(BuildRequest* request)
{
    std::unique_lock<std::mutex> lk(itsMutex);
    auto& r = request->request();
    auto it = std::find(itsQueue.begin(), itsQueue.end(), request);
    if (it != itsQueue.end())
    {
        itsQueue.erase(it);
        r.state = BuildRequest::Request::FAILED;
        r.error = "the session was destroyed before its build ran";
    }
    itsDone.wait(lk, [&] { return r.state != BuildRequest::Request::RUNNING; });
    r.owner = nullptr;
}

static double percentile(std::vector<double> v, double q)
{
    if (v.empty())
        return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const size_t lo  = static_cast<size_t>(pos);
    const size_t hi  = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}

std::vector<double> ServiceCore::stats()
{
    std::lock_guard<std::mutex> lk(itsMutex);
    double mean = 0, meanWait = 0, maxLatency = 0;
    for (double v : itsLatencies)
    {
        mean += v;
        maxLatency = std::max(maxLatency, v);
    }
    for (double v : itsWaits)
        meanWait += v;
    if (!itsLatencies.empty())
    {
        mean /= itsLatencies.size();
        meanWait /= itsWaits.size();
    }
    const double done = itsCompleted ? static_cast<double>(itsCompleted) : 1.0;
    return {static_cast<double>(itsSubmitted), static_cast<double>(itsCompleted), static_cast<double>(itsFailed),
            static_cast<double>(itsQueue.size()), static_cast<double>(runningLocked()),
            static_cast<double>(itsMaxQueue), percentile(itsLatencies, 0.50), percentile(itsLatencies, 0.99), mean,
            static_cast<double>(reservedLocked()), static_cast<double>(itsBudget), static_cast<double>(itsActive),
            static_cast<double>(itsWorkers), meanWait, maxLatency, static_cast<double>(itsPeakReserved),
            itsPhaseMs[0] / done, itsPhaseMs[1] / done, itsPhaseMs[2] / done, static_cast<double>(itsBatches),
            itsBatches ? static_cast<double>(itsBatchedRequests) / itsBatches : 0.0};
}

void ServiceCore::resetStats()
{
    std::lock_guard<std::mutex> lk(itsMutex);
    itsSubmitted = itsCompleted = itsFailed = 0;
    itsMaxQueue                             = itsQueue.size();
    itsPeakReserved                         = 0;
    itsBatches = itsBatchedRequests = 0;
    itsLatencies.clear();
    itsWaits.clear();
    itsPhaseMs[0] = itsPhaseMs[1] = itsPhaseMs[2] = 0;
}

// ---- slot service -------------------------------------------------------------------------------

BuildService::BuildService(int device, uint32_t slots, uint64_t memoryBudget)
    : ServiceCore(memoryBudget, slots)
{
    if (slots == 0 || slots > 64)
        throw std::runtime_error("a build service needs 1 to 64 slots");
    if (device < 0 || device >= deviceCount())
        throw std::runtime_error("device " + std::to_string(device) + " does not exist");
    itsPinned = std::make_shared<PinnedPool>(size_t(16) << 30);
    for (uint32_t i = 0; i < slots; ++i)
    {
        auto slot     = std::make_unique<Slot>();
        slot->builder = makeSlotBuilder(device);
        itsSlots.push_back(std::move(slot));
    }
    for (uint32_t i = 0; i < slots; ++i)
        itsSlots[i]->worker = std::thread(&BuildService::run, this, i);
}

BuildService::~BuildService()
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        itsStop = true;
        for (BuildRequest* s : itsQueue)
        {
            s->request().state = BuildRequest::Request::FAILED;
            s->request().error = "the build service was destroyed before the request ran";
            s->request().owner = nullptr;
        }
        itsQueue.clear();
    }
    itsWork.notify_all();
    itsDone.notify_all();
    for (auto& slot : itsSlots)
        if (slot->worker.joinable())
            slot->worker.join();
}

uint64_t BuildService::reservedExcept(uint32_t index) const
{
    uint64_t total = 0;
    for (uint32_t i = 0; i < itsSlots.size(); ++i)
        if (i != index)
            total += itsSlots[i]->reserved;
    return total;
}

uint32_t BuildService::runningLocked() const
{
    uint32_t running = 0;
    for (auto& s : itsSlots)
        running += s->busy ? 1 : 0;
    return running;
}

void BuildService::submit(BuildRequest* request)
{
    request->validate();
    const uint64_t footprint =
            slotFootprint(request->cloudInfo(), request->processingInfo(), request->hostCloud() != nullptr);
    if (footprint > itsBudget)
        throw std::runtime_error("the build needs about " + std::to_string(footprint) +
                                 " device bytes, more than the service's memory budget of " +
                                 std::to_string(itsBudget));
    enqueue(request, footprint);
}

void BuildService::setActiveSlots(uint32_t active)
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        itsActive = std::max<uint32_t>(1, std::min<uint32_t>(active, static_cast<uint32_t>(itsSlots.size())));
    }
    itsWork.notify_all();
}

// A slot builds one request while the snapshot copies of the one before it finish: after a
// build, its copies are enqueued on the builder's copy stream and the slot goes on with the next
// request (the cloud upload and the count stage overlap the copies; the build waits for them
// before its merge). The finished request is handed to its session as soon as its copies have
// landed: the slot looks after its setup and between the stages of the next build, and at the
// latest when the slot has nothing else to do.
void BuildService::run(uint32_t index)
{
    Slot& slot = *itsSlots[index];

    struct Done
    {
        BuildRequest* session = nullptr;
        std::unique_ptr<HostBuildResult> result;  // valid once the builder's snapshot copies are done
        std::string error;
        double phase[3] = {0, 0, 0};  // setup, generate, capture (ms)
    } pending;

    // publishes `pending` (call without the lock; the lock is taken for the bookkeeping)
    auto publish = [&](bool stillBusy) {
        if (!pending.session)
            return;
        if (pending.result)
        {
            const auto t = Clock::now();
            try
            {
                slot.builder->waitSnapshot();
                pending.phase[2] += milliseconds(t, Clock::now());
                pending.session->adoptResult(std::move(pending.result));
            }
            catch (const std::exception& e)
            {
                pending.error = e.what();
                slot.builder->drainCopies();  // nothing may still write into the block when it is freed
                pending.result.reset();
            }
        }
        std::lock_guard<std::mutex> g(itsMutex);
        finishLocked(pending.session, pending.error, pending.phase);
        pending = Done{};
        if (!stillBusy)
        {
            slot.busy     = false;
            slot.reserved = slot.builder->arenaBytes();
        }
        itsDone.notify_all();
        itsWork.notify_all();
    };
    // the pending request, if its copies have landed (between the stages of the next build)
    auto publishIfCopied = [&]() {
        if (pending.session && (!pending.result || slot.builder->snapshotDone()))
            publish(true);
    };
    slot.builder->setProgress(publishIfCopied);

    std::unique_lock<std::mutex> lk(itsMutex);
    for (;;)
    {
        if (itsStop)
        {
            lk.unlock();
            publish(false);
            return;
        }
        if (index >= itsActive && (slot.reserved > 0 || pending.session))
        {
            // an inactive slot keeps no memory
            lk.unlock();
            publish(false);
            slot.builder->releaseArena();
            lk.lock();
            slot.reserved = 0;
            itsWork.notify_all();
            continue;
        }
        BuildRequest* session = nullptr;
        bool releaseFirst     = false;
        if (index < itsActive && !itsQueue.empty())
        {
            BuildRequest* head  = itsQueue.front();
            const uint64_t need = head->request().footprint;
            const uint64_t held = slot.builder->arenaBytes();
            const uint64_t rest = reservedExcept(index);
            if (rest + std::max(held, need) <= itsBudget)
            {
                session = head;
            }
            else if (held > 0 && rest + need <= itsBudget)
            {
                session      = head;
                releaseFirst = true;
            }
            else if (held > 0)
            {
                // the head does not fit while this idle slot holds memory: give it back
                lk.unlock();
                publish(false);
                slot.builder->releaseArena();
                lk.lock();
                slot.reserved = 0;
                itsWork.notify_all();
                continue;
            }
            if (session)
            {
                itsQueue.pop_front();
                slot.busy     = true;
                slot.reserved = releaseFirst ? need : std::max(held, need);
                session->request().state   = BuildRequest::Request::RUNNING;
                session->request().started = Clock::now();
                itsPeakReserved            = std::max(itsPeakReserved, rest + slot.reserved);
            }
        }
        if (!session)
        {
            if (pending.session)
            {
                lk.unlock();
                publish(false);
                lk.lock();
                continue;
            }
            itsWork.wait(lk);
            continue;
        }
        lk.unlock();

        Done next;
        next.session = session;
        try
        {
            if (releaseFirst)
            {
                publish(true);  // the copies read the arena that is about to be freed
                slot.builder->releaseArena();
            }
            const auto t0 = Clock::now();
            slot.builder->setup(session->cloudInfo(), session->processingInfo(), session->hostCloud(),
                                session->deviceCloud());
            next.phase[0] = milliseconds(t0, Clock::now());
        }
        catch (const std::exception& e)
        {
            next.error = e.what();
        }
        catch (...)
        {
            next.error = "unknown error";
        }
        publishIfCopied();
        if (next.error.empty())
        {
            try
            {
                const auto t1 = Clock::now();
                slot.builder->generate();
                next.phase[1] = milliseconds(t1, Clock::now());
            }
            catch (const std::exception& e)
            {
                next.error = e.what();
            }
            catch (...)
            {
                next.error = "unknown error";
            }
        }
        publish(true);  // the previous request, if the build did not already (its copies are done by now)
        if (next.error.empty())
        {
            try
            {
                const auto t2  = Clock::now();
                uint8_t* block = nullptr;
                next.result    = HostBuildResult::prepare(slot.builder->lastBuild(), itsPinned.get(), &block);
                slot.builder->snapshotAsync(block);
                next.phase[2] = milliseconds(t2, Clock::now());
            }
            catch (const std::exception& e)
            {
                next.error = e.what();
                slot.builder->drainCopies();  // copies enqueued before the failure may still write the block
                next.result = nullptr;
            }
            catch (...)
            {
                next.error = "unknown error";
                slot.builder->drainCopies();
                next.result = nullptr;
            }
        }
        pending = std::move(next);
        lk.lock();
    }
}

// ---- batch service -------------------------------------------------------------------------------

BatchService::BatchService(int device, uint32_t workers, uint32_t maxBatch, uint64_t memoryBudget)
    : ServiceCore(memoryBudget, workers), itsMaxBatch(maxBatch)
{
    if (workers == 0 || workers > 16)
        throw std::runtime_error("a batch service needs 1 to 16 workers");
    if (maxBatch == 0)
        throw std::runtime_error("the batch size must be at least 1");
    if (device < 0 || device >= deviceCount())
        throw std::runtime_error("device " + std::to_string(device) + " does not exist");
    itsPinned = std::make_shared<PinnedPool>(size_t(32) << 30);
    for (uint32_t i = 0; i < workers; ++i)
    {
        auto w     = std::make_unique<Worker>();
        w->builder = makeBatchSlotBuilder(device, itsPinned);
        itsWorkersList.push_back(std::move(w));
    }
    for (uint32_t i = 0; i < workers; ++i)
        itsWorkersList[i]->thread = std::thread(&BatchService::run, this, i);
}

BatchService::~BatchService()
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        itsStop = true;
        for (BuildRequest* s : itsQueue)
        {
            s->request().state = BuildRequest::Request::FAILED;
            s->request().error = "the build service was destroyed before the request ran";
            s->request().owner = nullptr;
        }
        itsQueue.clear();
    }
    itsWork.notify_all();
    itsDone.notify_all();
    for (auto& w : itsWorkersList)
        if (w->thread.joinable())
            w->thread.join();
}

uint64_t BatchService::reservedExcept(uint32_t index) const
{
    uint64_t total = 0;
    for (uint32_t i = 0; i < itsWorkersList.size(); ++i)
        if (i != index)
            total += itsWorkersList[i]->reserved;
    return total;
}

uint32_t BatchService::runningLocked() const
{
    uint32_t running = 0;
    for (auto& w : itsWorkersList)
        running += w->busy ? 1 : 0;
    return running;
}

static BatchRequest batchRequestOf(BuildRequest* r)
{
    BatchRequest b;
    b.cloud       = r->cloudInfo();
    b.proc        = r->processingInfo();
    b.hostCloud   = r->hostCloud();
    b.deviceCloud = r->deviceCloud();
    return b;
}

void BatchService::submit(BuildRequest* request)
{
    request->validate();
    const uint64_t footprint = batchFootprint({batchRequestOf(request)});
    if (footprint > itsBudget)
        throw std::runtime_error("the build needs about " + std::to_string(footprint) +
                                 " device bytes, more than the service's memory budget of " +
                                 std::to_string(itsBudget));
    enqueue(request, footprint);
}

void BatchService::setActiveSlots(uint32_t active)
{
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        itsActive = std::max<uint32_t>(1, std::min<uint32_t>(active, static_cast<uint32_t>(itsWorkersList.size())));
    }
    itsWork.notify_all();
}

// A free worker takes, from the head of the queue, every request whose footprint still fits the
// budget next to what the other workers hold (up to the batch size), builds them together and
// hands every result to its session. Requests that arrive meanwhile wait for the next batch, which
// a second worker may build while the first one's batch is still on the device.
void BatchService::run(uint32_t index)
{
    Worker& w = *itsWorkersList[index];
    std::unique_lock<std::mutex> lk(itsMutex);
    for (;;)
    {
        if (itsStop)
            return;
        if (index >= itsActive && w.reserved > 0)
        {
            lk.unlock();
            w.builder->releaseArena();
            lk.lock();
            w.reserved = 0;
            itsWork.notify_all();
            continue;
        }
        if (index >= itsActive || itsQueue.empty())
        {
            itsWork.wait(lk);
            continue;
        }
        const uint64_t held = w.builder->arenaBytes();
        const uint64_t rest = reservedExcept(index);
        std::vector<BuildRequest*> taken;
        uint64_t need = 0;
        while (!itsQueue.empty() && taken.size() < itsMaxBatch)
        {
            const uint64_t f = itsQueue.front()->request().footprint;
            if (rest + std::max(held, need + f) > itsBudget)
                break;
            need += f;
            taken.push_back(itsQueue.front());
            itsQueue.pop_front();
        }
        bool releaseFirst = false;
        if (taken.empty())
        {
            const uint64_t f = itsQueue.front()->request().footprint;
            if (held > 0 && rest + f <= itsBudget)
            {
                releaseFirst = true;  // the head fits once this worker's arena is given back
                need         = f;
                taken.push_back(itsQueue.front());
                itsQueue.pop_front();
            }
            else if (held > 0)
            {
                lk.unlock();
                w.builder->releaseArena();
                lk.lock();
                w.reserved = 0;
                itsWork.notify_all();
                continue;
            }
            else
            {
                itsWork.wait(lk);  // other workers hold the memory the head needs
                continue;
            }
        }
        const auto now = Clock::now();
        for (BuildRequest* r : taken)
        {
            r->request().state   = BuildRequest::Request::RUNNING;
            r->request().started = now;
        }
        w.busy     = true;
        w.reserved = releaseFirst ? need : std::max(held, need);
        itsPeakReserved = std::max(itsPeakReserved, rest + w.reserved);
        itsBatches++;
        itsBatchedRequests += taken.size();
        lk.unlock();

        std::vector<BatchRequest> requests;
        for (BuildRequest* r : taken)
            requests.push_back(batchRequestOf(r));
        std::vector<BatchOutcome> outcomes(taken.size());
        std::string failure;
        const auto t0 = Clock::now();
        try
        {
            if (releaseFirst)
                w.builder->releaseArena();
            outcomes = w.builder->build(requests);
        }
        catch (const std::exception& e)
        {
            failure = e.what();
        }
        catch (...)
        {
            failure = "unknown error";
        }
        const double buildMs = milliseconds(t0, Clock::now());
        for (size_t i = 0; i < taken.size(); ++i)
            if (failure.empty() && outcomes[i].result)
                taken[i]->adoptResult(std::move(outcomes[i].result));

        lk.lock();
        for (size_t i = 0; i < taken.size(); ++i)
        {
            std::string error = failure;
            if (error.empty() && !outcomes[i].error.empty())
                error = outcomes[i].error;
            const double phase[3] = {0.0, buildMs, 0.0};
            finishLocked(taken[i], error, phase);
        }
        w.busy     = false;
        w.reserved = w.builder->arenaBytes();
        itsDone.notify_all();
        itsWork.notify_all();
    }
}

}  // namespace oc
