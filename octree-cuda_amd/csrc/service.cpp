#include "service.h"

#include <hip/hip_runtime.h>

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

BuildService* BuildService::fromHandle(void* handle)
{
    auto* s = static_cast<BuildService*>(handle);
    if (!s)
        throw std::runtime_error("No BuildService is currently initialized!");
    return s;
}

BuildService::BuildService(int device, uint32_t slots, uint64_t memoryBudget)
    : itsDevice(device), itsBudget(memoryBudget), itsActive(slots)
{
    if (slots == 0 || slots > 64)
        throw std::runtime_error("a build service needs 1 to 64 slots");
    itsPinned = std::make_shared<PinnedPool>(size_t(16) << 30);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        throw std::runtime_error("device " + std::to_string(device) + " does not exist");
    for (uint32_t i = 0; i < slots; ++i)
    {
        auto slot     = std::make_unique<Slot>();
        slot->builder = std::make_unique<OctreeBuilder>(device);
        slot->builder->setDetailedTimings(false);
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
        for (Session* s : itsQueue)
        {
            s->request().state = Session::Request::FAILED;
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

void BuildService::submit(Session* session)
{
    session->validate();
    const uint64_t footprint =
            OctreeBuilder::estimateFootprint(session->cloudInfo(), session->processingInfo(),
                                             session->hostCloud() != nullptr);
    if (footprint > itsBudget)
        throw std::runtime_error("the build needs about " + std::to_string(footprint) +
                                 " device bytes, more than the service's memory budget of " +
                                 std::to_string(itsBudget));
    {
        std::lock_guard<std::mutex> lk(itsMutex);
        auto& r = session->request();
        if (r.state == Session::Request::QUEUED || r.state == Session::Request::RUNNING)
            throw std::runtime_error("the session is already submitted");
        r           = Session::RequestInfo{};
        r.owner     = this;
        r.state     = Session::Request::QUEUED;
        r.footprint = footprint;
        r.submitted = Clock::now();
        itsQueue.push_back(session);
        itsSubmitted++;
        itsMaxQueue = std::max<uint64_t>(itsMaxQueue, itsQueue.size());
    }
    itsWork.notify_all();
}

bool BuildService::wait(Session* session, uint32_t timeoutMs)
{
    std::unique_lock<std::mutex> lk(itsMutex);
    auto& r = session->request();
    if (r.state == Session::Request::NONE)
        throw std::runtime_error("the session was not submitted to this service");
    const bool finished = itsDone.wait_for(lk, std::chrono::milliseconds(timeoutMs), [&] {
        return r.state == Session::Request::DONE || r.state == Session::Request::FAILED;
    });
    if (!finished)
        return false;
    if (r.state == Session::Request::FAILED)
        throw std::runtime_error(r.error);
    return true;
}

void BuildService::forget(Session* session)
{
    std::unique_lock<std::mutex> lk(itsMutex);
    auto& r = session->request();
    auto it = std::find(itsQueue.begin(), itsQueue.end(), session);
    if (it != itsQueue.end())
    {
        itsQueue.erase(it);
        r.state = Session::Request::FAILED;
        r.error = "the session was destroyed before its build ran";
    }
    itsDone.wait(lk, [&] { return r.state != Session::Request::RUNNING; });
    r.owner = nullptr;
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
// before its merge). The finished request is published after that next build, when its copies
// are long done, or as soon as the slot has nothing else to do.
void BuildService::run(uint32_t index)
{
    Slot& slot = *itsSlots[index];
    (void)hipSetDevice(itsDevice);

    struct Done
    {
        Session* session = nullptr;
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
            }
        }
        std::lock_guard<std::mutex> g(itsMutex);
        auto& r    = pending.session->request();
        r.finished = Clock::now();
        if (pending.error.empty())
        {
            r.state = Session::Request::DONE;
            itsCompleted++;
            itsLatencies.push_back(milliseconds(r.submitted, r.finished));
            itsWaits.push_back(milliseconds(r.submitted, r.started));
            for (int p = 0; p < 3; ++p)
                itsPhaseMs[p] += pending.phase[p];
        }
        else
        {
            r.state = Session::Request::FAILED;
            r.error = pending.error;
            itsFailed++;
            log(LOG_ERROR, "service build failed: " + pending.error);
        }
        pending = Done{};
        if (!stillBusy)
        {
            slot.busy     = false;
            slot.reserved = slot.builder->arenaBytes();
        }
        itsDone.notify_all();
        itsWork.notify_all();
    };

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
        Session* session  = nullptr;
        bool releaseFirst = false;
        if (index < itsActive && !itsQueue.empty())
        {
            Session* head       = itsQueue.front();
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
                session->request().state   = Session::Request::RUNNING;
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
        publish(true);  // the previous request: its copies finished before this build's merge
        if (next.error.empty())
        {
            try
            {
                const auto t2  = Clock::now();
                uint8_t* block = nullptr;
                next.result    = HostBuildResult::prepare(*slot.builder, itsPinned.get(), &block);
                slot.builder->snapshotAsync(block);
                next.phase[2] = milliseconds(t2, Clock::now());
            }
            catch (const std::exception& e)
            {
                next.error  = e.what();
                next.result = nullptr;
            }
            catch (...)
            {
                next.error  = "unknown error";
                next.result = nullptr;
            }
        }
        pending = std::move(next);
        lk.lock();
    }
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

std::vector<double> BuildService::stats()
{
    std::lock_guard<std::mutex> lk(itsMutex);
    uint64_t running = 0, reserved = 0;
    for (auto& s : itsSlots)
    {
        running += s->busy ? 1 : 0;
        reserved += s->reserved;
    }
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
            static_cast<double>(itsQueue.size()), static_cast<double>(running), static_cast<double>(itsMaxQueue),
            percentile(itsLatencies, 0.50), percentile(itsLatencies, 0.99), mean, static_cast<double>(reserved),
            static_cast<double>(itsBudget), static_cast<double>(itsActive), static_cast<double>(itsSlots.size()),
            meanWait, maxLatency, static_cast<double>(itsPeakReserved), itsPhaseMs[0] / done,
            itsPhaseMs[1] / done, itsPhaseMs[2] / done};
}

void BuildService::resetStats()
{
    std::lock_guard<std::mutex> lk(itsMutex);
    itsSubmitted = itsCompleted = itsFailed = 0;
    itsMaxQueue                             = itsQueue.size();
    itsPeakReserved                         = 0;
    itsLatencies.clear();
    itsWaits.clear();
    itsPhaseMs[0] = itsPhaseMs[1] = itsPhaseMs[2] = 0;
}

}  // namespace oc
