// CPU harness of the build services (csrc/service.cpp) with their pinned-buffer pool and result
// snapshots (csrc/pinned_pool.cpp, csrc/build_result.cpp), built with clang++ under ThreadSanitizer or
// AddressSanitizer by tests/test_service_harness.py. The GPU builders are replaced by stand-ins
// through the seams of slot_builder.h:
//   StubSlotBuilder   setup / generate take a little time and call the progress hook between
//                     stages; snapshotAsync starts a thread that fills the snapshot block some time
//                     later, the way the device's copy engine writes a pinned block while the host
//                     goes on; snapshotDone / waitSnapshot / drainCopies observe or join it
//   StubBatchBuilder  builds a batch of requests, one snapshot each
//   pinned memory     malloc / free
// and the threaded host copy that moves large results into callers' memory is run by several
// callers at once.
// Every result's content is a function of its request, so a result handed to the wrong session,
// read before its copy finished or written after its block was recycled shows up as a wrong value
// or a sanitizer report. Requests fail on purpose (in validation, in generate, by budget), slots
// are switched on and off, sessions are destroyed while queued, and the memory budget forces
// arenas to be released and requests to wait.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "build_request.h"
#include "build_result.h"
#include "host_copy.h"
#include "pinned_pool.h"
#include "service.h"
#include "slot_builder.h"

using namespace oc;

namespace {

std::atomic<int> gFailures{0};
std::mutex gLogMutex;

void fail(const std::string& what)
{
    std::lock_guard<std::mutex> lk(gLogMutex);
    std::fprintf(stderr, "harness: %s\n", what.c_str());
    gFailures++;
}

void pause(uint32_t microseconds)
{
    std::this_thread::sleep_for(std::chrono::microseconds(microseconds));
}

// The request id travels in the cloud description: points = id + 1; the "result" of request id has
// m = 1 + id % 5 nodes and used = 3 + id % 11 records, all fields derived from id.
uint32_t idOf(const CloudInfo& c)
{
    return c.pointAmount - 1;
}

class StubResult final : public BuildResult
{
public:
    CloudInfo itsCloud;
    ProcessingInfo itsProc;
    std::vector<float> itsTimings = std::vector<float>(7, 0.f);
    uint32_t m() const { return 1 + idOf(itsCloud) % 5; }
    const CloudInfo& cloud() const override { return itsCloud; }
    const ProcessingInfo& processing() const override { return itsProc; }
    uint32_t depth() const override { return 3; }
    uint32_t denseNodes() const override { return 73; }
    uint32_t sparseNodes() const override { return m(); }
    uint64_t capacity() const override { return 64; }
    uint64_t used() const override { return 3 + idOf(itsCloud) % 11; }
    const std::vector<float>& stageTimings() const override { return itsTimings; }
    void copyNodes(HostOctree&) const override { throw std::runtime_error("not used"); }
    void copyLut(uint32_t*, uint64_t) const override { throw std::runtime_error("not used"); }
    void copyOutputRange(uint8_t*, uint64_t, uint64_t) const override { throw std::runtime_error("not used"); }
};

// The snapshot bytes of request id, written as the device's copies would.
void fillSnapshot(uint8_t* dst, const StubResult& r)
{
    const uint32_t id = idOf(r.itsCloud);
    const SnapshotLayout l(r.sparseNodes(), r.used());
    for (uint32_t i = 0; i < r.sparseNodes(); ++i)
    {
        const uint64_t data = 1000ull * id + i;
        std::memcpy(dst + l.dataIdx + 8 * i, &data, 8);
        const uint32_t key = 7 * id + i;
        std::memcpy(dst + l.denseKey + 4 * i, &key, 4);
        std::memcpy(dst + l.count + 4 * i, &key, 4);
        std::memcpy(dst + l.pointCount + 4 * i, &key, 4);
        const int32_t parent = -1;
        std::memcpy(dst + l.parent + 4 * i, &parent, 4);
        for (int k = 0; k < 8; ++k)
            std::memcpy(dst + l.children + 32 * i + 4 * k, &parent, 4);
        dst[l.flags + i] = static_cast<uint8_t>(id);
    }
    for (uint64_t j = 0; j < r.used(); ++j)
    {
        const uint32_t lut = id * 100 + static_cast<uint32_t>(j);
        std::memcpy(dst + l.lut + 4 * j, &lut, 4);
        for (int b = 0; b < 18; ++b)
            dst[l.out + 18 * j + b] = static_cast<uint8_t>(id + j + b);
    }
}

// Checks a finished result against its request id.
void checkResult(uint32_t id, const BuildResult& r)
{
    HostOctree o;
    r.copyNodes(o);
    const uint32_t m = 1 + id % 5;
    const uint64_t used = 3 + id % 11;
    if (o.dataIdx.size() != m || r.used() != used)
    {
        fail("request " + std::to_string(id) + ": wrong sizes");
        return;
    }
    for (uint32_t i = 0; i < m; ++i)
        if (o.dataIdx[i] != 1000ull * id + i || o.denseKey[i] != 7 * id + i || o.flags[i] != static_cast<uint8_t>(id))
            fail("request " + std::to_string(id) + ": node " + std::to_string(i) + " holds another request's data");
    std::vector<uint32_t> lut(used);
    std::vector<uint8_t> out(18 * used);
    r.copyLut(lut.data(), used);
    r.copyOutputRange(out.data(), 0, used);
    for (uint64_t j = 0; j < used; ++j)
    {
        if (lut[j] != id * 100 + j)
            fail("request " + std::to_string(id) + ": LUT entry " + std::to_string(j));
        for (int b = 0; b < 18; ++b)
            if (out[18 * j + b] != static_cast<uint8_t>(id + j + b))
                fail("request " + std::to_string(id) + ": output byte");
    }
}

class StubSlotBuilder final : public SlotBuilder
{
public:
    ~StubSlotBuilder() override { drainCopies(); }

    void setup(const CloudInfo& cloud, const ProcessingInfo& proc, const uint8_t* hostCloud,
               const uint8_t*) override
    {
        if (!hostCloud)
            throw std::runtime_error("no cloud");
        pause(20 + idOf(cloud) % 50);
        // growing the arena while the last snapshot's copies read it would free what they read:
        // the GPU builder waits for them first, and so does this one
        const uint64_t need = 1000 + 100 * (idOf(cloud) % 7);
        if (need > itsArena)
        {
            waitSnapshot();
            itsArena = need;
        }
        itsBuilding.itsCloud = cloud;
        itsBuilding.itsProc  = proc;
    }

    void generate() override
    {
        const uint32_t id = idOf(itsBuilding.itsCloud);
        for (int stage = 0; stage < 4; ++stage)
        {
            pause(30 + (id * 13 + stage * 7) % 120);
            if (itsProgress)
                itsProgress();
        }
        if (id % 17 == 5)
            throw std::runtime_error("the subsampled points exceed the estimated output size (N + I > E)");
        itsBuilt = itsBuilding;
    }

    const BuildResult& lastBuild() const override { return itsBuilt; }

    void snapshotAsync(uint8_t* dst) override
    {
        drainCopies();  // one snapshot in flight at a time, as on the copy stream
        itsDone          = false;
        const StubResult r = itsBuilt;
        const uint32_t delay = 50 + (idOf(r.itsCloud) * 31) % 400;
        itsCopy = std::thread([this, dst, r, delay] {
            pause(delay);
            fillSnapshot(dst, r);
            itsDone.store(true, std::memory_order_release);
        });
    }

    bool snapshotDone() override { return itsDone.load(std::memory_order_acquire); }

    void waitSnapshot() override
    {
        if (itsCopy.joinable())
            itsCopy.join();
    }

    void drainCopies() noexcept override
    {
        if (itsCopy.joinable())
            itsCopy.join();
    }

    uint64_t arenaBytes() const override { return itsArena; }

    void releaseArena() override
    {
        drainCopies();
        itsArena = 0;
    }

    void setProgress(std::function<void()> hook) override { itsProgress = std::move(hook); }

private:
    StubResult itsBuilding, itsBuilt;
    std::thread itsCopy;
    std::atomic<bool> itsDone{true};
    uint64_t itsArena = 0;
    std::function<void()> itsProgress;
};

class StubBatchBuilder final : public BatchSlotBuilder
{
public:
    explicit StubBatchBuilder(std::shared_ptr<PinnedPool> pinned) : itsPinned(std::move(pinned)) {}

    std::vector<BatchOutcome> build(const std::vector<BatchRequest>& requests) override
    {
        std::vector<BatchOutcome> out(requests.size());
        uint64_t need = 0;
        for (const BatchRequest& r : requests)
            need += 1000 + 100 * (idOf(r.cloud) % 7);
        itsArena = std::max(itsArena, need);
        pause(100 + 20 * static_cast<uint32_t>(requests.size()));
        // the batch's results share one block, as the batched builder's do
        std::vector<StubResult> results(requests.size());
        std::vector<uint64_t> offsets(requests.size());
        uint64_t total = 0;
        for (size_t i = 0; i < requests.size(); ++i)
        {
            results[i].itsCloud = requests[i].cloud;
            results[i].itsProc  = requests[i].proc;
            offsets[i]          = total;
            total += (SnapshotLayout(results[i].sparseNodes(), results[i].used()).bytes + 256) & ~uint64_t(255);
        }
        std::shared_ptr<uint8_t> block = itsPinned->acquire(total);
        std::thread copier([&] {
            for (size_t i = 0; i < requests.size(); ++i)
                if (idOf(requests[i].cloud) % 17 != 5)
                    fillSnapshot(block.get() + offsets[i], results[i]);
        });
        copier.join();
        for (size_t i = 0; i < requests.size(); ++i)
        {
            if (idOf(requests[i].cloud) % 17 == 5)
            {
                out[i].error = "the subsampled points exceed the estimated output size (N + I > E)";
                continue;
            }
            const StubResult& r = results[i];
            out[i].result = HostBuildResult::inBlock(r.itsCloud, r.itsProc, r.depth(), r.denseNodes(),
                                                     r.sparseNodes(), r.capacity(), r.used(), r.itsTimings, block,
                                                     block.get() + offsets[i]);
        }
        return out;
    }

    uint64_t arenaBytes() const override { return itsArena; }
    void releaseArena() override { itsArena = 0; }

private:
    std::shared_ptr<PinnedPool> itsPinned;
    uint64_t itsArena = 0;
};

class StubRequest final : public BuildRequest
{
public:
    explicit StubRequest(uint32_t id) : itsId(id), itsBytes(16, static_cast<uint8_t>(id))
    {
        itsCloud.pointAmount = id + 1;
        itsCloud.dataStride  = 15;
    }
    ~StubRequest() override
    {
        ServiceBase* owner  = itsRequest.owner.load();
        const Request state = itsRequest.state.load();
        if (owner && (state == Request::QUEUED || state == Request::RUNNING))
            owner->forget(this);
    }
    void validate() const override
    {
        if (itsId % 23 == 11)
            throw std::runtime_error("invalid configuration");
    }
    const CloudInfo& cloudInfo() const override { return itsCloud; }
    const ProcessingInfo& processingInfo() const override { return itsProc; }
    const uint8_t* hostCloud() const override { return itsBytes.data(); }
    const uint8_t* deviceCloud() const override { return nullptr; }
    void adoptResult(std::unique_ptr<HostBuildResult> result) override { itsResult = std::move(result); }

    uint32_t itsId;
    std::vector<uint8_t> itsBytes;
    CloudInfo itsCloud;
    ProcessingInfo itsProc;
    std::unique_ptr<HostBuildResult> itsResult;
};

}  // namespace

// seams of slot_builder.h
namespace oc {

std::unique_ptr<SlotBuilder> makeSlotBuilder(int)
{
    return std::make_unique<StubSlotBuilder>();
}

std::unique_ptr<BatchSlotBuilder> makeBatchSlotBuilder(int, std::shared_ptr<PinnedPool> pinned)
{
    return std::make_unique<StubBatchBuilder>(std::move(pinned));
}

uint64_t slotFootprint(const CloudInfo& cloud, const ProcessingInfo&, bool)
{
    return 1000 + 100 * (idOf(cloud) % 7);
}

uint64_t batchFootprint(const std::vector<BatchRequest>& requests)
{
    uint64_t total = 0;
    for (const BatchRequest& r : requests)
        total += 1000 + 100 * (idOf(r.cloud) % 7);
    return total;
}

int deviceCount()
{
    return 1;
}

void* pinnedAllocate(size_t bytes)
{
    void* p = std::malloc(bytes);
    if (!p)
        throw std::runtime_error("out of host memory");
    return p;
}

void pinnedRelease(void* p)
{
    std::free(p);
}

}  // namespace oc

namespace {

// Submits requests [first, first + count) from `threads` threads with random gaps, waits for each
// and checks its result; every 9th request is destroyed while it may still be queued.
void drive(ServiceBase& svc, uint32_t first, uint32_t count, uint32_t threads, uint32_t seed)
{
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            std::mt19937 rng(seed + t);
            std::vector<std::unique_ptr<StubRequest>> mine;
            for (uint32_t id = first + t; id < first + count; id += threads)
            {
                auto r = std::make_unique<StubRequest>(id);
                try
                {
                    svc.submit(r.get());
                }
                catch (const std::exception& e)
                {
                    if (id % 23 != 11)
                        fail("request " + std::to_string(id) + " refused: " + e.what());
                    continue;
                }
                if (id % 23 == 11)
                    fail("request " + std::to_string(id) + " should have been refused");
                if (id % 9 == 4)
                {
                    r.reset();  // destroyed while queued or running
                    continue;
                }
                mine.push_back(std::move(r));
                pause(rng() % 200);
            }
            for (auto& r : mine)
            {
                try
                {
                    if (!svc.wait(r.get(), 60000))
                        fail("request " + std::to_string(r->itsId) + " timed out");
                    else if (r->itsId % 17 == 5)
                        fail("request " + std::to_string(r->itsId) + " should have failed");
                    else if (!r->itsResult)
                        fail("request " + std::to_string(r->itsId) + " has no result");
                    else
                        checkResult(r->itsId, *r->itsResult);
                }
                catch (const std::exception& e)
                {
                    if (r->itsId % 17 != 5)
                        fail("request " + std::to_string(r->itsId) + " failed: " + e.what());
                }
            }
        });
    for (auto& t : pool)
        t.join();
}

}  // namespace

int main()
{
    {
        // slot service, room for every slot
        BuildService svc(0, 4, 1ull << 40);
        drive(svc, 0, 240, 4, 1);
        svc.setActiveSlots(2);
        drive(svc, 240, 120, 3, 2);
        svc.setActiveSlots(4);
        drive(svc, 360, 120, 4, 3);
        const std::vector<double> st = svc.stats();
        if (st[2] < 1)
            fail("no request failed in generate");
    }
    {
        // slot service, a budget of about two requests: slots release their arenas and wait
        BuildService svc(0, 4, 3000);
        drive(svc, 1000, 160, 4, 4);
        const std::vector<double> st = svc.stats();
        if (st[15] > 3000)
            fail("peak reserved memory above the budget");
    }
    {
        // a service destroyed with requests still queued
        auto svc = std::make_unique<BuildService>(0, 2, 1ull << 40);
        std::vector<std::unique_ptr<StubRequest>> reqs;
        for (uint32_t id = 2000; id < 2040; ++id)
        {
            reqs.push_back(std::make_unique<StubRequest>(id));
            if (id % 23 != 11)
                svc->submit(reqs.back().get());
        }
        svc.reset();
    }
    {
        // batch service
        BatchService svc(0, 2, 16, 1ull << 40);
        drive(svc, 3000, 300, 6, 5);
        const std::vector<double> st = svc.stats();
        if (st[19] < 2 || st[20] < 1.0)
            fail("no batches were built");
    }
    {
        // batch service under a budget of a few requests
        BatchService svc(0, 3, 8, 5000);
        drive(svc, 4000, 150, 4, 6);
        const std::vector<double> st = svc.stats();
        if (st[15] > 5000)
            fail("peak reserved memory above the budget");
    }
    {
        // the threaded host copy of large results (host_copy.cpp): three callers at once, sizes
        // that are not multiples of its parts, into fresh destinations
        std::vector<std::thread> callers;
        for (uint32_t t = 0; t < 3; ++t)
            callers.emplace_back([t] {
                for (uint32_t k = 0; k < 4; ++k)
                {
                    const size_t bytes = (9u << 20) + 4099u * (t + 1) * (k + 1);
                    std::vector<uint8_t> src(bytes);
                    for (size_t i = 0; i < bytes; ++i)
                        src[i] = static_cast<uint8_t>((i * 131u + t * 7u + k) >> 3);
                    std::unique_ptr<uint8_t[]> dst(new uint8_t[bytes]);
                    parallelMemcpy(dst.get(), src.data(), bytes);
                    if (std::memcmp(dst.get(), src.data(), bytes) != 0)
                        fail("threaded host copy differs from its source");
                }
            });
        for (auto& c : callers)
            c.join();
    }
    std::fprintf(stderr, "harness: %d failures\n", gFailures.load());
    return gFailures.load() == 0 ? 0 : 1;
}
