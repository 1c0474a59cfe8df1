// SlotBuilder: the builder a slot of the build service drives, one request at a time, with the
// result copies of one request overlapping the next build. The GPU builder (OctreeBuilder)
// implements it; the service's CPU test harness has a stand-in. The free functions are the link
// seam: the library defines them with the GPU builder, the harness with its stand-in.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "build_result.h"

namespace oc {

class SlotBuilder
{
public:
    virtual ~SlotBuilder() = default;

    // Prepares the next build (uploading a host cloud); may run while the last snapshot's copies are
    // in flight.
    virtual void setup(const CloudInfo& cloud, const ProcessingInfo& proc, const uint8_t* hostCloud,
                       const uint8_t* deviceCloud) = 0;
    virtual void generate()                       = 0;
    virtual const BuildResult& lastBuild() const  = 0;
    // The finished build's copies into dst (SnapshotLayout), enqueued behind the build.
    virtual void snapshotAsync(uint8_t* dst)      = 0;
    virtual bool snapshotDone()                   = 0;  // the last snapshot's copies have finished
    virtual void waitSnapshot()                   = 0;
    virtual void drainCopies() noexcept           = 0;  // waits for every copy in flight, ignoring errors
    virtual uint64_t arenaBytes() const           = 0;
    virtual void releaseArena()                   = 0;
    // Called by generate() between its stages (on the calling thread).
    virtual void setProgress(std::function<void()> hook) = 0;
};

// A batch of requests built together (a worker of the batch service); the GPU BatchBuilder
// implements it.
struct BatchRequest
{
    CloudInfo cloud;
    ProcessingInfo proc;
    const uint8_t* hostCloud   = nullptr;  // borrowed until build() returns
    const uint8_t* deviceCloud = nullptr;
};

struct BatchOutcome
{
    std::unique_ptr<HostBuildResult> result;  // null when the build failed
    std::string error;
};

class BatchSlotBuilder
{
public:
    virtual ~BatchSlotBuilder() = default;
    // One outcome per request, in order.
    virtual std::vector<BatchOutcome> build(const std::vector<BatchRequest>& requests) = 0;
    virtual uint64_t arenaBytes() const = 0;
    virtual void releaseArena()         = 0;
};

std::unique_ptr<SlotBuilder> makeSlotBuilder(int device);
std::unique_ptr<BatchSlotBuilder> makeBatchSlotBuilder(int device, std::shared_ptr<class PinnedPool> pinned);
// device bytes a batch of these requests needs (the sum of their parts of the batch arena)
uint64_t batchFootprint(const std::vector<BatchRequest>& requests);
// device bytes a build of this cloud may need (what the slot's arena grows to)
uint64_t slotFootprint(const CloudInfo& cloud, const ProcessingInfo& proc, bool hostCloud);
// the number of devices (a service checks its device against it)
int deviceCount();

// Host memory for result snapshots: page-locked in the library (hipHostMalloc / hipHostFree).
void* pinnedAllocate(size_t bytes);
void pinnedRelease(void* p);

}  // namespace oc
