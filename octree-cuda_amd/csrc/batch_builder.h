// BatchBuilder: many clouds built together, one launch per stage for all of them.
//
// The reference builds one cloud per process and launches every kernel per phase and per node
// (kernel_executor.cuh:8-28, rp_subsampling.cu:10-222, fp_subsampling.cu:9-185). A small cloud
// then costs a hundred or so launches and a dozen host round trips whatever its size. Here the
// clouds of a batch that share a record type, chunking grid, colour filter and subsampling
// method go through the pipeline together: each stage is one launch over the work items of all
// of them (kernels_batch.h), sizes that only the device knows are scanned on the device, and the
// host waits twice per batch (after the merge, to size the node arrays, and at the end). Each
// cloud keeps its own arrays and configuration, so its result is the same as a build of that
// cloud alone (node arrays, data offsets and record sets; distance-weighted colours within the
// float tolerance of the filter).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "build_result.h"
#include "builder.h"
#include "pinned_pool.h"
#include "slot_builder.h"

namespace oc {


// Launch shapes of the batched kernels, the autotuner's knobs: persistent workgroups per launch
// for the kernels with one item per thread (256 threads), one tile per workgroup (256 threads)
// and one sort chunk per workgroup (1024 threads).
struct BatchLaunch
{
    uint32_t threadGrid = 2048;
    uint32_t tileGrid   = 4096;
    uint32_t sortGrid   = 1024;
};

// Upper bounds over the clouds of a batch that size the batched scans.
struct BatchBounds
{
    uint64_t levelNodes[MAX_LEVELS + 1];
    uint64_t segments, words, cells;
};

struct BatchLevels;

class BatchBuilder final : public BatchSlotBuilder
{
public:
    explicit BatchBuilder(int device, std::shared_ptr<PinnedPool> pinned = nullptr);
    ~BatchBuilder();
    BatchBuilder(const BatchBuilder&) = delete;
    BatchBuilder& operator=(const BatchBuilder&) = delete;

    // Builds every request (validated by the caller) and returns one outcome per request, in order.
    std::vector<BatchOutcome> build(const std::vector<BatchRequest>& requests) override;

    void setLaunch(const BatchLaunch& launch) { itsLaunch = launch; }
    // Capture the two launch segments of a batch (before and after the host reads the node
    // counts) as HIP graphs, cached by everything their launches depend on, and replay them.
    void useGraphs(bool on) { itsUseGraphs = on; }
    uint32_t graphsCaptured() const { return itsGraphsCaptured; }
    const BatchLaunch& launch() const { return itsLaunch; }

    // Of the last build(), summed over its groups: device ms of [0] count [1] merge [2] distribute
    // [3] subsample and placement [4] the whole pipeline, and host ms of [5] the whole call (uploads,
    // pipeline and result copies) [6] the result copies.
    const std::vector<float>& timings() const { return itsTimings; }
    uint32_t groupsBuilt() const { return itsGroups; }

    uint64_t arenaBytes() const override
    {
        return itsArena.capacity() + itsNodeArena.capacity() + itsPacked.capacity() + itsPieces.capacity();
    }
    void releaseArena() override;
    // Device bytes a batch of these requests needs (the arena sizes build() allocates).
    static uint64_t estimateFootprint(const std::vector<BatchRequest>& requests);

private:
    template <typename T>
    void buildGroup(const std::vector<uint32_t>& members, const std::vector<BatchRequest>& requests,
                    std::vector<BatchOutcome>& outcomes);
    template <typename T, uint32_t MODE, bool RANDOM>
    void subsampleGroup(uint32_t B, const BatchLevels& levels, const BatchBounds& bounds);

    void runSegment(bool graphs, const std::vector<uint64_t>& key, const std::function<void()>& issue);
    void clearGraphs();

    int itsDevice;
    hipStream_t itsStream = nullptr;
    bool itsUseGraphs           = true;
    uint32_t itsGraphsCaptured  = 0;
    std::map<std::vector<uint64_t>, hipGraphExec_t> itsGraphs;
    std::shared_ptr<PinnedPool> itsPinned;
    BatchLaunch itsLaunch;
    DeviceBuffer itsArena, itsNodeArena, itsTable, itsPrefix, itsPacked, itsPieces;
    std::vector<uint8_t> itsHostTable;
    std::vector<uint32_t> itsHostPrefix;
    std::vector<uint64_t> itsHostScalars;
    hipEvent_t itsEvents[6] = {};
    std::vector<float> itsTimings;
    uint32_t itsGroups = 0;
};

}  // namespace oc
