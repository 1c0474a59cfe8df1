// OctreeBuilder: host orchestration of the level-batched GPU build on one HIP stream.
//
// Replaces the reference's OctreeProcessorImpl (octree_processor_impl.cu:12-186,
// octree_chunking.cu:18-124, rp_subsampling.cu, fp_subsampling.cu). Phases:
//   count       two-pass Morton bucket sort of the cloud (coarse LDS histograms, bucket scatter,
//               per-bucket LDS fine sort) that also yields the dense count pyramid up to the
//               bucket level (K1)
//   merge       rest of the count pyramid, canonical sparse ids by scan, node links and flags,
//               leaf data blocks by scan (K2-K4), cell -> leaf map
//   distribute  the records gathered through the permutation into leaf order (and, inside a
//               leaf, the cell order of its parent's subsampling grid), then each leaf's range of
//               that sorted cloud -> its data block: point index and 18-byte record, coalesced (K5)
//   subsample   one batch per octree level (all internal nodes of a level at once) with per-node
//               occupancy bitmaps and a popcount rank dictionary; winners by atomicMin, colours
//               by packed integer sums or a gathered distance-weighted filter; then DFS
//               post-order placement
// All buffers live in a grow-only per-session arena, so repeated builds allocate nothing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <array>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "build_result.h"
#include "slot_builder.h"
#include "types.h"

namespace oc {

class DeviceBuffer
{
public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    // Grows (never shrinks) to at least `bytes`; contents are not preserved on growth.
    void ensure(size_t bytes);
    void release();
    void* get() const { return itsPtr; }
    size_t capacity() const { return itsBytes; }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(itsPtr);
    }

    static size_t totalAllocated();
    static size_t peakAllocated();

private:
    void* itsPtr   = nullptr;
    size_t itsBytes = 0;
};

class OctreeBuilder final : public BuildResult, public SlotBuilder
{
public:
    // A builder owns one HIP stream and a grow-only device arena; setup() prepares it for one
    // cloud (uploading a host cloud or borrowing a device one) and may be called again for the
    // next cloud, reusing the arena.
    explicit OctreeBuilder(int device);
    OctreeBuilder(int device, const CloudInfo& cloud, const ProcessingInfo& proc, const uint8_t* hostCloud,
                  const uint8_t* deviceCloud);
    ~OctreeBuilder() override;
    OctreeBuilder(const OctreeBuilder&) = delete;
    OctreeBuilder& operator=(const OctreeBuilder&) = delete;

    void setup(const CloudInfo& cloud, const ProcessingInfo& proc, const uint8_t* hostCloud,
               const uint8_t* deviceCloud) override;
    void generate() override;
    const BuildResult& lastBuild() const override { return *this; }
    void setProgress(std::function<void()> hook) override { itsProgress = std::move(hook); }
    // Per-batch subsampling events (the insert / emit split of the stage timings); off, a build
    // makes fewer runtime calls, which matters when many builders share one process.
    void setDetailedTimings(bool on) { itsDetailedTimings = on; }

    // Device bytes held by the arena, and its release (the next setup grows it again).
    uint64_t arenaBytes() const override;
    void releaseArena() override;
    // Upper bound of the arena a build of this cloud and configuration needs, from the same sizes
    // setup() and the build allocate (the output capacity E = N * outputFactor drives most terms).
    static uint64_t estimateFootprint(const CloudInfo& cloud, const ProcessingInfo& proc, bool hostCloud);

    bool built() const { return itsBuilt; }
    const CloudInfo& cloud() const override { return itsCloud; }
    const ProcessingInfo& processing() const override { return itsProc; }
    uint32_t depth() const override { return itsDepth; }
    uint32_t denseNodes() const override { return itsDense; }
    uint32_t sparseNodes() const override { return itsM; }
    uint64_t capacity() const override { return itsE; }
    uint64_t used() const override { return itsUsed; }
    uint32_t levelStart(uint32_t level) const { return itsLevelStart[level]; }
    const std::vector<float>& stageTimings() const override { return itsTimings; }

    // Copies device results to host (valid after generate()).
    void copyNodes(HostOctree& out) const override;
    void copyLut(uint32_t* dst, uint64_t count) const override;
    void copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const override;
    // device bytes at src to host memory at dst (synchronous; staged for large pageable copies)
    void download(void* dst, const void* src, uint64_t bytes) const;
    // host bytes at src to device memory at dst on the builder's stream (staged, and then
    // synchronous, for large pageable sources)
    void upload(void* dst, const void* src, uint64_t bytes);
    void copySnapshot(uint8_t* dst) const override;
    // The snapshot copies on the builder's copy stream, after the build: the host goes on (the
    // next setup's upload overlaps them); the next setup / generate makes the build stream wait
    // for them before it overwrites anything they read. waitSnapshot blocks until they are done.
    void snapshotAsync(uint8_t* dst) override;
    void waitSnapshot() override;
    bool snapshotDone() override;
    void drainCopies() noexcept override;

private:
    template <typename T>
    void generateTyped();
    template <typename T>
    void countPoints();
    void mergeLevels();
    // whether the fine sort writes the level-0 occupancy bits (a bucket's level-0 rows are 32 cells)
    bool levelZeroBits() const;
    template <typename T>
    void distributePoints();
    template <typename T>
    bool orderLeafPoints();
    template <typename T>
    void subsampleLevels();
    void placeInternalNodes();
    void placeLevelByLevel();

    void syncStream();
    void progress();
    void waitForSnapshotOnStream();
    template <typename V>
    V readDevice(const V* devicePtr);
    std::array<uint64_t, 2> readDevicePair(const uint64_t* devicePtr);
    void waitMailbox(uint64_t seq);

    int itsDevice;
    hipStream_t itsStream = nullptr;
    hipStream_t itsCopyStream         = nullptr;  // snapshot copies
    hipEvent_t itsBuiltEvent          = nullptr;  // the build's last kernel
    hipEvent_t itsCopiedEvent         = nullptr;  // the snapshot's last copy
    hipEvent_t itsUploadedEvent       = nullptr;  // the cloud upload of the current setup
    bool itsSnapshotPending           = false;    // copies enqueued that the build stream has not waited for
    // Size readbacks: a one-thread kernel publishes {value, value, sequence} into coherent pinned
    // host memory and the host spins on the sequence (no staging copy, no blocking-sync wake-up).
    uint64_t* itsMailbox  = nullptr;
    uint64_t itsMailSeq   = 0;
    bool itsReady         = false;  // setup() has run
    bool itsDetailedTimings = true;
    std::function<void()> itsProgress;  // called between the stages of generate()
    CloudInfo itsCloud;
    ProcessingInfo itsProc;
    ColorMode itsColorMode = COLOR_NONE;

    // level tables (octree.cu:13-23)
    uint32_t itsDepth = 0;
    uint32_t itsDense = 0;
    uint32_t itsCoarse  = 0;  // bucket level of the sort
    uint32_t itsBuckets = 0;  // (chunkingGrid >> itsCoarse)^3
    uint32_t itsLevelGrid[MAX_LEVELS]   = {};
    uint32_t itsLevelOffset[MAX_LEVELS] = {};
    uint32_t itsLevelNodes[MAX_LEVELS]  = {};

    // results of the last build
    bool itsBuilt          = false;
    uint32_t itsM          = 0;
    uint64_t itsE          = 0;
    uint64_t itsUsed       = 0;
    uint32_t itsLevelStart[MAX_LEVELS + 1] = {};
    std::vector<float> itsTimings;

    // cloud
    const uint8_t* itsCloudDevice = nullptr;
    DeviceBuffer itsCloudOwned;

    // dense pyramid and sparse nodes
    DeviceBuffer itsCount, itsDenseToSparse, itsOccupancy;  // dense counts, sparse ids, occupancy bits
    DeviceBuffer itsNodeKey, itsNodeCount, itsNodePoints, itsNodeData, itsNodeParent, itsNodeChild, itsNodeFlags,
            itsNodeSorted, itsNodeAux, itsNodeStage, itsNodeSub, itsNodeRegion;

    // Morton bucket sort of the cloud
    DeviceBuffer itsBucketed, itsSorted, itsFineCode, itsPerm, itsBucketCount, itsBucketStart, itsBucketCursor;
    DeviceBuffer itsSplitChunks, itsSplitHist;  // the fine sort's split buckets: chunk list, per-chunk histograms

    // output region [0, E) (lut: point index, out: 18-byte record): leaf points at [0, N), internal
    // subsamples at [N, N + I) after the final placement. Staging (st*): subsampled points in
    // level order, read by the next level, placed at the end.
    DeviceBuffer itsLut, itsOut, itsStLut, itsStOut, itsStPts, itsStRgb;

    // per-level subsampling workspace: internal node list, input segments and tiles, per-node
    // occupancy bitmaps with their popcount prefix, per-cell (dense index) state, per-input state
    DeviceBuffer itsLeafList, itsInternalList, itsSegCount, itsSegSrc, itsSegTile, itsSegInput, itsNodeGrid,
            itsTileList;
    DeviceBuffer itsBitmap, itsPrefix, itsNodeFirst;
    DeviceBuffer itsWin, itsColor, itsCellCnt, itsCellStart, itsCellKey, itsCellNode;
    DeviceBuffer itsCellOf, itsRankOf, itsPayload;

    // small device scalars and scan scratch
    DeviceBuffer itsScalars, itsScanScratch;

    std::vector<DeviceBuffer*> arena();

    enum EventSlot
    {
        EV_START,
        EV_COUNT_BEGIN,
        EV_COUNT_END,
        EV_MERGE_END,
        EV_DIST_BEGIN,
        EV_DIST_END,
        EV_END,
        EV_SLOTS
    };
    hipEvent_t itsEvents[EV_SLOTS] = {};
    // per subsampling batch: evaluate begin/end, emit begin/end (when detailed timings are on)
    std::vector<std::array<hipEvent_t, 4>> itsBatchEvents;
    uint32_t itsBatchesRun = 0;
    std::array<hipEvent_t, 4>& batchEvents();
};

}  // namespace oc
