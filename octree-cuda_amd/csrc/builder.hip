// Level-batched LOD octree build on one MI355X (gfx950). See builder.h for the phase overview.
//
// Data layout in HBM (per session):
//   cloud          N x S bytes, packed AoS records (15 or 27 bytes), read twice (bucket
//                  histogram, bucket scatter)
//   bucketed       N sorted-point records (32 bytes for double clouds, 20 for float) grouped by
//                  coarse bucket, plus a u16 fine code per point
//   sorted         the same records in level-0 Morton order: every node is one contiguous range;
//                  the leaves' points are read from here by the subsampling
//   count          D x u32 dense count pyramid; after the merge its level-0 part maps cell -> leaf
//   denseToSparse  D x i32, set for occupied cells (non-zero dense count) only
//   nodes (SoA)    key, dense count, pointCount, dataIdx (u64), parent, children[8], flags, ...
//   output region  E records: lut (u32 point index), out (18-byte record). [0, N) holds the leaves,
//                  [N, N + I) the subsampled points in DFS post-order
//   staging        subsampled points in level order (index, record, coordinates), read by the
//                  next level and permuted into the output region at the end
#include "builder.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "device_common.h"
#include "host_staging.h"
#include "kernels_chunking.h"
#include "kernels_sort.h"
#include "kernels_subsample.h"
#include "scan.h"

namespace oc {

#define OC_HIP_CHECK(expr)                                                                                             \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw std::runtime_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));                      \
    } while (0)

#define OC_LAUNCH_CHECK() OC_HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------------------------------------
// Device arena
// ------------------------------------------------------------------------------------------------

static std::atomic<size_t> gAllocated{0};
static std::atomic<size_t> gPeak{0};

DeviceBuffer::~DeviceBuffer()
{
    release();
}

void DeviceBuffer::release()
{
    if (itsPtr)
    {
        (void)hipFree(itsPtr);
        gAllocated -= itsBytes;
    }
    itsPtr   = nullptr;
    itsBytes = 0;
}

void DeviceBuffer::ensure(size_t bytes)
{
    if (bytes <= itsBytes)
        return;
    release();
    size_t rounded = (bytes + 255) & ~size_t(255);
    void* p        = nullptr;
    hipError_t e   = hipMalloc(&p, rounded);
    if (e != hipSuccess)
        throw std::runtime_error("hipMalloc of " + std::to_string(rounded) + " bytes failed: " + hipGetErrorString(e));
    itsPtr   = p;
    itsBytes = rounded;
    size_t now = (gAllocated += rounded);
    size_t peak = gPeak.load();
    while (now > peak && !gPeak.compare_exchange_weak(peak, now))
    {
    }
}

size_t DeviceBuffer::totalAllocated()
{
    return gAllocated.load();
}

size_t DeviceBuffer::peakAllocated()
{
    return gPeak.load();
}

// ------------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------------

// ------------------------------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------------------------------

// Bytes of bitmap + prefix the nodes of one subsampling batch may use.
static constexpr uint64_t kBitmapBudget = 6ull << 30;

static uint32_t octreeDepth(uint32_t grid)  // tools.cu:63-88
{
    for (uint32_t d = 1; d <= 9; ++d)
        if (grid == (1u << d))
            return d;
    return 0;
}

OctreeBuilder::OctreeBuilder(int device) : itsDevice(device)
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    OC_HIP_CHECK(hipStreamCreateWithFlags(&itsStream, hipStreamNonBlocking));
    OC_HIP_CHECK(hipStreamCreateWithFlags(&itsCopyStream, hipStreamNonBlocking));
    OC_HIP_CHECK(hipEventCreateWithFlags(&itsBuiltEvent, hipEventDisableTiming));
    OC_HIP_CHECK(hipEventCreateWithFlags(&itsCopiedEvent, hipEventDisableTiming));
    OC_HIP_CHECK(hipEventCreateWithFlags(&itsUploadedEvent, hipEventDisableTiming));
    for (auto& ev : itsEvents)
        OC_HIP_CHECK(hipEventCreate(&ev));
    OC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&itsMailbox), 3 * sizeof(uint64_t), hipHostMallocCoherent));
    itsMailbox[0] = 0;
    itsMailbox[1] = 0;
    itsMailbox[2] = 0;
}

OctreeBuilder::OctreeBuilder(int device, const CloudInfo& cloud, const ProcessingInfo& proc,
                             const uint8_t* hostCloud, const uint8_t* deviceCloud)
    : OctreeBuilder(device)
{
    setup(cloud, proc, hostCloud, deviceCloud);
}

static uint64_t coordinateBytes(const CloudInfo& c)
{
    return c.cloudType == CLOUD_FLOAT_UINT8_T ? 12 : 24;
}

static uint64_t sortedPointBytes(const CloudInfo& c)
{
    return c.cloudType == CLOUD_FLOAT_UINT8_T ? sizeof(SortedPoint<float>) : sizeof(SortedPoint<double>);
}

void OctreeBuilder::setup(const CloudInfo& cloud, const ProcessingInfo& proc, const uint8_t* hostCloud,
                          const uint8_t* deviceCloud)
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    itsReady = false;
    itsBuilt = false;
    itsCloud = cloud;
    itsProc  = proc;
    itsDepth = octreeDepth(proc.chunkingGrid);
    itsDense = 0;
    for (uint32_t l = 0; l <= itsDepth; ++l)
    {
        uint32_t g          = proc.chunkingGrid >> l;
        itsLevelGrid[l]     = g;
        itsLevelOffset[l]   = itsDense;
        itsLevelNodes[l]    = g * g * g;
        itsDense += g * g * g;
    }
    itsCoarse  = std::min<uint32_t>(MAX_COARSE_LEVEL, itsDepth > 3 ? itsDepth - 3 : 0);
    itsBuckets = (proc.chunkingGrid >> itsCoarse) * (proc.chunkingGrid >> itsCoarse) * (proc.chunkingGrid >> itsCoarse);
    itsColorMode = proc.useIntraCellAvg ? COLOR_INTRA
                                        : (proc.useInterCellAvg ? (proc.useWeightingFunction ? COLOR_INTER_WEIGHTED
                                                                                            : COLOR_INTER)
                                                                : COLOR_NONE);
    itsE = static_cast<uint64_t>(static_cast<float>(cloud.pointAmount) * proc.outputFactor);  // point_cloud.cuh:24

    const uint64_t cloudBytes = static_cast<uint64_t>(cloud.pointAmount) * cloud.dataStride;
    if (deviceCloud)
    {
        itsCloudDevice = deviceCloud;
    }
    else
    {
        itsCloudOwned.ensure(cloudBytes);
        upload(itsCloudOwned.get(), hostCloud, cloudBytes);
        OC_HIP_CHECK(hipEventRecord(itsUploadedEvent, itsStream));
        itsCloudDevice = itsCloudOwned.as<uint8_t>();
    }

    // Fixed-size state: the dense pyramid, the node key/count arrays (bounded by min(D, levels*N))
    // and the output region.
    itsCount.ensure(sizeof(uint32_t) * static_cast<uint64_t>(itsDense));
    itsDenseToSparse.ensure(sizeof(int32_t) * static_cast<uint64_t>(itsDense));
    itsOccupancy.ensure(sizeof(uint32_t) * ((static_cast<uint64_t>(itsDense) + 31) / 32));
    uint64_t mBound = std::min<uint64_t>(itsDense, static_cast<uint64_t>(itsDepth + 1) * cloud.pointAmount);
    // The previous build's snapshot copies may still read the node keys and counts, the LUT and
    // the output: growing one of them frees it, so those copies are waited for first. (The other
    // arrays a snapshot reads are grown by the merge, after its stream has waited for the copies
    // and the host has read the node count back.)
    if (itsSnapshotPending && (sizeof(uint32_t) * mBound > itsNodeKey.capacity() ||
                               sizeof(uint32_t) * mBound > itsNodeCount.capacity() ||
                               sizeof(uint32_t) * itsE > itsLut.capacity() ||
                               sizeof(OutRecord) * itsE > itsOut.capacity()))
        waitSnapshot();
    itsNodeKey.ensure(sizeof(uint32_t) * mBound);
    itsNodeCount.ensure(sizeof(uint32_t) * mBound);
    itsLut.ensure(sizeof(uint32_t) * itsE);
    itsOut.ensure(sizeof(OutRecord) * itsE);
    const uint64_t coordBytes  = coordinateBytes(cloud);
    const uint64_t sortedBytes = sortedPointBytes(cloud);
    const uint64_t stagedCap   = std::max<uint64_t>(1, itsE - cloud.pointAmount);
    itsSorted.ensure(sortedBytes * cloud.pointAmount);
    itsBucketed.ensure(sortedBytes * cloud.pointAmount);
    itsFineCode.ensure(sizeof(uint16_t) * cloud.pointAmount);
    itsPerm.ensure(sizeof(uint32_t) * cloud.pointAmount);
    itsBucketCount.ensure(sizeof(uint32_t) * (itsBuckets + 1));
    itsBucketStart.ensure(sizeof(uint32_t) * (itsBuckets + 1));
    itsBucketCursor.ensure(sizeof(uint32_t) * (itsBuckets + 1));
    itsStLut.ensure(sizeof(uint32_t) * stagedCap);
    itsStOut.ensure(sizeof(OutRecord) * stagedCap);
    itsStPts.ensure(coordBytes * stagedCap);
    if (itsColorMode == COLOR_INTER_WEIGHTED)
        itsStRgb.ensure(sizeof(uint32_t) * stagedCap);
    itsScalars.ensure(sizeof(uint64_t) * SC_COUNT);
    // The output starts zeroed (the tail past N + I stays zero, A15). The fill goes on the copy
    // stream, behind the previous build's snapshot copies that still read the output; the build
    // waits for that stream before its first write to anything a snapshot reads (generate).
    OC_HIP_CHECK(hipMemsetAsync(itsOut.get(), 0, sizeof(OutRecord) * itsE, itsCopyStream));
    OC_HIP_CHECK(hipEventRecord(itsCopiedEvent, itsCopyStream));
    itsSnapshotPending = true;
    // the caller's cloud buffer is borrowed until this returns: wait for the upload only
    if (!deviceCloud)
        OC_HIP_CHECK(hipEventSynchronize(itsUploadedEvent));
    itsReady = true;
}

std::vector<DeviceBuffer*> OctreeBuilder::arena()
{
    return {&itsCloudOwned,  &itsCount,       &itsDenseToSparse, &itsOccupancy,
            &itsNodeKey,     &itsNodeCount,   &itsNodePoints,
            &itsNodeData,    &itsNodeParent,  &itsNodeChild,     &itsNodeFlags,    &itsNodeSorted,  &itsNodeAux,
            &itsNodeStage,   &itsNodeSub,     &itsNodeRegion,    &itsBucketed,     &itsSorted,      &itsFineCode,
            &itsPerm,
            &itsBucketCount, &itsBucketStart, &itsBucketCursor,  &itsLut,          &itsOut,
            &itsStLut,       &itsStOut,       &itsStPts,         &itsStRgb,        &itsInternalList, &itsSegCount,    &itsSegSrc,
            &itsSegTile,     &itsSegInput,    &itsNodeGrid,      &itsTileList,     &itsBitmap,      &itsPrefix,
            &itsNodeFirst,   &itsWin,         &itsColor,         &itsCellCnt,      &itsCellStart,   &itsCellKey,
            &itsCellNode,    &itsCellOf,      &itsRankOf,        &itsPayload,      &itsScalars,
            &itsScanScratch, &itsLeafList,   &itsSplitChunks,   &itsSplitHist};
}

uint64_t OctreeBuilder::arenaBytes() const
{
    uint64_t total = 0;
    for (DeviceBuffer* b : const_cast<OctreeBuilder*>(this)->arena())
        total += b->capacity();
    return total;
}

void OctreeBuilder::releaseArena()
{
    OC_HIP_CHECK(hipStreamSynchronize(itsCopyStream));
    itsSnapshotPending = false;
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
    for (DeviceBuffer* b : arena())
        b->release();
    itsReady = false;
    itsBuilt = false;
}

uint64_t OctreeBuilder::estimateFootprint(const CloudInfo& cloud, const ProcessingInfo& proc, bool hostCloud)
{
    const uint32_t depth = octreeDepth(proc.chunkingGrid);
    uint64_t dense       = 0;
    for (uint32_t l = 0; l <= depth; ++l)
    {
        const uint64_t g = proc.chunkingGrid >> l;
        dense += g * g * g;
    }
    const uint64_t n      = cloud.pointAmount;
    const uint64_t e      = static_cast<uint64_t>(static_cast<float>(cloud.pointAmount) * proc.outputFactor);
    const uint64_t staged = e > n ? e - n : 1;
    const uint64_t mBound = std::min<uint64_t>(dense, (depth + 1) * n);
    uint64_t bytes        = hostCloud ? n * cloud.dataStride : 0;
    bytes += 8 * dense + 4 * ((dense + 31) / 32);          // count pyramid, dense -> sparse, occupancy bits
    bytes += 8 * mBound + 77 * mBound + n / 1024;           // node keys / counts, node arrays, leaf chunks
    bytes += 22 * e;                                        // lut + output records
    bytes += (2 * sortedPointBytes(cloud) + 2 + 4) * n;     // bucketed, sorted, fine codes, permutation
    if (n > FINE_SPLIT)                                     // split buckets of the fine sort
        bytes += (2 * n / FINE_SPLIT + 1) * (8 + 4ull * MAX_FINE_CELLS);
    bytes += (4 + 18 + coordinateBytes(cloud)) * staged;     // staging
    bytes += 8 * (staged / PLACE_CHUNK + 1);                // placement work list
    if (proc.useInterCellAvg && proc.useWeightingFunction && !proc.useIntraCellAvg)
        bytes += 4 * staged;                                // staged colours of the weighted filter
    // per-level work: inputs (cell, then its dense index; rank; payload) and cells (winner, colour sum,
    // count, start, key, node), both bounded by E; the occupancy bitmaps and their prefix
    const uint64_t wordsPerNode = (rowsPerNode(proc.subsamplingGrid) * proc.subsamplingGrid + 31) / 32;
    const uint64_t levelNodes   = std::min<uint64_t>(mBound, n / std::max<uint32_t>(1, proc.mergingThreshold) + 8);
    bytes += (8 + sizeof(WeightPayload) + sizeof(uint64_t)) * e + 32 * staged;
    bytes += std::min<uint64_t>(kBitmapBudget, 8 * wordsPerNode * levelNodes);
    return bytes;
}

OctreeBuilder::~OctreeBuilder()
{
    if (itsCopyStream)
    {
        (void)hipStreamSynchronize(itsCopyStream);
        (void)hipStreamDestroy(itsCopyStream);
    }
    if (itsStream)
    {
        (void)hipStreamSynchronize(itsStream);
        (void)hipStreamDestroy(itsStream);
    }
    if (itsBuiltEvent)
        (void)hipEventDestroy(itsBuiltEvent);
    if (itsCopiedEvent)
        (void)hipEventDestroy(itsCopiedEvent);
    if (itsUploadedEvent)
        (void)hipEventDestroy(itsUploadedEvent);
    for (auto& ev : itsEvents)
        if (ev)
            (void)hipEventDestroy(ev);
    for (auto& quad : itsBatchEvents)
        for (auto& ev : quad)
            if (ev)
                (void)hipEventDestroy(ev);
    if (itsMailbox)
        (void)hipHostFree(itsMailbox);
}

std::array<hipEvent_t, 4>& OctreeBuilder::batchEvents()
{
    if (itsBatchesRun == itsBatchEvents.size())
    {
        std::array<hipEvent_t, 4> quad{};
        for (auto& ev : quad)
            OC_HIP_CHECK(hipEventCreate(&ev));
        itsBatchEvents.push_back(quad);
    }
    return itsBatchEvents[itsBatchesRun++];
}

void OctreeBuilder::syncStream()
{
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
}

// Publishes a device value to the host mailbox: the value first, then (after a system-scope
// release) the sequence number the host waits for.
template <typename V>
__global__ void kPublish(const V* __restrict__ src, uint64_t* box, uint64_t seq)
{
    if (threadIdx.x != 0)
        return;
    __hip_atomic_store(box, static_cast<uint64_t>(*src), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(box + 2, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Two consecutive device words through the mailbox (one wait for two sizes).
__global__ void kPublishPair(const uint64_t* __restrict__ src, uint64_t* box, uint64_t seq)
{
    if (threadIdx.x != 0)
        return;
    __hip_atomic_store(box, src[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(box + 1, src[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(box + 2, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A size the host needs to size the next launches. The stream is drained up to the publish: the
// host spins on the mailbox for up to 2 ms (the common case: the value is ready within microseconds
// of the last kernel) and then falls back to a blocking synchronisation, which also reports a
// failed kernel instead of waiting forever.
void OctreeBuilder::waitMailbox(uint64_t seq)
{
    const auto start = std::chrono::steady_clock::now();
    while (__atomic_load_n(itsMailbox + 2, __ATOMIC_ACQUIRE) != seq)
    {
        __builtin_ia32_pause();
        if (std::chrono::steady_clock::now() - start > std::chrono::milliseconds(2))
        {
            OC_HIP_CHECK(hipStreamSynchronize(itsStream));
            if (__atomic_load_n(itsMailbox + 2, __ATOMIC_ACQUIRE) != seq)
                throw std::runtime_error("device value was not published");
            break;
        }
    }
}

template <typename V>
V OctreeBuilder::readDevice(const V* devicePtr)
{
    const uint64_t seq = ++itsMailSeq;
    kPublish<V><<<1, 64, 0, itsStream>>>(devicePtr, itsMailbox, seq);
    OC_LAUNCH_CHECK();
    waitMailbox(seq);
    return static_cast<V>(__atomic_load_n(itsMailbox, __ATOMIC_ACQUIRE));
}

std::array<uint64_t, 2> OctreeBuilder::readDevicePair(const uint64_t* devicePtr)
{
    const uint64_t seq = ++itsMailSeq;
    kPublishPair<<<1, 64, 0, itsStream>>>(devicePtr, itsMailbox, seq);
    OC_LAUNCH_CHECK();
    waitMailbox(seq);
    return {__atomic_load_n(itsMailbox, __ATOMIC_ACQUIRE), __atomic_load_n(itsMailbox + 1, __ATOMIC_ACQUIRE)};
}

void OctreeBuilder::generate()
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    if (!itsReady)
        throw std::runtime_error("the builder has not been set up for a cloud");
    itsBuilt = false;
    if (itsCloud.cloudType == CLOUD_FLOAT_UINT8_T)
        generateTyped<float>();
    else
        generateTyped<double>();
    itsBuilt = true;
}

template <typename T>
void OctreeBuilder::generateTyped()
{
    uint64_t* sc = itsScalars.as<uint64_t>();
    kFill64<<<1, BLOCK, 0, itsStream>>>(sc, SC_COUNT, 0);
    OC_LAUNCH_CHECK();
    itsBatchesRun = 0;

    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_START], itsStream));
    countPoints<T>();  // writes nothing a snapshot reads, so it overlaps the previous one's copies
    progress();
    waitForSnapshotOnStream();
    mergeLevels();
    progress();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_MERGE_END], itsStream));
    distributePoints<T>();
    progress();
    subsampleLevels<T>();
    placeInternalNodes();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_END], itsStream));
    syncStream();

    uint32_t err = static_cast<uint32_t>(readDevice(sc + SC_ERROR));
    if (err & ERR_OUTSIDE)
        throw std::runtime_error("a point lies outside the cubic bounding box");
    if (err & ERR_CAPACITY)
        throw std::runtime_error("a node has more than 2^31 input points");

    auto ms = [&](int a, int b) {
        float v = 0.f;
        OC_HIP_CHECK(hipEventElapsedTime(&v, itsEvents[a], itsEvents[b]));
        return v;
    };
    itsTimings.assign(7, 0.f);
    itsTimings[0] = ms(EV_COUNT_BEGIN, EV_COUNT_END);
    itsTimings[1] = ms(EV_COUNT_END, EV_MERGE_END);
    itsTimings[2] = ms(EV_DIST_BEGIN, EV_DIST_END);
    itsTimings[3] = ms(EV_DIST_END, EV_END);
    itsTimings[4] = ms(EV_START, EV_END);
    for (uint32_t i = 0; i < itsBatchesRun; ++i)
    {
        float a = 0.f, b = 0.f;
        OC_HIP_CHECK(hipEventElapsedTime(&a, itsBatchEvents[i][0], itsBatchEvents[i][1]));
        OC_HIP_CHECK(hipEventElapsedTime(&b, itsBatchEvents[i][2], itsBatchEvents[i][3]));
        itsTimings[5] += a;
        itsTimings[6] += b;
    }
}

template <typename T>
void OctreeBuilder::countPoints()
{
    uint64_t* sc       = itsScalars.as<uint64_t>();
    const uint32_t n   = itsCloud.pointAmount;
    const uint32_t nb  = itsBuckets;
    uint32_t* bucketCount  = itsBucketCount.as<uint32_t>();
    uint32_t* bucketStart  = itsBucketStart.as<uint32_t>();
    uint32_t* bucketCursor = itsBucketCursor.as<uint32_t>();
    Gridding g{itsCloud.bbMin[0], itsCloud.bbMin[1], itsCloud.bbMin[2], itsCloud.bbMax[0] - itsCloud.bbMin[0],
               itsProc.chunkingGrid};
    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(nb) + 1));
    kFill32<<<gridFor(nb + 1), BLOCK, 0, itsStream>>>(bucketCount, nb + 1, 0);
    OC_LAUNCH_CHECK();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_COUNT_BEGIN], itsStream));

    const unsigned histBlocks = static_cast<unsigned>(std::min<uint64_t>(2048, (n + 8191) / 8192));
    kBucketHist<T><<<histBlocks, HIST_THREADS, 0, itsStream>>>(itsCloudDevice, n, itsCloud.dataStride, g,
                                                               3 * itsCoarse, nb, bucketCount,
                                                               reinterpret_cast<uint32_t*>(sc + SC_ERROR));
    OC_LAUNCH_CHECK();
    scanExclusive<uint32_t>(nb, ArrayIn{bucketCount}, BucketStartOut{bucketStart, bucketCursor},
                            itsScanScratch.as<uint32_t>(), bucketStart + nb, itsStream);
    OC_LAUNCH_CHECK();
    const unsigned scatterBlocks = static_cast<unsigned>((static_cast<uint64_t>(n) + STAGE_CHUNK - 1) / STAGE_CHUNK);
    kBucketScatter<T><<<scatterBlocks, SCATTER_THREADS, 0, itsStream>>>(itsCloudDevice, n, itsCloud.dataStride, g,
                                                                     3 * itsCoarse, nb, bucketCursor,
                                                                     itsBucketed.as<SortedPoint<T>>(),
                                                                     itsFineCode.as<uint16_t>());
    OC_LAUNCH_CHECK();
    FineSortParams fp{};
    fp.lc         = itsCoarse;
    fp.bucketGrid = itsProc.chunkingGrid >> itsCoarse;
    for (uint32_t l = 0; l <= itsDepth; ++l)
    {
        fp.levelOffset[l] = itsLevelOffset[l];
        fp.levelGrid[l]   = itsLevelGrid[l];
    }
    // the level-0 occupancy bits with the counts where a bucket row is a whole bitmap word
    uint32_t* occ0 = levelZeroBits() ? itsOccupancy.as<uint32_t>() : nullptr;
    kFineSort<<<nb, SORT_THREADS, 0, itsStream>>>(itsFineCode.as<uint16_t>(), bucketStart, fp,
                                                  itsCount.as<uint32_t>(), itsPerm.as<uint32_t>(), occ0);
    OC_LAUNCH_CHECK();
    if (n > FINE_SPLIT)
    {
        // buckets above FINE_SPLIT points, a chunk per workgroup (at most 2 n / FINE_SPLIT chunks)
        const uint64_t bound = 2ull * n / FINE_SPLIT + 1;
        itsSplitChunks.ensure(sizeof(uint64_t) * bound);
        itsSplitHist.ensure(sizeof(uint32_t) * MAX_FINE_CELLS * bound);
        uint32_t* chunkCount = reinterpret_cast<uint32_t*>(itsScalars.as<uint64_t>() + SC_SPLIT);
        scanExclusive<uint32_t>(nb, SplitChunksIn{bucketStart}, SplitChunksOut{itsSplitChunks.as<uint64_t>()},
                                itsScanScratch.as<uint32_t>(), chunkCount, itsStream);
        OC_LAUNCH_CHECK();
        const unsigned grid = static_cast<unsigned>((bound + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT);
        kFineChunkHist<<<grid, SORT_THREADS, 0, itsStream>>>(itsFineCode.as<uint16_t>(), bucketStart,
                                                             itsSplitChunks.as<uint64_t>(), chunkCount, fp.lc,
                                                             itsSplitHist.as<uint32_t>());
        OC_LAUNCH_CHECK();
        kFineSplitScan<<<grid, SORT_THREADS, 0, itsStream>>>(bucketStart, itsSplitChunks.as<uint64_t>(), chunkCount,
                                                             fp, itsCount.as<uint32_t>(), itsSplitHist.as<uint32_t>(),
                                                             occ0);
        OC_LAUNCH_CHECK();
        kFineChunkPerm<<<grid, SORT_THREADS, 0, itsStream>>>(itsFineCode.as<uint16_t>(), bucketStart,
                                                             itsSplitChunks.as<uint64_t>(), chunkCount, fp.lc,
                                                             itsSplitHist.as<uint32_t>(), itsPerm.as<uint32_t>());
        OC_LAUNCH_CHECK();
    }
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_COUNT_END], itsStream));  // the records move with the distribution
}

bool OctreeBuilder::levelZeroBits() const
{
    // 32-cell rows starting at multiples of 32, and level 1 starting at a multiple of 64
    return itsCoarse == MAX_COARSE_LEVEL && itsProc.chunkingGrid % 64 == 0;
}

void OctreeBuilder::mergeLevels()
{
    uint64_t* sc    = itsScalars.as<uint64_t>();
    uint32_t* count = itsCount.as<uint32_t>();
    for (uint32_t l = itsCoarse + 1; l <= itsDepth; ++l)  // levels 0..coarse come from the fine sort
    {
        kPropagate<<<gridFor(itsLevelNodes[l]), BLOCK, 0, itsStream>>>(count, itsLevelOffset[l], itsLevelGrid[l],
                                                                        itsLevelOffset[l - 1], itsLevelGrid[l - 1]);
        OC_LAUNCH_CHECK();
    }

    // occupancy bits of the levels the fine sort has not marked, then sparse ids (A2) + node keys /
    // counts + first id of every level, by one scan over the bitmap's words
    uint32_t* occ         = itsOccupancy.as<uint32_t>();
    const uint64_t marked = levelZeroBits() ? itsLevelOffset[1] : 0;
    kOccupancyBits<<<gridFor(itsDense - marked), BLOCK, 0, itsStream>>>(count, marked, itsDense, occ);
    OC_LAUNCH_CHECK();
    uint32_t* levelStart = reinterpret_cast<uint32_t*>(sc + SC_LEVEL_START);
    const uint64_t words = (static_cast<uint64_t>(itsDense) + 31) / 32;
    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(words) + 1));
    OccupiedWordOut occOut{};
    occOut.occ           = occ;
    occOut.denseToSparse = itsDenseToSparse.as<int32_t>();
    occOut.nodeKey       = itsNodeKey.as<uint32_t>();
    occOut.nodeCount     = itsNodeCount.as<uint32_t>();
    occOut.count         = count;
    scanExclusive<uint32_t>(words, OccupiedWordIn{occ}, occOut, itsScanScratch.as<uint32_t>(),
                            reinterpret_cast<uint32_t*>(sc + SC_M), itsStream);
    OC_LAUNCH_CHECK();
    LevelOffsets lo{};
    for (uint32_t l = 0; l <= itsDepth; ++l)
        lo.offset[l] = itsLevelOffset[l];
    lo.levels = itsDepth + 1;
    kLevelStarts<<<1, 64, 0, itsStream>>>(itsNodeKey.as<uint32_t>(), reinterpret_cast<uint32_t*>(sc + SC_M), lo,
                                          levelStart);
    OC_LAUNCH_CHECK();

    uint32_t hostStart[MAX_LEVELS + 1] = {};
    OC_HIP_CHECK(hipMemcpyAsync(hostStart, levelStart, sizeof(uint32_t) * (itsDepth + 1), hipMemcpyDeviceToHost,
                                itsStream));
    uint64_t mRaw = 0, errRaw = 0;
    OC_HIP_CHECK(hipMemcpyAsync(&mRaw, sc + SC_M, sizeof(uint64_t), hipMemcpyDeviceToHost, itsStream));
    OC_HIP_CHECK(hipMemcpyAsync(&errRaw, sc + SC_ERROR, sizeof(uint64_t), hipMemcpyDeviceToHost, itsStream));
    syncStream();
    if (errRaw & ERR_OUTSIDE)
        throw std::runtime_error("a point lies outside the cubic bounding box");
    itsM = static_cast<uint32_t>(mRaw);
    for (uint32_t l = 0; l <= itsDepth; ++l)
        itsLevelStart[l] = hostStart[l];
    itsLevelStart[itsDepth + 1] = itsM;
    if (itsM == 0)
        throw std::runtime_error("the cloud has no points inside the bounding box");

    const uint64_t m = itsM;
    itsNodePoints.ensure(sizeof(uint32_t) * m);
    itsNodeData.ensure(sizeof(uint64_t) * m);
    itsNodeParent.ensure(sizeof(int32_t) * m);
    itsNodeChild.ensure(sizeof(int32_t) * 8 * m);
    itsNodeFlags.ensure(m);
    itsNodeSorted.ensure(sizeof(uint32_t) * m);
    itsNodeAux.ensure(sizeof(uint32_t) * (m + 1));
    itsNodeStage.ensure(sizeof(uint32_t) * m);
    itsNodeSub.ensure(sizeof(uint32_t) * m);
    itsNodeRegion.ensure(sizeof(uint32_t) * m);
    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(m) + 1));

    LevelParams lp{};
    for (uint32_t l = 0; l <= itsDepth; ++l)
    {
        lp.offset[l] = itsLevelOffset[l];
        lp.grid[l]   = itsLevelGrid[l];
    }
    lp.depth     = itsDepth;
    lp.threshold = itsProc.mergingThreshold;
    kNodeLinks<<<gridFor(m), BLOCK, 0, itsStream>>>(
            itsM, lp, itsNodeKey.as<uint32_t>(), itsNodeCount.as<uint32_t>(), count, itsDenseToSparse.as<int32_t>(),
            itsNodePoints.as<uint32_t>(), itsNodeData.as<uint64_t>(), itsNodeParent.as<int32_t>(),
            itsNodeChild.as<int32_t>(), itsNodeFlags.as<uint8_t>(), itsNodeAux.as<uint32_t>(),
            itsNodeSub.as<uint32_t>(), occ);
    OC_LAUNCH_CHECK();

    // leaf data blocks (A4): exclusive scan of the block sizes in id order
    scanExclusive<uint32_t>(m, ArrayIn{itsNodeAux.as<uint32_t>()}, ExclusiveToArray{itsNodeAux.as<uint32_t>()},
                            itsScanScratch.as<uint32_t>(), reinterpret_cast<uint32_t*>(sc + SC_LEAF_TOTAL),
                            itsStream);
    OC_LAUNCH_CHECK();
    kLeafDataIdx<<<gridFor(m), BLOCK, 0, itsStream>>>(itsM, itsProc.mergingThreshold, itsNodeAux.as<uint32_t>(),
                                                       itsNodeFlags.as<uint8_t>(), itsNodeChild.as<int32_t>(),
                                                       itsNodePoints.as<uint32_t>(), itsNodeData.as<uint64_t>());
    OC_LAUNCH_CHECK();
    kCellLeaf<<<gridFor(itsLevelStart[1]), BLOCK, 0, itsStream>>>(itsLevelStart[1], itsNodeKey.as<uint32_t>(),
                                                                   itsNodeFlags.as<uint8_t>(),
                                                                   itsNodeParent.as<int32_t>(), count);
    OC_LAUNCH_CHECK();
}

template <typename T>
void OctreeBuilder::distributePoints()
{
    Gridding g{itsCloud.bbMin[0], itsCloud.bbMin[1], itsCloud.bbMin[2], itsCloud.bbMax[0] - itsCloud.bbMin[0],
               itsProc.chunkingGrid};
    // octree_chunking.cu:108-112
    double sx = 1.0 / itsCloud.scale[0], sy = 1.0 / itsCloud.scale[1], sz = 1.0 / itsCloud.scale[2];
    LevelParams lp{};
    for (uint32_t l = 0; l <= itsDepth; ++l)
    {
        lp.offset[l] = itsLevelOffset[l];
        lp.grid[l]   = itsLevelGrid[l];
    }
    lp.depth     = itsDepth;
    lp.threshold = itsProc.mergingThreshold;
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_DIST_BEGIN], itsStream));
    kSortedStart<<<gridFor(itsM), BLOCK, 0, itsStream>>>(itsM, lp, itsNodeKey.as<uint32_t>(),
                                                         itsNodeCount.as<uint32_t>(), itsNodeFlags.as<uint8_t>(),
                                                         itsNodeParent.as<int32_t>(), itsNodeChild.as<int32_t>(),
                                                         itsNodeSorted.as<uint32_t>());
    OC_LAUNCH_CHECK();
    // the records move to the sorted cloud: leaf and cell order when there are internal nodes,
    // plain Morton order otherwise; then the leaf data blocks are copied from it
    const uint32_t n = itsCloud.pointAmount;
    if (!orderLeafPoints<T>())
    {
        kGatherSorted<T><<<gatherGrid(n), BLOCK, 0, itsStream>>>(itsBucketed.as<SortedPoint<T>>(),
                                                                     itsPerm.as<uint32_t>(), n,
                                                                     itsSorted.as<SortedPoint<T>>());
        OC_LAUNCH_CHECK();
    }
    kDistributeSorted<T><<<gridFor(n, 16384), BLOCK, 0, itsStream>>>(
            itsSorted.as<SortedPoint<T>>(), n, g, itsCount.as<uint32_t>(), itsNodeSorted.as<uint32_t>(),
            itsNodeData.as<uint64_t>(), itsLut.as<uint32_t>(), itsOut.as<OutRecord>(), sx, sy, sz);
    OC_LAUNCH_CHECK();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_DIST_END], itsStream));
}

template <typename T, uint32_t MODE>
static void launchAccumulate(bool random, unsigned blocks, hipStream_t s, uint32_t tileBase, uint32_t nodeBase,
                             uint64_t words, const uint32_t* tileList, const uint32_t* segTile,
                             const uint32_t* segCount, const uint32_t* segSrc, const uint32_t* segInput,
                             const NodeGrid* nodeGrid, const uint32_t* bitmap, const uint32_t* prefix,
                             uint32_t* cellOf, const PointStore<T>& ps, uint64_t* win, uint64_t* color,
                             uint32_t* cellCnt, uint32_t* cellKey, uint32_t* cellNode, uint32_t* rankOf)
{
    if (random)
        kAccumulate<T, MODE, true><<<blocks, TILE, 0, s>>>(tileBase, nodeBase, words, tileList, segTile, segCount,
                                                            segSrc, segInput, nodeGrid, bitmap, prefix, cellOf, ps,
                                                            win, color, cellCnt, cellKey, cellNode, rankOf);
    else
        kAccumulate<T, MODE, false><<<blocks, TILE, 0, s>>>(tileBase, nodeBase, words, tileList, segTile, segCount,
                                                             segSrc, segInput, nodeGrid, bitmap, prefix, cellOf, ps,
                                                             win, color, cellCnt, cellKey, cellNode, rankOf);
}


// The sorted cloud in leaf and cell order (kLeafCellOrder): false when no leaf has a parent (the
// root is the only leaf), so nothing is subsampled and plain Morton order serves.
template <typename T>
bool OctreeBuilder::orderLeafPoints()
{
    if (itsProc.subsamplingGrid < 2 || itsM < 2)
        return false;
    uint64_t* sc = itsScalars.as<uint64_t>();
    // chunks per leaf are at most N / LEAF_CHUNK + (number of leaves) in total
    itsLeafList.ensure(sizeof(uint64_t) * (itsM + itsCloud.pointAmount / LEAF_CHUNK + 1));
    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(itsM) + 1));
    scanExclusive<uint32_t>(itsM, LeafIn{itsNodeFlags.as<uint8_t>(), itsNodeParent.as<int32_t>(),
                                         itsNodeCount.as<uint32_t>()},
                            LeafOut{itsLeafList.as<uint64_t>()}, itsScanScratch.as<uint32_t>(),
                            reinterpret_cast<uint32_t*>(sc + SC_TABLE_TOTAL), itsStream);
    OC_LAUNCH_CHECK();
    const uint32_t chunks = static_cast<uint32_t>(readDevice(sc + SC_TABLE_TOTAL));
    if (chunks == 0)
        return false;  // the root is the only leaf: nothing is subsampled
    const double cubicSize = itsCloud.bbMax[0] - itsCloud.bbMin[0];  // metadata.cuh:66-69
    LeafOrderParams p{};
    for (int a = 0; a < 3; ++a)
        p.cubicMin[a] = itsCloud.bbMin[a];
    for (uint32_t l = 0; l <= itsDepth; ++l)
    {
        p.side[l]      = cubicSize / std::pow(2.0, static_cast<double>(itsDepth - l));  // rp_subsampling.cu:17-18
        p.sideHalf[l]  = p.side[l] / 2.0;
        p.lp.offset[l] = itsLevelOffset[l];
        p.lp.grid[l]   = itsLevelGrid[l];
    }
    p.lp.depth     = itsDepth;
    p.lp.threshold = itsProc.mergingThreshold;
    p.g            = itsProc.subsamplingGrid;
    const uint32_t h    = p.g > 1 ? p.g / 2 : 1;  // cells per axis of a leaf's octant, as kLeafCellOrder
    const uint64_t keys = rowsPerNode(h) * ((h + 31) / 32);
    while (((keys - 1) >> p.keyShift) >= LEAF_KEYS)
        ++p.keyShift;
    kLeafCellOrder<T><<<chunks, LEAF_THREADS, 0, itsStream>>>(
            itsLeafList.as<uint64_t>(), p, itsNodeKey.as<uint32_t>(), itsNodeParent.as<int32_t>(),
            itsNodeCount.as<uint32_t>(), itsNodeSorted.as<uint32_t>(), itsBucketed.as<SortedPoint<T>>(),
            itsPerm.as<uint32_t>(), itsSorted.as<SortedPoint<T>>());
    OC_LAUNCH_CHECK();
    return true;
}

template <typename T>
void OctreeBuilder::subsampleLevels()
{
    uint64_t* sc     = itsScalars.as<uint64_t>();
    uint32_t* err    = reinterpret_cast<uint32_t*>(sc + SC_ERROR);
    const uint32_t N = itsCloud.pointAmount;
    const uint32_t g = itsProc.subsamplingGrid;
    const double cubicSize   = itsCloud.bbMax[0] - itsCloud.bbMin[0];  // metadata.cuh:66-69
    const uint64_t cellsPerNode = rowsPerNode(g) * g;
    const uint64_t wordsPerNode = (cellsPerNode + 31) / 32;
    const uint32_t batchNodes   = static_cast<uint32_t>(
            std::max<uint64_t>(1, kBitmapBudget / (2 * sizeof(uint32_t) * wordsPerNode)));
    const bool packed   = itsColorMode == COLOR_INTRA || itsColorMode == COLOR_INTER;
    const bool weighted = itsColorMode == COLOR_INTER_WEIGHTED;
    uint64_t cursor     = N;
    PointStore<T> ps{};
    ps.sorted = itsSorted.as<SortedPoint<T>>();
    ps.stLut  = itsStLut.as<uint32_t>();
    ps.stOut  = itsStOut.as<OutRecord>();
    ps.stPts  = itsStPts.as<T>();
    ps.stRgb  = weighted ? itsStRgb.as<uint32_t>() : nullptr;
    ps.N      = N;
    ps.cubic  = Gridding{itsCloud.bbMin[0], itsCloud.bbMin[1], itsCloud.bbMin[2], cubicSize, itsProc.chunkingGrid};
    ps.sx     = 1.0 / itsCloud.scale[0];
    ps.sy     = 1.0 / itsCloud.scale[1];
    ps.sz     = 1.0 / itsCloud.scale[2];

    SubsampleParams sp{};
    for (int a = 0; a < 3; ++a)
        sp.cubicMin[a] = itsCloud.bbMin[a];
    for (uint32_t L = 0; L <= itsDepth; ++L)
        sp.sideHalf[L] = (cubicSize / std::pow(2.0, static_cast<double>(itsDepth - L))) / 2.0;
    sp.depth = itsDepth;
    sp.g     = g;
    sp.seed  = itsProc.seed;

    for (uint32_t l = 1; l <= itsDepth; ++l)
    {
        const uint32_t ls          = itsLevelStart[l];
        const uint32_t levelNodes  = itsLevelStart[l + 1] - ls;
        if (levelNodes == 0)
            continue;
        sp.level       = l;
        sp.levelOffset = itsLevelOffset[l];
        sp.levelGrid   = itsLevelGrid[l];
        sp.side        = cubicSize / std::pow(2.0, static_cast<double>(itsDepth - l));  // rp_subsampling.cu:17-18

        // the internal nodes of the level, their input segments and griddings, the segment scan
        // and the tile list: one workgroup and one readback for small levels, four launches and
        // two readbacks otherwise (inputs <= cursor: leaf points plus everything subsampled so far)
        const bool smallLevel = levelNodes <= SMALL_LEVEL_NODES && cursor <= SMALL_LEVEL_POINTS;
        itsInternalList.ensure(sizeof(uint32_t) * levelNodes);
        itsSegCount.ensure(sizeof(uint32_t) * 8 * levelNodes);
        itsSegSrc.ensure(sizeof(uint32_t) * 8 * levelNodes);
        itsSegTile.ensure(sizeof(uint32_t) * 8 * levelNodes);
        itsSegInput.ensure(sizeof(uint32_t) * 8 * levelNodes);
        itsNodeGrid.ensure(sizeof(NodeGrid) * levelNodes);
        auto tables = [&]() {
            return LevelTables{itsNodeKey.as<uint32_t>(),     itsNodeFlags.as<uint8_t>(),   itsNodeChild.as<int32_t>(),
                               itsNodePoints.as<uint32_t>(),  itsNodeSorted.as<uint32_t>(), itsNodeStage.as<uint32_t>(),
                               itsInternalList.as<uint32_t>(), itsSegCount.as<uint32_t>(),  itsSegSrc.as<uint32_t>(),
                               itsSegTile.as<uint32_t>(),     itsSegInput.as<uint32_t>(),   itsTileList.as<uint32_t>(),
                               itsNodeGrid.as<NodeGrid>()};
        };
        uint32_t K        = 0;
        uint64_t segTotal = 0;
        if (smallLevel)
        {
            itsTileList.ensure(sizeof(uint32_t) * (cursor / TILE + 8ull * levelNodes + 1));
            kLevelSetup<<<1, SCAN_THREADS, 0, itsStream>>>(ls, levelNodes, sp, tables(),
                                                           reinterpret_cast<uint32_t*>(sc + SC_TABLE_TOTAL),
                                                           sc + SC_SEG_TOTAL);
            OC_LAUNCH_CHECK();
            const auto v = readDevicePair(sc + SC_SEG_TOTAL);  // segment total, internal-node count
            segTotal     = v[0];
            K            = static_cast<uint32_t>(v[1]);
            if (K == 0)
                continue;
        }
        else
        {
            itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(8ull * levelNodes) + 1));
            scanExclusive<uint32_t>(levelNodes, InternalIn{itsNodeFlags.as<uint8_t>(), ls},
                                    InternalOut{itsInternalList.as<uint32_t>(), ls}, itsScanScratch.as<uint32_t>(),
                                    reinterpret_cast<uint32_t*>(sc + SC_TABLE_TOTAL), itsStream);
            OC_LAUNCH_CHECK();
            K = static_cast<uint32_t>(readDevice(sc + SC_TABLE_TOTAL));
            if (K == 0)
                continue;
            kSegments<<<gridFor(K), BLOCK, 0, itsStream>>>(K, sp, tables());
            OC_LAUNCH_CHECK();
            scanExclusive<uint64_t>(8ull * K, SegmentIn{itsSegCount.as<uint32_t>()},
                                    SegmentOut{itsSegTile.as<uint32_t>(), itsSegInput.as<uint32_t>()},
                                    itsScanScratch.as<uint64_t>(), sc + SC_SEG_TOTAL, itsStream);
            OC_LAUNCH_CHECK();
            segTotal = readDevice(sc + SC_SEG_TOTAL);
            itsTileList.ensure(sizeof(uint32_t) * (segTotal >> 32));
            kTileList<<<gridFor(8ull * K), BLOCK, 0, itsStream>>>(8ull * K, itsSegCount.as<uint32_t>(),
                                                                  itsSegTile.as<uint32_t>(), itsTileList.as<uint32_t>());
            OC_LAUNCH_CHECK();
        }
        const uint32_t tiles  = static_cast<uint32_t>(segTotal >> 32);
        const uint32_t inputs = static_cast<uint32_t>(segTotal);
        if (inputs == 0)
            continue;

        itsCellOf.ensure(sizeof(uint32_t) * inputs);
        if (weighted)
        {
            itsRankOf.ensure(sizeof(uint32_t) * inputs);
            itsPayload.ensure((sizeof(WeightPayload) + sizeof(uint64_t)) * inputs);  // payloads, then winner keys
        }
        uint64_t* winKey =
                weighted ? reinterpret_cast<uint64_t*>(itsPayload.as<uint8_t>() + sizeof(WeightPayload) * inputs) : nullptr;

        for (uint32_t k0 = 0; k0 < K; k0 += batchNodes)
        {
            const uint32_t k1 = std::min<uint32_t>(K, k0 + batchNodes);
            const uint32_t Kb = k1 - k0;
            uint32_t tileLo = 0, tileHi = tiles;
            if (k0 > 0)
                tileLo = readDevice(itsSegTile.as<uint32_t>() + 8ull * k0);
            if (k1 < K)
                tileHi = readDevice(itsSegTile.as<uint32_t>() + 8ull * k1);
            const unsigned blocks = tileHi - tileLo;
            const uint64_t words  = static_cast<uint64_t>(Kb) * wordsPerNode;
            itsBitmap.ensure(sizeof(uint32_t) * words);
            itsPrefix.ensure(sizeof(uint32_t) * (words + 1));
            itsNodeFirst.ensure(sizeof(uint32_t) * (Kb + 1));
            itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(words) + 1));
            uint32_t* bitmap = itsBitmap.as<uint32_t>();
            uint32_t* prefix = itsPrefix.as<uint32_t>();
            std::array<hipEvent_t, 4> none{};
            auto& ev         = itsDetailedTimings ? batchEvents() : none;

            kFill32<<<gridFor(words), BLOCK, 0, itsStream>>>(bitmap, words, 0);
            OC_LAUNCH_CHECK();
            if (itsDetailedTimings)
                OC_HIP_CHECK(hipEventRecord(ev[0], itsStream));
            if (blocks > 0)
            {
                kMark<T><<<xcdRunGrid(blocks, XCD_TILE_RUN), TILE, 0, itsStream>>>(
                        tileLo, blocks, k0, wordsPerNode, itsTileList.as<uint32_t>(), itsSegTile.as<uint32_t>(),
                        itsSegCount.as<uint32_t>(), itsSegSrc.as<uint32_t>(), itsSegInput.as<uint32_t>(),
                        itsNodeGrid.as<NodeGrid>(), ps, bitmap, itsCellOf.as<uint32_t>(), err);
                OC_LAUNCH_CHECK();
            }
            // per-cell state, reset for the occupied cells: by the scan itself on small levels
            // (arrays sized by the inputs, an upper bound of the cells), else by kClearCompact
            auto sizeCells = [&](uint64_t n) {
                if (!weighted)  // the weighted filter takes its winners from the payload keys
                    itsWin.ensure(sizeof(uint64_t) * n);
                itsCellKey.ensure(sizeof(uint32_t) * n);
                itsCellNode.ensure(sizeof(uint32_t) * n);
                if (packed)
                    itsColor.ensure(sizeof(uint64_t) * n);
                if (weighted)
                {
                    itsCellCnt.ensure(sizeof(uint32_t) * n);
                    itsCellStart.ensure(sizeof(uint32_t) * (n + 1));
                    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(n) + 1));
                }
            };
            if (smallLevel)
            {
                sizeCells(inputs);
                scanExclusive<uint32_t>(words, PopcountIn{bitmap},
                                        PrefixClearOut{prefix, weighted ? nullptr : itsWin.as<uint64_t>(),
                                                       packed ? itsColor.as<uint64_t>() : nullptr,
                                                       weighted ? itsCellCnt.as<uint32_t>() : nullptr},
                                        itsScanScratch.as<uint32_t>(), prefix + words, itsStream);
            }
            else
                scanExclusive<uint32_t>(words, PopcountIn{bitmap}, ExclusiveToArray{prefix},
                                        itsScanScratch.as<uint32_t>(), prefix + words, itsStream);
            OC_LAUNCH_CHECK();
            kBatchNodes<<<gridFor(Kb + 1), BLOCK, 0, itsStream>>>(
                    k0, Kb, wordsPerNode, itsNodeGrid.as<NodeGrid>(), prefix, static_cast<uint32_t>(cursor),
                    itsNodePoints.as<uint32_t>(), itsNodeStage.as<uint32_t>(), itsNodeFirst.as<uint32_t>());
            OC_LAUNCH_CHECK();
            const uint32_t cells = readDevice(prefix + words);
            if (cursor + cells > itsE)  // A14
                throw std::runtime_error("the subsampled points exceed the estimated output size (N + I > E); "
                                         "raise the output factor");
            if (cells == 0)
                continue;
            if (!smallLevel)
            {
                sizeCells(cells);
                kClearCompact<<<gridFor(cells), BLOCK, 0, itsStream>>>(
                        cells, weighted ? nullptr : itsWin.as<uint64_t>(), packed ? itsColor.as<uint64_t>() : nullptr,
                        weighted ? itsCellCnt.as<uint32_t>() : nullptr);
                OC_LAUNCH_CHECK();
            }
            uint64_t* color   = packed ? itsColor.as<uint64_t>() : nullptr;
            uint32_t* cellCnt = weighted ? itsCellCnt.as<uint32_t>() : nullptr;


            if (blocks > 0)
            {
                switch (itsColorMode)
                {
#define OC_ACC(M)                                                                                                      \
    launchAccumulate<T, M>(itsProc.useRandomSubsampling, blocks, itsStream, tileLo, k0, wordsPerNode,                 \
                           itsTileList.as<uint32_t>(), itsSegTile.as<uint32_t>(), itsSegCount.as<uint32_t>(),         \
                           itsSegSrc.as<uint32_t>(), itsSegInput.as<uint32_t>(), itsNodeGrid.as<NodeGrid>(), bitmap,  \
                           prefix, itsCellOf.as<uint32_t>(), ps, itsWin.as<uint64_t>(), color, cellCnt,               \
                           itsCellKey.as<uint32_t>(), itsCellNode.as<uint32_t>(),                                     \
                        weighted ? itsRankOf.as<uint32_t>() : nullptr)
                case COLOR_NONE: OC_ACC(COLOR_NONE); break;
                case COLOR_INTRA: OC_ACC(COLOR_INTRA); break;
                case COLOR_INTER: OC_ACC(COLOR_INTER); break;
                default: OC_ACC(COLOR_INTER_WEIGHTED); break;
#undef OC_ACC
                }
                OC_LAUNCH_CHECK();
            }
            if (itsDetailedTimings)
                OC_HIP_CHECK(hipEventRecord(ev[1], itsStream));

            if (weighted)
            {
                scanExclusive<uint32_t>(cells, ArrayIn{cellCnt}, ExclusiveToArray{itsCellStart.as<uint32_t>()},
                                        itsScanScratch.as<uint32_t>(), itsCellStart.as<uint32_t>() + cells,
                                        itsStream);
                OC_LAUNCH_CHECK();
                if (blocks > 0)
                {
                    if (itsProc.useRandomSubsampling)
                        kScatterPayload<T, true><<<xcdRunGrid(blocks, XCD_TILE_RUN), TILE, 0, itsStream>>>(
                                tileLo, blocks, itsTileList.as<uint32_t>(), itsSegTile.as<uint32_t>(),
                                itsSegCount.as<uint32_t>(), itsSegSrc.as<uint32_t>(), itsSegInput.as<uint32_t>(),
                                itsNodeGrid.as<NodeGrid>(), itsCellKey.as<uint32_t>(), itsRankOf.as<uint32_t>(),
                                itsCellStart.as<uint32_t>(), itsCellOf.as<uint32_t>(), ps,
                                itsPayload.as<WeightPayload>(), winKey);
                    else
                        kScatterPayload<T, false><<<xcdRunGrid(blocks, XCD_TILE_RUN), TILE, 0, itsStream>>>(
                                tileLo, blocks, itsTileList.as<uint32_t>(), itsSegTile.as<uint32_t>(),
                                itsSegCount.as<uint32_t>(), itsSegSrc.as<uint32_t>(), itsSegInput.as<uint32_t>(),
                                itsNodeGrid.as<NodeGrid>(), itsCellKey.as<uint32_t>(), itsRankOf.as<uint32_t>(),
                                itsCellStart.as<uint32_t>(), itsCellOf.as<uint32_t>(), ps,
                                itsPayload.as<WeightPayload>(), winKey);
                    OC_LAUNCH_CHECK();
                }
            }

            if (itsDetailedTimings)
                OC_HIP_CHECK(hipEventRecord(ev[2], itsStream));
            if (weighted)
            {
                const uint32_t emitBlocks = xcdRunGrid((cells + BLOCK - 1) / BLOCK, EMIT_CHUNK);
                kEmitWeighted<T><<<emitBlocks, BLOCK, 0, itsStream>>>(
                        cells, k0, wordsPerNode, static_cast<uint32_t>(cursor), itsNodeGrid.as<NodeGrid>(), bitmap,
                        prefix, itsCellKey.as<uint32_t>(), itsCellNode.as<uint32_t>(), itsCellStart.as<uint32_t>(),
                        itsPayload.as<WeightPayload>(), winKey, ps);
                OC_LAUNCH_CHECK();
            }
            else
            {
                switch (itsColorMode)
                {
#define OC_EMIT(M)                                                                                                     \
    kEmit<T, M><<<gridFor(cells), BLOCK, 0, itsStream>>>(cells, k0, wordsPerNode, static_cast<uint32_t>(cursor),      \
                                                         itsNodeGrid.as<NodeGrid>(), bitmap, prefix,                  \
                                                         itsCellKey.as<uint32_t>(), itsCellNode.as<uint32_t>(),       \
                                                         itsWin.as<uint64_t>(), itsColor.as<uint64_t>(), ps)
                case COLOR_NONE: OC_EMIT(COLOR_NONE); break;
                case COLOR_INTRA: OC_EMIT(COLOR_INTRA); break;
                default: OC_EMIT(COLOR_INTER); break;
#undef OC_EMIT
                }
                OC_LAUNCH_CHECK();
            }
            if (itsDetailedTimings)
                OC_HIP_CHECK(hipEventRecord(ev[3], itsStream));
            cursor += cells;
        }
        progress();
    }
    itsUsed = cursor;
}

// Trees up to this many nodes are placed by one workgroup in one launch.
static constexpr uint32_t kPlaceOneGroupNodes = 1u << 16;

void OctreeBuilder::placeInternalNodes()
{
    const uint32_t N      = itsCloud.pointAmount;
    const uint64_t staged = itsUsed - N;
    if (itsM <= kPlaceOneGroupNodes)
    {
        PlaceParams pp{};
        for (uint32_t l = 0; l <= itsDepth + 1 && l <= MAX_LEVELS; ++l)
            pp.levelStart[l] = itsLevelStart[l];
        pp.depth = itsDepth;
        pp.m     = itsM;
        pp.root  = itsM - 1;
        pp.N     = N;
        kPlaceLevels<<<1, PLACE_THREADS, 0, itsStream>>>(pp, itsNodeFlags.as<uint8_t>(), itsNodeChild.as<int32_t>(),
                                                          itsNodePoints.as<uint32_t>(), itsNodeSub.as<uint32_t>(),
                                                          itsNodeRegion.as<uint32_t>(), itsNodeData.as<uint64_t>());
        OC_LAUNCH_CHECK();
    }
    else
        placeLevelByLevel();
    if (staged > 0)
    {
        // (internal node, chunk) work items; the leaf chunk list is free again after the distribution
        uint64_t* sc = itsScalars.as<uint64_t>();
        itsLeafList.ensure(sizeof(uint64_t) * (itsM + staged / PLACE_CHUNK + 1));
        itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(itsM) + 1));
        scanExclusive<uint32_t>(itsM, PlacedChunksIn{itsNodeFlags.as<uint8_t>(), itsNodePoints.as<uint32_t>()},
                                LeafOut{itsLeafList.as<uint64_t>()}, itsScanScratch.as<uint32_t>(),
                                reinterpret_cast<uint32_t*>(sc + SC_TABLE_TOTAL), itsStream);
        OC_LAUNCH_CHECK();
        const uint32_t chunks = static_cast<uint32_t>(readDevice(sc + SC_TABLE_TOTAL));
        if (chunks == 0)
            return;
        const Gridding cubic{itsCloud.bbMin[0], itsCloud.bbMin[1], itsCloud.bbMin[2],
                             itsCloud.bbMax[0] - itsCloud.bbMin[0], itsProc.chunkingGrid};
        const double sx = 1.0 / itsCloud.scale[0], sy = 1.0 / itsCloud.scale[1], sz = 1.0 / itsCloud.scale[2];
        const uint32_t* rgb = itsColorMode == COLOR_INTER_WEIGHTED ? itsStRgb.as<uint32_t>() : nullptr;
        if (itsCloud.cloudType == CLOUD_FLOAT_UINT8_T)
            kPermute<float><<<chunks, BLOCK, 0, itsStream>>>(
                    itsLeafList.as<uint64_t>(), N, itsNodeStage.as<uint32_t>(), itsNodeData.as<uint64_t>(),
                    itsNodePoints.as<uint32_t>(), itsStLut.as<uint32_t>(), itsStOut.as<OutRecord>(),
                    itsStPts.as<float>(), rgb, cubic, sx, sy, sz, itsLut.as<uint32_t>(), itsOut.as<OutRecord>());
        else
            kPermute<double><<<chunks, BLOCK, 0, itsStream>>>(
                    itsLeafList.as<uint64_t>(), N, itsNodeStage.as<uint32_t>(), itsNodeData.as<uint64_t>(),
                    itsNodePoints.as<uint32_t>(), itsStLut.as<uint32_t>(), itsStOut.as<OutRecord>(),
                    itsStPts.as<double>(), rgb, cubic, sx, sy, sz, itsLut.as<uint32_t>(), itsOut.as<OutRecord>());
        OC_LAUNCH_CHECK();
    }
}

void OctreeBuilder::placeLevelByLevel()
{
    const uint32_t N = itsCloud.pointAmount;
    kFill32<<<gridFor(itsM), BLOCK, 0, itsStream>>>(itsNodeSub.as<uint32_t>(), itsM, 0);
    OC_LAUNCH_CHECK();
    for (uint32_t l = 1; l <= itsDepth; ++l)
    {
        uint32_t ls = itsLevelStart[l], nodes = itsLevelStart[l + 1] - ls;
        if (nodes == 0)
            continue;
        kSubtree<<<gridFor(nodes), BLOCK, 0, itsStream>>>(ls, nodes, itsNodeFlags.as<uint8_t>(),
                                                          itsNodeChild.as<int32_t>(), itsNodePoints.as<uint32_t>(),
                                                          itsNodeSub.as<uint32_t>());
        OC_LAUNCH_CHECK();
    }
    for (uint32_t l = itsDepth; l >= 1; --l)
    {
        uint32_t ls = itsLevelStart[l], nodes = itsLevelStart[l + 1] - ls;
        if (nodes == 0)
            continue;
        kRegion<<<gridFor(nodes), BLOCK, 0, itsStream>>>(ls, nodes, itsM - 1, N, itsNodeFlags.as<uint8_t>(),
                                                         itsNodeChild.as<int32_t>(), itsNodeSub.as<uint32_t>(),
                                                         itsNodeRegion.as<uint32_t>(), itsNodeData.as<uint64_t>());
        OC_LAUNCH_CHECK();
    }
}

// Copies go through the builder's own stream, so slots of a build service copy concurrently.
void OctreeBuilder::copyNodes(HostOctree& o) const
{
    if (!itsBuilt)
        throw std::runtime_error("the octree has not been generated");
    const uint64_t m = itsM;
    o.denseKey.resize(m);
    o.count.resize(m);
    o.pointCount.resize(m);
    o.dataIdx.resize(m);
    o.parent.resize(m);
    o.children.resize(8 * m);
    o.flags.resize(m);
    auto copy = [&](void* dst, const void* src, uint64_t bytes) {
        OC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, itsStream));
    };
    copy(o.denseKey.data(), itsNodeKey.get(), 4 * m);
    copy(o.count.data(), itsNodeCount.get(), 4 * m);
    copy(o.pointCount.data(), itsNodePoints.get(), 4 * m);
    copy(o.dataIdx.data(), itsNodeData.get(), 8 * m);
    copy(o.parent.data(), itsNodeParent.get(), 4 * m);
    copy(o.children.data(), itsNodeChild.get(), 32 * m);
    copy(o.flags.data(), itsNodeFlags.get(), m);
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
}

void OctreeBuilder::progress()
{
    if (itsProgress)
        itsProgress();
}

bool OctreeBuilder::snapshotDone()
{
    const hipError_t e = hipEventQuery(itsCopiedEvent);
    if (e == hipErrorNotReady)
        return false;
    OC_HIP_CHECK(e);
    return true;
}

void OctreeBuilder::drainCopies() noexcept
{
    (void)hipStreamSynchronize(itsCopyStream);
    (void)hipGetLastError();
}

std::unique_ptr<SlotBuilder> makeSlotBuilder(int device)
{
    auto b = std::make_unique<OctreeBuilder>(device);
    b->setDetailedTimings(false);
    return b;
}

uint64_t slotFootprint(const CloudInfo& cloud, const ProcessingInfo& proc, bool hostCloud)
{
    return OctreeBuilder::estimateFootprint(cloud, proc, hostCloud);
}

void OctreeBuilder::waitForSnapshotOnStream()
{
    if (itsSnapshotPending)
    {
        OC_HIP_CHECK(hipStreamWaitEvent(itsStream, itsCopiedEvent, 0));
        itsSnapshotPending = false;
    }
}

void OctreeBuilder::copySnapshot(uint8_t* dst) const
{
    auto* self = const_cast<OctreeBuilder*>(this);
    self->snapshotAsync(dst);
    self->waitSnapshot();
}

void OctreeBuilder::waitSnapshot()
{
    OC_HIP_CHECK(hipEventSynchronize(itsCopiedEvent));
}

void OctreeBuilder::snapshotAsync(uint8_t* dst)
{
    if (!itsBuilt)
        throw std::runtime_error("the octree has not been generated");
    const SnapshotLayout l(itsM, itsUsed);
    const uint64_t m = itsM;
    OC_HIP_CHECK(hipEventRecord(itsBuiltEvent, itsStream));
    OC_HIP_CHECK(hipStreamWaitEvent(itsCopyStream, itsBuiltEvent, 0));
    auto copy = [&](uint64_t offset, const void* src, uint64_t bytes) {
        if (bytes)
            OC_HIP_CHECK(hipMemcpyAsync(dst + offset, src, bytes, hipMemcpyDeviceToHost, itsCopyStream));
    };
    copy(l.dataIdx, itsNodeData.get(), 8 * m);
    copy(l.denseKey, itsNodeKey.get(), 4 * m);
    copy(l.count, itsNodeCount.get(), 4 * m);
    copy(l.pointCount, itsNodePoints.get(), 4 * m);
    copy(l.parent, itsNodeParent.get(), 4 * m);
    copy(l.children, itsNodeChild.get(), 32 * m);
    copy(l.flags, itsNodeFlags.get(), m);
    copy(l.lut, itsLut.get(), 4 * itsUsed);
    copy(l.out, itsOut.get(), sizeof(OutRecord) * itsUsed);
    OC_HIP_CHECK(hipEventRecord(itsCopiedEvent, itsCopyStream));
    itsSnapshotPending = true;
}

// Device to host, synchronous; large copies into pageable memory are staged (host_staging.h).
void OctreeBuilder::download(void* dst, const void* src, uint64_t bytes) const
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    downloadStaged(dst, src, bytes, itsStream);
}

// Host to device on the builder's stream; large copies from pageable memory are staged and have
// left the host memory when this returns (host_staging.h).
void OctreeBuilder::upload(void* dst, const void* src, uint64_t bytes)
{
    const HostSegment seg{dst, src, bytes};
    uploadSegments(&seg, 1, itsStream);
}

void OctreeBuilder::copyLut(uint32_t* dst, uint64_t count) const
{
    if (count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    download(dst, itsLut.get(), 4 * count);
}

void OctreeBuilder::copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const
{
    if (first + count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    download(dst, itsOut.as<OutRecord>() + first, sizeof(OutRecord) * count);
}

}  // namespace oc
