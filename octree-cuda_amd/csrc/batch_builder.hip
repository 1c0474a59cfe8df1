// Batched build of many clouds on one MI355X (gfx950). See batch_builder.h for the overview and
// kernels_batch.h for the kernels.
//
// Device memory of a batch (grow-only, reused by the next batch):
//   arena       per cloud, carved in turn: its scalar slots (all clouds' slots first, so one copy
//               reads them back), the uploaded cloud, bucket tables, sorted records, the dense
//               pyramid, node keys and counts, the output region, staging and the per-level work
//               arrays (bounded by the output capacity E and by N / threshold internal nodes)
//   node arena  per cloud, after the merge: the M sparse nodes' arrays and the chunk lists
//   table       one BatchCloud entry per cloud
//   prefix      work prefixes, PREFIX_SLOTS arrays of B + 1 entries
#include "batch_builder.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>

#include "host_staging.h"
#include "kernels_batch.h"

namespace oc {

#define OC_HIP_CHECK(expr)                                                                                             \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw std::runtime_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));                      \
    } while (0)

#define OC_LAUNCH_CHECK() OC_HIP_CHECK(hipGetLastError())

namespace {

// work prefix slots
enum PrefixSlot : uint32_t
{
    P_HIST,
    P_SCATTER,
    P_POINTS,
    P_DIST,
    P_SPLIT,
    P_NODES,
    P_LEVEL0,
    P_LEAFCHUNK,
    P_K,
    P_SEG,
    P_TILES,
    P_WORDS,
    P_K1,
    P_CELLS,
    P_LIVE_TILES,
    P_PLACE,
    P_LEVEL,  // + level: nodes of that level
    PREFIX_SLOTS = P_LEVEL + MAX_LEVELS
};

uint32_t octreeDepthOf(uint32_t grid)  // tools.cu:63-88
{
    for (uint32_t d = 1; d <= 9; ++d)
        if (grid == (1u << d))
            return d;
    return 0;
}

BatchLevels levelsOf(uint32_t G)
{
    BatchLevels L{};
    L.depth = octreeDepthOf(G);
    for (uint32_t l = 0; l <= L.depth; ++l)
    {
        const uint32_t g = G >> l;
        L.grid[l]        = g;
        L.offset[l]      = L.dense;
        L.nodes[l]       = g * g * g;
        L.dense += g * g * g;
    }
    L.lc      = std::min<uint32_t>(MAX_COARSE_LEVEL, L.depth > 3 ? L.depth - 3 : 0);
    const uint32_t bg = G >> L.lc;
    L.buckets = bg * bg * bg;
    return L;
}

ColorMode colorModeOf(const ProcessingInfo& p)
{
    return p.useIntraCellAvg ? COLOR_INTRA
                             : (p.useInterCellAvg ? (p.useWeightingFunction ? COLOR_INTER_WEIGHTED : COLOR_INTER)
                                                  : COLOR_NONE);
}

// Sizes of one cloud's part of the arena.
struct CloudSizes
{
    uint64_t n, E, staged, mBound, kBound, tileBound, wordsPerNode;
    uint32_t keyShift;
};

CloudSizes sizesOf(const BatchRequest& r, const BatchLevels& L)
{
    CloudSizes s{};
    s.n      = r.cloud.pointAmount;
    s.E      = static_cast<uint64_t>(static_cast<float>(r.cloud.pointAmount) * r.proc.outputFactor);  // point_cloud.cuh:24
    s.staged = std::max<uint64_t>(1, s.E > s.n ? s.E - s.n : 0);
    s.mBound = std::min<uint64_t>(L.dense, static_cast<uint64_t>(L.depth + 1) * s.n);
    // internal nodes of one level hold disjoint point sets of at least `threshold` points each
    const uint64_t T = r.proc.mergingThreshold;
    s.kBound       = std::max<uint64_t>(1, std::min<uint64_t>(L.depth >= 1 ? L.nodes[1] : 1, T ? s.n / T : s.n));
    s.tileBound    = s.E / TILE + 8 * s.kBound + 1;
    const uint32_t g = r.proc.subsamplingGrid;
    s.wordsPerNode = (rowsPerNode(g) * g + 31) / 32;
    const uint32_t h = g > 1 ? g / 2 : 1;
    const uint64_t keys = rowsPerNode(h) * ((h + 31) / 32);
    s.keyShift          = 0;
    while (((keys - 1) >> s.keyShift) >= LEAF_KEYS)
        ++s.keyShift;
    return s;
}

// Sequential carving of a device region: pass 1 (base null) sums the sizes, pass 2 sets pointers.
struct Carver
{
    uint8_t* base = nullptr;
    uint64_t off  = 0;
    template <typename P>
    void take(P*& p, uint64_t count)
    {
        off = (off + 255) & ~uint64_t(255);
        p   = base ? reinterpret_cast<P*>(base + off) : nullptr;
        off += count * sizeof(P);
    }
};

template <typename T>
void carveCloud(Carver& cv, BatchCloud<T>& cc, const BatchRequest& r, const CloudSizes& s, const BatchLevels& L,
                ColorMode mode, uint8_t** cloudCopy)
{
    const uint64_t n = s.n;
    if (!r.deviceCloud)
        cv.take(*cloudCopy, n * r.cloud.dataStride);
    cv.take(cc.bucketCount, L.buckets + 1);
    cv.take(cc.bucketStart, L.buckets + 1);
    cv.take(cc.bucketCursor, L.buckets + 1);
    cv.take(cc.bucketed, n);
    cv.take(cc.sorted, n);
    cv.take(cc.fineCode, n);
    cv.take(cc.perm, n);
    if (n > FINE_SPLIT)
    {
        const uint64_t bound = 2 * n / FINE_SPLIT + 1;
        cv.take(cc.splitChunks, bound);
        cv.take(cc.splitHist, bound * MAX_FINE_CELLS);
    }
    else
    {
        cc.splitChunks = nullptr;
        cc.splitHist   = nullptr;
    }
    cv.take(cc.count, L.dense);
    cv.take(cc.denseToSparse, L.dense);
    cv.take(cc.nodeKey, s.mBound);
    cv.take(cc.nodeCount, s.mBound);
    cv.take(cc.scanSums, BATCH_SCAN_X);
    cv.take(cc.lut, s.E);
    cv.take(cc.out, s.E);
    cv.take(cc.stLut, s.staged);
    cv.take(cc.stOut, s.staged);
    cv.take(cc.stPts, 3 * s.staged);
    if (mode == COLOR_INTER_WEIGHTED)
        cv.take(cc.stRgb, s.staged);
    else
        cc.stRgb = nullptr;
    cv.take(cc.internalList, s.kBound);
    cv.take(cc.segCount, 8 * s.kBound);
    cv.take(cc.segSrc, 8 * s.kBound);
    cv.take(cc.segTile, 8 * s.kBound);
    cv.take(cc.segInput, 8 * s.kBound);
    cv.take(cc.nodeGrid, s.kBound);
    cv.take(cc.tileList, s.tileBound);
    cv.take(cc.bitmap, s.kBound * s.wordsPerNode);
    cv.take(cc.prefix, s.kBound * s.wordsPerNode + 1);
    cv.take(cc.nodeFirst, s.kBound + 1);
    cv.take(cc.cellOf, s.E);
    cv.take(cc.cellKey, s.E);
    cv.take(cc.cellNode, s.E);
    cc.win = cc.color = nullptr;
    cc.cellCnt = cc.cellStart = cc.rankOf = nullptr;
    cc.payload = nullptr;
    cc.winKey  = nullptr;
    if (mode == COLOR_INTER_WEIGHTED)
    {
        cv.take(cc.cellCnt, s.E);
        cv.take(cc.cellStart, s.E + 1);
        cv.take(cc.rankOf, s.E);
        cv.take(cc.payload, s.E);
        cv.take(cc.winKey, s.E);
    }
    else
    {
        cv.take(cc.win, s.E);
        if (mode != COLOR_NONE)
            cv.take(cc.color, s.E);
    }
}

template <typename T>
void carveNodes(Carver& cv, BatchCloud<T>& cc, uint64_t m, const CloudSizes& s)
{
    cv.take(cc.nodePoints, m);
    cv.take(cc.nodeData, m);
    cv.take(cc.nodeParent, m);
    cv.take(cc.nodeChild, 8 * m);
    cv.take(cc.nodeFlags, m);
    cv.take(cc.nodeSorted, m);
    cv.take(cc.nodeAux, m + 1);
    cv.take(cc.nodeStage, m);
    cv.take(cc.nodeSub, m);
    cv.take(cc.nodeRegion, m);
    cv.take(cc.keyOut, m);
    cv.take(cc.countOut, m);
    cv.take(cc.chunkList, std::max<uint64_t>(m + s.n / LEAF_CHUNK + 1, m + s.staged / PLACE_CHUNK + 1));
}

}  // namespace

void BatchBuilder::runSegment(bool graphs, const std::vector<uint64_t>& key, const std::function<void()>& issue)
{
    if (!graphs)
    {
        issue();
        return;
    }
    auto it = itsGraphs.find(key);
    if (it == itsGraphs.end())
    {
        // the launches of a segment only depend on the key; every size the device decides and every
        // address of the batch is read from the cloud table and the prefix arrays at run time
        OC_HIP_CHECK(hipStreamBeginCapture(itsStream, hipStreamCaptureModeThreadLocal));
        try
        {
            issue();
        }
        catch (...)
        {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(itsStream, &broken);
            if (broken)
                (void)hipGraphDestroy(broken);
            throw;
        }
        hipGraph_t graph = nullptr;
        OC_HIP_CHECK(hipStreamEndCapture(itsStream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e  = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        OC_HIP_CHECK(e);
        it = itsGraphs.emplace(key, exec).first;
        itsGraphsCaptured++;
    }
    OC_HIP_CHECK(hipGraphLaunch(it->second, itsStream));
}

void BatchBuilder::clearGraphs()
{
    for (auto& kv : itsGraphs)
        (void)hipGraphExecDestroy(kv.second);
    itsGraphs.clear();
}

std::unique_ptr<BatchSlotBuilder> makeBatchSlotBuilder(int device, std::shared_ptr<PinnedPool> pinned)
{
    return std::make_unique<BatchBuilder>(device, std::move(pinned));
}

uint64_t batchFootprint(const std::vector<BatchRequest>& requests)
{
    return BatchBuilder::estimateFootprint(requests);
}

BatchBuilder::BatchBuilder(int device, std::shared_ptr<PinnedPool> pinned) : itsDevice(device), itsPinned(pinned)
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    OC_HIP_CHECK(hipStreamCreateWithFlags(&itsStream, hipStreamNonBlocking));
    for (auto& ev : itsEvents)
        OC_HIP_CHECK(hipEventCreate(&ev));
    if (!itsPinned)
        itsPinned = std::make_shared<PinnedPool>(size_t(16) << 30);
}

BatchBuilder::~BatchBuilder()
{
    clearGraphs();
    if (itsStream)
    {
        (void)hipStreamSynchronize(itsStream);
        (void)hipStreamDestroy(itsStream);
    }
    for (auto& ev : itsEvents)
        if (ev)
            (void)hipEventDestroy(ev);
}

void BatchBuilder::releaseArena()
{
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
    clearGraphs();  // they hold the addresses of the table and the prefix arrays
    itsArena.release();
    itsNodeArena.release();
    itsPacked.release();
    itsPieces.release();
    itsTable.release();
    itsPrefix.release();
}

uint64_t BatchBuilder::estimateFootprint(const std::vector<BatchRequest>& requests)
{
    uint64_t bytes = 0;
    for (const BatchRequest& r : requests)
    {
        const BatchLevels L = levelsOf(r.proc.chunkingGrid);
        const CloudSizes s  = sizesOf(r, L);
        Carver cv;
        const ColorMode mode = colorModeOf(r.proc);
        uint8_t* copy        = nullptr;
        if (r.cloud.cloudType == CLOUD_FLOAT_UINT8_T)
        {
            BatchCloud<float> cc{};
            carveCloud(cv, cc, r, s, L, mode, &copy);
            // node arrays: at most the dense bound, in practice about 1.2 N; admission uses 2 N
            carveNodes(cv, cc, std::min<uint64_t>(s.mBound, 2 * s.n + 16), s);
        }
        else
        {
            BatchCloud<double> cc{};
            carveCloud(cv, cc, r, s, L, mode, &copy);
            carveNodes(cv, cc, std::min<uint64_t>(s.mBound, 2 * s.n + 16), s);
        }
        bytes += cv.off + BS_SLOTS * 8 + 4096;
    }
    return bytes;
}

std::vector<BatchOutcome> BatchBuilder::build(const std::vector<BatchRequest>& requests)
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<BatchOutcome> outcomes(requests.size());
    itsTimings.assign(7, 0.f);
    itsGroups = 0;
    // groups: record type, chunking grid, colour filter, method
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, std::vector<uint32_t>> groups;
    for (uint32_t i = 0; i < requests.size(); ++i)
    {
        const BatchRequest& r = requests[i];
        groups[{r.cloud.cloudType, r.proc.chunkingGrid, static_cast<uint32_t>(colorModeOf(r.proc)),
                r.proc.useRandomSubsampling ? 1u : 0u}]
                .push_back(i);
    }
    for (auto& kv : groups)
    {
        const std::vector<uint32_t>& all = kv.second;
        for (size_t first = 0; first < all.size(); first += MAX_BATCH)
        {
            std::vector<uint32_t> members(all.begin() + first,
                                          all.begin() + std::min<size_t>(all.size(), first + MAX_BATCH));
            if (std::get<0>(kv.first) == CLOUD_FLOAT_UINT8_T)
                buildGroup<float>(members, requests, outcomes);
            else
                buildGroup<double>(members, requests, outcomes);
            itsGroups++;
        }
    }
    itsTimings[5] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return outcomes;
}

template <typename T>
void BatchBuilder::buildGroup(const std::vector<uint32_t>& members, const std::vector<BatchRequest>& requests,
                              std::vector<BatchOutcome>& outcomes)
{
    const BatchRequest& first = requests[members[0]];
    const BatchLevels L       = levelsOf(first.proc.chunkingGrid);
    const ColorMode mode      = colorModeOf(first.proc);
    const bool random         = first.proc.useRandomSubsampling;
    const uint32_t B0         = static_cast<uint32_t>(members.size());
    hipStream_t s             = itsStream;

    // ---- sizes and the arena ------------------------------------------------------------------
    std::vector<CloudSizes> sizes(B0);
    std::vector<BatchCloud<T>> table(B0);
    std::vector<uint8_t*> cloudCopy(B0, nullptr);
    for (uint32_t i = 0; i < B0; ++i)
        sizes[i] = sizesOf(requests[members[i]], L);
    auto carveAll = [&](uint8_t* base) {
        Carver cv;
        cv.base = base;
        uint64_t* scalars;
        cv.take(scalars, static_cast<uint64_t>(B0) * BS_SLOTS);
        for (uint32_t i = 0; i < B0; ++i)
        {
            table[i].sc = scalars ? scalars + static_cast<uint64_t>(i) * BS_SLOTS : nullptr;
            carveCloud(cv, table[i], requests[members[i]], sizes[i], L, mode, &cloudCopy[i]);
        }
        return cv.off;
    };
    itsArena.ensure(carveAll(nullptr));
    carveAll(itsArena.as<uint8_t>());
    uint64_t* scalarsBase = table[0].sc;

    std::vector<HostSegment> uploads;
    for (uint32_t i = 0; i < B0; ++i)
    {
        const BatchRequest& r   = requests[members[i]];
        const CloudSizes& sz    = sizes[i];
        BatchCloud<T>& cc       = table[i];
        const double cubicSize  = r.cloud.bbMax[0] - r.cloud.bbMin[0];  // metadata.cuh:66-69
        cc.n                    = r.cloud.pointAmount;
        cc.stride               = r.cloud.dataStride;
        cc.grid                 = Gridding{r.cloud.bbMin[0], r.cloud.bbMin[1], r.cloud.bbMin[2], cubicSize,
                                           r.proc.chunkingGrid};
        cc.sx                   = 1.0 / r.cloud.scale[0];  // octree_chunking.cu:108-112
        cc.sy                   = 1.0 / r.cloud.scale[1];
        cc.sz                   = 1.0 / r.cloud.scale[2];
        for (uint32_t l = 0; l < MAX_LEVELS; ++l)
        {
            const double side = l <= L.depth ? cubicSize / std::pow(2.0, static_cast<double>(L.depth - l)) : 0.0;
            cc.side[l]        = side;  // rp_subsampling.cu:17-18
            cc.sideHalf[l]    = side / 2.0;
        }
        cc.E            = sz.E;
        cc.seed         = r.proc.seed;
        cc.wordsPerNode = sz.wordsPerNode;
        cc.threshold    = r.proc.mergingThreshold;
        cc.g            = r.proc.subsamplingGrid;
        cc.keyShift     = sz.keyShift;
        cc.kBound       = static_cast<uint32_t>(sz.kBound);
        cc.tileBound    = static_cast<uint32_t>(sz.tileBound);
        cc.m            = 0;
        if (r.deviceCloud)
            cc.cloud = r.deviceCloud;
        else
        {
            uploads.push_back({cloudCopy[i], r.hostCloud, static_cast<uint64_t>(sz.n) * r.cloud.dataStride});
            cc.cloud = cloudCopy[i];
        }
    }
    // the group's host clouds, staged together
    uploadSegments(uploads.data(), uploads.size(), s);

    // ---- tables ------------------------------------------------------------------------------
    itsTable.ensure(sizeof(BatchCloud<T>) * MAX_BATCH);
    itsPrefix.ensure(sizeof(uint32_t) * PREFIX_SLOTS * (MAX_BATCH + 1));
    uint32_t* prefix = itsPrefix.as<uint32_t>();
    auto slot        = [&](uint32_t k) { return prefix + static_cast<uint64_t>(k) * (MAX_BATCH + 1); };
    itsHostPrefix.assign(static_cast<size_t>(PREFIX_SLOTS) * (MAX_BATCH + 1), 0);
    auto hostPrefix = [&](uint32_t k, uint32_t B, auto count) {
        uint32_t* p = itsHostPrefix.data() + static_cast<uint64_t>(k) * (MAX_BATCH + 1);
        uint64_t acc = 0;
        for (uint32_t i = 0; i < B; ++i)
        {
            p[i] = static_cast<uint32_t>(acc);
            acc += count(i);
        }
        if (acc >= (1ull << 32))
            throw std::runtime_error("a batch has more than 2^32 work items in one stage");
        p[B] = static_cast<uint32_t>(acc);
    };
    auto uploadTables = [&](uint32_t B) {
        itsHostTable.resize(sizeof(BatchCloud<T>) * B);
        std::memcpy(itsHostTable.data(), table.data(), sizeof(BatchCloud<T>) * B);
        OC_HIP_CHECK(hipMemcpyAsync(itsTable.get(), itsHostTable.data(), sizeof(BatchCloud<T>) * B,
                                    hipMemcpyHostToDevice, s));
        OC_HIP_CHECK(hipMemcpyAsync(prefix, itsHostPrefix.data(), sizeof(uint32_t) * itsHostPrefix.size(),
                                    hipMemcpyHostToDevice, s));
    };
    hostPrefix(P_HIST, B0, [&](uint32_t i) { return (sizes[i].n + HIST_CHUNK - 1) / HIST_CHUNK; });
    hostPrefix(P_SCATTER, B0, [&](uint32_t i) { return (sizes[i].n + STAGE_CHUNK - 1) / STAGE_CHUNK; });
    uploadTables(B0);
    const BatchCloud<T>* ctx = itsTable.as<BatchCloud<T>>();
    const BatchLaunch& LG    = itsLaunch;

    bool split = false;
    for (const CloudSizes& sz : sizes)
        split |= sz.n > FINE_SPLIT;
    FineSortParams fp{};
    fp.lc         = L.lc;
    fp.bucketGrid = first.proc.chunkingGrid >> L.lc;
    for (uint32_t l = 0; l <= L.depth; ++l)
    {
        fp.levelOffset[l] = L.offset[l];
        fp.levelGrid[l]   = L.grid[l];
    }
    LevelOffsets lo{};
    for (uint32_t l = 0; l <= L.depth; ++l)
        lo.offset[l] = L.offset[l];
    lo.levels = L.depth + 1;
    const bool graphs = itsUseGraphs;
    auto mark         = [&](int k) {
        if (!graphs)
            OC_HIP_CHECK(hipEventRecord(itsEvents[k], s));
    };

    // ---- segment 1: count, pyramid, canonical ids, level starts ---------------------------------
    auto segment1 = [&]() {
    kBatchInit<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B0, L.buckets);
    OC_LAUNCH_CHECK();
    const uint32_t fineBits = 3 * L.lc;
    kBucketHistB<T><<<LG.sortGrid, HIST_THREADS, 0, s>>>(ctx, B0, slot(P_HIST), fineBits, L.buckets);
    OC_LAUNCH_CHECK();
    kBucketScanB<T><<<std::min<uint32_t>(B0, 1024), SCAN_THREADS, 0, s>>>(ctx, B0, L.buckets);
    OC_LAUNCH_CHECK();
    kBucketScatterB<T><<<LG.sortGrid, SCATTER_THREADS, 0, s>>>(ctx, B0, slot(P_SCATTER), fineBits);
    OC_LAUNCH_CHECK();
    kFineSortB<T><<<LG.sortGrid, SORT_THREADS, 0, s>>>(ctx, B0, fp, L.buckets);
    OC_LAUNCH_CHECK();
    if (split)
    {
        kSplitChunksB<T><<<std::min<uint32_t>(B0, 1024), SCAN_THREADS, 0, s>>>(ctx, B0, L.buckets);
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(ScalarCount<T>{ctx, BS_SPLIT, false}, B0, slot(P_SPLIT));
        kFineChunkHistB<T><<<LG.sortGrid, SORT_THREADS, 0, s>>>(ctx, B0, slot(P_SPLIT), L.lc);
        kFineSplitScanB<T><<<LG.sortGrid, SORT_THREADS, 0, s>>>(ctx, B0, slot(P_SPLIT), fp);
        kFineChunkPermB<T><<<LG.sortGrid, SORT_THREADS, 0, s>>>(ctx, B0, slot(P_SPLIT), L.lc);
        OC_LAUNCH_CHECK();
    }
    mark(1);
    for (uint32_t l = L.lc + 1; l <= L.depth; ++l)
    {
        kPropagateB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B0, L.offset[l], L.grid[l], L.offset[l - 1],
                                                        L.grid[l - 1]);
        OC_LAUNCH_CHECK();
    }
    scanBatch<uint32_t>(OccupancyJob<T>{ctx, L.dense}, B0, L.dense, s);
    OC_LAUNCH_CHECK();
    kLevelStartsB<T><<<B0, 64, 0, s>>>(ctx, lo);
    OC_LAUNCH_CHECK();
    };
    OC_HIP_CHECK(hipEventRecord(itsEvents[0], s));
    runSegment(graphs, {1, sizeof(T), first.proc.chunkingGrid, B0, split ? 1u : 0u, LG.threadGrid, LG.tileGrid,
                        LG.sortGrid},
               segment1);
    if (graphs)
        OC_HIP_CHECK(hipEventRecord(itsEvents[1], s));
    // the one host wait before the nodes
    itsHostScalars.resize(static_cast<size_t>(B0) * BS_SLOTS);
    OC_HIP_CHECK(hipMemcpyAsync(itsHostScalars.data(), scalarsBase, sizeof(uint64_t) * itsHostScalars.size(),
                                hipMemcpyDeviceToHost, s));
    OC_HIP_CHECK(hipStreamSynchronize(s));

    // clouds that failed here leave the batch: the rest goes on without them
    std::vector<uint32_t> live;           // index into members
    std::vector<std::array<uint32_t, MAX_LEVELS + 1>> levelStart(B0);
    for (uint32_t i = 0; i < B0; ++i)
    {
        const uint64_t* sc = itsHostScalars.data() + static_cast<uint64_t>(i) * BS_SLOTS;
        const uint32_t m   = static_cast<uint32_t>(sc[BS_M]);
        std::memcpy(levelStart[i].data(), sc + BS_LEVEL_START, sizeof(uint32_t) * (L.depth + 1));
        levelStart[i][L.depth + 1] = m;
        if (sc[BS_ERROR] & ERR_OUTSIDE)
            outcomes[members[i]].error = "a point lies outside the cubic bounding box";
        else if (m == 0)
            outcomes[members[i]].error = "the cloud has no points inside the bounding box";
        else
        {
            table[i].m = m;
            live.push_back(i);
        }
    }
    if (live.empty())
        return;
    const uint32_t B = static_cast<uint32_t>(live.size());
    std::vector<BatchCloud<T>> liveTable(B);
    std::vector<CloudSizes> liveSizes(B);
    for (uint32_t j = 0; j < B; ++j)
    {
        liveTable[j] = table[live[j]];
        liveSizes[j] = sizes[live[j]];
    }
    table.swap(liveTable);
    auto carveNodeArena = [&](uint8_t* base) {
        Carver cv;
        cv.base = base;
        for (uint32_t j = 0; j < B; ++j)
            carveNodes(cv, table[j], table[j].m, liveSizes[j]);
        return cv.off;
    };
    itsNodeArena.ensure(carveNodeArena(nullptr));
    carveNodeArena(itsNodeArena.as<uint8_t>());
    auto lsOf = [&](uint32_t j) -> const std::array<uint32_t, MAX_LEVELS + 1>& { return levelStart[live[j]]; };
    uint64_t maxM = 0;
    BatchBounds bounds{};
    for (uint32_t j = 0; j < B; ++j)
    {
        maxM = std::max<uint64_t>(maxM, table[j].m);
        for (uint32_t l = 1; l <= L.depth; ++l)
            bounds.levelNodes[l] = std::max<uint64_t>(bounds.levelNodes[l], lsOf(j)[l + 1] - lsOf(j)[l]);
        bounds.segments = std::max<uint64_t>(bounds.segments, 8 * liveSizes[j].kBound);
        bounds.words    = std::max<uint64_t>(bounds.words, liveSizes[j].kBound * liveSizes[j].wordsPerNode);
        bounds.cells    = std::max<uint64_t>(bounds.cells, liveSizes[j].E);
    }
    hostPrefix(P_POINTS, B, [&](uint32_t j) { return liveSizes[j].n; });
    hostPrefix(P_DIST, B, [&](uint32_t j) { return (liveSizes[j].n + DIST_CHUNK - 1) / DIST_CHUNK; });
    hostPrefix(P_NODES, B, [&](uint32_t j) { return table[j].m; });
    hostPrefix(P_LEVEL0, B, [&](uint32_t j) { return lsOf(j)[1]; });
    for (uint32_t l = 1; l <= L.depth; ++l)
        hostPrefix(P_LEVEL + l, B, [&](uint32_t j) { return lsOf(j)[l + 1] - lsOf(j)[l]; });
    uploadTables(B);

    // ---- segment 2: nodes, distribution, subsampling, placement ---------------------------------
    auto segment2 = [&]() {
    kNodeLinksB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_NODES), L);
    OC_LAUNCH_CHECK();
    scanBatch<uint32_t>(LeafBlockJob<T>{ctx}, B, maxM, s);
    kLeafDataIdxB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_NODES));
    kCellLeafB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_LEVEL0));
    OC_LAUNCH_CHECK();
    mark(2);

    // ---- distribution: leaf and cell order of the sorted clouds, then the leaf blocks ----------
    kSortedStartB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_NODES), L);
    scanBatch<uint32_t>(LeafChunkJob<T>{ctx}, B, maxM, s);
    kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(ScalarCount<T>{ctx, BS_TABLE_TOTAL, false}, B, slot(P_LEAFCHUNK));
    kLeafCellOrderB<T><<<LG.sortGrid, LEAF_THREADS, 0, s>>>(ctx, B, slot(P_LEAFCHUNK), L);
    kGatherSortedB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_POINTS));
    kDistributeSortedB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_DIST));
    OC_LAUNCH_CHECK();
    mark(3);

    // ---- subsampling, level by level for all clouds --------------------------------------------
    switch (mode)
    {
#define OC_SUB(M)                                                                                                      \
    if (random)                                                                                                        \
        subsampleGroup<T, M, true>(B, L, bounds);                                                                      \
    else                                                                                                               \
        subsampleGroup<T, M, false>(B, L, bounds);                                                                     \
    break;
    case COLOR_NONE: OC_SUB(COLOR_NONE)
    case COLOR_INTRA: OC_SUB(COLOR_INTRA)
    case COLOR_INTER: OC_SUB(COLOR_INTER)
    default: OC_SUB(COLOR_INTER_WEIGHTED)
#undef OC_SUB
    }

    // ---- placement: DFS post-order of the internal nodes (A6) ------------------------------------
    kZeroSubtreeB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_NODES));
    for (uint32_t l = 1; l <= L.depth; ++l)
        kSubtreeB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_LEVEL + l), l);
    for (uint32_t l = L.depth; l >= 1; --l)
        kRegionB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_LEVEL + l), l);
    scanBatch<uint32_t>(PlacedChunkJob<T>{ctx}, B, maxM, s);
    kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(ScalarCount<T>{ctx, BS_TABLE_TOTAL, true}, B, slot(P_PLACE));
    kPermuteB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_PLACE));
    OC_LAUNCH_CHECK();
    };
    if (graphs)
        OC_HIP_CHECK(hipEventRecord(itsEvents[2], s));
    runSegment(graphs,
               {2, sizeof(T), first.proc.chunkingGrid, B, static_cast<uint64_t>(mode), random ? 1u : 0u,
                LG.threadGrid, LG.tileGrid, LG.sortGrid, scanWidth(maxM), scanWidth(bounds.segments),
                scanWidth(bounds.words), scanWidth(bounds.cells), scanWidth(bounds.levelNodes[1]),
                scanWidth(bounds.levelNodes[2]), scanWidth(bounds.levelNodes[3]), scanWidth(bounds.levelNodes[4]),
                scanWidth(bounds.levelNodes[5]), scanWidth(bounds.levelNodes[6]), scanWidth(bounds.levelNodes[7]),
                scanWidth(bounds.levelNodes[8]), scanWidth(bounds.levelNodes[9])},
               segment2);
    OC_HIP_CHECK(hipEventRecord(itsEvents[4], s));

    // ---- results -------------------------------------------------------------------------------
    OC_HIP_CHECK(hipMemcpyAsync(itsHostScalars.data(), scalarsBase, sizeof(uint64_t) * itsHostScalars.size(),
                                hipMemcpyDeviceToHost, s));
    OC_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    if (graphs)
    {
        // segment 1 (count and pyramid) and segment 2 (nodes to placement)
        OC_HIP_CHECK(hipEventElapsedTime(&ms, itsEvents[0], itsEvents[1]));
        itsTimings[0] += ms;
        OC_HIP_CHECK(hipEventElapsedTime(&ms, itsEvents[2], itsEvents[4]));
        itsTimings[3] += ms;
    }
    else
    {
        const int stageEv[5] = {0, 1, 2, 3, 4};
        for (int k = 0; k < 4; ++k)
        {
            OC_HIP_CHECK(hipEventElapsedTime(&ms, itsEvents[stageEv[k]], itsEvents[stageEv[k + 1]]));
            itsTimings[k] += ms;
        }
    }
    OC_HIP_CHECK(hipEventElapsedTime(&ms, itsEvents[0], itsEvents[4]));
    itsTimings[4] += ms;

    const auto tc = std::chrono::steady_clock::now();
    std::vector<uint64_t> used(B, 0), offsets(B, 0);
    uint64_t total = 0;
    for (uint32_t j = 0; j < B; ++j)
    {
        const uint64_t* sc = itsHostScalars.data() + static_cast<uint64_t>(live[j]) * BS_SLOTS;
        BatchOutcome& o    = outcomes[members[live[j]]];
        if (sc[BS_ERROR] & ERR_OVERFLOW)
            o.error = "the subsampled points exceed the estimated output size (N + I > E); raise the output factor";
        else if (sc[BS_ERROR] & ERR_OUTSIDE)
            o.error = "a point lies outside the cubic bounding box";
        else if (sc[BS_ERROR] & ERR_CAPACITY)
            o.error = "a level of the batched build needs more work entries than its arrays hold";
        if (!o.error.empty())
            continue;
        used[j]    = sc[BS_CURSOR];
        offsets[j] = total;
        total += (SnapshotLayout(table[j].m, used[j]).bytes + 256) & ~uint64_t(255);
    }
    if (total == 0)
        return;
    // the snapshots are packed on the device in their host layout and come home in one copy
    std::vector<PackPiece> pieces;
    for (uint32_t j = 0; j < B; ++j)
    {
        if (!outcomes[members[live[j]]].error.empty())
            continue;
        const BatchCloud<T>& cc = table[j];
        const uint64_t m        = cc.m;
        const SnapshotLayout l(cc.m, used[j]);
        auto piece = [&](uint64_t offset, const void* src, uint64_t bytes) {
            for (uint64_t at = 0; at < bytes; at += PACK_CHUNK)
                pieces.push_back(PackPiece{static_cast<const uint8_t*>(src) + at, offsets[j] + offset + at,
                                           std::min<uint64_t>(PACK_CHUNK, bytes - at)});
        };
        piece(l.dataIdx, cc.nodeData, 8 * m);
        piece(l.denseKey, cc.keyOut, 4 * m);
        piece(l.count, cc.countOut, 4 * m);
        piece(l.pointCount, cc.nodePoints, 4 * m);
        piece(l.parent, cc.nodeParent, 4 * m);
        piece(l.children, cc.nodeChild, 32 * m);
        piece(l.flags, cc.nodeFlags, m);
        piece(l.lut, cc.lut, 4 * used[j]);
        piece(l.out, cc.out, sizeof(OutRecord) * used[j]);
    }
    itsPacked.ensure(total);
    itsPieces.ensure(sizeof(PackPiece) * pieces.size());
    OC_HIP_CHECK(hipMemcpyAsync(itsPieces.get(), pieces.data(), sizeof(PackPiece) * pieces.size(),
                                hipMemcpyHostToDevice, s));
    kPackResults<<<std::min<size_t>(pieces.size(), 8192), BLOCK, 0, s>>>(itsPieces.as<PackPiece>(),
                                                                          static_cast<uint32_t>(pieces.size()),
                                                                          itsPacked.as<uint8_t>());
    OC_LAUNCH_CHECK();
    std::shared_ptr<uint8_t> block = itsPinned->acquire(total);
    OC_HIP_CHECK(hipMemcpyAsync(block.get(), itsPacked.get(), total, hipMemcpyDeviceToHost, s));
    OC_HIP_CHECK(hipStreamSynchronize(s));
    for (uint32_t j = 0; j < B; ++j)
    {
        BatchOutcome& o = outcomes[members[live[j]]];
        if (!o.error.empty())
            continue;
        const BatchRequest& r = requests[members[live[j]]];
        o.result = HostBuildResult::inBlock(r.cloud, r.proc, L.depth, L.dense, table[j].m, table[j].E, used[j],
                                            std::vector<float>(7, 0.f), block, block.get() + offsets[j]);
    }
    itsTimings[6] += std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - tc).count();
}

template <typename T, uint32_t MODE, bool RANDOM>
void BatchBuilder::subsampleGroup(uint32_t B, const BatchLevels& L, const BatchBounds& bounds)
{
    const uint32_t depth     = L.depth;
    const BatchCloud<T>* ctx = itsTable.as<BatchCloud<T>>();
    hipStream_t s            = itsStream;
    const BatchLaunch& LG    = itsLaunch;
    uint32_t* prefix         = itsPrefix.as<uint32_t>();
    auto slot                = [&](uint32_t k) { return prefix + static_cast<uint64_t>(k) * (MAX_BATCH + 1); };
    for (uint32_t l = 1; l <= depth; ++l)
    {
        scanBatch<uint32_t>(InternalJob<T>{ctx, l}, B, bounds.levelNodes[l], s);
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(LevelNodeCount<T>{ctx, 1, 0}, B, slot(P_K));
        kSegmentsB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_K), L, l);
        scanBatch<uint64_t>(SegmentJob<T>{ctx}, B, bounds.segments, s);
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(LevelNodeCount<T>{ctx, 8, 0}, B, slot(P_SEG));
        kTileListB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_SEG));
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(TileCount<T>{ctx}, B, slot(P_TILES));
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(WordCount<T>{ctx}, B, slot(P_WORDS));
        kClearBitmapsB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_WORDS));
        kMarkB<T><<<LG.tileGrid, TILE, 0, s>>>(ctx, B, slot(P_TILES));
        scanBatch<uint32_t>(PopcountJob<T>{ctx}, B, bounds.words, s);
        kLevelCellsB<T><<<(B + 255) / 256, 256, 0, s>>>(ctx, B);
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(LevelNodeCount<T>{ctx, 1, 1}, B, slot(P_K1));
        kBatchNodesB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_K1));
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(ScalarCount<T>{ctx, BS_CELLS, true}, B, slot(P_CELLS));
        kClearCellsB<T><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_CELLS), MODE);
        kWorkPrefix<<<1, SCAN_THREADS, 0, s>>>(TileCount<T>{ctx}, B, slot(P_LIVE_TILES));
        kAccumulateB<T, MODE, RANDOM><<<LG.tileGrid, TILE, 0, s>>>(ctx, B, slot(P_LIVE_TILES));
        if constexpr (MODE == COLOR_INTER_WEIGHTED)
        {
            scanBatch<uint32_t>(CellCountJob<T>{ctx}, B, bounds.cells, s);
            kScatterPayloadB<T, RANDOM><<<LG.tileGrid, TILE, 0, s>>>(ctx, B, slot(P_LIVE_TILES));
        }
        kEmitB<T, MODE><<<LG.threadGrid, BLOCK, 0, s>>>(ctx, B, slot(P_CELLS));
        kAdvanceB<T><<<(B + 255) / 256, 256, 0, s>>>(ctx, B);
        OC_LAUNCH_CHECK();
    }
}

}  // namespace oc
