// Kernels of the batched build: many clouds in one launch per stage.
//
// Every cloud of a batch keeps its own arrays (a BatchCloud entry of a device table holds their
// addresses and the cloud's configuration), so each cloud's results are exactly those of a build
// of that cloud alone. A stage's work items of all clouds form one index space: item i belongs to
// cloud c with prefix[c] <= i < prefix[c + 1], where prefix is an exclusive scan of the clouds'
// item counts kept in device memory. Counts that only the device knows (internal nodes, segments,
// tiles, cells, chunks of a level) are scanned on the device, so a batch needs the host only once,
// after the merge (to size the node arrays), and every other launch has a fixed grid: persistent
// workgroups that stride over the items. The per-item work is the single-cloud builder's own
// (the device functions of kernels_sort.h, kernels_chunking.h and kernels_subsample.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device_common.h"
#include "kernels_chunking.h"
#include "kernels_sort.h"
#include "kernels_subsample.h"
#include "scan.h"
#include "types.h"

namespace oc {

constexpr uint32_t MAX_BATCH      = 1024;   // clouds per batched launch
constexpr uint32_t HIST_CHUNK     = 16384;  // points per work item of the batched bucket histogram
constexpr uint32_t DIST_CHUNK     = 1024;   // sorted points per work item of the batched distribution
constexpr uint32_t BATCH_SCAN_X   = 256;    // workgroups per cloud of a batched multi-workgroup scan
constexpr uint32_t ERR_OVERFLOW   = 4u;     // N + I > E (A14)

// per-cloud scalar slots (u64)
enum BatchSlot : int
{
    BS_ERROR       = 0,
    BS_M           = 1,
    BS_LEAF_TOTAL  = 2,
    BS_SEG_TOTAL   = 3,  // (tiles << 32 | inputs) of the current level
    BS_TABLE_TOTAL = 4,  // internal nodes of the level / chunk counts
    BS_SPLIT       = 5,
    BS_CURSOR      = 6,  // N + subsampled points so far
    BS_CELLS       = 7,  // occupied cells of the current level (0 once the cloud has failed)
    BS_DEAD        = 8,  // non-zero: the cloud's build failed; its later stages do no work
    BS_LEVEL_START = 9,  // u32[MAX_LEVELS + 1] packed into slots 9..14
    BS_SLOTS       = 16
};

template <typename T>
struct BatchCloud
{
    // input and configuration
    const uint8_t* cloud;
    uint32_t n, stride;
    Gridding grid;      // chunking gridding: cubic minimum, side, chunking grid
    double sx, sy, sz;  // 1 / scale
    double sideHalf[MAX_LEVELS];
    double side[MAX_LEVELS];
    uint64_t E;
    uint64_t seed;
    uint64_t wordsPerNode;
    uint32_t threshold;
    uint32_t g;          // subsampling grid
    uint32_t keyShift;   // leaf-order word keys >> keyShift < LEAF_KEYS
    uint32_t kBound;     // internal nodes of a level the level arrays hold
    uint32_t tileBound;  // entries of the tile list
    uint32_t m;          // sparse nodes (after the merge)
    uint64_t* sc;
    // count and merge
    uint32_t *bucketCount, *bucketStart, *bucketCursor;
    SortedPoint<T>*bucketed, *sorted;
    uint16_t* fineCode;
    uint32_t* perm;
    uint64_t* splitChunks;
    uint32_t* splitHist;
    uint32_t* count;
    int32_t* denseToSparse;
    uint32_t *nodeKey, *nodeCount;  // bounded by min(D, (depth + 1) N)
    // nodes
    uint32_t* nodePoints;
    uint64_t* nodeData;
    int32_t *nodeParent, *nodeChild;
    uint8_t* nodeFlags;
    uint32_t *nodeSorted, *nodeAux, *nodeStage, *nodeSub, *nodeRegion;
    uint32_t *keyOut, *countOut;  // packed copies of the node keys and dense counts for the result
    uint64_t* chunkList;
    // output and staging
    uint32_t* lut;
    OutRecord* out;
    uint32_t* stLut;
    OutRecord* stOut;
    T* stPts;
    uint32_t* stRgb;
    // per-level work
    uint32_t *internalList, *segCount, *segSrc, *segTile, *segInput, *tileList;
    NodeGrid* nodeGrid;
    uint32_t *bitmap, *prefix, *nodeFirst;
    uint64_t *win, *color;
    uint32_t *cellCnt, *cellStart, *cellKey, *cellNode, *cellOf, *rankOf;
    WeightPayload* payload;
    uint64_t* winKey;
    uint64_t* scanSums;  // BATCH_SCAN_X block sums of the batched scans

    __device__ __forceinline__ uint32_t* levelStart() const { return reinterpret_cast<uint32_t*>(sc + BS_LEVEL_START); }
    __device__ __forceinline__ bool dead() const { return sc[BS_DEAD] != 0; }
    __device__ __forceinline__ PointStore<T> store() const
    {
        PointStore<T> ps;
        ps.sorted = sorted;
        ps.stLut  = stLut;
        ps.stOut  = stOut;
        ps.stPts  = stPts;
        ps.stRgb  = stRgb;
        ps.N      = n;
        ps.cubic  = grid;
        ps.sx     = sx;
        ps.sy     = sy;
        ps.sz     = sz;
        return ps;
    }
};

// Octree levels of the batch (one chunking grid per batch).
struct BatchLevels
{
    uint32_t depth, lc, buckets;
    uint32_t offset[MAX_LEVELS];
    uint32_t grid[MAX_LEVELS];
    uint32_t nodes[MAX_LEVELS];
    uint32_t dense;
};

// The cloud of item i: the last c with prefix[c] <= i (prefix[B] = total > i).
__device__ __forceinline__ uint32_t findCloud(const uint32_t* prefix, uint32_t B, uint32_t i)
{
    uint32_t lo = 0, hi = B;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// the clouds' work prefix in LDS, for kernels with one item per thread
__device__ __forceinline__ void loadPrefix(const uint32_t* __restrict__ prefix, uint32_t B, uint32_t* lds)
{
    for (uint32_t i = threadIdx.x; i <= B; i += blockDim.x)
        lds[i] = prefix[i];
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Work prefixes: an exclusive scan of per-cloud counts, computed by one workgroup.
// ------------------------------------------------------------------------------------------------

template <typename Count>
static __global__ __launch_bounds__(SCAN_THREADS) void kWorkPrefix(Count count, uint32_t B, uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t lds[SCAN_ITEMS * SCAN_WAVES];
    blockScanAll<uint32_t>(count, ExclusiveToArray{prefix}, B, prefix + B, lds);
}

// per-cloud counts read from the table
template <typename T>
struct ScalarCount  // sc[slot] (low 32 bits), or 0 for a failed cloud when live is set
{
    const BatchCloud<T>* ctx;
    int slot;
    bool live;
    __device__ uint32_t operator()(uint64_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        if (live && cc.dead())
            return 0u;
        return static_cast<uint32_t>(cc.sc[slot]);
    }
};

template <typename T>
struct TileCount  // tiles of the current level (high half of BS_SEG_TOTAL), 0 once failed
{
    const BatchCloud<T>* ctx;
    __device__ uint32_t operator()(uint64_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        return cc.dead() ? 0u : static_cast<uint32_t>(cc.sc[BS_SEG_TOTAL] >> 32);
    }
};

template <typename T>
struct LevelNodeCount  // internal nodes x factor (+ extra), 0 once failed
{
    const BatchCloud<T>* ctx;
    uint32_t factor, extra;
    __device__ uint32_t operator()(uint64_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        if (cc.dead())
            return 0u;
        return static_cast<uint32_t>(cc.sc[BS_TABLE_TOTAL]) * factor + extra;
    }
};

template <typename T>
struct WordCount  // bitmap words of the level's internal nodes
{
    const BatchCloud<T>* ctx;
    __device__ uint32_t operator()(uint64_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        return cc.dead() ? 0u : static_cast<uint32_t>(cc.sc[BS_TABLE_TOTAL] * cc.wordsPerNode);
    }
};

// ------------------------------------------------------------------------------------------------
// Batched exclusive scans: one independent scan per cloud, grid (BATCH_SCAN_X, B), in the three
// launches of scanExclusive. A job binds cloud c to {in, out, n, total, sums}.
// ------------------------------------------------------------------------------------------------

template <typename V, typename Job>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanReduceB(Job job)
{
    __shared__ V lds[SCAN_WAVES];
    const auto j = job.bind(blockIdx.y);
    uint64_t first, last;
    tileRange(scanTiles(j.n), first, last);
    V s = 0;
    for (uint64_t tile = first; tile < last; ++tile)
    {
        V v[SCAN_ITEMS];
        loadTile<V>(j.in, tile * SCAN_TILE, j.n, v);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            s += v[k];
    }
    V total;
    (void)blockExclusiveScan(s, total, lds);
    if (threadIdx.x == 0)
        reinterpret_cast<V*>(j.sums)[blockIdx.x] = total;
}

template <typename V, typename Job>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanTilesB(Job job, uint32_t blocks)
{
    __shared__ V lds[SCAN_ITEMS * SCAN_WAVES];
    const auto j = job.bind(blockIdx.y);
    V* sums      = reinterpret_cast<V*>(j.sums);
    V carry      = 0;
    for (uint64_t base = 0; base < blocks; base += SCAN_TILE)
    {
        V v[SCAN_ITEMS], pre[SCAN_ITEMS];
        loadTile<V>([sums](uint64_t i) { return sums[i]; }, base, blocks, v);
        const V chunk = tileExclusiveScan(v, pre, lds);
        storeTile<V>([sums](uint64_t i, V p, V) { sums[i] = p; }, base, blocks, v, pre, carry);
        carry += chunk;
    }
    if (threadIdx.x == 0)
        j.setTotal(carry);
}

template <typename V, typename Job>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanDownB(Job job)
{
    __shared__ V lds[SCAN_ITEMS * SCAN_WAVES];
    const auto j = job.bind(blockIdx.y);
    uint64_t first, last;
    tileRange(scanTiles(j.n), first, last);
    V carry = reinterpret_cast<const V*>(j.sums)[blockIdx.x];
    for (uint64_t tile = first; tile < last; ++tile)
    {
        V v[SCAN_ITEMS], pre[SCAN_ITEMS];
        loadTile<V>(j.in, tile * SCAN_TILE, j.n, v);
        const V chunk = tileExclusiveScan(v, pre, lds);
        storeTile<V>(j.out, tile * SCAN_TILE, j.n, v, pre, carry);
        carry += chunk;
    }
}

// maxItems: an upper bound of any cloud's item count; it sizes the workgroups per cloud (each at
// least 8 tiles of the largest cloud), so small scans do not launch mostly idle workgroups.
inline uint32_t scanWidth(uint64_t maxItems)
{
    const uint64_t tiles = scanTiles(maxItems);
    return static_cast<uint32_t>(std::min<uint64_t>(BATCH_SCAN_X, std::max<uint64_t>(1, (tiles + 7) / 8)));
}

template <typename V, typename Job>
void scanBatch(Job job, uint32_t B, uint64_t maxItems, hipStream_t stream)
{
    const uint32_t X = scanWidth(maxItems);
    const dim3 grid(X, B);
    kScanReduceB<V, Job><<<grid, SCAN_THREADS, 0, stream>>>(job);
    kScanTilesB<V, Job><<<dim3(1, B), SCAN_THREADS, 0, stream>>>(job, X);
    kScanDownB<V, Job><<<grid, SCAN_THREADS, 0, stream>>>(job);
}

// A bound job: the cloud's functors, item count, block sums and where the total goes.
template <typename In, typename Out>
struct BoundScan
{
    In in;
    Out out;
    uint64_t n;
    uint64_t* sums;
    uint64_t* total;  // the cloud's scalar slot (or nullptr)
    uint32_t* total32;
    template <typename V>
    __device__ void setTotal(V v) const
    {
        if (total)
            *total = static_cast<uint64_t>(v);
        if (total32)
            *total32 = static_cast<uint32_t>(v);
    }
};

template <typename In, typename Out>
__device__ __forceinline__ BoundScan<In, Out> bound(In in, Out out, uint64_t n, uint64_t* sums, uint64_t* total,
                                                   uint32_t* total32 = nullptr)
{
    return BoundScan<In, Out>{in, out, n, sums, total, total32};
}

// canonical sparse ids over the dense pyramid (A2)
template <typename T>
struct OccupancyJob
{
    const BatchCloud<T>* ctx;
    uint64_t dense;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        return bound(OccupiedIn{cc.count}, OccupiedOut{cc.denseToSparse, cc.nodeKey, cc.nodeCount, cc.count}, dense,
                     cc.scanSums, cc.sc + BS_M);
    }
};

// leaf data blocks (A4): scan of the block sizes in id order, in place
template <typename T>
struct LeafBlockJob
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        return bound(ArrayIn{cc.nodeAux}, ExclusiveToArray{cc.nodeAux}, cc.m, cc.scanSums, cc.sc + BS_LEAF_TOTAL);
    }
};

// chunks of the leaves the subsampling reads (none when the cloud subsamples nothing)
template <typename T>
struct LeafChunkJob
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t n        = (cc.g >= 2 && cc.m >= 2) ? cc.m : 0;
        return bound(LeafIn{cc.nodeFlags, cc.nodeParent, cc.nodeCount}, LeafOut{cc.chunkList}, n, cc.scanSums,
                     cc.sc + BS_TABLE_TOTAL);
    }
};

// internal nodes of level l -> list; a list longer than the level arrays fails the cloud
struct BoundedInternalOut
{
    uint32_t* list;
    uint32_t levelStart;
    uint32_t bound;
    uint64_t* sc;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t v) const
    {
        if (!v)
            return;
        if (prefix < bound)
            list[prefix] = levelStart + static_cast<uint32_t>(i);
        else
        {
            atomicOr(reinterpret_cast<unsigned long long*>(sc + BS_ERROR), ERR_CAPACITY);
            atomicOr(reinterpret_cast<unsigned long long*>(sc + BS_DEAD), 1ull);
        }
    }
};

template <typename T>
struct InternalJob
{
    const BatchCloud<T>* ctx;
    uint32_t level;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t* ls      = cc.levelStart();
        const uint32_t start    = ls[level];
        const uint32_t end      = ls[level + 1] < cc.m ? ls[level + 1] : cc.m;
        const uint64_t n        = (cc.dead() || end < start) ? 0 : end - start;
        return bound(InternalIn{cc.nodeFlags, start}, BoundedInternalOut{cc.internalList, start, cc.kBound, cc.sc}, n,
                     cc.scanSums, cc.sc + BS_TABLE_TOTAL);
    }
};

template <typename T>
struct SegmentJob
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t n        = cc.dead() ? 0 : 8ull * cc.sc[BS_TABLE_TOTAL];
        return bound(SegmentIn{cc.segCount}, SegmentOut{cc.segTile, cc.segInput}, n, cc.scanSums,
                     cc.sc + BS_SEG_TOTAL);
    }
};

template <typename T>
struct PopcountJob
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t words    = cc.dead() ? 0 : cc.sc[BS_TABLE_TOTAL] * cc.wordsPerNode;
        return bound(PopcountIn{cc.bitmap}, ExclusiveToArray{cc.prefix}, words, cc.scanSums, nullptr,
                     cc.prefix + words);
    }
};

template <typename T>
struct CellCountJob  // weighted filter: cell starts of the payload runs
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t cells    = cc.sc[BS_CELLS];
        return bound(ArrayIn{cc.cellCnt}, ExclusiveToArray{cc.cellStart}, cells, cc.scanSums, nullptr,
                     cc.cellStart + cells);
    }
};

template <typename T>
struct PlacedChunkJob
{
    const BatchCloud<T>* ctx;
    __device__ auto bind(uint32_t c) const
    {
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t n        = cc.dead() ? 0 : cc.m;
        return bound(PlacedChunksIn{cc.nodeFlags, cc.nodePoints}, LeafOut{cc.chunkList}, n, cc.scanSums,
                     cc.sc + BS_TABLE_TOTAL);
    }
};

// ------------------------------------------------------------------------------------------------
// Count stage
// ------------------------------------------------------------------------------------------------

// zero scalars and bucket counts of every cloud
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kBatchInit(const BatchCloud<T>* __restrict__ ctx, uint32_t B, uint32_t buckets)
{
    const uint64_t per = BS_SLOTS + buckets + 1;
    for (uint64_t i = globalThread(); i < per * B; i += globalStride())
    {
        const BatchCloud<T>& cc = ctx[i / per];
        const uint32_t k        = static_cast<uint32_t>(i % per);
        if (k < BS_SLOTS)
            cc.sc[k] = (k == BS_CURSOR) ? cc.n : 0;
        else
            cc.bucketCount[k - BS_SLOTS] = 0;
    }
}

// items: (cloud, chunk of HIST_CHUNK points)
template <typename T>
static __global__ __launch_bounds__(HIST_THREADS) void kBucketHistB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                              const uint32_t* __restrict__ prefix, uint32_t fineBits,
                                                              uint32_t buckets)
{
    __shared__ uint32_t h[MAX_BUCKETS];
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        for (uint32_t b = threadIdx.x; b < buckets; b += HIST_THREADS)
            h[b] = 0;
        __syncthreads();
        const uint64_t lo = static_cast<uint64_t>(item - prefix[c]) * HIST_CHUNK;
        const uint64_t hi = lo + HIST_CHUNK < cc.n ? lo + HIST_CHUNK : cc.n;
        bool outside      = false;
        for (uint64_t i = lo + threadIdx.x; i < hi; i += HIST_THREADS)
        {
            double x, y, z;
            loadXYZ<T>(cc.cloud + i * cc.stride, x, y, z);
            const uint32_t m = pointMorton(x, y, z, cc.grid);
            if (m == ~0u)
            {
                outside = true;
                continue;
            }
            atomicAdd(h + (m >> fineBits), 1u);
        }
        if (outside)
            atomicOr(reinterpret_cast<unsigned long long*>(cc.sc + BS_ERROR), ERR_OUTSIDE);
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < buckets; b += HIST_THREADS)
            if (h[b])
                atomicAdd(cc.bucketCount + b, h[b]);
        __syncthreads();
    }
}

// bucket starts and cursors, one workgroup per cloud
template <typename T>
static __global__ __launch_bounds__(SCAN_THREADS) void kBucketScanB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                             uint32_t buckets)
{
    __shared__ uint32_t lds[SCAN_ITEMS * SCAN_WAVES];
    for (uint32_t c = blockIdx.x; c < B; c += gridDim.x)
    {
        const BatchCloud<T>& cc = ctx[c];
        blockScanAll<uint32_t>(ArrayIn{cc.bucketCount}, BucketStartOut{cc.bucketStart, cc.bucketCursor}, buckets,
                               cc.bucketStart + buckets, lds);
        __syncthreads();
    }
}

// items: (cloud, chunk of STAGE_CHUNK points)
template <typename T>
static __global__ __launch_bounds__(SCATTER_THREADS) void kBucketScatterB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                                 const uint32_t* __restrict__ prefix,
                                                                 uint32_t fineBits)
{
    __shared__ ScatterLds<T> lds;
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        bucketScatterChunk<T>(item - prefix[c], cc.cloud, cc.n, cc.stride, cc.grid, fineBits, cc.bucketCursor,
                              cc.bucketed, cc.fineCode, lds);
    }
}

// items: (cloud, bucket)
template <typename T>
static __global__ __launch_bounds__(SORT_THREADS) void kFineSortB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                            FineSortParams fp, uint32_t buckets)
{
    __shared__ uint32_t h[MAX_FINE_CELLS + 1];
    __shared__ uint32_t scratch[SORT_WAVES];
    const uint32_t total = B * buckets;
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const BatchCloud<T>& cc = ctx[item / buckets];
        fineSortBucket(item % buckets, cc.fineCode, cc.bucketStart, fp, cc.count, cc.perm, h, scratch);
        __syncthreads();
    }
}

// split path: the chunk list of every cloud's buckets above FINE_SPLIT points (one workgroup per cloud)
template <typename T>
static __global__ __launch_bounds__(SCAN_THREADS) void kSplitChunksB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                              uint32_t buckets)
{
    __shared__ uint32_t lds[SCAN_ITEMS * SCAN_WAVES];
    for (uint32_t c = blockIdx.x; c < B; c += gridDim.x)
    {
        const BatchCloud<T>& cc = ctx[c];
        uint32_t* total         = reinterpret_cast<uint32_t*>(cc.sc + BS_SPLIT);
        if (cc.n > FINE_SPLIT)
            blockScanAll<uint32_t>(SplitChunksIn{cc.bucketStart}, SplitChunksOut{cc.splitChunks}, buckets, total,
                                   lds);
        else if (threadIdx.x == 0)
            *total = 0;
        __syncthreads();
    }
}

template <typename T>
static __global__ __launch_bounds__(SORT_THREADS) void kFineChunkHistB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                                 const uint32_t* __restrict__ prefix, uint32_t lc)
{
    __shared__ uint32_t h[MAX_FINE_CELLS];
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        fineChunkHist(item - prefix[c], cc.fineCode, cc.bucketStart, cc.splitChunks, lc, cc.splitHist, h);
        __syncthreads();
    }
}

template <typename T>
static __global__ __launch_bounds__(SORT_THREADS) void kFineSplitScanB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                                 const uint32_t* __restrict__ prefix,
                                                                 FineSortParams fp)
{
    __shared__ uint32_t h[MAX_FINE_CELLS + 1];
    __shared__ uint32_t scratch[SORT_WAVES];
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t chunk    = item - prefix[c];
        if (static_cast<uint32_t>(cc.splitChunks[chunk]) != 0)
            continue;  // not the first chunk of its bucket (uniform over the workgroup)
        fineSplitScan(chunk, cc.bucketStart, cc.splitChunks, fp, cc.count, cc.splitHist, h, scratch);
        __syncthreads();
    }
}

template <typename T>
static __global__ __launch_bounds__(SORT_THREADS) void kFineChunkPermB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                                 const uint32_t* __restrict__ prefix, uint32_t lc)
{
    __shared__ uint32_t h[MAX_FINE_CELLS];
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        fineChunkPerm(item - prefix[c], cc.fineCode, cc.bucketStart, cc.splitChunks, lc, cc.splitHist, cc.perm, h);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------
// Merge
// ------------------------------------------------------------------------------------------------

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kPropagateB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                      uint32_t offset, uint32_t grid, uint32_t childOffset,
                                                      uint32_t childGrid)
{
    const uint64_t per = static_cast<uint64_t>(grid) * grid * grid;
    for (uint64_t i = globalThread(); i < per * B; i += globalStride())
        propagateOne(static_cast<uint32_t>(i % per), ctx[i / per].count, offset, grid, childOffset, childGrid);
}

template <typename T>
static __global__ void kLevelStartsB(const BatchCloud<T>* __restrict__ ctx, LevelOffsets lo)
{
    const BatchCloud<T>& cc = ctx[blockIdx.x];
    const uint32_t* m       = reinterpret_cast<const uint32_t*>(cc.sc + BS_M);
    kLevelStartsBody(cc.nodeKey, m, lo, cc.levelStart());
    if (threadIdx.x == 0)
        cc.levelStart()[lo.levels] = *m;  // the end of the top level
}

__device__ __forceinline__ LevelParams levelParams(const BatchLevels& L, uint32_t threshold)
{
    LevelParams lp;
    for (uint32_t l = 0; l < MAX_LEVELS; ++l)
    {
        lp.offset[l] = L.offset[l];
        lp.grid[l]   = L.grid[l];
    }
    lp.depth     = L.depth;
    lp.threshold = threshold;
    return lp;
}

// per-thread items over the sparse nodes of all clouds (prefix = exclusive scan of m)
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kNodeLinksB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                      const uint32_t* __restrict__ prefix, BatchLevels L)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t id       = static_cast<uint32_t>(i) - wp[c];
        nodeLinksOf(id, levelParams(L, cc.threshold), cc.nodeKey, cc.nodeCount, cc.count, cc.denseToSparse, cc.nodePoints,
                    cc.nodeData, cc.nodeParent, cc.nodeChild, cc.nodeFlags, cc.nodeAux, cc.nodeSub);
        cc.keyOut[id]   = cc.nodeKey[id];
        cc.countOut[id] = cc.nodeCount[id];
    }
}

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kLeafDataIdxB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                        const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        leafDataIdxOf(static_cast<uint32_t>(i) - wp[c], cc.m, cc.nodeAux, cc.nodeFlags, cc.nodeChild, cc.nodePoints,
                      cc.nodeData);
    }
}

// items: level-0 nodes of all clouds
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kCellLeafB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                     const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        cellLeafOf(static_cast<uint32_t>(i) - wp[c], cc.nodeKey, cc.nodeFlags, cc.nodeParent, cc.count);
    }
}

// ------------------------------------------------------------------------------------------------
// Distribution
// ------------------------------------------------------------------------------------------------

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kSortedStartB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                        const uint32_t* __restrict__ prefix, BatchLevels L)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        sortedStartOf(static_cast<uint32_t>(i) - wp[c], levelParams(L, cc.threshold), cc.nodeKey, cc.nodeCount,
                      cc.nodeFlags, cc.nodeParent, cc.nodeChild, cc.nodeSorted);
    }
}

// items: leaf chunks of all clouds (prefix over the clouds' chunk counts)
template <typename T>
static __global__ __launch_bounds__(LEAF_THREADS) void kLeafCellOrderB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                                 const uint32_t* __restrict__ prefix, BatchLevels L)
{
    __shared__ LeafOrderLds lds;
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        // the cubic minimum is Gridding's first three members
        const LeafOrderView p{&cc.grid.minX, cc.sideHalf, cc.side, L.offset, L.grid, L.depth, cc.g, cc.keyShift};
        leafCellOrderChunk<T>(cc.chunkList[item - prefix[c]], p, cc.nodeKey, cc.nodeParent, cc.nodeCount,
                              cc.nodeSorted, cc.bucketed, cc.perm, cc.sorted, lds);
    }
}

// Morton-order gather for the clouds that subsample nothing (no leaf chunk); items: points
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kGatherSortedB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                         const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        if (cc.sc[BS_TABLE_TOTAL] != 0)
            continue;  // ordered by kLeafCellOrderB
        const uint32_t j = static_cast<uint32_t>(i) - wp[c];
        cc.sorted[j]     = cc.bucketed[cc.perm[j]];
    }
}

// K5 over the sorted clouds; items: (cloud, chunk of DIST_CHUNK sorted points)
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kDistributeSortedB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                             const uint32_t* __restrict__ prefix)
{
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        const uint64_t lo       = static_cast<uint64_t>(item - prefix[c]) * DIST_CHUNK;
        const uint64_t hi       = lo + DIST_CHUNK < cc.n ? lo + DIST_CHUNK : cc.n;
        const Gridding& g       = cc.grid;
        for (uint64_t s = lo + threadIdx.x; s < hi; s += BLOCK)
        {
            const SortedPoint<T> p = cc.sorted[s];
            uint64_t ix, iy, iz;
            gridCoords(static_cast<double>(p.x), static_cast<double>(p.y), static_cast<double>(p.z), g, ix, iy, iz);
            if (ix >= g.gridSize || iy >= g.gridSize || iz >= g.gridSize)
                continue;  // only in-grid points are sorted
            const uint32_t cell = static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize);
            const uint32_t leaf = cc.count[cell];
            const uint64_t pos  = cc.nodeData[leaf] + (s - cc.nodeSorted[leaf]);
            cc.lut[pos]         = p.idx;
            storeRecord(cc.out + pos, quantizePoint(p.x, p.y, p.z, p.rgb, g, cc.sx, cc.sy, cc.sz));
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Subsampling, one level of all clouds per launch
// ------------------------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ SubsampleParams subsampleParams(const BatchCloud<T>& cc, const BatchLevels& L,
                                                           uint32_t level)
{
    SubsampleParams sp;
    sp.cubicMin[0] = cc.grid.minX;
    sp.cubicMin[1] = cc.grid.minY;
    sp.cubicMin[2] = cc.grid.minZ;
    for (uint32_t l = 0; l < MAX_LEVELS; ++l)
        sp.sideHalf[l] = cc.sideHalf[l];
    sp.side        = cc.side[level];
    sp.level       = level;
    sp.depth       = L.depth;
    sp.levelOffset = L.offset[level];
    sp.levelGrid   = L.grid[level];
    sp.g           = cc.g;
    sp.seed        = cc.seed;
    return sp;
}

template <typename T>
__device__ __forceinline__ LevelTables levelTables(const BatchCloud<T>& cc)
{
    return LevelTables{cc.nodeKey,   cc.nodeFlags,  cc.nodeChild,  cc.nodePoints, cc.nodeSorted,
                       cc.nodeStage, cc.internalList, cc.segCount, cc.segSrc,     cc.segTile,
                       cc.segInput,  cc.tileList,   cc.nodeGrid};
}

// items: internal nodes of the level (prefix over K)
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kSegmentsB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                     const uint32_t* __restrict__ prefix, BatchLevels L, uint32_t level)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        segmentsOf(static_cast<uint32_t>(i) - wp[c], subsampleParams(cc, L, level), levelTables(cc));
    }
}

// the tile list (items: segments, prefix over 8 K); a list longer than its array fails the cloud
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kTileListB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                     const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        if ((cc.sc[BS_SEG_TOTAL] >> 32) > cc.tileBound)
        {
            if (static_cast<uint32_t>(i) == wp[c])
            {
                atomicOr(reinterpret_cast<unsigned long long*>(cc.sc + BS_ERROR), ERR_CAPACITY);
                atomicOr(reinterpret_cast<unsigned long long*>(cc.sc + BS_DEAD), 1ull);
            }
            continue;
        }
        tileListOf(static_cast<uint32_t>(i) - wp[c], cc.segCount, cc.segTile, cc.tileList);
    }
}

// zero the bitmaps (items: words, prefix over K x words per node)
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kClearBitmapsB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                         const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c = findCloud(wp, B, static_cast<uint32_t>(i));
        ctx[c].bitmap[static_cast<uint32_t>(i) - wp[c]] = 0;
    }
}

// items: tiles (one per workgroup and step)
template <typename T>
static __global__ __launch_bounds__(TILE) void kMarkB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                const uint32_t* __restrict__ prefix)
{
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        markTile<T>(item - prefix[c], 0, cc.wordsPerNode, cc.tileList, cc.segTile, cc.segCount, cc.segSrc,
                    cc.segInput, cc.nodeGrid, cc.store(), cc.bitmap, cc.cellOf,
                    reinterpret_cast<uint32_t*>(cc.sc + BS_ERROR));
    }
}

// the level's occupied cells; N + I > E fails the cloud (A14) and leaves it no further work
template <typename T>
static __global__ void kLevelCellsB(const BatchCloud<T>* __restrict__ ctx, uint32_t B)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < B; c += gridDim.x * blockDim.x)
    {
        const BatchCloud<T>& cc = ctx[c];
        uint64_t cells          = 0;
        if (!cc.dead())
        {
            const uint64_t words = cc.sc[BS_TABLE_TOTAL] * cc.wordsPerNode;
            cells                = cc.prefix[words];
            if (cc.sc[BS_CURSOR] + cells > cc.E)
            {
                cc.sc[BS_ERROR] |= ERR_OVERFLOW;
                cc.sc[BS_DEAD] = 1;
                cells          = 0;
            }
        }
        cc.sc[BS_CELLS] = cells;
    }
}

// items: internal nodes + 1 per cloud
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kBatchNodesB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                       const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t K        = static_cast<uint32_t>(cc.sc[BS_TABLE_TOTAL]);
        batchNodeOf(static_cast<uint32_t>(i) - wp[c], 0, K, cc.wordsPerNode, cc.nodeGrid, cc.prefix,
                    static_cast<uint32_t>(cc.sc[BS_CURSOR]), cc.nodePoints, cc.nodeStage, cc.nodeFirst);
    }
}

// per-cell state reset (items: cells)
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kClearCellsB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                       const uint32_t* __restrict__ prefix, uint32_t mode)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t k        = static_cast<uint32_t>(i) - wp[c];
        if (mode == COLOR_INTER_WEIGHTED)
            cc.cellCnt[k] = 0;
        else
        {
            cc.win[k] = ~0ull;
            if (mode != COLOR_NONE)
                cc.color[k] = 0;
        }
    }
}

template <typename T, uint32_t MODE, bool RANDOM>
static __global__ __launch_bounds__(TILE) void kAccumulateB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                      const uint32_t* __restrict__ prefix)
{
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        accumulateTile<T, MODE, RANDOM>(item - prefix[c], 0, cc.wordsPerNode, cc.tileList, cc.segTile, cc.segCount,
                                        cc.segSrc, cc.segInput, cc.nodeGrid, cc.bitmap, cc.prefix, cc.cellOf,
                                        cc.store(), cc.win, cc.color, cc.cellCnt, cc.cellKey, cc.cellNode,
                                        cc.rankOf);
    }
}

template <typename T, bool RANDOM>
static __global__ __launch_bounds__(TILE) void kScatterPayloadB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                          const uint32_t* __restrict__ prefix)
{
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        scatterPayloadTile<T, RANDOM>(item - prefix[c], cc.tileList, cc.segTile, cc.segCount, cc.segSrc,
                                      cc.segInput, cc.nodeGrid, cc.cellKey, cc.rankOf, cc.cellStart, cc.cellOf,
                                      cc.store(), cc.payload, cc.winKey);
    }
}

// items: cells
template <typename T, uint32_t MODE>
static __global__ __launch_bounds__(BLOCK) void kEmitB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                 const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        const uint32_t k        = static_cast<uint32_t>(i) - wp[c];
        const uint32_t cursor   = static_cast<uint32_t>(cc.sc[BS_CURSOR]);
        if constexpr (MODE == COLOR_INTER_WEIGHTED)
            emitWeightedCell<T>(k, 0, cc.wordsPerNode, cursor, cc.nodeGrid, cc.bitmap, cc.prefix, cc.cellKey,
                                cc.cellNode, cc.cellStart, cc.payload, cc.winKey, cc.store());
        else
            emitCell<T, MODE>(k, 0, cc.wordsPerNode, cursor, cc.nodeGrid, cc.bitmap, cc.prefix, cc.cellKey,
                              cc.cellNode, cc.win, cc.color, cc.store());
    }
}

template <typename T>
static __global__ void kAdvanceB(const BatchCloud<T>* __restrict__ ctx, uint32_t B)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < B; c += gridDim.x * blockDim.x)
        ctx[c].sc[BS_CURSOR] += ctx[c].sc[BS_CELLS];
}

// ------------------------------------------------------------------------------------------------
// Placement
// ------------------------------------------------------------------------------------------------

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kZeroSubtreeB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                        const uint32_t* __restrict__ prefix)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c = findCloud(wp, B, static_cast<uint32_t>(i));
        ctx[c].nodeSub[static_cast<uint32_t>(i) - wp[c]] = 0;
    }
}

// items: nodes of one level (prefix over the clouds' level sizes); first = the level's first id
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kSubtreeB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                    const uint32_t* __restrict__ prefix, uint32_t level)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        if (cc.dead())
            continue;
        subtreeOf(cc.levelStart()[level] + static_cast<uint32_t>(i) - wp[c], cc.nodeFlags, cc.nodeChild,
                  cc.nodePoints, cc.nodeSub);
    }
}

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kRegionB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                   const uint32_t* __restrict__ prefix, uint32_t level)
{
    __shared__ uint32_t wp[MAX_BATCH + 1];
    loadPrefix(prefix, B, wp);
    for (uint64_t i = globalThread(); i < wp[B]; i += globalStride())
    {
        const uint32_t c        = findCloud(wp, B, static_cast<uint32_t>(i));
        const BatchCloud<T>& cc = ctx[c];
        if (cc.dead())
            continue;
        regionOf(cc.levelStart()[level] + static_cast<uint32_t>(i) - wp[c], cc.m - 1, cc.n, cc.nodeFlags,
                 cc.nodeChild, cc.nodeSub, cc.nodeRegion, cc.nodeData);
    }
}

// items: (internal node, chunk) of all clouds
template <typename T>
static __global__ __launch_bounds__(BLOCK) void kPermuteB(const BatchCloud<T>* __restrict__ ctx, uint32_t B,
                                                    const uint32_t* __restrict__ prefix)
{
    const uint32_t total = prefix[B];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x)
    {
        const uint32_t c        = findCloud(prefix, B, item);
        const BatchCloud<T>& cc = ctx[c];
        permuteChunk<T>(cc.chunkList[item - prefix[c]], cc.n, cc.nodeStage, cc.nodeData, cc.nodePoints, cc.stLut,
                        cc.stOut, cc.stPts, cc.stRgb, cc.grid, cc.sx, cc.sy, cc.sz, cc.lut, cc.out);
    }
}

// ------------------------------------------------------------------------------------------------
// Results: every cloud's snapshot packed into one device buffer, so one copy brings them home
// ------------------------------------------------------------------------------------------------

constexpr uint64_t PACK_CHUNK = 256 * 1024;  // bytes per work item of the packer

// A piece of at most PACK_CHUNK bytes: src and dst are 16-byte aligned.
struct PackPiece
{
    const uint8_t* src;
    uint64_t dst;
    uint64_t bytes;
};

static __global__ __launch_bounds__(BLOCK) void kPackResults(const PackPiece* __restrict__ pieces, uint32_t count,
                                                              uint8_t* __restrict__ dst)
{
    for (uint32_t p = blockIdx.x; p < count; p += gridDim.x)
    {
        const PackPiece piece = pieces[p];
        const uint64_t words  = piece.bytes / 16;
        const uint4* s        = reinterpret_cast<const uint4*>(piece.src);
        uint4* d              = reinterpret_cast<uint4*>(dst + piece.dst);
        for (uint64_t w = threadIdx.x; w < words; w += BLOCK)
            d[w] = s[w];
        for (uint64_t b = words * 16 + threadIdx.x; b < piece.bytes; b += BLOCK)
            dst[piece.dst + b] = piece.src[b];
    }
}

}  // namespace oc
