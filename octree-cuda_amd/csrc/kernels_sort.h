// Counting and distribution as a two-pass Morton bucket sort of the cloud.
//
// The cloud arrives in arbitrary spatial order, so counting it straight into the level-0 grid and
// scattering every point to its leaf costs one random global atomic and a handful of random
// partial-line stores per point. Here the points are sorted by the Morton code of their level-0
// cell instead (child slot order: x bit 4, y bit 2, z bit 1, as point_count_propagation.cuh:62-70),
// in two passes that only use LDS atomics per point:
//   kBucketHist     per-workgroup LDS histogram of coarse buckets (level-lc cells) -> bucket sizes
//   kBucketScatter  each workgroup ranks its chunk of 4096 points per bucket in LDS, reserves one
//                   range per bucket with a single global atomic, stages the chunk in bucket order
//                   in LDS and writes 32-byte sorted-point records (coordinates, point index,
//                   colour) plus the 3*lc-bit fine code as contiguous runs
//   kFineSort       one workgroup per bucket: LDS histogram over the bucket's (2^lc)^3 level-0
//                   cells, exclusive scan, dense counts of levels 0..lc of the count pyramid,
//                   then the permutation into full Morton order (4-byte indices)
// In Morton order every octree node is one contiguous range, so after the merge a leaf's points
// are the range [sortedStart, sortedStart + count). The records move once more, then: gathered
// through the permutation and ordered within every leaf by its parent's subsampling cells
// (kLeafCellOrder, with the subsampling kernels), or in plain Morton order (kGatherSorted) when
// the root is the only leaf. The distribution is a coalesced copy of the leaf ranges of that
// sorted cloud into the leaf data blocks (kDistributeSorted).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"
#include "kernels_chunking.h"
#include "scan.h"
#include "types.h"

namespace oc {

constexpr uint32_t MAX_COARSE_LEVEL = 5;                               // (2^5)^3 level-0 cells per bucket
constexpr uint32_t MAX_FINE_CELLS   = 1u << (3 * MAX_COARSE_LEVEL);    // 32768
constexpr uint32_t MAX_BUCKETS      = 4096;                            // (512 >> 5)^3
constexpr int SORT_THREADS          = 1024;
constexpr int SORT_WAVES            = SORT_THREADS / 64;
constexpr int SCATTER_THREADS       = 1024;
constexpr int STAGE_ITEMS           = 4;
constexpr uint32_t STAGE_CHUNK      = SCATTER_THREADS * STAGE_ITEMS;   // points per scatter workgroup
constexpr int HIST_THREADS          = 512;
constexpr uint32_t XCD_COUNT        = 8;  // workgroup b runs on XCD b mod 8, each XCD with its own L2

// The item of a work list of `count` entries that this workgroup takes when XCD x is to take the
// x-th eighth of the list, front to back (~0u: none). Neighbouring items then share one L2 instead
// of being spread over eight. The grid must hold at least count rounded up to a multiple of 8.
__device__ __forceinline__ uint32_t xcdItem(uint32_t count)
{
    const uint32_t per = (count + XCD_COUNT - 1) / XCD_COUNT;
    const uint32_t at  = blockIdx.x / XCD_COUNT;
    if (at >= per)
        return ~0u;
    const uint32_t item = (blockIdx.x % XCD_COUNT) * per + at;
    return item < count ? item : ~0u;
}

// A point of the sorted cloud: the original coordinates, its index in the input cloud and its
// colour (r | g << 8 | b << 16).
template <typename T>
struct alignas(sizeof(T) == 8 ? 16 : 4) SortedPoint
{
    T x, y, z;
    uint32_t idx;
    uint32_t rgb;
};
static_assert(sizeof(SortedPoint<double>) == 32, "double sorted point is 32 bytes");
static_assert(sizeof(SortedPoint<float>) == 20, "float sorted point is 20 bytes");
static_assert(STAGE_CHUNK <= 65535, "staged slots are 16-bit");

// Bit interleaving for coordinates below 2^10 (part1by2 / compact1by2).
__host__ __device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x000003FFu;
    v = (v ^ (v << 16)) & 0xFF0000FFu;
    v = (v ^ (v << 8)) & 0x0300F00Fu;
    v = (v ^ (v << 4)) & 0x030C30C3u;
    v = (v ^ (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ __forceinline__ uint32_t compact3(uint32_t v)
{
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0xFF0000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

// Morton code whose 3-bit groups are child slots (x -> 4, y -> 2, z -> 1).
__host__ __device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

// Exclusive scan over a 1024-thread workgroup; `total` receives the sum. Contains barriers.
__device__ __forceinline__ uint32_t sortBlockExclusiveScan(uint32_t v, uint32_t& total, uint32_t* lds)
{
    const uint32_t inc = waveInclusiveScan(v);
    const int w        = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63)
        lds[w] = inc;
    __syncthreads();
    uint32_t off = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < SORT_WAVES; ++i)
    {
        const uint32_t s = lds[i];
        off += (i < w) ? s : 0u;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return inc - v + off;
}

// Exclusive scan over a workgroup of THREADS threads (a multiple of 64); `total` receives the sum.
// Contains barriers.
template <int THREADS>
__device__ __forceinline__ uint32_t sortBlockExclusiveScanN(uint32_t v, uint32_t& total, uint32_t* lds)
{
    const uint32_t inc = waveInclusiveScan(v);
    const int w        = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63)
        lds[w] = inc;
    __syncthreads();
    uint32_t off = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < THREADS / 64; ++i)
    {
        const uint32_t s = lds[i];
        off += (i < w) ? s : 0u;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return inc - v + off;
}

// Level-0 Morton code of a point, or ~0u when it lies outside the grid.
__device__ __forceinline__ uint32_t pointMorton(double x, double y, double z, const Gridding& g)
{
    uint64_t ix, iy, iz;
    gridCoords(x, y, z, g, ix, iy, iz);
    if (ix >= g.gridSize || iy >= g.gridSize || iz >= g.gridSize)
        return ~0u;
    return morton3(static_cast<uint32_t>(ix), static_cast<uint32_t>(iy), static_cast<uint32_t>(iz));
}

template <typename T>
static __global__ __launch_bounds__(HIST_THREADS) void kBucketHist(const uint8_t* __restrict__ cloud, uint32_t n,
                                                             uint32_t stride, Gridding g, uint32_t fineBits,
                                                             uint32_t buckets, uint32_t* __restrict__ bucketCount,
                                                             uint32_t* __restrict__ err)
{
    __shared__ uint32_t h[MAX_BUCKETS];
    for (uint32_t b = threadIdx.x; b < buckets; b += HIST_THREADS)
        h[b] = 0;
    __syncthreads();
    bool outside = false;
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        double x, y, z;
        loadXYZ<T>(cloud + i * stride, x, y, z);
        const uint32_t m = pointMorton(x, y, z, g);
        if (m == ~0u)
        {
            outside = true;
            continue;
        }
        atomicAdd(h + (m >> fineBits), 1u);
    }
    if (outside)
        atomicOr(err, ERR_OUTSIDE);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < buckets; b += HIST_THREADS)
        if (h[b])
            atomicAdd(bucketCount + b, h[b]);
}

// bucket sizes -> bucket starts (exclusive) and the scatter cursors
struct BucketStartOut
{
    uint32_t* start;
    uint32_t* cursor;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t) const
    {
        start[i]  = prefix;
        cursor[i] = prefix;
    }
};

// Each workgroup takes a chunk of STAGE_CHUNK consecutive points, ranks them per bucket in LDS,
// reserves one range per occupied bucket with a single global atomic, stages the records in LDS
// in bucket order and writes them out so that consecutive lanes write consecutive records of one
// bucket.
// LDS of one bucket-scatter workgroup
template <typename T>
struct ScatterLds
{
    SortedPoint<T> stage[STAGE_CHUNK];
    uint32_t delta[MAX_BUCKETS];           // counts, then global start - local start
    uint16_t localStart[MAX_BUCKETS];      // first staged slot of every bucket
    uint32_t scratch[SCATTER_THREADS / 64];
};

// One chunk of STAGE_CHUNK points (all threads of the workgroup; ends with the LDS free again).
template <typename T>
__device__ __forceinline__ void bucketScatterChunk(uint32_t chunk, const uint8_t* __restrict__ cloud, uint32_t n,
                                                   uint32_t stride, const Gridding& g, uint32_t fineBits,
                                                   uint32_t* __restrict__ cursor, SortedPoint<T>* __restrict__ dst,
                                                   uint16_t* __restrict__ fineCode, ScatterLds<T>& lds)
{
    SortedPoint<T>* stage = lds.stage;
    uint32_t* delta       = lds.delta;
    uint16_t* localStart  = lds.localStart;
    uint32_t* scratch     = lds.scratch;
    for (uint32_t b = threadIdx.x; b < MAX_BUCKETS; b += SCATTER_THREADS)
        delta[b] = 0;
    __syncthreads();

    const uint64_t base = static_cast<uint64_t>(chunk) * STAGE_CHUNK;
    uint32_t code[STAGE_ITEMS], rank[STAGE_ITEMS];
    SortedPoint<T> p[STAGE_ITEMS];
#pragma unroll
    for (int j = 0; j < STAGE_ITEMS; ++j)
    {
        const uint64_t i = base + static_cast<uint64_t>(j) * SCATTER_THREADS + threadIdx.x;
        code[j]          = ~0u;
        rank[j]          = 0;
        if (i < n)
        {
            const uint8_t* rec = cloud + i * stride;
            __builtin_memcpy(&p[j].x, rec, 3 * sizeof(T));
            const uint8_t* col = rec + 3 * sizeof(T);
            p[j].idx           = static_cast<uint32_t>(i);
            p[j].rgb = static_cast<uint32_t>(col[0]) | (static_cast<uint32_t>(col[1]) << 8) |
                       (static_cast<uint32_t>(col[2]) << 16);
            code[j] = pointMorton(static_cast<double>(p[j].x), static_cast<double>(p[j].y),
                                  static_cast<double>(p[j].z), g);
            if (code[j] != ~0u)
                rank[j] = atomicAdd(delta + (code[j] >> fineBits), 1u);
        }
    }
    __syncthreads();

    // exclusive scan of the bucket counts (MAX_BUCKETS / SCATTER_THREADS consecutive per thread),
    // one global reservation per occupied bucket
    constexpr int PER = MAX_BUCKETS / SCATTER_THREADS;
    uint32_t cnt[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k)
    {
        cnt[k] = delta[PER * threadIdx.x + k];
        sum += cnt[k];
    }
    uint32_t total;
    uint32_t off = sortBlockExclusiveScanN<SCATTER_THREADS>(sum, total, scratch);
#pragma unroll
    for (int k = 0; k < PER; ++k)
    {
        const uint32_t b = PER * threadIdx.x + k;
        localStart[b]    = static_cast<uint16_t>(off);
        if (cnt[k])
            delta[b] = atomicAdd(cursor + b, cnt[k]) - off;
        off += cnt[k];
    }
    __syncthreads();
    // records into LDS in bucket order; the fine code straight to its final place (this lane
    // already has the point's Morton code)
    const uint32_t fineMask = (1u << fineBits) - 1u;
#pragma unroll
    for (int j = 0; j < STAGE_ITEMS; ++j)
        if (code[j] != ~0u)
        {
            const uint32_t b    = code[j] >> fineBits;
            const uint32_t slot = localStart[b] + rank[j];
            stage[slot]         = p[j];
            const uint32_t d    = slot + delta[b];
            if (d < n)
                fineCode[d] = static_cast<uint16_t>(code[j] & fineMask);
        }
    __syncthreads();

    // a staged record's bucket is recomputed from its coordinates (a few FP64 operations) rather
    // than searched for among the buckets' first slots (a dozen dependent LDS reads)
    for (uint32_t j = threadIdx.x; j < total; j += SCATTER_THREADS)
    {
        const SortedPoint<T> r = stage[j];
        const uint32_t m       = pointMorton(static_cast<double>(r.x), static_cast<double>(r.y),
                                             static_cast<double>(r.z), g);  // in the grid: it was staged
        const uint32_t d       = j + delta[m >> fineBits];
        if (d < n)  // always: the histogram and this pass compute the same codes
            dst[d] = r;
    }
    __syncthreads();
}

template <typename T>
static __global__ __launch_bounds__(SCATTER_THREADS) void kBucketScatter(const uint8_t* __restrict__ cloud, uint32_t n,
                                                                uint32_t stride, Gridding g, uint32_t fineBits,
                                                                uint32_t buckets, uint32_t* __restrict__ cursor,
                                                                SortedPoint<T>* __restrict__ dst,
                                                                uint16_t* __restrict__ fineCode)
{
    __shared__ ScatterLds<T> lds;
    bucketScatterChunk<T>(blockIdx.x, cloud, n, stride, g, fineBits, cursor, dst, fineCode, lds);
}

struct FineSortParams
{
    uint32_t lc;                   // coarse level: a bucket is one level-lc cell
    uint32_t bucketGrid;           // buckets per axis (chunking grid >> lc)
    uint32_t levelOffset[MAX_LEVELS];
    uint32_t levelGrid[MAX_LEVELS];
};

// Dense counts of bucket b's cells on levels 0..lc (kernelPropagatePointCounts,
// point_count_propagation.cuh:34-86, as differences of the Morton prefix h over the bucket's
// fine cells; h null: an empty bucket, all zero). With occ and 32-cell rows (lc = 5), the
// level-0 occupancy bits instead of the level-0 zero counts: a wave writes two whole rows, 32
// consecutive cells from a multiple of 32, so each row is one word of the bitmap, written whole,
// no atomics; the count of a level-0 cell is then written only where it is not zero (the later
// passes read level-0 counts of occupied cells only, and test occupancy on the bitmap).
__device__ __forceinline__ void bucketCounts(uint32_t b, const FineSortParams& fp, const uint32_t* h,
                                             uint32_t* __restrict__ count, uint32_t* __restrict__ occ = nullptr)
{
    const uint32_t bx = compact3(b >> 2), by = compact3(b >> 1), bz = compact3(b);
    for (uint32_t k = 0; k <= fp.lc; ++k)
    {
        const uint32_t side   = 1u << (fp.lc - k);
        const uint32_t n3     = side * side * side;
        const uint64_t gk     = fp.levelGrid[k];
        const uint64_t offset = fp.levelOffset[k];
        const bool bits       = k == 0 && occ && side == 32;
        if (bits && !h)
        {
            // an empty bucket: its 32 x 32 rows of level 0 are zero words
            for (uint32_t r = threadIdx.x; r < side * side; r += SORT_THREADS)
            {
                const uint64_t gy = static_cast<uint64_t>(by) * side + (r % side),
                               gz = static_cast<uint64_t>(bz) * side + (r / side);
                occ[(offset + static_cast<uint64_t>(bx) * side + gy * gk + gz * gk * gk) >> 5] = 0;
            }
            continue;
        }
        for (uint32_t j = threadIdx.x; j < n3; j += SORT_THREADS)
        {
            const uint32_t lx = j & (side - 1), ly = (j >> (fp.lc - k)) & (side - 1), lz = j >> (2 * (fp.lc - k));
            uint32_t v = 0;
            if (h)
            {
                const uint32_t m = morton3(lx, ly, lz);
                v                = h[(m + 1) << (3 * k)] - h[m << (3 * k)];
            }
            const uint64_t gx = static_cast<uint64_t>(bx) * side + lx, gy = static_cast<uint64_t>(by) * side + ly,
                           gz = static_cast<uint64_t>(bz) * side + lz;
            const uint64_t d = offset + gx + gy * gk + gz * gk * gk;
            if (!bits || v)
                count[d] = v;
            if (bits)  // uniform; every lane of the wave is here (n3 % SORT_THREADS == 0)
            {
                const uint64_t set = __ballot(v != 0);
                if ((threadIdx.x & 31) == 0)
                    occ[d >> 5] = static_cast<uint32_t>(set >> (threadIdx.x & 32));
            }
        }
    }
}

// In place: the exclusive prefix of h[0 .. cells), the total in h[cells]. All threads; cells is a
// power of two (8^lc). Each thread scans a run of R = cells / SORT_THREADS consecutive entries
// around one workgroup scan of the runs' sums (a workgroup scan per SORT_THREADS entries would
// take 32 of them, 64 barriers, per 32^3-cell bucket). A thread walks its run starting at entry
// t mod R, so the lanes of a wave read different banks (runs of 32 words start in one bank).
__device__ __forceinline__ void bucketScan(uint32_t* h, uint32_t cells, uint32_t* scratch)
{
    if (cells <= SORT_THREADS)
    {
        const uint32_t c = threadIdx.x;
        const uint32_t v = c < cells ? h[c] : 0u;
        uint32_t total;
        const uint32_t ex = sortBlockExclusiveScan(v, total, scratch);
        if (c < cells)
            h[c] = ex;
        if (threadIdx.x == 0)
            h[cells] = total;
        __syncthreads();
        return;
    }
    const uint32_t R    = cells / SORT_THREADS;
    uint32_t* run       = h + threadIdx.x * R;
    const uint32_t from = threadIdx.x & (R - 1);
    // the run's sum, and the sum of its entries before `from` (read last)
    uint32_t sum = 0, before = 0;
    for (uint32_t k = 0; k < R; ++k)
    {
        const uint32_t i = (from + k) & (R - 1);
        const uint32_t v = run[i];
        sum += v;
        before += i < from ? v : 0u;
    }
    uint32_t total;
    const uint32_t off = sortBlockExclusiveScan(sum, total, scratch);
    // the same walk: entries from `from` on start at off + before, the wrapped ones at off
    uint32_t acc = off + before;
    for (uint32_t k = 0; k < R; ++k)
    {
        const uint32_t i = (from + k) & (R - 1);
        if (i == 0)
            acc = off;
        const uint32_t v = run[i];
        run[i]           = acc;
        acc += v;
    }
    if (threadIdx.x == 0)
        h[cells] = total;
    __syncthreads();
}

// Buckets above FINE_SPLIT points are sorted by several workgroups (kFineChunkHist,
// kFineSplitScan, kFineChunkPerm): one workgroup's time grows with its bucket, and a dense
// cluster puts more than a million points in one bucket while the median holds a quarter of that.
constexpr uint32_t FINE_SPLIT = 196608;

// Fine codes per thread and step of the fine-sort loops: all of a step's loads are in flight
// before its LDS atomics (a load -> atomic -> load chain waits out the memory latency per code).
constexpr uint32_t FINE_ITEMS = 8;

// LDS histogram of fineCode[lo, hi) into h.
__device__ __forceinline__ void fineHistogram(const uint16_t* __restrict__ fineCode, uint32_t lo, uint32_t hi,
                                              uint32_t* h)
{
    uint32_t i = lo + threadIdx.x;
    for (; i + (FINE_ITEMS - 1) * SORT_THREADS < hi; i += FINE_ITEMS * SORT_THREADS)
    {
        uint32_t c[FINE_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            c[k] = fineCode[i + k * SORT_THREADS];
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            atomicAdd(h + c[k], 1u);
    }
    for (; i < hi; i += SORT_THREADS)
        atomicAdd(h + fineCode[i], 1u);
}

// perm[base + slot] = i for every i in [lo, hi), slot taken from the running starts h[code].
__device__ __forceinline__ void finePermute(const uint16_t* __restrict__ fineCode, uint32_t lo, uint32_t hi,
                                            uint32_t* h, uint32_t base, uint32_t* __restrict__ perm)
{
    uint32_t i = lo + threadIdx.x;
    for (; i + (FINE_ITEMS - 1) * SORT_THREADS < hi; i += FINE_ITEMS * SORT_THREADS)
    {
        uint32_t c[FINE_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            c[k] = fineCode[i + k * SORT_THREADS];
        uint32_t slot[FINE_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            slot[k] = atomicAdd(h + c[k], 1u);
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            perm[base + slot[k]] = i + k * SORT_THREADS;
    }
    for (; i < hi; i += SORT_THREADS)
    {
        const uint32_t slot = atomicAdd(h + fineCode[i], 1u);
        perm[base + slot]   = i;
    }
}

// One bucket (all threads of the workgroup; h holds MAX_FINE_CELLS + 1 values). Ends with a barrier.
__device__ __forceinline__ void fineSortBucket(uint32_t b, const uint16_t* __restrict__ fineCode,
                                               const uint32_t* __restrict__ bucketStart, const FineSortParams& fp,
                                               uint32_t* __restrict__ count, uint32_t* __restrict__ perm, uint32_t* h,
                                               uint32_t* scratch, uint32_t* __restrict__ occ = nullptr)
{
    const uint32_t cells = 1u << (3 * fp.lc);
    const uint32_t lo = bucketStart[b], hi = bucketStart[b + 1];
    if (hi - lo > FINE_SPLIT)
        return;  // sorted by the split path
    if (lo == hi)
    {
        // an empty bucket (most of them in a sparse cloud): zero counts, no histogram, no scan
        bucketCounts(b, fp, nullptr, count, occ);
        return;
    }

    for (uint32_t c = threadIdx.x; c <= cells; c += SORT_THREADS)
        h[c] = 0;
    __syncthreads();
    fineHistogram(fineCode, lo, hi, h);
    __syncthreads();
    bucketScan(h, cells, scratch);
    bucketCounts(b, fp, h, count, occ);
    __syncthreads();

    // the permutation into Morton order: 4-byte scattered stores instead of 32-byte records
    finePermute(fineCode, lo, hi, h, lo, perm);
}

static __global__ __launch_bounds__(SORT_THREADS) void kFineSort(const uint16_t* __restrict__ fineCode,
                                                           const uint32_t* __restrict__ bucketStart, FineSortParams fp,
                                                           uint32_t* __restrict__ count, uint32_t* __restrict__ perm,
                                                           uint32_t* __restrict__ occ)
{
    __shared__ uint32_t h[MAX_FINE_CELLS + 1];
    __shared__ uint32_t scratch[SORT_WAVES];
    fineSortBucket(blockIdx.x, fineCode, bucketStart, fp, count, perm, h, scratch, occ);
}

// chunks of the buckets above FINE_SPLIT points: entry = first chunk of the bucket's points,
// bucket << 32 | part
struct SplitChunksIn
{
    const uint32_t* bucketStart;
    __device__ uint32_t operator()(uint64_t b) const
    {
        const uint32_t n = bucketStart[b + 1] - bucketStart[b];
        return n > FINE_SPLIT ? (n + FINE_SPLIT - 1) / FINE_SPLIT : 0u;
    }
};

struct SplitChunksOut
{
    uint64_t* list;
    __device__ void operator()(uint64_t b, uint32_t prefix, uint32_t v) const
    {
        for (uint32_t k = 0; k < v; ++k)
            list[prefix + k] = (b << 32) | k;
    }
};

// Split path, 1: the histogram of one chunk over the bucket's fine cells, to chunkHist[chunk].
__device__ __forceinline__ void fineChunkHist(uint32_t chunk, const uint16_t* __restrict__ fineCode,
                                              const uint32_t* __restrict__ bucketStart,
                                              const uint64_t* __restrict__ chunks, uint32_t lc,
                                              uint32_t* __restrict__ chunkHist, uint32_t* h)
{
    const uint32_t cells = 1u << (3 * lc);
    const uint64_t e     = chunks[chunk];
    const uint32_t b = static_cast<uint32_t>(e >> 32), part = static_cast<uint32_t>(e);
    const uint32_t lo = bucketStart[b] + part * FINE_SPLIT;
    const uint32_t hi = min(bucketStart[b + 1], lo + FINE_SPLIT);
    for (uint32_t c = threadIdx.x; c < cells; c += SORT_THREADS)
        h[c] = 0;
    __syncthreads();
    fineHistogram(fineCode, lo, hi, h);
    __syncthreads();
    uint32_t* dst = chunkHist + static_cast<uint64_t>(chunk) * MAX_FINE_CELLS;
    for (uint32_t c = threadIdx.x; c < cells; c += SORT_THREADS)
        dst[c] = h[c];
}

static __global__ __launch_bounds__(SORT_THREADS) void kFineChunkHist(const uint16_t* __restrict__ fineCode,
                                                                const uint32_t* __restrict__ bucketStart,
                                                                const uint64_t* __restrict__ chunks,
                                                                const uint32_t* __restrict__ chunkCount, uint32_t lc,
                                                                uint32_t* __restrict__ chunkHist)
{
    __shared__ uint32_t h[MAX_FINE_CELLS];
    if (blockIdx.x >= *chunkCount)
        return;  // the whole workgroup
    fineChunkHist(blockIdx.x, fineCode, bucketStart, chunks, lc, chunkHist, h);
}

// Split path, 2: per split bucket (the workgroup of its part-0 chunk), the bucket's histogram as
// the sum of its chunks', its scan and dense counts. The permutation is then written by
// partitions of the bucket's fine cells, one per chunk: partition k starts at the first cell
// whose points start at or after k * FINE_SPLIT in the bucket, so each writes one contiguous
// range of the permutation (chunks of points would interleave their runs in every cell). The
// first chunk's slot receives the cells' starts; word 0 of chunk k's slot (k >= 1) the first cell
// of partition k.
// (all threads; chunk is the first chunk of its bucket)
__device__ __forceinline__ void fineSplitScan(uint32_t chunk, const uint32_t* __restrict__ bucketStart,
                                              const uint64_t* __restrict__ chunks, const FineSortParams& fp,
                                              uint32_t* __restrict__ count, uint32_t* __restrict__ chunkHist,
                                              uint32_t* h, uint32_t* scratch, uint32_t* __restrict__ occ = nullptr)
{
    const uint32_t b     = static_cast<uint32_t>(chunks[chunk] >> 32);
    const uint32_t parts = (bucketStart[b + 1] - bucketStart[b] + FINE_SPLIT - 1) / FINE_SPLIT;
    const uint32_t cells = 1u << (3 * fp.lc);
    uint32_t* first      = chunkHist + static_cast<uint64_t>(chunk) * MAX_FINE_CELLS;
    for (uint32_t c = threadIdx.x; c < cells; c += SORT_THREADS)
    {
        uint32_t v = 0;
        for (uint32_t k = 0; k < parts; ++k)
            v += first[static_cast<uint64_t>(k) * MAX_FINE_CELLS + c];
        h[c] = v;
    }
    __syncthreads();
    bucketScan(h, cells, scratch);  // every chunk histogram has been read
    bucketCounts(b, fp, h, count, occ);
    for (uint32_t c = threadIdx.x; c < cells; c += SORT_THREADS)
        first[c] = h[c];
    for (uint32_t k = 1 + threadIdx.x; k < parts; k += SORT_THREADS)
    {
        // the first cell c with h[c] >= k * FINE_SPLIT (k * FINE_SPLIT < h[cells], so c <= cells)
        const uint32_t target = k * FINE_SPLIT;
        uint32_t lo = 0, hi = cells;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi) >> 1;
            if (h[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        first[static_cast<uint64_t>(k) * MAX_FINE_CELLS] = lo;
    }
}

static __global__ __launch_bounds__(SORT_THREADS) void kFineSplitScan(const uint32_t* __restrict__ bucketStart,
                                                                const uint64_t* __restrict__ chunks,
                                                                const uint32_t* __restrict__ chunkCount,
                                                                FineSortParams fp, uint32_t* __restrict__ count,
                                                                uint32_t* __restrict__ chunkHist,
                                                                uint32_t* __restrict__ occ)
{
    __shared__ uint32_t h[MAX_FINE_CELLS + 1];
    __shared__ uint32_t scratch[SORT_WAVES];
    if (blockIdx.x >= *chunkCount || static_cast<uint32_t>(chunks[blockIdx.x]) != 0)
        return;  // the whole workgroup: not the first chunk of a bucket
    fineSplitScan(blockIdx.x, bucketStart, chunks, fp, count, chunkHist, h, scratch, occ);
}

// Split path, 3: partition `part` of the bucket's fine cells: every point of the bucket is read,
// those in the partition's cells [cLo, cHi) take their slots from the cells' running starts.
__device__ __forceinline__ void fineChunkPerm(uint32_t chunk, const uint16_t* __restrict__ fineCode,
                                              const uint32_t* __restrict__ bucketStart,
                                              const uint64_t* __restrict__ chunks, uint32_t lc,
                                              const uint32_t* __restrict__ chunkHist, uint32_t* __restrict__ perm,
                                              uint32_t* h)
{
    const uint32_t cells = 1u << (3 * lc);
    const uint64_t e     = chunks[chunk];
    const uint32_t b = static_cast<uint32_t>(e >> 32), part = static_cast<uint32_t>(e);
    const uint32_t base  = bucketStart[b];
    const uint32_t end   = bucketStart[b + 1];
    const uint32_t parts = (end - base + FINE_SPLIT - 1) / FINE_SPLIT;
    const uint32_t* start = chunkHist + static_cast<uint64_t>(chunk - part) * MAX_FINE_CELLS;
    const uint32_t cLo    = part ? chunkHist[static_cast<uint64_t>(chunk) * MAX_FINE_CELLS] : 0u;
    const uint32_t cHi    = part + 1 < parts ? chunkHist[static_cast<uint64_t>(chunk + 1) * MAX_FINE_CELLS] : cells;
    const uint32_t span   = cHi - cLo;
    for (uint32_t c = threadIdx.x; c < span; c += SORT_THREADS)
        h[c] = start[cLo + c];
    __syncthreads();
    uint32_t i = base + threadIdx.x;
    for (; i + (FINE_ITEMS - 1) * SORT_THREADS < end; i += FINE_ITEMS * SORT_THREADS)
    {
        uint32_t c[FINE_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            c[k] = static_cast<uint32_t>(fineCode[i + k * SORT_THREADS]) - cLo;  // wraps below the partition
#pragma unroll
        for (uint32_t k = 0; k < FINE_ITEMS; ++k)
            if (c[k] < span)
                perm[base + atomicAdd(h + c[k], 1u)] = i + k * SORT_THREADS;
    }
    for (; i < end; i += SORT_THREADS)
    {
        const uint32_t c = static_cast<uint32_t>(fineCode[i]) - cLo;
        if (c < span)
            perm[base + atomicAdd(h + c, 1u)] = i;
    }
}

static __global__ __launch_bounds__(SORT_THREADS) void kFineChunkPerm(const uint16_t* __restrict__ fineCode,
                                                                const uint32_t* __restrict__ bucketStart,
                                                                const uint64_t* __restrict__ chunks,
                                                                const uint32_t* __restrict__ chunkCount, uint32_t lc,
                                                                const uint32_t* __restrict__ chunkHist,
                                                                uint32_t* __restrict__ perm)
{
    __shared__ uint32_t h[MAX_FINE_CELLS];
    // the partitions of one bucket read the same fine codes: on one XCD, from one L2
    const uint32_t chunk = xcdItem(*chunkCount);
    if (chunk == ~0u)
        return;  // the whole workgroup
    fineChunkPerm(chunk, fineCode, bucketStart, chunks, lc, chunkHist, perm, h);
}

// The sorted cloud: gathered records, written in order. The permutation only moves a record
// inside its bucket, so the reads of one output range hit one bucket's lines. Workgroup b runs on
// XCD b mod 8, and XCD x gathers the x-th eighth of the output front to back: each bucket's lines
// are fetched into a single L2 and reused by all of its points instead of by up to eight L2s.
constexpr int GATHER_ITEMS      = 4;
constexpr uint32_t GATHER_TILE  = BLOCK * GATHER_ITEMS;

// Workgroups of kGatherSorted for n points (a multiple of XCD_COUNT).
inline unsigned gatherGrid(uint32_t n)
{
    const uint64_t tiles = (static_cast<uint64_t>(n) + GATHER_TILE - 1) / GATHER_TILE;
    const uint64_t perXcd = (tiles + XCD_COUNT - 1) / XCD_COUNT;
    return static_cast<unsigned>((perXcd ? perXcd : 1) * XCD_COUNT);
}

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kGatherSorted(const SortedPoint<T>* __restrict__ bucketed,
                                                        const uint32_t* __restrict__ perm, uint32_t n,
                                                        SortedPoint<T>* __restrict__ sorted)
{
    const uint32_t perXcd = gridDim.x / XCD_COUNT;
    const uint64_t tile   = static_cast<uint64_t>(blockIdx.x % XCD_COUNT) * perXcd + blockIdx.x / XCD_COUNT;
    const uint64_t base   = tile * GATHER_TILE + threadIdx.x;
#pragma unroll
    for (int k = 0; k < GATHER_ITEMS; ++k)
    {
        const uint64_t j = base + static_cast<uint64_t>(k) * BLOCK;
        if (j < n)
            sorted[j] = bucketed[perm[j]];
    }
}

// Start of every leaf's range in the sorted cloud: the points of all nodes that precede it in
// Morton order, i.e. of the lower-slot siblings of the leaf and of each of its ancestors.
__device__ __forceinline__ void sortedStartOf(uint32_t id, const LevelParams& lp, const uint32_t* __restrict__ nodeKey,
                                              const uint32_t* __restrict__ nodeCount,
                                              const uint8_t* __restrict__ flags, const int32_t* __restrict__ parent,
                                              const int32_t* __restrict__ children,
                                              uint32_t* __restrict__ sortedStart)
{
    if ((flags[id] & (NODE_FINISHED | NODE_INTERNAL)) != NODE_FINISHED)
        return;
    uint32_t start = 0;
    int32_t cur    = static_cast<int32_t>(id);
    int32_t p      = parent[cur];
    while (p >= 0)
    {
        const uint32_t key = nodeKey[cur];
        const uint32_t l   = levelOfKey(key, lp);
        const uint32_t g   = lp.grid[l];
        const uint32_t lin = key - lp.offset[l];
        const uint32_t x = lin % g, y = (lin / g) % g, z = lin / (g * g);
        const int slot     = static_cast<int>(((x & 1) << 2) | ((y & 1) << 1) | (z & 1));
        for (int k = 0; k < slot; ++k)
        {
            const int32_t c = children[8 * static_cast<uint64_t>(p) + k];
            if (c >= 0)
                start += nodeCount[c];
        }
        cur = p;
        p   = parent[cur];
    }
    sortedStart[id] = start;
}

static __global__ __launch_bounds__(BLOCK) void kSortedStart(uint32_t m, LevelParams lp, const uint32_t* __restrict__ nodeKey,
                                                       const uint32_t* __restrict__ nodeCount,
                                                       const uint8_t* __restrict__ flags,
                                                       const int32_t* __restrict__ parent,
                                                       const int32_t* __restrict__ children,
                                                       uint32_t* __restrict__ sortedStart)
{
    for (uint64_t i = globalThread(); i < m; i += globalStride())
        sortedStartOf(static_cast<uint32_t>(i), lp, nodeKey, nodeCount, flags, parent, children, sortedStart);
}

// K5: kernelDistributePoints (point_distributing.cuh:31-76) over the sorted cloud. The leaf of a
// point is the one of its level-0 cell (cellLeaf); its slot is its offset in the leaf's sorted
// range, so consecutive threads write consecutive records of one leaf block.
constexpr int DIST_ITEMS = 4;

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kDistributeSorted(const SortedPoint<T>* __restrict__ sorted, uint32_t n,
                                                            Gridding g, const uint32_t* __restrict__ cellLeaf,
                                                            const uint32_t* __restrict__ sortedStart,
                                                            const uint64_t* __restrict__ dataIdx,
                                                            uint32_t* __restrict__ lut, OutRecord* __restrict__ out,
                                                            double sx, double sy, double sz)
{
    // DIST_ITEMS records per thread and step, each level of the load chain (record, its leaf,
    // the leaf's placement) issued for all of them before the next
    const uint64_t stride = globalStride();
    auto cellOf = [&](const SortedPoint<T>& p, uint32_t& cell) {
        uint64_t ix, iy, iz;
        gridCoords(static_cast<double>(p.x), static_cast<double>(p.y), static_cast<double>(p.z), g, ix, iy, iz);
        cell = static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize);
        return ix < g.gridSize && iy < g.gridSize && iz < g.gridSize;  // only in-grid points are sorted
    };
    for (uint64_t s0 = globalThread(); s0 < n; s0 += stride * DIST_ITEMS)
    {
        if (s0 + (DIST_ITEMS - 1) * stride < n)
        {
            SortedPoint<T> p[DIST_ITEMS];
            uint32_t leaf[DIST_ITEMS];
            bool in[DIST_ITEMS];
#pragma unroll
            for (int k = 0; k < DIST_ITEMS; ++k)
                p[k] = sorted[s0 + k * stride];
            // the lookups go to cell 0 for a point outside the grid (none is sorted), so every
            // load is unconditional and only the stores are predicated
#pragma unroll
            for (int k = 0; k < DIST_ITEMS; ++k)
            {
                uint32_t cell;
                in[k]   = cellOf(p[k], cell);
                leaf[k] = cellLeaf[in[k] ? cell : 0u];
            }
            uint64_t pos[DIST_ITEMS];
#pragma unroll
            for (int k = 0; k < DIST_ITEMS; ++k)
                pos[k] = dataIdx[leaf[k]] + (s0 + k * stride - sortedStart[leaf[k]]);
#pragma unroll
            for (int k = 0; k < DIST_ITEMS; ++k)
                if (in[k])
                {
                    lut[pos[k]] = p[k].idx;
                    storeRecord(out + pos[k], quantizePoint(p[k].x, p[k].y, p[k].z, p[k].rgb, g, sx, sy, sz));
                }
            continue;
        }
        for (uint64_t s = s0; s < n; s += stride)
        {
            const SortedPoint<T> p = sorted[s];
            uint32_t cell;
            if (!cellOf(p, cell))
                continue;
            const uint32_t leaf = cellLeaf[cell];
            const uint64_t pos  = dataIdx[leaf] + (s - sortedStart[leaf]);
            lut[pos]            = p.idx;
            storeRecord(out + pos, quantizePoint(p.x, p.y, p.z, p.rgb, g, sx, sy, sz));
        }
    }
}

}  // namespace oc
