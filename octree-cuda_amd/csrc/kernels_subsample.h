// Kernels of the level-batched subsampling and the final DFS post-order placement.
//
// All internal nodes of one octree level are processed together (in memory-bounded batches of
// nodes). Each node of a batch owns an occupancy bitmap over its g^3 subsampling grid and the
// exclusive popcount prefix of that bitmap's words, i.e. a rank dictionary: the occupied cells of
// the batch get dense indices in (node, cell index) order; the cell index groups the grid's x-rows
// into 4 x 4 pencils (rowIndex). Per batch:
//   kMark        input point -> cell (the reference's mapPointToGrid with the node's gridding),
//                set the cell's bit
//   prefix scan  popcounts of all bitmap words -> dense index of every occupied cell
//   kAccumulate  winner per cell (atomicMin of rank<<32 | input position), packed colour sums,
//                per-cell point counts for the weighted filter
//   kScatterPayload (weighted only) each point's offset from its cell centre + colour, and its
//                winner key, in cell order
//   kEmit        one output record per occupied cell, in ascending cell order per node
//                (kEmitWeighted for the distance-weighted filter)
// Leaf points come from the sorted cloud, whose leaf ranges kLeafCellOrder has put in the cell
// order of each leaf's parent grid, so a level's first reads of them follow the bitmap.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"
#include "kernels_chunking.h"
#include "kernels_sort.h"
#include "types.h"

namespace oc {

constexpr uint32_t TILE = 256;  // inputs per workgroup in the per-input kernels

// Cell index of a node grid. The x-rows (fixed y, z) are grouped into 4 x 4 pencils, pencils in
// (z / 4, y / 4) order and rows inside a pencil in (z % 4, y % 4) order; a cell is row * g + x.
// A row stays one contiguous run of bits (and of dense ranks), and the 9 rows a cell's 3x3x3
// neighbourhood touches lie a few rows apart instead of a whole z-plane apart.
constexpr uint32_t PENCIL = 4;
constexpr uint32_t EMIT_CHUNK = 32;  // consecutive cell blocks of the weighted emit per XCD turn
constexpr uint32_t XCD_TILE_RUN = 32;  // consecutive input tiles of kMark / kScatterPayload per XCD turn

// Workgroup b runs on XCD b mod 8, each XCD with its own L2. The work item of workgroup b when
// every XCD is to take runs of `run` consecutive items: rounds of 8 * run items, XCD x taking
// items x * run .. x * run + run - 1 of each round. Grids are rounded up to whole rounds.
__device__ __forceinline__ uint32_t xcdRunItem(uint32_t b, uint32_t run)
{
    return (b / (XCD_COUNT * run)) * (XCD_COUNT * run) + (b % XCD_COUNT) * run + (b / XCD_COUNT) % run;
}

__host__ __forceinline__ uint32_t xcdRunGrid(uint64_t items, uint32_t run)
{
    const uint64_t round = static_cast<uint64_t>(XCD_COUNT) * run;
    return static_cast<uint32_t>((items + round - 1) / round * round);
}

__host__ __device__ __forceinline__ uint32_t pencilsPerAxis(uint32_t g)
{
    return (g + PENCIL - 1) / PENCIL;
}

__host__ __device__ __forceinline__ uint64_t rowsPerNode(uint32_t g)
{
    const uint64_t p = pencilsPerAxis(g);
    return p * p * PENCIL * PENCIL;
}

__host__ __device__ __forceinline__ uint32_t rowIndex(uint32_t y, uint32_t z, uint32_t g)
{
    const uint32_t p = pencilsPerAxis(g);
    return (((z / PENCIL) * p + (y / PENCIL)) * PENCIL + (z % PENCIL)) * PENCIL + (y % PENCIL);
}

__host__ __device__ __forceinline__ void rowCoords(uint32_t row, uint32_t g, uint32_t& y, uint32_t& z)
{
    const uint32_t p = pencilsPerAxis(g), pencil = row / (PENCIL * PENCIL), inside = row % (PENCIL * PENCIL);
    y = (pencil % p) * PENCIL + inside % PENCIL;
    z = (pencil / p) * PENCIL + inside / PENCIL;
}

// The points a subsampling level reads: leaf points at positions [0, N) of the sorted cloud (a
// leaf's points are one contiguous range there) and already subsampled points of lower
// levels at positions [N, N + I) (staging, in level order until the final DFS post-order
// placement). A leaf point's output record is derived from its sorted record exactly as the
// distribution writes it.
template <typename T>
struct PointStore
{
    const SortedPoint<T>* sorted;
    uint32_t* stLut;
    OutRecord* stOut;
    T* stPts;
    uint32_t* stRgb;  // distance-weighted filter: staged colours r | g << 8 | b << 16 (8 bits each)
    uint32_t N;
    Gridding cubic;
    double sx, sy, sz;
    __device__ __forceinline__ uint32_t index(uint32_t pos) const
    {
        return pos < N ? sorted[pos].idx : stLut[pos - N];
    }
    __device__ __forceinline__ OutRecord record(uint32_t pos) const
    {
        if (pos < N)
        {
            const SortedPoint<T>& p = sorted[pos];
            return quantizePoint(p.x, p.y, p.z, p.rgb, cubic, sx, sy, sz);
        }
        return loadRecord(stOut + (pos - N));
    }
    // the point's colour r | g << 8 | b << 16, the low 8 bits of each channel (distance-weighted
    // filter only: a staged point's colour comes from stRgb, 4 bytes instead of its 18-byte record)
    __device__ __forceinline__ uint32_t rgb(uint32_t pos) const
    {
        return pos < N ? sorted[pos].rgb : stRgb[pos - N];
    }
    __device__ __forceinline__ void xyz(uint32_t pos, T& x, T& y, T& z) const
    {
        if (pos < N)
        {
            const SortedPoint<T>& p = sorted[pos];
            x                        = p.x;
            y                        = p.y;
            z                        = p.z;
            return;
        }
        const T* q = stPts + 3 * static_cast<uint64_t>(pos - N);
        x          = q[0];
        y          = q[1];
        z          = q[2];
    }
};

struct SubsampleParams
{
    double cubicMin[3];
    double sideHalf[MAX_LEVELS];  // side of level L / 2 (the child offset when descending from L)
    double side;                  // side of the current level
    uint32_t level;
    uint32_t depth;
    uint32_t levelOffset;
    uint32_t levelGrid;
    uint32_t g;  // subsampling grid
    uint64_t seed;
};

// internal nodes of a level -> compact list
struct InternalIn
{
    const uint8_t* flags;
    uint32_t levelStart;
    __device__ uint32_t operator()(uint64_t i) const { return (flags[levelStart + i] & NODE_INTERNAL) ? 1u : 0u; }
};

struct InternalOut
{
    uint32_t* list;
    uint32_t levelStart;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t v) const
    {
        if (v)
            list[prefix] = levelStart + static_cast<uint32_t>(i);
    }
};

// Subsampling grid of the node with linear index lin at level `level` (grid gl per axis): the
// node minimum accumulated top-down exactly like the reference's DFS (calculateChildMinBB,
// tools.cu:29-61; rp_subsampling.cu:16-60): x = min.x; for every step down whose slot has the x
// bit, x = x + side/2 of that step; the extent is bbSize = (min.x + side) - min.x.
__device__ __forceinline__ Gridding nodeGridding(uint32_t lin, uint32_t gl, uint32_t level, uint32_t depth,
                                                 const double* cubicMin, const double* sideHalf, double side,
                                                 uint32_t g)
{
    const uint32_t x = lin % gl, y = (lin / gl) % gl, z = lin / (gl * gl);
    double mx = cubicMin[0], my = cubicMin[1], mz = cubicMin[2];
    for (uint32_t L = depth; L > level; --L)
    {
        const uint32_t b = L - 1 - level;
        if ((x >> b) & 1)
            mx = mx + sideHalf[L];
        if ((y >> b) & 1)
            my = my + sideHalf[L];
        if ((z >> b) & 1)
            mz = mz + sideHalf[L];
    }
    return Gridding{mx, my, mz, (mx + side) - mx, g};
}

// Per internal node k of the level: its eight input segments (children in slot order, the leaf's
// range of the sorted cloud or the child's staged subsample) and its gridding (nodeGridding).
struct LevelTables
{
    const uint32_t* nodeKey;
    const uint8_t* flags;
    const int32_t* children;
    const uint32_t* pointCount;
    const uint32_t* sortedStart;
    const uint32_t* stage;
    uint32_t* internalList;
    uint32_t* segCount;
    uint32_t* segSrc;
    uint32_t* segTile;
    uint32_t* segInput;
    uint32_t* tileList;
    NodeGrid* nodeGrid;
};

__device__ __forceinline__ void segmentsOf(uint64_t k, const SubsampleParams& sp, const LevelTables& lt)
{
    const uint32_t id = lt.internalList[k];
    for (int s = 0; s < 8; ++s)
    {
        const int32_t c = lt.children[8 * static_cast<uint64_t>(id) + s];
        uint32_t cnt = 0, src = 0;
        if (c >= 0)
        {
            cnt = lt.pointCount[c];
            src = (lt.flags[c] & NODE_INTERNAL) ? lt.stage[c] : lt.sortedStart[c];
        }
        lt.segCount[8 * k + s] = cnt;
        lt.segSrc[8 * k + s]   = src;
    }
    const uint32_t key = lt.nodeKey[id];
    NodeGrid ng;
    ng.grid        = nodeGridding(key - sp.levelOffset, sp.levelGrid, sp.level, sp.depth, sp.cubicMin, sp.sideHalf,
                                  sp.side, sp.g);
    const double t = ng.grid.bbSize / sp.g;
    ng.diag        = sqrt((t * 3) * (t * 3) * 3) / 2.0;
    ng.salt        = nodeSalt(sp.seed, key);
    ng.id          = id;
    lt.nodeGrid[k] = ng;
}

static __global__ __launch_bounds__(BLOCK) void kSegments(uint32_t nodes, SubsampleParams sp, LevelTables lt)
{
    for (uint64_t k = globalThread(); k < nodes; k += globalStride())
        segmentsOf(k, sp, lt);
}

// The sorted cloud in leaf and cell order. The fine sort's permutation puts the points in Morton
// order of their finest cells, so every leaf is one range; within a leaf that order says nothing
// about the parent's subsampling grid (a leaf spans 64^3 of the parent's 128^3 cells), and
// subsampling would see consecutive points in unrelated bitmap words and cells, making every
// atomic, rank lookup and payload store of the level random. So the records are gathered through
// the permutation leaf by leaf and reordered by bitmap word in the parent grid (the cell rows of
// the leaf's octant in the bitmap's pencil order, 32 cells per word): every leaf range is cut into
// chunks of at most LEAF_CHUNK points, one workgroup sorts a chunk with a counting sort over an
// LDS histogram and stores every record at its place in the chunk's range. Chunks bound the work
// of one workgroup (a dense cluster puts 10^5 points in one finest cell). Points within a leaf
// are unordered by definition (node point sets are compared), so the order is free to choose;
// the key only orders, so a point the parent grid clamps lands in an edge key.
constexpr int LEAF_THREADS    = 1024;
constexpr int LEAF_ITEMS      = 4;                          // points per thread, held in registers
constexpr uint32_t LEAF_KEYS  = 8192;                       // word keys (histogram entries)
constexpr uint32_t LEAF_CHUNK = LEAF_THREADS * LEAF_ITEMS;  // points per workgroup

struct LeafOrderParams
{
    double cubicMin[3];
    double sideHalf[MAX_LEVELS];
    double side[MAX_LEVELS];  // node side per level
    LevelParams lp;
    uint32_t g;         // subsampling grid (>= 2)
    uint32_t keyShift;  // word keys >> keyShift < LEAF_KEYS
};

// What one leaf chunk reads of the level tables and node sides, by pointer: into the kernel
// argument (single cloud) or into the cloud's table in device memory (batches), so that no copy
// of the per-level arrays is made per chunk.
struct LeafOrderView
{
    const double* cubicMin;  // 3
    const double* sideHalf;  // per level
    const double* side;      // per level
    const uint32_t* offset;  // per level
    const uint32_t* grid;    // per level
    uint32_t depth, g, keyShift;
};

__device__ __forceinline__ LeafOrderView leafOrderView(const LeafOrderParams& p)
{
    return LeafOrderView{p.cubicMin, p.sideHalf, p.side, p.lp.offset, p.lp.grid, p.lp.depth, p.g, p.keyShift};
}

// Keys a leaf chunk can take: the bitmap words of a leaf's octant of the parent grid, shifted.
__host__ __device__ __forceinline__ uint32_t leafKeyCount(uint32_t g, uint32_t keyShift)
{
    const uint32_t h    = g > 1 ? g / 2 : 1;
    const uint64_t keys = rowsPerNode(h) * ((h + 31) / 32);
    return static_cast<uint32_t>(((keys - 1) >> keyShift) + 1);
}

// chunks of a leaf that a subsampling level reads (finished, not internal, with a parent)
struct LeafIn
{
    const uint8_t* flags;
    const int32_t* parent;
    const uint32_t* count;
    __device__ uint32_t operator()(uint64_t i) const
    {
        return ((flags[i] & (NODE_FINISHED | NODE_INTERNAL)) == NODE_FINISHED && parent[i] >= 0)
                       ? (count[i] + LEAF_CHUNK - 1) / LEAF_CHUNK
                       : 0u;
    }
};

// work list entry per chunk: chunk number << 32 | leaf id
struct LeafOut
{
    uint64_t* list;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t v) const
    {
        for (uint32_t c = 0; c < v; ++c)
            list[prefix + c] = (static_cast<uint64_t>(c) << 32) | i;
    }
};

// histogram slot of a key: one pad word per 32 keys, so the per-thread runs of the scan fall in
// distinct banks
__device__ __forceinline__ uint32_t leafSlot(uint32_t key)
{
    return key + (key >> 5);
}

__device__ __forceinline__ uint32_t leafLocal(uint64_t i, uint32_t origin, uint32_t h)
{
    return i < origin ? 0u : static_cast<uint32_t>(i - origin < h ? i - origin : h - 1);
}

// Word key of a point: its cell in the parent grid (multiplied by 1 / t instead of divided by t:
// the key only orders, so a point next to a cell face may take either side) relative to the leaf's
// octant, as (pencil row, x / 32) in the bitmap's order.
struct LeafKeyer
{
    Gridding grid;
    double inv;  // gridSize / bbSize
    uint32_t ox, oy, oz, h, wordsPerRow, shift;
    __device__ __forceinline__ uint32_t operator()(double x, double y, double z) const
    {
        const double top = grid.gridSize - 1.0;
        const uint64_t ix = static_cast<uint64_t>(static_cast<int64_t>(fmin((x - grid.minX) * inv, top)));
        const uint64_t iy = static_cast<uint64_t>(static_cast<int64_t>(fmin((y - grid.minY) * inv, top)));
        const uint64_t iz = static_cast<uint64_t>(static_cast<int64_t>(fmin((z - grid.minZ) * inv, top)));
        const uint32_t lx = leafLocal(ix, ox, h), ly = leafLocal(iy, oy, h), lz = leafLocal(iz, oz, h);
        return (rowIndex(ly, lz, h) * wordsPerRow + (lx >> 5)) >> shift;
    }
};

struct LeafOrderLds
{
    uint32_t hist[LEAF_KEYS + LEAF_KEYS / 32];
    uint32_t scanLds[LEAF_THREADS / 64];
};

// One chunk of one leaf (all threads; the LDS is free again on return).
template <typename T>
__device__ __forceinline__ void leafCellOrderChunk(uint64_t entry, const LeafOrderView& p,
                                                   const uint32_t* __restrict__ nodeKey,
                                                   const int32_t* __restrict__ parent,
                                                   const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ sortedStart,
                                                   const SortedPoint<T>* __restrict__ bucketed,
                                                   const uint32_t* __restrict__ perm, SortedPoint<T>* __restrict__ out,
                                                   LeafOrderLds& lds)
{
    uint32_t* hist    = lds.hist;
    uint32_t* scanLds = lds.scanLds;
    const uint32_t leaf  = static_cast<uint32_t>(entry);
    const uint32_t first = static_cast<uint32_t>(entry >> 32) * LEAF_CHUNK;
    const uint32_t total = count[leaf];
    const uint32_t src   = sortedStart[leaf] + first;
    const uint32_t cnt   = total - first < LEAF_CHUNK ? total - first : LEAF_CHUNK;
    const uint32_t pkey  = nodeKey[static_cast<uint32_t>(parent[leaf])];
    uint32_t pl = 0;
    for (uint32_t k = 1; k <= p.depth; ++k)
        if (pkey >= p.offset[k])
            pl = k;
    LeafKeyer keyer;
    keyer.grid = nodeGridding(pkey - p.offset[pl], p.grid[pl], pl, p.depth, p.cubicMin, p.sideHalf, p.side[pl], p.g);
    keyer.inv  = keyer.grid.gridSize / keyer.grid.bbSize;
    // the leaf's octant of the parent grid: the low bit of its coordinates one level down
    const uint32_t ll = pl - 1, lgl = p.grid[ll], llin = nodeKey[leaf] - p.offset[ll];
    keyer.h           = p.g > 1 ? p.g / 2 : 1;  // cells per axis of the leaf's octant (1 for a 1-cell grid)
    keyer.ox          = ((llin % lgl) & 1u) * keyer.h;
    keyer.oy          = (((llin / lgl) % lgl) & 1u) * keyer.h;
    keyer.oz          = ((llin / (lgl * lgl)) & 1u) * keyer.h;
    keyer.wordsPerRow = (keyer.h + 31) / 32;
    keyer.shift       = p.keyShift;

    // only the keys this grid can produce take part (1024 of the 8192 at g = 64)
    const uint32_t keys = leafKeyCount(p.g, p.keyShift);
    for (uint32_t i = threadIdx.x; i < leafSlot(keys); i += LEAF_THREADS)
        hist[i] = 0;
    __syncthreads();
    // all loads in flight first, then the keys and their ranks within the key
    // every record is gathered once, kept in registers and stored at its place in the chunk's
    // range (a few hundred KB, so the scattered stores merge into whole lines in L2)
    uint32_t key[LEAF_ITEMS], rank[LEAF_ITEMS];
    SortedPoint<T> rec[LEAF_ITEMS];
    {
        uint32_t from[LEAF_ITEMS];
#pragma unroll
        for (int k = 0; k < LEAF_ITEMS; ++k)
        {
            const uint32_t i = threadIdx.x + k * LEAF_THREADS;
            if (i < cnt)
                from[k] = perm[src + i];
        }
#pragma unroll
        for (int k = 0; k < LEAF_ITEMS; ++k)
            if (threadIdx.x + k * LEAF_THREADS < cnt)
                rec[k] = bucketed[from[k]];
#pragma unroll
        for (int k = 0; k < LEAF_ITEMS; ++k)
            if (threadIdx.x + k * LEAF_THREADS < cnt)
            {
                key[k]  = keyer(static_cast<double>(rec[k].x), static_cast<double>(rec[k].y),
                                static_cast<double>(rec[k].z));
                rank[k] = atomicAdd(hist + leafSlot(key[k]), 1u);
            }
    }
    __syncthreads();

    // exclusive scan of the histogram, up to LEAF_KEYS / LEAF_THREADS consecutive keys per thread
    constexpr uint32_t PER = LEAF_KEYS / LEAF_THREADS;
    const uint32_t base    = threadIdx.x * PER;
    uint32_t sum           = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k)
        if (base + k < keys)
            sum += hist[leafSlot(base + k)];
    uint32_t all;
    uint32_t off = sortBlockExclusiveScanN<LEAF_THREADS>(sum, all, scanLds);
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k)
        if (base + k < keys)
        {
            const uint32_t v         = hist[leafSlot(base + k)];
            hist[leafSlot(base + k)] = off;
            off += v;
        }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < LEAF_ITEMS; ++k)
        if (threadIdx.x + k * LEAF_THREADS < cnt)
            out[src + hist[leafSlot(key[k])] + rank[k]] = rec[k];
    __syncthreads();
}

template <typename T>
static __global__ __launch_bounds__(LEAF_THREADS) void kLeafCellOrder(const uint64_t* __restrict__ chunkList, LeafOrderParams p,
                                                                const uint32_t* __restrict__ nodeKey,
                                                                const int32_t* __restrict__ parent,
                                                                const uint32_t* __restrict__ count,
                                                                const uint32_t* __restrict__ sortedStart,
                                                                const SortedPoint<T>* __restrict__ bucketed,
                                                                const uint32_t* __restrict__ perm,
                                                                SortedPoint<T>* __restrict__ out)
{
    __shared__ LeafOrderLds lds;
    leafCellOrderChunk<T>(chunkList[blockIdx.x], leafOrderView(p), nodeKey, parent, count, sortedStart, bucketed, perm,
                          out, lds);
}

// packed (tiles << 32 | inputs) per segment, so one scan yields both prefix sums
struct SegmentIn
{
    const uint32_t* segCount;
    __device__ uint64_t operator()(uint64_t s) const
    {
        uint64_t c = segCount[s];
        return (((c + TILE - 1) / TILE) << 32) | c;
    }
};

struct SegmentOut
{
    uint32_t* segTile;
    uint32_t* segInput;
    __device__ void operator()(uint64_t s, uint64_t prefix, uint64_t) const
    {
        segTile[s]  = static_cast<uint32_t>(prefix >> 32);
        segInput[s] = static_cast<uint32_t>(prefix);
    }
};

__device__ __forceinline__ void tileListOf(uint64_t s, const uint32_t* __restrict__ segCount,
                                           const uint32_t* __restrict__ segTile, uint32_t* __restrict__ tileList)
{
    uint32_t tiles = (segCount[s] + TILE - 1) / TILE;
    uint32_t first = segTile[s];
    for (uint32_t t = 0; t < tiles; ++t)
        tileList[first + t] = static_cast<uint32_t>(s);
}

static __global__ __launch_bounds__(BLOCK) void kTileList(uint64_t segments, const uint32_t* __restrict__ segCount,
                                                    const uint32_t* __restrict__ segTile, uint32_t* __restrict__ tileList)
{
    for (uint64_t s = globalThread(); s < segments; s += globalStride())
        tileListOf(s, segCount, segTile, tileList);
}

// A whole level's setup in one workgroup and one launch, for small levels (up to
// SMALL_LEVEL_NODES nodes and SMALL_LEVEL_POINTS candidate inputs; a small build is bound by its
// launch count): the internal nodes of the level, their segments and griddings, the segment scan
// and the tile list. Writes K to *internalTotal and (tiles << 32 | inputs) to *segTotal.
constexpr uint32_t SMALL_LEVEL_NODES  = 1u << 16;
constexpr uint64_t SMALL_LEVEL_POINTS = 1u << 22;

static __global__ __launch_bounds__(SCAN_THREADS) void kLevelSetup(uint32_t levelStart, uint32_t levelNodes,
                                                            SubsampleParams sp, LevelTables lt,
                                                            uint32_t* __restrict__ internalTotal,
                                                            uint64_t* __restrict__ segTotal)
{
    __shared__ uint32_t lds32[SCAN_ITEMS * SCAN_WAVES];
    __shared__ uint64_t lds64[SCAN_ITEMS * SCAN_WAVES];
    blockScanAll<uint32_t>(InternalIn{lt.flags, levelStart}, InternalOut{lt.internalList, levelStart}, levelNodes,
                           internalTotal, lds32);
    __syncthreads();
    const uint32_t K = *internalTotal;
    for (uint32_t k = threadIdx.x; k < K; k += SCAN_THREADS)
        segmentsOf(k, sp, lt);
    __syncthreads();
    const uint64_t segments = 8ull * K;
    blockScanAll<uint64_t>(SegmentIn{lt.segCount}, SegmentOut{lt.segTile, lt.segInput}, segments, segTotal, lds64);
    __syncthreads();
    for (uint64_t s = threadIdx.x; s < segments; s += SCAN_THREADS)
    {
        const uint32_t tiles = (lt.segCount[s] + TILE - 1) / TILE;
        const uint32_t first = lt.segTile[s];
        for (uint32_t t = 0; t < tiles; ++t)
            lt.tileList[first + t] = static_cast<uint32_t>(s);
    }
}

// Identifies the input of this thread: tile -> segment -> (node k, input position, input index q).
struct InputRef
{
    uint32_t seg, local, k, pos, q;
    bool valid;
};

__device__ __forceinline__ InputRef inputOf(uint32_t tile, const uint32_t* tileList, const uint32_t* segTile,
                                            const uint32_t* segCount, const uint32_t* segSrc,
                                            const uint32_t* segInput)
{
    InputRef r;
    r.seg   = tileList[tile];
    r.local = (tile - segTile[r.seg]) * TILE + threadIdx.x;
    r.valid = r.local < segCount[r.seg];
    r.k     = r.seg >> 3;
    r.pos   = segSrc[r.seg] + r.local;
    r.q     = segInput[r.seg] + r.local;
    return r;
}

// Evaluate, part 1 (rp_subsample_evaluation.cuh:37-203, fp_subsample_evaluation.cuh:22-110):
// the cell of every input point in its node's grid; marks the cell occupied.
template <typename T>
__device__ __forceinline__ void markTile(uint32_t tile, uint32_t nodeBase, uint64_t wordsPerNode,
                                         const uint32_t* __restrict__ tileList, const uint32_t* __restrict__ segTile,
                                         const uint32_t* __restrict__ segCount, const uint32_t* __restrict__ segSrc,
                                         const uint32_t* __restrict__ segInput, const NodeGrid* __restrict__ nodeGrid,
                                         const PointStore<T>& ps, uint32_t* __restrict__ bitmap,
                                         uint32_t* __restrict__ cellOf, uint32_t* __restrict__ err)
{
    // All lanes take part in the wave-level merge of bit sets bound for the same bitmap word.
    InputRef in   = inputOf(tile, tileList, segTile, segCount, segSrc, segInput);
    bool valid    = in.valid;
    uint32_t cell = ~0u;
    if (valid)
    {
        const NodeGrid ng = nodeGrid[in.k];
        T px, py, pz;
        ps.xyz(in.pos, px, py, pz);
        uint64_t ix, iy, iz;
        gridCoords(static_cast<double>(px), static_cast<double>(py), static_cast<double>(pz), ng.grid, ix, iy, iz);
        const uint32_t gs = ng.grid.gridSize;
        if (ix >= gs || iy >= gs || iz >= gs)
        {
            atomicOr(err, ERR_OUTSIDE);  // skipped by the later passes; the build reports the error
            cellOf[in.q] = ~0u;
            valid        = false;
        }
        else
        {
            cell = rowIndex(static_cast<uint32_t>(iy), static_cast<uint32_t>(iz), gs) * gs + static_cast<uint32_t>(ix);
            cellOf[in.q] = cell;
        }
    }
    const uint64_t word = valid ? (in.k - nodeBase) * wordsPerNode + (cell >> 5) : ~0ull - threadIdx.x;
    const WaveRun run   = waveRun(word);
    const uint32_t bits = runScan<uint32_t>(valid ? 1u << (cell & 31) : 0u, run,
                                            [](uint32_t x, uint32_t y) { return x | y; });
    if (valid && run.lane == run.end)
        atomicOr(bitmap + word, bits);
}

template <typename T>
static __global__ __launch_bounds__(TILE) void kMark(uint32_t tileBase, uint32_t tiles, uint32_t nodeBase,
                                               uint64_t wordsPerNode,
                                               const uint32_t* __restrict__ tileList,
                                               const uint32_t* __restrict__ segTile,
                                               const uint32_t* __restrict__ segCount,
                                               const uint32_t* __restrict__ segSrc,
                                               const uint32_t* __restrict__ segInput,
                                               const NodeGrid* __restrict__ nodeGrid, PointStore<T> ps,
                                               uint32_t* __restrict__ bitmap, uint32_t* __restrict__ cellOf,
                                               uint32_t* __restrict__ err)
{
    // consecutive tiles share bitmap words: runs of them per XCD
    const uint32_t t = xcdRunItem(blockIdx.x, XCD_TILE_RUN);
    if (t >= tiles)
        return;
    markTile<T>(tileBase + t, nodeBase, wordsPerNode, tileList, segTile, segCount, segSrc, segInput,
                nodeGrid, ps, bitmap, cellOf, err);
}

struct PopcountIn
{
    const uint32_t* bitmap;
    __device__ uint32_t operator()(uint64_t w) const { return __popc(bitmap[w]); }
};

// dense index of a set cell: exclusive popcount prefix of its word + set bits below it
__device__ __forceinline__ uint32_t cellRank(const uint32_t* bitmap, const uint32_t* prefix, uint64_t word,
                                             uint32_t bit)
{
    return prefix[word] + __popc(bitmap[word] & ((1u << bit) - 1u));
}

// Per node of the batch: subsample size (occupied cells), staging position, first dense index.
__device__ __forceinline__ void batchNodeOf(uint64_t kl, uint32_t nodeBase, uint32_t nodes, uint64_t wordsPerNode,
                                            const NodeGrid* __restrict__ nodeGrid,
                                            const uint32_t* __restrict__ prefix, uint32_t cursor,
                                            uint32_t* __restrict__ pointCount, uint32_t* __restrict__ stage,
                                            uint32_t* __restrict__ nodeFirst)
{
    uint32_t b    = prefix[kl * wordsPerNode];
    nodeFirst[kl] = b;
    if (kl == nodes)
        return;
    uint32_t e     = prefix[(kl + 1) * wordsPerNode];
    uint32_t id    = nodeGrid[nodeBase + kl].id;
    pointCount[id] = e - b;
    stage[id]      = cursor + b;
}

static __global__ __launch_bounds__(BLOCK) void kBatchNodes(uint32_t nodeBase, uint32_t nodes, uint64_t wordsPerNode,
                                                      const NodeGrid* __restrict__ nodeGrid,
                                                      const uint32_t* __restrict__ prefix, uint32_t cursor,
                                                      uint32_t* __restrict__ pointCount, uint32_t* __restrict__ stage,
                                                      uint32_t* __restrict__ nodeFirst)
{
    for (uint64_t kl = globalThread(); kl <= nodes; kl += globalStride())
        batchNodeOf(kl, nodeBase, nodes, wordsPerNode, nodeGrid, prefix, cursor, pointCount, stage, nodeFirst);
}

static __global__ __launch_bounds__(BLOCK) void kClearCompact(uint64_t n, uint64_t* __restrict__ win,
                                                        uint64_t* __restrict__ color, uint32_t* __restrict__ cellCnt)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        if (win)
            win[i] = ~0ull;
        if (color)
            color[i] = 0;
        if (cellCnt)
            cellCnt[i] = 0;
    }
}

// Output of the popcount scan for small levels (one launch less than a separate clear; on large
// levels the per-word loops cost more than kClearCompact): the word's prefix, and the per-cell
// state of the word's occupied
// cells (dense indices prefix .. prefix + popcount - 1) reset for the accumulation: no winner,
// no colour sum, no points. Arrays that the colour mode does not use are null.
struct PrefixClearOut
{
    uint32_t* prefix;
    uint64_t* win;
    uint64_t* color;
    uint32_t* cellCnt;
    __device__ void operator()(uint64_t w, uint32_t p, uint32_t cells) const
    {
        prefix[w] = p;
        for (uint32_t k = 0; k < cells; ++k)
        {
            if (win)
                win[p + k] = ~0ull;
            if (color)
                color[p + k] = 0;
            if (cellCnt)
                cellCnt[p + k] = 0;
        }
    }
};

// Evaluate, part 2: winner (A7/A8: atomicMin of rank << 32 | input position, rank = point index
// for the implicit method or the keyed bijective hash of it for the explicit one), the packed
// colour sum of the input records (intra-cell, and the per-cell sums the inter-cell filter is
// built from), and per-cell point counts + each point's rank in its cell (weighted filter).
template <typename T, uint32_t MODE, bool RANDOM>
__device__ __forceinline__ void accumulateTile(uint32_t tile, uint32_t nodeBase, uint64_t wordsPerNode,
                                               const uint32_t* __restrict__ tileList,
                                               const uint32_t* __restrict__ segTile,
                                               const uint32_t* __restrict__ segCount,
                                               const uint32_t* __restrict__ segSrc,
                                               const uint32_t* __restrict__ segInput,
                                               const NodeGrid* __restrict__ nodeGrid,
                                               const uint32_t* __restrict__ bitmap,
                                               const uint32_t* __restrict__ prefix, uint32_t* __restrict__ cellOf,
                                               const PointStore<T>& ps, uint64_t* __restrict__ win,
                                               uint64_t* __restrict__ color, uint32_t* __restrict__ cellCnt,
                                               uint32_t* __restrict__ cellKey, uint32_t* __restrict__ cellNode,
                                               uint32_t* __restrict__ rankOf)
{
    // All lanes take part: runs of lanes in the same cell make one atomic per run.
    InputRef in         = inputOf(tile, tileList, segTile, segCount, segSrc, segInput);
    const uint32_t cell = in.valid ? cellOf[in.q] : ~0u;
    const bool valid    = cell != ~0u;  // an invalid cell: the point lay outside its node's grid
    uint32_t c          = 0;
    if (valid)
    {
        const uint64_t word = (in.k - nodeBase) * wordsPerNode + (cell >> 5);
        c                   = cellRank(bitmap, prefix, word, cell & 31);
    }
    const WaveRun run = waveRun(valid ? c : ~0ull - threadIdx.x);
    if (valid && run.lane == run.start)
    {
        cellKey[c]  = cell;
        cellNode[c] = in.k - nodeBase;
    }
    if (MODE == COLOR_INTER_WEIGHTED)
    {
        // the winner is taken from the cell's payload run in the emit; the input's cell index is
        // replaced by the cell's dense index (the payload scatter reads the cell back from cellKey)
        uint32_t base = 0;
        if (valid && run.lane == run.end)
            base = atomicAdd(cellCnt + c, static_cast<uint32_t>(run.end - run.start + 1));
        base = __shfl(base, run.end, 64);
        if (valid)
        {
            cellOf[in.q] = c;
            rankOf[in.q] = base + static_cast<uint32_t>(run.lane - run.start);
        }
        return;
    }
    uint64_t key = ~0ull, col = 0;
    if (valid)
    {
        const uint32_t idx  = ps.index(in.pos);
        const uint32_t rank = RANDOM ? fmix32(idx ^ nodeGrid[in.k].salt) : idx;
        key                 = (static_cast<uint64_t>(rank) << 32) | in.pos;
        if (MODE == COLOR_INTRA || MODE == COLOR_INTER)
        {
            const OutRecord r = ps.record(in.pos);
            col               = encodeColors(r.r, r.g, r.b);
        }
    }
    key = runScan<uint64_t>(key, run, [](uint64_t x, uint64_t y) { return x < y ? x : y; });
    if (MODE == COLOR_INTRA || MODE == COLOR_INTER)
        col = runScan<uint64_t>(col, run, [](uint64_t x, uint64_t y) { return x + y; });
    if (valid && run.lane == run.end)
    {
        atomicMin(reinterpret_cast<unsigned long long*>(win + c), static_cast<unsigned long long>(key));
        if (MODE == COLOR_INTRA || MODE == COLOR_INTER)
            atomicAdd(reinterpret_cast<unsigned long long*>(color + c), static_cast<unsigned long long>(col));
    }
}

template <typename T, uint32_t MODE, bool RANDOM>
static __global__ __launch_bounds__(TILE) void kAccumulate(uint32_t tileBase, uint32_t nodeBase, uint64_t wordsPerNode,
                                                     const uint32_t* __restrict__ tileList,
                                                     const uint32_t* __restrict__ segTile,
                                                     const uint32_t* __restrict__ segCount,
                                                     const uint32_t* __restrict__ segSrc,
                                                     const uint32_t* __restrict__ segInput,
                                                     const NodeGrid* __restrict__ nodeGrid,
                                                     const uint32_t* __restrict__ bitmap,
                                                     const uint32_t* __restrict__ prefix,
                                                     uint32_t* __restrict__ cellOf, PointStore<T> ps,
                                                     uint64_t* __restrict__ win,
                                                     uint64_t* __restrict__ color, uint32_t* __restrict__ cellCnt,
                                                     uint32_t* __restrict__ cellKey, uint32_t* __restrict__ cellNode,
                                                     uint32_t* __restrict__ rankOf)
{
    accumulateTile<T, MODE, RANDOM>(tileBase + blockIdx.x, nodeBase, wordsPerNode, tileList, segTile, segCount,
                                    segSrc, segInput, nodeGrid, bitmap, prefix, cellOf, ps, win, color, cellCnt,
                                    cellKey, cellNode, rankOf);
}

// Per-point data of the weighted filter, in cell order so that a cell's points are one contiguous
// run (cellStart[cell] + arrival rank), in two arrays:
//   WeightPayload: the point's offset from its own cell's centre, (p - centre) rounded to float,
//     and its colour r | g << 8 | b << 16 with the cell's x coordinate mod 4 in bits 24-25 (a
//     neighbour's x differs from the selecting cell's by -1, 0 or 1, which x mod 4 identifies).
//     16 bytes: the select's loop reads one record per lane from lines that differ between lanes,
//     and every load instruction costs the memory pipeline a pass per line touched.
//   winner key: rank << 32 | source position, the rank being the point index (implicit method) or
//     its hash (explicit). Ranks are bijective, so the minimum key of a cell names its winner.
struct alignas(16) WeightPayload
{
    float ox, oy, oz;
    uint32_t rgbx;
};
static_assert(sizeof(WeightPayload) == 16, "weighted payload is one 16-byte load");

// Centre of cell (ix, iy, iz) of a node grid, in the select's expression order (min + i t + t / 2).
__device__ __forceinline__ void cellCentre(const Gridding& g, uint32_t ix, uint32_t iy, uint32_t iz, double& x,
                                           double& y, double& z)
{
    const double t = g.bbSize / g.gridSize;
    x              = g.minX + ix * t + (t * 0.5);
    y              = g.minY + iy * t + (t * 0.5);
    z              = g.minZ + iz * t + (t * 0.5);
}

template <typename T, bool RANDOM>
__device__ __forceinline__ void scatterPayloadTile(uint32_t tile, const uint32_t* __restrict__ tileList,
                                                   const uint32_t* __restrict__ segTile,
                                                   const uint32_t* __restrict__ segCount,
                                                   const uint32_t* __restrict__ segSrc,
                                                   const uint32_t* __restrict__ segInput,
                                                   const NodeGrid* __restrict__ nodeGrid,
                                                   const uint32_t* __restrict__ cellKey,
                                                   const uint32_t* __restrict__ rankOf,
                                                   const uint32_t* __restrict__ cellStart,
                                                   const uint32_t* __restrict__ cellOf, const PointStore<T>& ps,
                                                   WeightPayload* __restrict__ payload, uint64_t* __restrict__ winKey)
{
    InputRef in = inputOf(tile, tileList, segTile, segCount, segSrc, segInput);
    if (!in.valid)
        return;
    const uint32_t c = cellOf[in.q];  // the dense index kAccumulate left there
    if (c == ~0u)
        return;
    const uint32_t cell = cellKey[c];
    const NodeGrid ng   = nodeGrid[in.k];
    const uint32_t gs   = ng.grid.gridSize;
    uint32_t iy, iz;
    rowCoords(cell / gs, gs, iy, iz);
    const uint32_t ix = cell % gs;
    double cx, cy, cz;
    cellCentre(ng.grid, ix, iy, iz, cx, cy, cz);
    T px, py, pz;
    ps.xyz(in.pos, px, py, pz);
    WeightPayload pl;
    pl.ox   = static_cast<float>(static_cast<double>(px) - cx);
    pl.oy   = static_cast<float>(static_cast<double>(py) - cy);
    pl.oz   = static_cast<float>(static_cast<double>(pz) - cz);
    pl.rgbx = ps.rgb(in.pos) | ((ix & 3u) << 24);
    const uint32_t idx  = ps.index(in.pos);
    const uint32_t rank = RANDOM ? fmix32(idx ^ ng.salt) : idx;
    const uint32_t slot = cellStart[c] + rankOf[in.q];
    payload[slot]       = pl;
    winKey[slot]        = (static_cast<uint64_t>(rank) << 32) | in.pos;
}

template <typename T, bool RANDOM>
static __global__ __launch_bounds__(TILE) void kScatterPayload(uint32_t tileBase, uint32_t tiles,
                                                         const uint32_t* __restrict__ tileList,
                                                         const uint32_t* __restrict__ segTile,
                                                         const uint32_t* __restrict__ segCount,
                                                         const uint32_t* __restrict__ segSrc,
                                                         const uint32_t* __restrict__ segInput,
                                                         const NodeGrid* __restrict__ nodeGrid,
                                                         const uint32_t* __restrict__ cellKey,
                                                         const uint32_t* __restrict__ rankOf,
                                                         const uint32_t* __restrict__ cellStart,
                                                         const uint32_t* __restrict__ cellOf, PointStore<T> ps,
                                                         WeightPayload* __restrict__ payload,
                                                         uint64_t* __restrict__ winKey)
{
    // consecutive tiles store to neighbouring payload slots: runs of them per XCD
    const uint32_t t = xcdRunItem(blockIdx.x, XCD_TILE_RUN);
    if (t >= tiles)
        return;
    scatterPayloadTile<T, RANDOM>(tileBase + t, tileList, segTile, segCount, segSrc, segInput, nodeGrid,
                                  cellKey, rankOf, cellStart, cellOf, ps, payload, winKey);
}

// Select (rp_subsampling.cuh:70-321, fp_subsampling.cuh:30-306): one output record per occupied
// cell, written to staging at its dense index (node by node, cells ascending). The record keeps the
// winner's global quantised position; the colour is the winner's (no filter), the decoded intra
// sum, or the decoded sum of the per-cell sums over the 3x3x3 neighbourhood (inter; equal to the
// reference's per-point scatter into occupied neighbours because integer sums commute). The
// neighbourhood is walked as 9 x-rows: the occupied cells among x-1..x+1 of one row have
// consecutive dense indices, so each row is one contiguous run of per-cell sums.
template <typename T, uint32_t MODE>
__device__ __forceinline__ void emitCell(uint32_t c, uint32_t nodeBase, uint64_t wordsPerNode, uint32_t cursor,
                                         const NodeGrid* __restrict__ nodeGrid, const uint32_t* __restrict__ bitmap,
                                         const uint32_t* __restrict__ prefix, const uint32_t* __restrict__ cellKey,
                                         const uint32_t* __restrict__ cellNode, const uint64_t* __restrict__ win,
                                         const uint64_t* __restrict__ color, const PointStore<T>& ps)
{
    {
        const uint32_t kl = cellNode[c];
        const NodeGrid ng = nodeGrid[nodeBase + kl];
        OutRecord rec{};

        if (MODE == COLOR_INTRA)
        {
            decodeColors(color[c], rec.r, rec.g, rec.b);
        }
        else if (MODE == COLOR_INTER)
        {
            const uint32_t key    = cellKey[c];
            const uint32_t gs     = ng.grid.gridSize;
            const uint64_t words  = static_cast<uint64_t>(kl) * wordsPerNode;
            uint32_t ry, rz;
            rowCoords(key / gs, gs, ry, rz);
            const int32_t cx = static_cast<int32_t>(key % gs), cy = static_cast<int32_t>(ry),
                          cz = static_cast<int32_t>(rz);
            const int32_t x0 = cx > 0 ? cx - 1 : 0;
            const int32_t x1 = cx + 1 < static_cast<int32_t>(gs) ? cx + 1 : cx;
            const uint32_t span = static_cast<uint32_t>(x1 - x0 + 1);
            uint64_t sum        = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r)
            {
                const int32_t ny = cy + (r / 3) - 1, nz = cz + (r % 3) - 1;
                if (ny < 0 || nz < 0 || ny >= static_cast<int32_t>(gs) || nz >= static_cast<int32_t>(gs))
                    continue;
                const uint64_t L0 = static_cast<uint64_t>(x0) +
                                    static_cast<uint64_t>(rowIndex(static_cast<uint32_t>(ny), static_cast<uint32_t>(nz), gs)) * gs;
                const uint64_t w0 = words + (L0 >> 5);
                const uint32_t b0 = static_cast<uint32_t>(L0 & 31);
                uint64_t v        = bitmap[w0];
                if (b0 + span > 32)
                    v |= static_cast<uint64_t>(bitmap[w0 + 1]) << 32;
                const uint32_t m     = static_cast<uint32_t>(v >> b0) & ((1u << span) - 1u);
                const uint32_t first = prefix[w0] + __popc(static_cast<uint32_t>(v) & ((1u << b0) - 1u));
                // at most 3 occupied cells (x - 1 .. x + 1): predicated loads, no loop, so the
                // rows' loads are all in flight together
                const uint32_t n = __popc(m);
                if (n > 0)
                    sum += color[first];
                if (n > 1)
                    sum += color[first + 1];
                if (n > 2)
                    sum += color[first + 2];
            }
            decodeColors(sum, rec.r, rec.g, rec.b);
        }
        const uint32_t wpos = static_cast<uint32_t>(win[c] & 0xFFFFFFFFull);
        OutRecord wrec      = ps.record(wpos);
        if (MODE == COLOR_NONE)
        {
            rec = wrec;
        }
        else
        {
            rec.x = wrec.x;
            rec.y = wrec.y;
            rec.z = wrec.z;
        }
        const uint32_t st = cursor - ps.N + c;  // staging index
        storeRecord(ps.stOut + st, rec);
        ps.stLut[st] = ps.index(wpos);
        T* d         = ps.stPts + 3 * static_cast<uint64_t>(st);
        ps.xyz(wpos, d[0], d[1], d[2]);
    }
}

template <typename T, uint32_t MODE>
static __global__ __launch_bounds__(BLOCK) void kEmit(uint32_t cells, uint32_t nodeBase, uint64_t wordsPerNode,
                                                uint32_t cursor, const NodeGrid* __restrict__ nodeGrid,
                                                const uint32_t* __restrict__ bitmap,
                                                const uint32_t* __restrict__ prefix,
                                                const uint32_t* __restrict__ cellKey,
                                                const uint32_t* __restrict__ cellNode,
                                                const uint64_t* __restrict__ win, const uint64_t* __restrict__ color,
                                                PointStore<T> ps)
{
    for (uint64_t ci = globalThread(); ci < cells; ci += globalStride())
        emitCell<T, MODE>(static_cast<uint32_t>(ci), nodeBase, wordsPerNode, cursor, nodeGrid, bitmap, prefix, cellKey,
                          cellNode, win, color, ps);
}

// Select with the distance-weighted inter-cell filter (kernel_inter_cell_filtering.cuh:109-194,
// weighted decode rp_subsampling.cuh:222-230): the gathered average over the points of the 3x3x3
// neighbourhood, with the weight of a point p for this cell c
//   res = float(sqrt(|centre(c) - p|^2) / diag);  w = expf(-(res * res) * 10)
// and colour u16(sum c w / sum w). One thread per cell; the winner is the minimum key of the
// cell's own run, its exact coordinates and index read back through the key's source position.
// The neighbourhood is walked as 9 x-rows, each one contiguous run of payloads. With the point in
// neighbour cell m = c + d (d in {-1, 0, 1}^3) and its offset o from m's centre,
//   centre(c) - p = -d t - o,
// evaluated in float, and w = 2^(|centre(c) - p|^2 * (-10 log2(e) / diag^2)), which is the same
// function as the reference's expression. The float evaluation moves a weight by a few parts in
// 10^6, so a channel can differ by one where the average lies that close to an integer, the
// same tolerance the float sums' order dependence already needs (the reference adds with
// atomics). Two records per step, both loads in flight before either is weighed.
template <typename T>
__device__ __forceinline__ void emitWeightedCell(uint32_t c, uint32_t nodeBase, uint64_t wordsPerNode,
                                                 uint32_t cursor, const NodeGrid* __restrict__ nodeGrid,
                                                 const uint32_t* __restrict__ bitmap,
                                                 const uint32_t* __restrict__ prefix,
                                                 const uint32_t* __restrict__ cellKey,
                                                 const uint32_t* __restrict__ cellNode,
                                                 const uint32_t* __restrict__ cellStart,
                                                 const WeightPayload* __restrict__ payload,
                                                 const uint64_t* __restrict__ winKey, const PointStore<T>& ps)
{
    const uint32_t kl    = cellNode[c];
    const NodeGrid ng    = nodeGrid[nodeBase + kl];
    const uint32_t key   = cellKey[c];
    const uint32_t gs    = ng.grid.gridSize;
    const uint64_t words = static_cast<uint64_t>(kl) * wordsPerNode;
    uint32_t ry, rz;
    rowCoords(key / gs, gs, ry, rz);
    const int32_t cx = static_cast<int32_t>(key % gs), cy = static_cast<int32_t>(ry), cz = static_cast<int32_t>(rz);
    const int32_t x0 = cx > 0 ? cx - 1 : 0;
    const int32_t x1 = cx + 1 < static_cast<int32_t>(gs) ? cx + 1 : cx;
    const uint32_t span = static_cast<uint32_t>(x1 - x0 + 1);

    // first dense index and number of occupied cells of every row
    uint32_t first[9], count[9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
    {
        const int32_t ny = cy + (r / 3) - 1, nz = cz + (r % 3) - 1;
        count[r] = 0;
        first[r] = 0;
        if (ny < 0 || nz < 0 || ny >= static_cast<int32_t>(gs) || nz >= static_cast<int32_t>(gs))
            continue;
        const uint64_t L0 = static_cast<uint64_t>(x0) +
                            static_cast<uint64_t>(rowIndex(static_cast<uint32_t>(ny), static_cast<uint32_t>(nz), gs)) * gs;
        const uint64_t w0 = words + (L0 >> 5);
        const uint32_t b0 = static_cast<uint32_t>(L0 & 31);
        uint64_t v        = bitmap[w0];
        if (b0 + span > 32)
            v |= static_cast<uint64_t>(bitmap[w0 + 1]) << 32;
        const uint32_t m = static_cast<uint32_t>(v >> b0) & ((1u << span) - 1u);
        count[r]         = __popc(m);
        first[r]         = prefix[w0] + __popc(static_cast<uint32_t>(v) & ((1u << b0) - 1u));
    }
    uint32_t lo[9], hi[9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
    {
        lo[r] = count[r] ? cellStart[first[r]] : 0;
        hi[r] = count[r] ? cellStart[first[r] + count[r]] : 0;
    }

    // the winner: minimum key of the cell's own run
    uint64_t best = ~0ull;
    for (uint32_t q = cellStart[c], e = cellStart[c + 1]; q < e; ++q)
    {
        const uint64_t k = winKey[q];
        best             = k < best ? k : best;
    }
    const uint32_t wpos = static_cast<uint32_t>(best);
    T bx, by, bz;
    ps.xyz(wpos, bx, by, bz);
    const uint32_t bestIdx = ps.index(wpos);

    const double t     = ng.grid.bbSize / gs;
    const float tf     = static_cast<float>(t);
    const float kexp   = static_cast<float>((-10.0 * 1.4426950408889634) / (ng.diag * ng.diag));
    const uint32_t cxm = (static_cast<uint32_t>(cx) + 3u) & 3u;  // (c.x - 1) mod 4: d.x = ((m.x - cxm) & 3) - 1
    float acc[4]       = {0.f, 0.f, 0.f, 0.f};
    // payloads as 16-byte vectors (they stay in registers under conditional loads)
    const uint4* pv = reinterpret_cast<const uint4*>(payload);
    auto weigh      = [&](const uint4& pl, float uyc, float uzc) {
        const int32_t dx = static_cast<int32_t>(((pl.w >> 24) + 4u - cxm) & 3u) - 1;
        const float ux   = __builtin_fmaf(-static_cast<float>(dx), tf, -__uint_as_float(pl.x));
        const float uy   = uyc - __uint_as_float(pl.y);
        const float uz   = uzc - __uint_as_float(pl.z);
        float d2         = ux * ux;
        d2               = __builtin_fmaf(uy, uy, d2);
        d2               = __builtin_fmaf(uz, uz, d2);
        const float w    = __builtin_amdgcn_exp2f(d2 * kexp);
        acc[0]           = __builtin_fmaf(static_cast<float>(pl.w & 0xFF), w, acc[0]);
        acc[1]           = __builtin_fmaf(static_cast<float>((pl.w >> 8) & 0xFF), w, acc[1]);
        acc[2]           = __builtin_fmaf(static_cast<float>((pl.w >> 16) & 0xFF), w, acc[2]);
        acc[3] += w;
    };
    // The first two payloads of every row are loaded one row ahead: a row's loop is short (a few
    // payloads), so without this each row would wait out a full memory latency of its own.
    uint4 a0{}, a1{};
    bool v0 = lo[0] < hi[0], v1 = lo[0] + 1 < hi[0];
    if (v0)
        a0 = pv[lo[0]];
    if (v1)
        a1 = pv[lo[0] + 1];
#pragma unroll
    for (int r = 0; r < 9; ++r)
    {
        uint4 b0{}, b1{};
        bool u0 = false, u1 = false;
        if (r + 1 < 9)
        {
            u0 = lo[r + 1] < hi[r + 1];
            u1 = lo[r + 1] + 1 < hi[r + 1];
            if (u0)
                b0 = pv[lo[r + 1]];
            if (u1)
                b1 = pv[lo[r + 1] + 1];
        }
        const float uyc = -static_cast<float>((r / 3) - 1) * tf;  // -d.y t
        const float uzc = -static_cast<float>((r % 3) - 1) * tf;  // -d.z t
        if (v0)
            weigh(a0, uyc, uzc);
        if (v1)
            weigh(a1, uyc, uzc);
        uint32_t q = lo[r] + 2;
        for (; q + 1 < hi[r]; q += 2)
        {
            const uint4 p0 = pv[q];
            const uint4 p1 = pv[q + 1];
            weigh(p0, uyc, uzc);
            weigh(p1, uyc, uzc);
        }
        if (q < hi[r])
            weigh(pv[q], uyc, uzc);
        a0 = b0;
        a1 = b1;
        v0 = u0;
        v1 = u1;
    }
    // the colour (averages of 8-bit channels: at most 255); the record itself is quantised from
    // the staged coordinates by the final placement (kPermute)
    const uint32_t cr = static_cast<uint16_t>(acc[0] / acc[3]);
    const uint32_t cg = static_cast<uint16_t>(acc[1] / acc[3]);
    const uint32_t cb = static_cast<uint16_t>(acc[2] / acc[3]);
    const uint32_t st = cursor - ps.N + c;  // staging index
    ps.stLut[st] = bestIdx;
    ps.stRgb[st] = (cr & 0xFF) | ((cg & 0xFF) << 8) | ((cb & 0xFF) << 16);
    T* d         = ps.stPts + 3 * static_cast<uint64_t>(st);
    d[0]         = bx;
    d[1]         = by;
    d[2]         = bz;
}

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kEmitWeighted(uint32_t cells, uint32_t nodeBase, uint64_t wordsPerNode,
                                                        uint32_t cursor, const NodeGrid* __restrict__ nodeGrid,
                                                        const uint32_t* __restrict__ bitmap,
                                                        const uint32_t* __restrict__ prefix,
                                                        const uint32_t* __restrict__ cellKey,
                                                        const uint32_t* __restrict__ cellNode,
                                                        const uint32_t* __restrict__ cellStart,
                                                        const WeightPayload* __restrict__ payload,
                                                        const uint64_t* __restrict__ winKey, PointStore<T> ps)
{
    // each XCD takes runs of EMIT_CHUNK consecutive cell blocks
    const uint32_t c = xcdRunItem(blockIdx.x, EMIT_CHUNK) * BLOCK + threadIdx.x;
    if (c >= cells)
        return;
    emitWeightedCell<T>(c, nodeBase, wordsPerNode, cursor, nodeGrid, bitmap, prefix, cellKey, cellNode, cellStart,
                        payload, winKey, ps);
}

// DFS post-order placement (A6, kernel_helpers.cuh:55-59): subtree sums of subsample sizes
// bottom-up, then region starts top-down; an internal node's data follows its children's
// subtrees: dataIdx = N + regionStart + sum of its children's subtree sizes. kSubtree / kRegion
// take one level per launch; kPlaceLevels does all levels in one workgroup for small trees (a
// small build is bound by its launch count, not by these few thousand nodes).
__device__ __forceinline__ void subtreeOf(uint32_t id, const uint8_t* __restrict__ flags,
                                          const int32_t* __restrict__ children,
                                          const uint32_t* __restrict__ pointCount, uint32_t* __restrict__ subtree)
{
    if (!(flags[id] & NODE_INTERNAL))
        return;
    uint32_t s = pointCount[id];
    for (int k = 0; k < 8; ++k)
    {
        const int32_t c = children[8 * static_cast<uint64_t>(id) + k];
        if (c >= 0)
            s += subtree[c];
    }
    subtree[id] = s;
}

__device__ __forceinline__ void regionOf(uint32_t id, uint32_t root, uint32_t N, const uint8_t* __restrict__ flags,
                                         const int32_t* __restrict__ children, const uint32_t* __restrict__ subtree,
                                         uint32_t* __restrict__ region, uint64_t* __restrict__ dataIdx)
{
    if (!(flags[id] & NODE_INTERNAL))
        return;
    uint32_t r = (id == root) ? 0u : region[id];
    for (int k = 0; k < 8; ++k)
    {
        const int32_t c = children[8 * static_cast<uint64_t>(id) + k];
        if (c >= 0)
        {
            region[c] = r;
            r += subtree[c];
        }
    }
    dataIdx[id] = static_cast<uint64_t>(N) + r;
}

static __global__ __launch_bounds__(BLOCK) void kSubtree(uint32_t levelStart, uint32_t nodes,
                                                   const uint8_t* __restrict__ flags,
                                                   const int32_t* __restrict__ children,
                                                   const uint32_t* __restrict__ pointCount,
                                                   uint32_t* __restrict__ subtree)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
        subtreeOf(levelStart + static_cast<uint32_t>(j), flags, children, pointCount, subtree);
}

static __global__ __launch_bounds__(BLOCK) void kRegion(uint32_t levelStart, uint32_t nodes, uint32_t root, uint32_t N,
                                                  const uint8_t* __restrict__ flags,
                                                  const int32_t* __restrict__ children,
                                                  const uint32_t* __restrict__ subtree,
                                                  uint32_t* __restrict__ region, uint64_t* __restrict__ dataIdx)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
        regionOf(levelStart + static_cast<uint32_t>(j), root, N, flags, children, subtree, region, dataIdx);
}

constexpr int PLACE_THREADS = 1024;

struct PlaceParams
{
    uint32_t levelStart[MAX_LEVELS + 1];  // sparse ids of level l: [levelStart[l], levelStart[l + 1])
    uint32_t depth, m, root, N;
};

// One workgroup: subtree sizes zeroed, then summed level by level bottom-up, then regions and data
// offsets top-down, a barrier between levels (the workgroup's writes are visible to it after one).
static __global__ __launch_bounds__(PLACE_THREADS) void kPlaceLevels(PlaceParams pp, const uint8_t* __restrict__ flags,
                                                              const int32_t* __restrict__ children,
                                                              const uint32_t* __restrict__ pointCount,
                                                              uint32_t* __restrict__ subtree,
                                                              uint32_t* __restrict__ region,
                                                              uint64_t* __restrict__ dataIdx)
{
    for (uint32_t i = threadIdx.x; i < pp.m; i += PLACE_THREADS)
        subtree[i] = 0;
    __syncthreads();
    for (uint32_t l = 1; l <= pp.depth; ++l)
    {
        for (uint32_t id = pp.levelStart[l] + threadIdx.x; id < pp.levelStart[l + 1]; id += PLACE_THREADS)
            subtreeOf(id, flags, children, pointCount, subtree);
        __syncthreads();
    }
    for (uint32_t l = pp.depth; l >= 1; --l)
    {
        for (uint32_t id = pp.levelStart[l] + threadIdx.x; id < pp.levelStart[l + 1]; id += PLACE_THREADS)
            regionOf(id, pp.root, pp.N, flags, children, subtree, region, dataIdx);
        __syncthreads();
    }
}

// Placement of the subsampled records: the staged records of an internal node are one contiguous
// range, and so is its final range [dataIdx, dataIdx + pointCount), so every (node, chunk of
// PLACE_CHUNK records) work item is a contiguous copy.
constexpr uint32_t PLACE_CHUNK = 4096;

// chunks of an internal node's placement
struct PlacedChunksIn
{
    const uint8_t* flags;
    const uint32_t* pointCount;
    __device__ uint32_t operator()(uint64_t i) const
    {
        return (flags[i] & NODE_INTERNAL) ? (pointCount[i] + PLACE_CHUNK - 1) / PLACE_CHUNK : 0u;
    }
};

template <typename T>
__device__ __forceinline__ void permuteChunk(uint64_t entry, uint32_t N, const uint32_t* __restrict__ stage,
                                             const uint64_t* __restrict__ dataIdx,
                                             const uint32_t* __restrict__ pointCount,
                                             const uint32_t* __restrict__ stLut, const OutRecord* __restrict__ stOut,
                                             const T* __restrict__ stPts, const uint32_t* __restrict__ stRgb,
                                             const Gridding& cubic, double sx, double sy, double sz,
                                             uint32_t* __restrict__ lut, OutRecord* __restrict__ out)
{
    const uint32_t id    = static_cast<uint32_t>(entry);
    const uint32_t first = static_cast<uint32_t>(entry >> 32) * PLACE_CHUNK;
    const uint32_t total = pointCount[id];
    const uint32_t cnt   = total - first < PLACE_CHUNK ? total - first : PLACE_CHUNK;
    const uint64_t src   = static_cast<uint64_t>(stage[id]) - N + first;  // staging index of the chunk
    const uint64_t dst   = dataIdx[id] + first;
    for (uint32_t j = threadIdx.x; j < cnt; j += BLOCK)
    {
        const uint64_t s = src + j;
        lut[dst + j]     = stLut[s];
        if (stRgb)  // distance-weighted filter: the record from the staged coordinates and colour
        {
            const T* q = stPts + 3 * s;
            storeRecord(out + dst + j, quantizePoint(q[0], q[1], q[2], stRgb[s], cubic, sx, sy, sz));
        }
        else
            storeRecord(out + dst + j, loadRecord(stOut + s));
    }
}

template <typename T>
static __global__ __launch_bounds__(BLOCK) void kPermute(const uint64_t* __restrict__ chunkList, uint32_t N,
                                                   const uint32_t* __restrict__ stage,
                                                   const uint64_t* __restrict__ dataIdx,
                                                   const uint32_t* __restrict__ pointCount,
                                                   const uint32_t* __restrict__ stLut,
                                                   const OutRecord* __restrict__ stOut, const T* __restrict__ stPts,
                                                   const uint32_t* __restrict__ stRgb, Gridding cubic, double sx,
                                                   double sy, double sz, uint32_t* __restrict__ lut,
                                                   OutRecord* __restrict__ out)
{
    permuteChunk<T>(chunkList[blockIdx.x], N, stage, dataIdx, pointCount, stLut, stOut, stPts, stRgb, cubic, sx, sy,
                    sz, lut, out);
}

}  // namespace oc
