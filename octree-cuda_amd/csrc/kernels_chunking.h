// Kernels of the merging phase: count pyramid, canonical sparse ids, node links and flags, leaf
// data blocks and the level-0 cell -> leaf map.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"
#include "types.h"

namespace oc {

constexpr int BLOCK    = 256;

// scalar slots (u64) in itsScalars
enum ScalarSlot : int
{
    SC_ERROR       = 0,
    SC_M           = 1,
    SC_LEAF_TOTAL  = 2,
    SC_SEG_TOTAL   = 3,  // packed (tiles << 32 | inputs)
    SC_TABLE_TOTAL = 4,
    SC_SPLIT       = 5,  // chunks of the fine sort's split buckets
    SC_LEVEL_START = 8,  // u32[MAX_LEVELS + 1] packed into the following slots
    SC_COUNT       = 16
};

constexpr uint32_t ERR_OUTSIDE  = 1u;
constexpr uint32_t ERR_CAPACITY = 2u;

inline unsigned gridFor(uint64_t n, uint64_t cap = 65536)
{
    uint64_t b = (n + BLOCK - 1) / BLOCK;
    if (b == 0)
        b = 1;
    return static_cast<unsigned>(b < cap ? b : cap);
}

static __global__ __launch_bounds__(BLOCK) void kFill32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
        p[i] = v;
}

static __global__ __launch_bounds__(BLOCK) void kFill64(uint64_t* p, uint64_t n, uint64_t v)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
        p[i] = v;
}

// K2 (counts only): kernelPropagatePointCounts (point_count_propagation.cuh:34-86). Child slot k
// of parent (x, y, z) is (2x + (k>>2 & 1), 2y + (k>>1 & 1), 2z + (k & 1)).
__device__ __forceinline__ void propagateOne(uint32_t idx, uint32_t* __restrict__ count, uint32_t offset,
                                             uint32_t grid, uint32_t childOffset, uint32_t childGrid)
{
    const uint32_t cxy = childGrid * childGrid;
    uint32_t xy        = grid * grid;
    uint32_t z = idx / xy, y = (idx - z * xy) / grid, x = (idx - z * xy) % grid;
    const uint32_t* c = count + childOffset + 2 * z * cxy + 2 * y * childGrid + 2 * x;
    uint32_t sum      = c[0] + c[1] + c[childGrid] + c[childGrid + 1] + c[cxy] + c[cxy + 1] + c[cxy + childGrid] +
                   c[cxy + childGrid + 1];
    count[offset + idx] = sum;
}

static __global__ __launch_bounds__(BLOCK) void kPropagate(uint32_t* __restrict__ count, uint32_t offset, uint32_t grid,
                                                     uint32_t childOffset, uint32_t childGrid)
{
    const uint64_t n = static_cast<uint64_t>(grid) * grid * grid;
    for (uint64_t i = globalThread(); i < n; i += globalStride())
        propagateOne(static_cast<uint32_t>(i), count, offset, grid, childOffset, childGrid);
}

// Occupancy bitmap of the dense cells [from, dense) from their counts (bit d: cell d holds
// points). `from` is a multiple of 64 and every wave covers 64 consecutive cells, two whole words.
static __global__ __launch_bounds__(BLOCK) void kOccupancyBits(const uint32_t* __restrict__ count, uint64_t from,
                                                         uint64_t dense, uint32_t* __restrict__ occ)
{
    const uint64_t n = (dense - from + 63) / 64 * 64;  // whole waves: the loop is uniform per wave
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        const uint64_t d    = from + i;
        const uint64_t bits = __ballot(d < dense && count[d] != 0);
        if ((threadIdx.x & 31) == 0 && d < dense)
            occ[d >> 5] = static_cast<uint32_t>(bits >> (threadIdx.x & 32));
    }
}

// Sparse ids (A2) from the occupancy bitmap, one scan item per word: the exclusive rank of an
// occupied dense cell; its node key and dense count are written in the same pass.
struct OccupiedWordIn
{
    const uint32_t* occ;
    __device__ uint32_t operator()(uint64_t w) const { return __popc(occ[w]); }
};

struct OccupiedWordOut
{
    const uint32_t* occ;
    int32_t* denseToSparse;
    uint32_t* nodeKey;
    uint32_t* nodeCount;
    const uint32_t* count;
    __device__ void operator()(uint64_t w, uint32_t prefix, uint32_t v) const
    {
        uint32_t bits = v ? occ[w] : 0u;
        while (bits)
        {
            const uint64_t d  = 32 * w + static_cast<uint32_t>(__builtin_ctz(bits));
            bits &= bits - 1u;
            denseToSparse[d]  = static_cast<int32_t>(prefix);
            nodeKey[prefix]   = static_cast<uint32_t>(d);
            nodeCount[prefix] = count[d];
            ++prefix;
        }
    }
};

struct OccupiedIn
{
    const uint32_t* count;
    __device__ uint32_t operator()(uint64_t d) const { return count[d] ? 1u : 0u; }
};

// Sparse id = exclusive rank of the occupied dense cell (A2); the node's key and dense count are
// written in the same pass.
// denseToSparse is written for occupied cells only (a cell is occupied where its dense count is
// not zero): an array of D entries is never filled with -1.
struct OccupiedOut
{
    int32_t* denseToSparse;
    uint32_t* nodeKey;
    uint32_t* nodeCount;
    const uint32_t* count;
    __device__ void operator()(uint64_t d, uint32_t prefix, uint32_t v) const
    {
        if (v)
        {
            denseToSparse[d]  = static_cast<int32_t>(prefix);
            nodeKey[prefix]   = static_cast<uint32_t>(d);
            nodeCount[prefix] = count[d];
        }
    }
};

struct LevelOffsets
{
    uint32_t offset[MAX_LEVELS];
    uint32_t levels;
};

// First sparse id of every level: ids ascend with the dense index and a level is one dense range,
// so it is the number of node keys below the level's dense offset (a lower bound over the m keys).
__device__ __forceinline__ void kLevelStartsBody(const uint32_t* __restrict__ nodeKey, const uint32_t* __restrict__ m,
                                                 const LevelOffsets& lo, uint32_t* __restrict__ levelStart)
{
    const uint32_t l = threadIdx.x;
    if (l >= lo.levels)
        return;
    uint32_t a = 0, b = *m;
    while (a < b)
    {
        const uint32_t mid = a + (b - a) / 2;
        if (nodeKey[mid] < lo.offset[l])
            a = mid + 1;
        else
            b = mid;
    }
    levelStart[l] = a;
}

static __global__ void kLevelStarts(const uint32_t* __restrict__ nodeKey, const uint32_t* __restrict__ m,
                             LevelOffsets lo, uint32_t* __restrict__ levelStart)
{
    kLevelStartsBody(nodeKey, m, lo, levelStart);
}

struct LevelParams
{
    uint32_t offset[MAX_LEVELS];
    uint32_t grid[MAX_LEVELS];
    uint32_t depth;
    uint32_t threshold;
};

__device__ __forceinline__ uint32_t levelOfKey(uint32_t key, const LevelParams& lp)
{
    uint32_t l = 0;
    for (uint32_t k = 1; k <= lp.depth; ++k)
        if (key >= lp.offset[k])
            l = k;
    return l;
}

// Node links and flags: kernelInitLeafNodes (octree_initialization.cuh:24-52) and the structural
// part of kernelMergeHierarchical (hierarchical_merging.cuh:35-122):
//   internal  = level >= 1 && count >= threshold
//   finished  = parent is internal (the root is always finished, A13)
//   pointCount = internal ? 0 : count
// and the size of the leaf data block an internal node hands out to its leaf children.
__device__ __forceinline__ void nodeLinksOf(uint32_t id, const LevelParams& lp, const uint32_t* __restrict__ nodeKey,
                                            const uint32_t* __restrict__ nodeCount,
                                            const uint32_t* __restrict__ denseCount,
                                            const int32_t* __restrict__ denseToSparse,
                                            uint32_t* __restrict__ pointCount, uint64_t* __restrict__ dataIdx,
                                            int32_t* __restrict__ parent, int32_t* __restrict__ children,
                                            uint8_t* __restrict__ flags, uint32_t* __restrict__ blockSize,
                                            uint32_t* __restrict__ subtree, const uint32_t* __restrict__ occ = nullptr)
{
    {
        uint32_t key = nodeKey[id];
        uint32_t l   = levelOfKey(key, lp);
        uint32_t g   = lp.grid[l];
        uint32_t lin = key - lp.offset[l];
        uint32_t x = lin % g, y = (lin / g) % g, z = lin / (g * g);
        uint32_t cnt  = nodeCount[id];
        bool internal = l >= 1 && cnt >= lp.threshold;

        int32_t p = -1;
        if (l < lp.depth)
        {
            uint32_t pg = lp.grid[l + 1];  // the parent of an occupied cell is occupied
            p           = denseToSparse[lp.offset[l + 1] + (x >> 1) + (y >> 1) * pg + (z >> 1) * pg * pg];
        }
        bool finished = (p < 0) ? true : (nodeCount[p] >= lp.threshold);

        uint32_t blk = 0;
        if (l >= 1)
        {
            uint32_t cg = lp.grid[l - 1];
            for (int k = 0; k < 8; ++k)
            {
                uint32_t cx = 2 * x + ((k >> 2) & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k & 1);
                const uint32_t cd = lp.offset[l - 1] + cx + cy * cg + cz * cg * cg;
                // occupancy from the bitmap where there is one (level-0 counts of empty cells are
                // not written then); denseToSparse is set for occupied cells only
                const bool occupied = occ ? ((occ[cd >> 5] >> (cd & 31)) & 1u) != 0 : denseCount[cd] != 0;
                int32_t c           = occupied ? denseToSparse[cd] : -1;
                children[8 * static_cast<uint64_t>(id) + k] = c;
                if (internal && c >= 0)
                {
                    uint32_t cc        = nodeCount[c];
                    bool childInternal = (l - 1) >= 1 && cc >= lp.threshold;
                    if (!childInternal)
                        blk += cc;
                }
            }
        }
        else
        {
            for (int k = 0; k < 8; ++k)
                children[8 * static_cast<uint64_t>(id) + k] = -1;
        }
        pointCount[id] = internal ? 0u : cnt;
        dataIdx[id]    = 0;
        parent[id]     = p;
        flags[id]      = static_cast<uint8_t>((finished ? NODE_FINISHED : 0) | (internal ? NODE_INTERNAL : 0));
        blockSize[id]  = blk;
        subtree[id]    = 0;
    }
}

static __global__ __launch_bounds__(BLOCK) void kNodeLinks(uint32_t m, LevelParams lp, const uint32_t* __restrict__ nodeKey,
                                                     const uint32_t* __restrict__ nodeCount,
                                                     const uint32_t* __restrict__ denseCount,
                                                     const int32_t* __restrict__ denseToSparse,
                                                     uint32_t* __restrict__ pointCount, uint64_t* __restrict__ dataIdx,
                                                     int32_t* __restrict__ parent,
                                                     int32_t* __restrict__ children, uint8_t* __restrict__ flags,
                                                     uint32_t* __restrict__ blockSize,
                                                     uint32_t* __restrict__ subtree, const uint32_t* __restrict__ occ)
{
    for (uint64_t i = globalThread(); i < m; i += globalStride())
        nodeLinksOf(static_cast<uint32_t>(i), lp, nodeKey, nodeCount, denseCount, denseToSparse, pointCount, dataIdx,
                    parent, children, flags, blockSize, subtree, occ);
}

struct ArrayIn
{
    const uint32_t* v;
    __device__ uint32_t operator()(uint64_t i) const { return v[i]; }
};

struct ExclusiveToArray
{
    uint32_t* dst;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t) const { dst[i] = prefix; }
};

// Leaf data offsets (A4): an internal node's block starts at the exclusive scan of block sizes in
// sparse-id order (= parent level, parent linear index); children follow in slot order.
__device__ __forceinline__ void leafDataIdxOf(uint32_t id, uint32_t m, const uint32_t* __restrict__ blockStart,
                                              const uint8_t* __restrict__ flags,
                                              const int32_t* __restrict__ children,
                                              const uint32_t* __restrict__ pointCount,
                                              uint64_t* __restrict__ dataIdx)
{
    {
        if (!(flags[id] & NODE_INTERNAL))
        {
            if (id == m - 1)
                dataIdx[id] = 0;  // root as the only leaf (A13)
            return;
        }
        uint64_t r = blockStart[id];
        for (int k = 0; k < 8; ++k)
        {
            int32_t c = children[8 * static_cast<uint64_t>(id) + k];
            if (c >= 0 && !(flags[c] & NODE_INTERNAL))
            {
                dataIdx[c] = r;
                r += pointCount[c];
            }
        }
    }
}

static __global__ __launch_bounds__(BLOCK) void kLeafDataIdx(uint32_t m, uint32_t threshold,
                                                       const uint32_t* __restrict__ blockStart,
                                                       const uint8_t* __restrict__ flags,
                                                       const int32_t* __restrict__ children,
                                                       const uint32_t* __restrict__ pointCount,
                                                       uint64_t* __restrict__ dataIdx)
{
    for (uint64_t i = globalThread(); i < m; i += globalStride())
        leafDataIdxOf(static_cast<uint32_t>(i), m, blockStart, flags, children, pointCount, dataIdx);
}

// Level-0 cell -> its leaf (the first finished ancestor-or-self, point_distributing.cuh:56-60),
// written over the level-0 counts so the distribution needs one lookup per point.
__device__ __forceinline__ void cellLeafOf(uint32_t i, const uint32_t* __restrict__ nodeKey,
                                           const uint8_t* __restrict__ flags, const int32_t* __restrict__ parent,
                                           uint32_t* __restrict__ count)
{
    int32_t s = static_cast<int32_t>(i);
    while (!(flags[s] & NODE_FINISHED))
        s = parent[s];
    count[nodeKey[i]] = static_cast<uint32_t>(s);
}

static __global__ __launch_bounds__(BLOCK) void kCellLeaf(uint32_t level0Nodes, const uint32_t* __restrict__ nodeKey,
                                                    const uint8_t* __restrict__ flags,
                                                    const int32_t* __restrict__ parent, uint32_t* __restrict__ count)
{
    for (uint64_t i = globalThread(); i < level0Nodes; i += globalStride())
        cellLeafOf(static_cast<uint32_t>(i), nodeKey, flags, parent, count);
}

}  // namespace oc
