// Level-batched LOD octree build on one MI355X (gfx950). See builder.h for the phase overview.
//
// Data layout in HBM (per session):
//   cloud          N x S bytes, packed AoS records (15 or 27 bytes), read twice (count, distribute)
//   count          D x u32 dense count pyramid; after the merge its level-0 part maps cell -> leaf
//   denseToSparse  D x i32
//   nodes (SoA)    key, dense count, pointCount, dataIdx (u64), parent, children[8], flags, ...
//   output region  E records: lut (u32 point index), out (18-byte record), pts (3 x coordinate
//                  type, the original coordinates). [0, N) holds the leaves, [N, N + I) the
//                  subsampled points, first in level order (staging) and finally in DFS
//                  post-order.
//   level tables   per subsampling level: open-addressing cell tables, one slice per internal node
#include "builder.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "device_common.h"
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

constexpr int BLOCK    = 256;
constexpr uint32_t TILE = 256;  // inputs per workgroup in the subsampling insert

// scalar slots (u64) in itsScalars
enum ScalarSlot : int
{
    SC_ERROR       = 0,
    SC_M           = 1,
    SC_LEAF_TOTAL  = 2,
    SC_SEG_TOTAL   = 3,  // packed (tiles << 32 | inputs)
    SC_TABLE_TOTAL = 4,
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

__global__ __launch_bounds__(BLOCK) void kFill32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
        p[i] = v;
}

__global__ __launch_bounds__(BLOCK) void kFill64(uint64_t* p, uint64_t n, uint64_t v)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
        p[i] = v;
}

// K1: kernelPointCounting (point_counting.cuh:30-53) - dense level-0 histogram. Sparse ids are
// assigned afterwards by a scan (canonical order), so no global id counter is touched here.
template <typename T>
__global__ __launch_bounds__(BLOCK) void kCount(const uint8_t* __restrict__ cloud, uint32_t n, uint32_t stride,
                                                 Gridding g, uint32_t* __restrict__ count, uint32_t* __restrict__ err)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        double x, y, z;
        loadXYZ<T>(cloud + i * stride, x, y, z);
        uint64_t ix, iy, iz;
        gridCoords(x, y, z, g, ix, iy, iz);
        if (ix >= g.gridSize || iy >= g.gridSize || iz >= g.gridSize)
        {
            atomicOr(err, ERR_OUTSIDE);
            continue;
        }
        atomicAdd(count + static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize), 1u);
    }
}

// K2 (counts only): kernelPropagatePointCounts (point_count_propagation.cuh:34-86). Child slot k
// of parent (x, y, z) is (2x + (k>>2 & 1), 2y + (k>>1 & 1), 2z + (k & 1)).
__global__ __launch_bounds__(BLOCK) void kPropagate(uint32_t* __restrict__ count, uint32_t offset, uint32_t grid,
                                                     uint32_t childOffset, uint32_t childGrid)
{
    const uint64_t n  = static_cast<uint64_t>(grid) * grid * grid;
    const uint32_t cxy = childGrid * childGrid;
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        uint32_t idx = static_cast<uint32_t>(i);
        uint32_t xy  = grid * grid;
        uint32_t z = idx / xy, y = (idx - z * xy) / grid, x = (idx - z * xy) % grid;
        const uint32_t* c = count + childOffset + 2 * z * cxy + 2 * y * childGrid + 2 * x;
        uint32_t sum      = c[0] + c[1] + c[childGrid] + c[childGrid + 1] + c[cxy] + c[cxy + 1] + c[cxy + childGrid] +
                       c[cxy + childGrid + 1];
        count[offset + idx] = sum;
    }
}

struct OccupiedIn
{
    const uint32_t* count;
    __device__ uint32_t operator()(uint64_t d) const { return count[d] ? 1u : 0u; }
};

// Sparse id = exclusive rank of the occupied dense cell (A2); the first id of every level is
// recorded, and the node's key and dense count are written in the same pass.
struct OccupiedOut
{
    int32_t* denseToSparse;
    uint32_t* nodeKey;
    uint32_t* nodeCount;
    const uint32_t* count;
    uint32_t* levelStart;
    uint32_t offsets[MAX_LEVELS];
    uint32_t levels;
    __device__ void operator()(uint64_t d, uint32_t prefix, uint32_t v) const
    {
        denseToSparse[d] = v ? static_cast<int32_t>(prefix) : -1;
        if (v)
        {
            nodeKey[prefix]   = static_cast<uint32_t>(d);
            nodeCount[prefix] = count[d];
        }
        for (uint32_t l = 0; l < levels; ++l)
            if (d == offsets[l])
                levelStart[l] = prefix;
    }
};

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
__global__ __launch_bounds__(BLOCK) void kNodeLinks(uint32_t m, LevelParams lp, const uint32_t* __restrict__ nodeKey,
                                                     const uint32_t* __restrict__ nodeCount,
                                                     const int32_t* __restrict__ denseToSparse,
                                                     uint32_t* __restrict__ pointCount, uint64_t* __restrict__ dataIdx,
                                                     int32_t* __restrict__ parent,
                                                     int32_t* __restrict__ children, uint8_t* __restrict__ flags,
                                                     uint32_t* __restrict__ fill, uint32_t* __restrict__ blockSize,
                                                     uint32_t* __restrict__ subtree)
{
    for (uint64_t i = globalThread(); i < m; i += globalStride())
    {
        uint32_t id  = static_cast<uint32_t>(i);
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
            uint32_t pg = lp.grid[l + 1];
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
                int32_t c   = denseToSparse[lp.offset[l - 1] + cx + cy * cg + cz * cg * cg];
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
        fill[id]       = 0;
        blockSize[id]  = blk;
        subtree[id]    = 0;
    }
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
__global__ __launch_bounds__(BLOCK) void kLeafDataIdx(uint32_t m, uint32_t threshold,
                                                       const uint32_t* __restrict__ blockStart,
                                                       const uint8_t* __restrict__ flags,
                                                       const int32_t* __restrict__ children,
                                                       const uint32_t* __restrict__ pointCount,
                                                       uint64_t* __restrict__ dataIdx)
{
    for (uint64_t i = globalThread(); i < m; i += globalStride())
    {
        uint32_t id = static_cast<uint32_t>(i);
        if (!(flags[id] & NODE_INTERNAL))
        {
            if (id == m - 1)
                dataIdx[id] = 0;  // root as the only leaf (A13)
            continue;
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

// Level-0 cell -> its leaf (the first finished ancestor-or-self, point_distributing.cuh:56-60),
// written over the level-0 counts so the distribution needs one lookup per point.
__global__ __launch_bounds__(BLOCK) void kCellLeaf(uint32_t level0Nodes, const uint32_t* __restrict__ nodeKey,
                                                    const uint8_t* __restrict__ flags,
                                                    const int32_t* __restrict__ parent, uint32_t* __restrict__ count)
{
    for (uint64_t i = globalThread(); i < level0Nodes; i += globalStride())
    {
        int32_t s = static_cast<int32_t>(i);
        while (!(flags[s] & NODE_FINISHED))
            s = parent[s];
        count[nodeKey[i]] = static_cast<uint32_t>(s);
    }
}

// K5: kernelDistributePoints (point_distributing.cuh:31-76). Slot within the leaf by atomic
// (order unconstrained, A5); writes the point index, the quantised record and the coordinates.
template <typename T>
__global__ __launch_bounds__(BLOCK) void kDistribute(const uint8_t* __restrict__ cloud, uint32_t n, uint32_t stride,
                                                      Gridding g, const uint32_t* __restrict__ cellLeaf,
                                                      uint32_t* __restrict__ fill,
                                                      const uint64_t* __restrict__ dataIdx,
                                                      uint32_t* __restrict__ lut, OutRecord* __restrict__ out,
                                                      T* __restrict__ pts, double sx, double sy, double sz)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        const uint8_t* rec = cloud + i * stride;
        T v[3];
        __builtin_memcpy(v, rec, sizeof(v));
        double x = v[0], y = v[1], z = v[2];
        uint64_t ix, iy, iz;
        gridCoords(x, y, z, g, ix, iy, iz);
        if (ix >= g.gridSize || iy >= g.gridSize || iz >= g.gridSize)
            continue;  // reported by the counting pass
        uint32_t cell = static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize);
        uint32_t leaf = cellLeaf[cell];
        uint32_t slot = atomicAdd(fill + leaf, 1u);
        uint32_t pos  = static_cast<uint32_t>(dataIdx[leaf]) + slot;
        lut[pos]      = static_cast<uint32_t>(i);
        OutRecord o;
        o.x = static_cast<int32_t>(floor((x - g.minX) * sx));
        o.y = static_cast<int32_t>(floor((y - g.minY) * sy));
        o.z = static_cast<int32_t>(floor((z - g.minZ) * sz));
        const uint8_t* col = rec + 3 * sizeof(T);
        o.r = col[0];
        o.g = col[1];
        o.b = col[2];
        storeRecord(out + pos, o);
        __builtin_memcpy(pts + 3 * static_cast<uint64_t>(pos), v, sizeof(v));
    }
}

// ---- subsampling ----------------------------------------------------------------------------

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

// Per internal node of the level: its eight input segments (children in slot order, leaf data
// or the child's staged subsample), its table capacity, and its gridding. The node minimum is
// accumulated top-down exactly like the reference's DFS (calculateChildMinBB, tools.cu:29-61;
// rp_subsampling.cu:16-60): x = min.x; for every step down whose slot has the x bit,
// x = x + side/2 of that step. bbSize = (min.x + side) - min.x.
__global__ __launch_bounds__(BLOCK) void kSegments(uint32_t levelStart, uint32_t nodes, SubsampleParams sp,
                                                    const uint32_t* __restrict__ nodeKey,
                                                    const uint8_t* __restrict__ flags,
                                                    const int32_t* __restrict__ children,
                                                    const uint32_t* __restrict__ pointCount,
                                                    const uint64_t* __restrict__ dataIdx,
                                                    const uint32_t* __restrict__ stage,
                                                    uint32_t* __restrict__ segCount, uint32_t* __restrict__ segSrc,
                                                    uint32_t* __restrict__ nodeCap, NodeGrid* __restrict__ nodeGrid,
                                                    uint32_t* __restrict__ err)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
    {
        uint32_t id = levelStart + static_cast<uint32_t>(j);
        if (!(flags[id] & NODE_INTERNAL))
        {
            for (int k = 0; k < 8; ++k)
                segCount[8 * j + k] = 0;
            nodeCap[j] = 0;
            continue;
        }
        uint64_t inputs = 0;
        for (int k = 0; k < 8; ++k)
        {
            int32_t c     = children[8 * static_cast<uint64_t>(id) + k];
            uint32_t cnt  = 0, src = 0;
            if (c >= 0)
            {
                cnt = pointCount[c];
                src = (flags[c] & NODE_INTERNAL) ? stage[c] : static_cast<uint32_t>(dataIdx[c]);
            }
            segCount[8 * j + k] = cnt;
            segSrc[8 * j + k]   = src;
            inputs += cnt;
        }
        if (inputs > 0x7FFFFFF0ull)
        {
            atomicOr(err, ERR_CAPACITY);
            inputs = 0;
        }
        nodeCap[j] = static_cast<uint32_t>(2 * inputs + 4);

        uint32_t key = nodeKey[id];
        uint32_t lin = key - sp.levelOffset;
        uint32_t gl  = sp.levelGrid;
        uint32_t x = lin % gl, y = (lin / gl) % gl, z = lin / (gl * gl);
        double mx = sp.cubicMin[0], my = sp.cubicMin[1], mz = sp.cubicMin[2];
        for (uint32_t L = sp.depth; L > sp.level; --L)
        {
            uint32_t b = L - 1 - sp.level;
            if ((x >> b) & 1)
                mx = mx + sp.sideHalf[L];
            if ((y >> b) & 1)
                my = my + sp.sideHalf[L];
            if ((z >> b) & 1)
                mz = mz + sp.sideHalf[L];
        }
        NodeGrid ng;
        ng.grid.minX     = mx;
        ng.grid.minY     = my;
        ng.grid.minZ     = mz;
        ng.grid.bbSize   = (mx + sp.side) - mx;
        ng.grid.gridSize = sp.g;
        double t         = ng.grid.bbSize / sp.g;
        ng.diag          = sqrt((t * 3) * (t * 3) * 3) / 2.0;
        ng.salt          = nodeSalt(sp.seed, key);
        ng.id            = id;
        nodeGrid[j]      = ng;
    }
}

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

struct CapIn
{
    const uint32_t* cap;
    __device__ uint64_t operator()(uint64_t j) const { return cap[j]; }
};

struct CapOut
{
    uint64_t* base;
    __device__ void operator()(uint64_t j, uint64_t prefix, uint64_t) const { base[j] = prefix; }
};

__global__ __launch_bounds__(BLOCK) void kTileList(uint64_t segments, const uint32_t* __restrict__ segCount,
                                                    const uint32_t* __restrict__ segTile, uint32_t* __restrict__ tileList)
{
    for (uint64_t s = globalThread(); s < segments; s += globalStride())
    {
        uint32_t tiles = (segCount[s] + TILE - 1) / TILE;
        uint32_t first = segTile[s];
        for (uint32_t t = 0; t < tiles; ++t)
            tileList[first + t] = static_cast<uint32_t>(s);
    }
}

__global__ __launch_bounds__(BLOCK) void kClearTable(uint64_t n, uint32_t* __restrict__ keys,
                                                      uint64_t* __restrict__ win, uint64_t* __restrict__ color,
                                                      uint32_t* __restrict__ cellCnt, uint32_t* __restrict__ cellFill)
{
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        keys[i] = EMPTY_KEY;
        win[i]  = ~0ull;
        if (color)
            color[i] = 0;
        if (cellCnt)
        {
            cellCnt[i]  = 0;
            cellFill[i] = 0;
        }
    }
}

// Evaluate (rp_subsample_evaluation.cuh:37-203, fp_subsample_evaluation.cuh:22-110) for every
// input point of every internal node of the level: one workgroup per tile of TILE consecutive
// inputs of one segment. Winner (A7/A8): atomicMin of (rank << 32 | input position), rank = point
// index (implicit) or the keyed bijective hash of it (explicit). Colours: packed sum of the input
// record colour (intra, and the per-cell sums the inter filter is built from); the weighted filter
// counts points per cell for a gather in the emit step.
template <typename T, uint32_t MODE, bool RANDOM>
__global__ __launch_bounds__(TILE) void kInsert(const uint32_t* __restrict__ tileList,
                                                 const uint32_t* __restrict__ segTile,
                                                 const uint32_t* __restrict__ segCount,
                                                 const uint32_t* __restrict__ segSrc,
                                                 const uint32_t* __restrict__ segInput,
                                                 const NodeGrid* __restrict__ nodeGrid,
                                                 const uint64_t* __restrict__ tableBase,
                                                 const uint32_t* __restrict__ nodeCap,
                                                 const T* __restrict__ pts, const uint32_t* __restrict__ lut,
                                                 const OutRecord* __restrict__ out, uint32_t* __restrict__ keys,
                                                 uint64_t* __restrict__ win, uint64_t* __restrict__ color,
                                                 uint32_t* __restrict__ cellCnt, uint32_t* __restrict__ entryOf,
                                                 uint32_t* __restrict__ posOf, uint32_t* __restrict__ err)
{
    const uint32_t tile  = blockIdx.x;
    const uint32_t seg   = tileList[tile];
    const uint32_t local = (tile - segTile[seg]) * TILE + threadIdx.x;
    if (local >= segCount[seg])
        return;
    const uint32_t j    = seg >> 3;
    const uint32_t pos  = segSrc[seg] + local;
    const NodeGrid ng   = nodeGrid[j];
    const uint64_t base = tableBase[j];
    const uint32_t cap  = nodeCap[j];

    const T* p = pts + 3 * static_cast<uint64_t>(pos);
    uint64_t ix, iy, iz;
    gridCoords(static_cast<double>(p[0]), static_cast<double>(p[1]), static_cast<double>(p[2]), ng.grid, ix, iy, iz);
    const uint32_t gs = ng.grid.gridSize;
    if (ix >= gs || iy >= gs || iz >= gs)
    {
        atomicOr(err, ERR_OUTSIDE);
        return;
    }
    const uint32_t cell = static_cast<uint32_t>(ix + iy * gs + iz * gs * gs);
    bool inserted;
    const uint64_t e    = base + tableInsert(keys + base, cap, cell, inserted);
    const uint32_t idx  = lut[pos];
    const uint32_t rank = RANDOM ? fmix32(idx ^ ng.salt) : idx;
    atomicMin(reinterpret_cast<unsigned long long*>(win + e),
              static_cast<unsigned long long>((static_cast<uint64_t>(rank) << 32) | pos));
    if (MODE == COLOR_INTRA || MODE == COLOR_INTER)
    {
        OutRecord r = loadRecord(out + pos);
        atomicAdd(reinterpret_cast<unsigned long long*>(color + e),
                  static_cast<unsigned long long>(encodeColors(r.r, r.g, r.b)));
    }
    if (MODE == COLOR_INTER_WEIGHTED)
    {
        atomicAdd(cellCnt + e, 1u);
        const uint32_t q = segInput[seg] + local;
        entryOf[q]       = static_cast<uint32_t>(e);
        posOf[q]         = pos;
    }
}

struct OccIn
{
    const uint32_t* keys;
    __device__ uint32_t operator()(uint64_t e) const { return keys[e] != EMPTY_KEY ? 1u : 0u; }
};

// Per node: subsample size (= occupied cells) and its staging position.
__global__ __launch_bounds__(BLOCK) void kNodeSubsampleCount(uint32_t nodes, const NodeGrid* __restrict__ nodeGrid,
                                                              const uint32_t* __restrict__ nodeCap,
                                                              const uint64_t* __restrict__ tableBase,
                                                              const uint32_t* __restrict__ slot, uint64_t tableSize,
                                                              uint32_t cursor, uint32_t* __restrict__ pointCount,
                                                              uint32_t* __restrict__ stage)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
    {
        uint32_t cap = nodeCap[j];
        if (cap == 0)
            continue;
        uint64_t b    = tableBase[j];
        uint64_t e    = b + cap;
        uint32_t sb   = slot[b];
        uint32_t se   = slot[e < tableSize ? e : tableSize];
        uint32_t id   = nodeGrid[j].id;
        pointCount[id] = se - sb;
        stage[id]      = cursor + sb;
    }
}

__global__ __launch_bounds__(BLOCK) void kScatterCells(uint32_t inputs, const uint32_t* __restrict__ entryOf,
                                                        const uint32_t* __restrict__ posOf,
                                                        const uint32_t* __restrict__ cellStart,
                                                        uint32_t* __restrict__ cellFill, uint32_t* __restrict__ csr)
{
    for (uint64_t q = globalThread(); q < inputs; q += globalStride())
    {
        uint32_t e    = entryOf[q];
        uint32_t slot = atomicAdd(cellFill + e, 1u);
        csr[cellStart[e] + slot] = posOf[q];
    }
}

__device__ __forceinline__ uint32_t nodeOfEntry(const uint64_t* tableBase, uint32_t nodes, uint64_t e)
{
    // last j with tableBase[j] <= e (nodes with an empty slice share the next node's base)
    uint32_t lo = 0, hi = nodes;  // answer in [lo, hi)
    while (hi - lo > 1)
    {
        uint32_t mid = (lo + hi) >> 1;
        if (tableBase[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Select (rp_subsampling.cuh:70-321, fp_subsampling.cuh:30-306): one output record per occupied
// cell, at the node's staging position + the cell's rank among the level's occupied slots. The
// record keeps the winner's global quantised position; the colour is the winner's (no filter),
// the decoded intra sum, the decoded 3x3x3 sum of per-cell sums (inter; equal to the reference's
// per-point scatter into occupied neighbours because integer sums commute), or the gathered
// distance-weighted average (kernel_inter_cell_filtering.cuh:109-194).
template <typename T, uint32_t MODE>
__global__ __launch_bounds__(BLOCK) void kEmit(uint64_t tableSize, uint32_t nodes, uint32_t cursor, uint32_t N,
                                                const uint32_t* __restrict__ keys, const uint64_t* __restrict__ win,
                                                const uint64_t* __restrict__ color,
                                                const uint32_t* __restrict__ cellCnt,
                                                const uint32_t* __restrict__ cellStart,
                                                const uint32_t* __restrict__ csr,
                                                const uint32_t* __restrict__ slot,
                                                const uint64_t* __restrict__ tableBase,
                                                const uint32_t* __restrict__ nodeCap,
                                                const NodeGrid* __restrict__ nodeGrid, uint32_t* __restrict__ lut,
                                                OutRecord* __restrict__ out, T* __restrict__ pts,
                                                uint32_t* __restrict__ stageNode)
{
    for (uint64_t e = globalThread(); e < tableSize; e += globalStride())
    {
        const uint32_t key = keys[e];
        if (key == EMPTY_KEY)
            continue;
        const uint32_t j      = nodeOfEntry(tableBase, nodes, e);
        const uint32_t wpos   = static_cast<uint32_t>(win[e] & 0xFFFFFFFFull);
        OutRecord rec         = loadRecord(out + wpos);
        const uint32_t dst    = cursor + slot[e];
        const NodeGrid ng     = nodeGrid[j];

        if (MODE == COLOR_INTRA)
        {
            decodeColors(color[e], rec.r, rec.g, rec.b);
        }
        else if (MODE == COLOR_INTER || MODE == COLOR_INTER_WEIGHTED)
        {
            const uint32_t gs   = ng.grid.gridSize;
            const uint64_t base = tableBase[j];
            const uint32_t cap  = nodeCap[j];
            const int32_t cx = static_cast<int32_t>(key % gs), cy = static_cast<int32_t>((key / gs) % gs),
                          cz = static_cast<int32_t>(key / (gs * gs));
            uint64_t sum = 0;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            const double t = ng.grid.bbSize / gs;
            // centre of this (target) cell, as the reference computes it for a neighbour target
            const double ctrX = ng.grid.minX + cx * t + (t * 0.5);
            const double ctrY = ng.grid.minY + cy * t + (t * 0.5);
            const double ctrZ = ng.grid.minZ + cz * t + (t * 0.5);
            for (int ox = -1; ox <= 1; ++ox)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int oz = -1; oz <= 1; ++oz)
                    {
                        const int32_t nx = cx + ox, ny = cy + oy, nz = cz + oz;
                        if (nx < 0 || ny < 0 || nz < 0 || nx >= static_cast<int32_t>(gs) ||
                            ny >= static_cast<int32_t>(gs) || nz >= static_cast<int32_t>(gs))
                            continue;
                        const uint32_t ncell = static_cast<uint32_t>(nx) + static_cast<uint32_t>(ny) * gs +
                                               static_cast<uint32_t>(nz) * gs * gs;
                        const uint32_t f = tableFind(keys + base, cap, ncell);
                        if (f == EMPTY_KEY)
                            continue;
                        if (MODE == COLOR_INTER)
                        {
                            sum += color[base + f];
                        }
                        else
                        {
                            const uint32_t first = cellStart[base + f];
                            const uint32_t cnt   = cellCnt[base + f];
                            for (uint32_t q = 0; q < cnt; ++q)
                            {
                                const uint32_t ppos = csr[first + q];
                                const T* p          = pts + 3 * static_cast<uint64_t>(ppos);
                                const double uX     = ctrX - static_cast<double>(p[0]);
                                const double uY     = ctrY - static_cast<double>(p[1]);
                                const double uZ     = ctrZ - static_cast<double>(p[2]);
                                float res = static_cast<float>(sqrt(uX * uX + uY * uY + uZ * uZ) / ng.diag);
                                res       = expf(-((res * res) * 10.f));
                                OutRecord pr = loadRecord(out + ppos);
                                acc[0] += static_cast<float>(pr.r) * res;
                                acc[1] += static_cast<float>(pr.g) * res;
                                acc[2] += static_cast<float>(pr.b) * res;
                                acc[3] += res;
                            }
                        }
                    }
            if (MODE == COLOR_INTER)
            {
                decodeColors(sum, rec.r, rec.g, rec.b);
            }
            else
            {
                rec.r = static_cast<uint16_t>(acc[0] / acc[3]);
                rec.g = static_cast<uint16_t>(acc[1] / acc[3]);
                rec.b = static_cast<uint16_t>(acc[2] / acc[3]);
            }
        }
        storeRecord(out + dst, rec);
        lut[dst] = lut[wpos];
        const T* src = pts + 3 * static_cast<uint64_t>(wpos);
        T* d         = pts + 3 * static_cast<uint64_t>(dst);
        d[0]         = src[0];
        d[1]         = src[1];
        d[2]         = src[2];
        stageNode[dst - N] = ng.id;
    }
}

// DFS post-order placement (A6, kernel_helpers.cuh:55-59): subtree sums of subsample sizes
// bottom-up, then region starts top-down; an internal node's data follows its children's
// subtrees: dataIdx = N + regionStart + sum of its children's subtree sizes.
__global__ __launch_bounds__(BLOCK) void kSubtree(uint32_t levelStart, uint32_t nodes,
                                                   const uint8_t* __restrict__ flags,
                                                   const int32_t* __restrict__ children,
                                                   const uint32_t* __restrict__ pointCount,
                                                   uint32_t* __restrict__ subtree)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
    {
        uint32_t id = levelStart + static_cast<uint32_t>(j);
        if (!(flags[id] & NODE_INTERNAL))
            continue;
        uint32_t s = pointCount[id];
        for (int k = 0; k < 8; ++k)
        {
            int32_t c = children[8 * static_cast<uint64_t>(id) + k];
            if (c >= 0)
                s += subtree[c];
        }
        subtree[id] = s;
    }
}

__global__ __launch_bounds__(BLOCK) void kRegion(uint32_t levelStart, uint32_t nodes, uint32_t root, uint32_t N,
                                                  const uint8_t* __restrict__ flags,
                                                  const int32_t* __restrict__ children,
                                                  const uint32_t* __restrict__ subtree,
                                                  uint32_t* __restrict__ region, uint64_t* __restrict__ dataIdx)
{
    for (uint64_t j = globalThread(); j < nodes; j += globalStride())
    {
        uint32_t id = levelStart + static_cast<uint32_t>(j);
        if (!(flags[id] & NODE_INTERNAL))
            continue;
        uint32_t r = (id == root) ? 0u : region[id];
        for (int k = 0; k < 8; ++k)
        {
            int32_t c = children[8 * static_cast<uint64_t>(id) + k];
            if (c >= 0)
            {
                region[c] = r;
                r += subtree[c];
            }
        }
        dataIdx[id] = static_cast<uint64_t>(N) + r;
    }
}

__global__ __launch_bounds__(BLOCK) void kPermute(uint32_t staged, uint32_t N, const uint32_t* __restrict__ stageNode,
                                                   const uint32_t* __restrict__ stage,
                                                   const uint64_t* __restrict__ dataIdx,
                                                   const uint32_t* __restrict__ tmpLut,
                                                   const OutRecord* __restrict__ tmpOut, uint32_t* __restrict__ lut,
                                                   OutRecord* __restrict__ out)
{
    for (uint64_t s = globalThread(); s < staged; s += globalStride())
    {
        uint32_t id  = stageNode[s];
        uint64_t dst = dataIdx[id] + (static_cast<uint64_t>(N) + s - stage[id]);
        lut[dst]     = tmpLut[s];
        storeRecord(out + dst, loadRecord(tmpOut + s));
    }
}

// ------------------------------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------------------------------

static uint32_t octreeDepth(uint32_t grid)  // tools.cu:63-88
{
    for (uint32_t d = 1; d <= 9; ++d)
        if (grid == (1u << d))
            return d;
    return 0;
}

OctreeBuilder::OctreeBuilder(int device, const CloudInfo& cloud, const ProcessingInfo& proc,
                             const uint8_t* hostCloud, const uint8_t* deviceCloud)
    : itsDevice(device), itsCloud(cloud), itsProc(proc)
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
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
    itsColorMode = proc.useIntraCellAvg ? COLOR_INTRA
                                        : (proc.useInterCellAvg ? (proc.useWeightingFunction ? COLOR_INTER_WEIGHTED
                                                                                            : COLOR_INTER)
                                                                : COLOR_NONE);
    itsE = static_cast<uint64_t>(static_cast<float>(cloud.pointAmount) * proc.outputFactor);  // point_cloud.cuh:24

    OC_HIP_CHECK(hipStreamCreateWithFlags(&itsStream, hipStreamNonBlocking));
    for (auto& ev : itsEvents)
        OC_HIP_CHECK(hipEventCreate(&ev));
    for (auto& lev : itsLevelEvents)
        for (auto& ev : lev)
            OC_HIP_CHECK(hipEventCreate(&ev));

    const uint64_t cloudBytes = static_cast<uint64_t>(cloud.pointAmount) * cloud.dataStride;
    if (deviceCloud)
    {
        itsCloudDevice = deviceCloud;
    }
    else
    {
        itsCloudOwned.ensure(cloudBytes);
        OC_HIP_CHECK(hipMemcpy(itsCloudOwned.get(), hostCloud, cloudBytes, hipMemcpyHostToDevice));
        itsCloudDevice = itsCloudOwned.as<uint8_t>();
    }

    // Fixed-size state: the dense pyramid, the node key/count arrays (bounded by min(D, levels*N))
    // and the output region.
    itsCount.ensure(sizeof(uint32_t) * static_cast<uint64_t>(itsDense));
    itsDenseToSparse.ensure(sizeof(int32_t) * static_cast<uint64_t>(itsDense));
    uint64_t mBound = std::min<uint64_t>(itsDense, static_cast<uint64_t>(itsDepth + 1) * cloud.pointAmount);
    itsNodeKey.ensure(sizeof(uint32_t) * mBound);
    itsNodeCount.ensure(sizeof(uint32_t) * mBound);
    itsLut.ensure(sizeof(uint32_t) * itsE);
    itsOut.ensure(sizeof(OutRecord) * itsE);
    itsPts.ensure((cloud.cloudType == CLOUD_FLOAT_UINT8_T ? 12 : 24) * itsE);
    itsStageNode.ensure(sizeof(uint32_t) * std::max<uint64_t>(1, itsE - cloud.pointAmount));
    itsScalars.ensure(sizeof(uint64_t) * SC_COUNT);
    OC_HIP_CHECK(hipMemsetAsync(itsOut.get(), 0, sizeof(OutRecord) * itsE, itsStream));
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
}

OctreeBuilder::~OctreeBuilder()
{
    if (itsStream)
    {
        (void)hipStreamSynchronize(itsStream);
        (void)hipStreamDestroy(itsStream);
    }
    for (auto& ev : itsEvents)
        if (ev)
            (void)hipEventDestroy(ev);
    for (auto& lev : itsLevelEvents)
        for (auto& ev : lev)
            if (ev)
                (void)hipEventDestroy(ev);
}

void OctreeBuilder::syncStream()
{
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
}

template <typename V>
V OctreeBuilder::readDevice(const V* devicePtr)
{
    V v{};
    OC_HIP_CHECK(hipMemcpyAsync(&v, devicePtr, sizeof(V), hipMemcpyDeviceToHost, itsStream));
    OC_HIP_CHECK(hipStreamSynchronize(itsStream));
    return v;
}

void OctreeBuilder::generate()
{
    OC_HIP_CHECK(hipSetDevice(itsDevice));
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
    itsLevelsRun = 0;

    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_START], itsStream));
    countPoints<T>();
    mergeLevels();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_MERGE_END], itsStream));
    distributePoints<T>();
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
    for (uint32_t i = 0; i < itsLevelsRun; ++i)
    {
        float a = 0.f, b = 0.f;
        OC_HIP_CHECK(hipEventElapsedTime(&a, itsLevelEvents[i][0], itsLevelEvents[i][1]));
        OC_HIP_CHECK(hipEventElapsedTime(&b, itsLevelEvents[i][2], itsLevelEvents[i][3]));
        itsTimings[5] += a;
        itsTimings[6] += b;
    }
}

template <typename T>
void OctreeBuilder::countPoints()
{
    uint64_t* sc    = itsScalars.as<uint64_t>();
    uint32_t* count = itsCount.as<uint32_t>();
    kFill32<<<gridFor(itsLevelNodes[0]), BLOCK, 0, itsStream>>>(count, itsLevelNodes[0], 0);
    OC_LAUNCH_CHECK();
    Gridding g{itsCloud.bbMin[0], itsCloud.bbMin[1], itsCloud.bbMin[2], itsCloud.bbMax[0] - itsCloud.bbMin[0],
               itsProc.chunkingGrid};
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_COUNT_BEGIN], itsStream));
    kCount<T><<<gridFor(itsCloud.pointAmount, 16384), BLOCK, 0, itsStream>>>(
            itsCloudDevice, itsCloud.pointAmount, itsCloud.dataStride, g, count,
            reinterpret_cast<uint32_t*>(sc + SC_ERROR));
    OC_LAUNCH_CHECK();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_COUNT_END], itsStream));
}

void OctreeBuilder::mergeLevels()
{
    uint64_t* sc    = itsScalars.as<uint64_t>();
    uint32_t* count = itsCount.as<uint32_t>();
    for (uint32_t l = 1; l <= itsDepth; ++l)
    {
        kPropagate<<<gridFor(itsLevelNodes[l]), BLOCK, 0, itsStream>>>(count, itsLevelOffset[l], itsLevelGrid[l],
                                                                        itsLevelOffset[l - 1], itsLevelGrid[l - 1]);
        OC_LAUNCH_CHECK();
    }

    // sparse ids (A2) + node keys / counts + first id of every level
    uint32_t* levelStart = reinterpret_cast<uint32_t*>(sc + SC_LEVEL_START);
    itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(itsDense) + 1));
    OccupiedOut occOut{};
    occOut.denseToSparse = itsDenseToSparse.as<int32_t>();
    occOut.nodeKey       = itsNodeKey.as<uint32_t>();
    occOut.nodeCount     = itsNodeCount.as<uint32_t>();
    occOut.count         = count;
    occOut.levelStart    = levelStart;
    for (uint32_t l = 0; l <= itsDepth; ++l)
        occOut.offsets[l] = itsLevelOffset[l];
    occOut.levels = itsDepth + 1;
    scanExclusive<uint32_t>(itsDense, OccupiedIn{count}, occOut, itsScanScratch.as<uint32_t>(),
                            reinterpret_cast<uint32_t*>(sc + SC_M), itsStream);
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
    itsNodeFill.ensure(sizeof(uint32_t) * m);
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
            itsM, lp, itsNodeKey.as<uint32_t>(), itsNodeCount.as<uint32_t>(), itsDenseToSparse.as<int32_t>(),
            itsNodePoints.as<uint32_t>(), itsNodeData.as<uint64_t>(), itsNodeParent.as<int32_t>(),
            itsNodeChild.as<int32_t>(),
            itsNodeFlags.as<uint8_t>(), itsNodeFill.as<uint32_t>(), itsNodeAux.as<uint32_t>(),
            itsNodeSub.as<uint32_t>());
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
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_DIST_BEGIN], itsStream));
    kDistribute<T><<<gridFor(itsCloud.pointAmount, 16384), BLOCK, 0, itsStream>>>(
            itsCloudDevice, itsCloud.pointAmount, itsCloud.dataStride, g, itsCount.as<uint32_t>(),
            itsNodeFill.as<uint32_t>(), itsNodeData.as<uint64_t>(), itsLut.as<uint32_t>(), itsOut.as<OutRecord>(),
            itsPts.as<T>(), sx, sy, sz);
    OC_LAUNCH_CHECK();
    OC_HIP_CHECK(hipEventRecord(itsEvents[EV_DIST_END], itsStream));
}

template <typename T, uint32_t MODE>
static void launchInsert(bool random, unsigned tiles, hipStream_t s, const uint32_t* tileList, const uint32_t* segTile,
                         const uint32_t* segCount, const uint32_t* segSrc, const uint32_t* segInput,
                         const NodeGrid* nodeGrid, const uint64_t* tableBase, const uint32_t* nodeCap, const T* pts,
                         const uint32_t* lut, const OutRecord* out, uint32_t* keys, uint64_t* win, uint64_t* color,
                         uint32_t* cellCnt, uint32_t* entryOf, uint32_t* posOf, uint32_t* err)
{
    if (random)
        kInsert<T, MODE, true><<<tiles, TILE, 0, s>>>(tileList, segTile, segCount, segSrc, segInput, nodeGrid,
                                                       tableBase, nodeCap, pts, lut, out, keys, win, color, cellCnt,
                                                       entryOf, posOf, err);
    else
        kInsert<T, MODE, false><<<tiles, TILE, 0, s>>>(tileList, segTile, segCount, segSrc, segInput, nodeGrid,
                                                        tableBase, nodeCap, pts, lut, out, keys, win, color, cellCnt,
                                                        entryOf, posOf, err);
}

template <typename T>
void OctreeBuilder::subsampleLevels()
{
    uint64_t* sc     = itsScalars.as<uint64_t>();
    uint32_t* err    = reinterpret_cast<uint32_t*>(sc + SC_ERROR);
    const uint32_t N = itsCloud.pointAmount;
    const uint32_t g = itsProc.subsamplingGrid;
    const double cubicSize = itsCloud.bbMax[0] - itsCloud.bbMin[0];  // metadata.cuh:66-69
    uint64_t cursor  = N;
    itsInternal      = 0;

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
        const uint32_t ls    = itsLevelStart[l];
        const uint32_t nodes = itsLevelStart[l + 1] - ls;
        if (nodes == 0)
            continue;
        const uint64_t segments = 8ull * nodes;
        sp.level       = l;
        sp.levelOffset = itsLevelOffset[l];
        sp.levelGrid   = itsLevelGrid[l];
        sp.side        = cubicSize / std::pow(2.0, static_cast<double>(itsDepth - l));  // rp_subsampling.cu:17-18

        itsSegCount.ensure(sizeof(uint32_t) * segments);
        itsSegSrc.ensure(sizeof(uint32_t) * segments);
        itsSegTile.ensure(sizeof(uint32_t) * segments);
        itsSegInput.ensure(sizeof(uint32_t) * segments);
        itsNodeGrid.ensure(sizeof(NodeGrid) * nodes);
        itsNodeCap.ensure(sizeof(uint32_t) * nodes);
        itsTableBase.ensure(sizeof(uint64_t) * (nodes + 1));
        itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(segments) + 1));

        kSegments<<<gridFor(nodes), BLOCK, 0, itsStream>>>(
                ls, nodes, sp, itsNodeKey.as<uint32_t>(), itsNodeFlags.as<uint8_t>(), itsNodeChild.as<int32_t>(),
                itsNodePoints.as<uint32_t>(), itsNodeData.as<uint64_t>(), itsNodeStage.as<uint32_t>(),
                itsSegCount.as<uint32_t>(), itsSegSrc.as<uint32_t>(), itsNodeCap.as<uint32_t>(),
                itsNodeGrid.as<NodeGrid>(), err);
        OC_LAUNCH_CHECK();
        scanExclusive<uint64_t>(segments, SegmentIn{itsSegCount.as<uint32_t>()},
                                SegmentOut{itsSegTile.as<uint32_t>(), itsSegInput.as<uint32_t>()},
                                itsScanScratch.as<uint64_t>(), sc + SC_SEG_TOTAL, itsStream);
        OC_LAUNCH_CHECK();
        scanExclusive<uint64_t>(nodes, CapIn{itsNodeCap.as<uint32_t>()}, CapOut{itsTableBase.as<uint64_t>()},
                                itsScanScratch.as<uint64_t>(), itsTableBase.as<uint64_t>() + nodes, itsStream);
        OC_LAUNCH_CHECK();

        uint64_t totals[2] = {0, 0};
        OC_HIP_CHECK(hipMemcpyAsync(&totals[0], sc + SC_SEG_TOTAL, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                    itsStream));
        OC_HIP_CHECK(hipMemcpyAsync(&totals[1], itsTableBase.as<uint64_t>() + nodes, sizeof(uint64_t),
                                    hipMemcpyDeviceToHost, itsStream));
        syncStream();
        const uint32_t tiles     = static_cast<uint32_t>(totals[0] >> 32);
        const uint32_t inputs    = static_cast<uint32_t>(totals[0]);
        const uint64_t tableSize = totals[1];
        if (inputs == 0)
            continue;

        itsTileList.ensure(sizeof(uint32_t) * tiles);
        itsKeys.ensure(sizeof(uint32_t) * tableSize);
        itsWin.ensure(sizeof(uint64_t) * tableSize);
        itsSlot.ensure(sizeof(uint32_t) * (tableSize + 1));
        const bool packed   = itsColorMode == COLOR_INTRA || itsColorMode == COLOR_INTER;
        const bool weighted = itsColorMode == COLOR_INTER_WEIGHTED;
        if (packed)
            itsColor.ensure(sizeof(uint64_t) * tableSize);
        if (weighted)
        {
            itsCellCnt.ensure(sizeof(uint32_t) * tableSize);
            itsCellStart.ensure(sizeof(uint32_t) * (tableSize + 1));
            itsCellFill.ensure(sizeof(uint32_t) * tableSize);
            itsEntryOf.ensure(sizeof(uint32_t) * inputs);
            itsPosOf.ensure(sizeof(uint32_t) * inputs);
            itsCsr.ensure(sizeof(uint32_t) * inputs);
        }
        itsScanScratch.ensure(sizeof(uint64_t) * (scanTiles(tableSize) + 1));
        uint64_t* color    = packed ? itsColor.as<uint64_t>() : nullptr;
        uint32_t* cellCnt  = weighted ? itsCellCnt.as<uint32_t>() : nullptr;
        uint32_t* cellFill = weighted ? itsCellFill.as<uint32_t>() : nullptr;

        kClearTable<<<gridFor(tableSize), BLOCK, 0, itsStream>>>(tableSize, itsKeys.as<uint32_t>(),
                                                                 itsWin.as<uint64_t>(), color, cellCnt, cellFill);
        OC_LAUNCH_CHECK();
        kTileList<<<gridFor(segments), BLOCK, 0, itsStream>>>(segments, itsSegCount.as<uint32_t>(),
                                                              itsSegTile.as<uint32_t>(), itsTileList.as<uint32_t>());
        OC_LAUNCH_CHECK();

        hipEvent_t* lev = itsLevelEvents[itsLevelsRun];
        OC_HIP_CHECK(hipEventRecord(lev[0], itsStream));
        switch (itsColorMode)
        {
#define OC_INSERT(M)                                                                                                   \
    launchInsert<T, M>(itsProc.useRandomSubsampling, tiles, itsStream, itsTileList.as<uint32_t>(),                    \
                       itsSegTile.as<uint32_t>(), itsSegCount.as<uint32_t>(), itsSegSrc.as<uint32_t>(),               \
                       itsSegInput.as<uint32_t>(), itsNodeGrid.as<NodeGrid>(), itsTableBase.as<uint64_t>(),           \
                       itsNodeCap.as<uint32_t>(), itsPts.as<T>(), itsLut.as<uint32_t>(), itsOut.as<OutRecord>(),      \
                       itsKeys.as<uint32_t>(), itsWin.as<uint64_t>(), color, cellCnt,                                 \
                       weighted ? itsEntryOf.as<uint32_t>() : nullptr, weighted ? itsPosOf.as<uint32_t>() : nullptr,  \
                       err)
        case COLOR_NONE: OC_INSERT(COLOR_NONE); break;
        case COLOR_INTRA: OC_INSERT(COLOR_INTRA); break;
        case COLOR_INTER: OC_INSERT(COLOR_INTER); break;
        default: OC_INSERT(COLOR_INTER_WEIGHTED); break;
#undef OC_INSERT
        }
        OC_LAUNCH_CHECK();
        OC_HIP_CHECK(hipEventRecord(lev[1], itsStream));

        // slots: rank of each occupied entry among the level's occupied entries
        uint32_t* slot = itsSlot.as<uint32_t>();
        scanExclusive<uint32_t>(tableSize, OccIn{itsKeys.as<uint32_t>()}, ExclusiveToArray{slot},
                                itsScanScratch.as<uint32_t>(), slot + tableSize, itsStream);
        OC_LAUNCH_CHECK();
        kNodeSubsampleCount<<<gridFor(nodes), BLOCK, 0, itsStream>>>(
                nodes, itsNodeGrid.as<NodeGrid>(), itsNodeCap.as<uint32_t>(), itsTableBase.as<uint64_t>(), slot,
                tableSize, static_cast<uint32_t>(cursor), itsNodePoints.as<uint32_t>(), itsNodeStage.as<uint32_t>());
        OC_LAUNCH_CHECK();
        const uint32_t emitted = readDevice(slot + tableSize);
        if (cursor + emitted > itsE)  // A14
            throw std::runtime_error("the subsampled points exceed the estimated output size (N + I > E); "
                                     "raise the output factor");

        if (weighted)
        {
            scanExclusive<uint32_t>(tableSize, ArrayIn{cellCnt}, ExclusiveToArray{itsCellStart.as<uint32_t>()},
                                    itsScanScratch.as<uint32_t>(), itsCellStart.as<uint32_t>() + tableSize,
                                    itsStream);
            OC_LAUNCH_CHECK();
            kScatterCells<<<gridFor(inputs), BLOCK, 0, itsStream>>>(inputs, itsEntryOf.as<uint32_t>(),
                                                                    itsPosOf.as<uint32_t>(),
                                                                    itsCellStart.as<uint32_t>(), cellFill,
                                                                    itsCsr.as<uint32_t>());
            OC_LAUNCH_CHECK();
        }

        OC_HIP_CHECK(hipEventRecord(lev[2], itsStream));
        switch (itsColorMode)
        {
#define OC_EMIT(M)                                                                                                     \
    kEmit<T, M><<<gridFor(tableSize), BLOCK, 0, itsStream>>>(                                                          \
            tableSize, nodes, static_cast<uint32_t>(cursor), N, itsKeys.as<uint32_t>(), itsWin.as<uint64_t>(), color, \
            cellCnt, weighted ? itsCellStart.as<uint32_t>() : nullptr, weighted ? itsCsr.as<uint32_t>() : nullptr,    \
            slot, itsTableBase.as<uint64_t>(), itsNodeCap.as<uint32_t>(), itsNodeGrid.as<NodeGrid>(),                 \
            itsLut.as<uint32_t>(), itsOut.as<OutRecord>(), itsPts.as<T>(), itsStageNode.as<uint32_t>())
        case COLOR_NONE: OC_EMIT(COLOR_NONE); break;
        case COLOR_INTRA: OC_EMIT(COLOR_INTRA); break;
        case COLOR_INTER: OC_EMIT(COLOR_INTER); break;
        default: OC_EMIT(COLOR_INTER_WEIGHTED); break;
#undef OC_EMIT
        }
        OC_LAUNCH_CHECK();
        OC_HIP_CHECK(hipEventRecord(lev[3], itsStream));
        itsLevelsRun++;
        cursor += emitted;
    }
    itsUsed = cursor;
}

void OctreeBuilder::placeInternalNodes()
{
    const uint32_t N      = itsCloud.pointAmount;
    const uint64_t staged = itsUsed - N;
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
    if (staged > 0)
    {
        itsTmpLut.ensure(sizeof(uint32_t) * staged);
        itsTmpOut.ensure(sizeof(OutRecord) * staged);
        OC_HIP_CHECK(hipMemcpyAsync(itsTmpLut.get(), itsLut.as<uint32_t>() + N, sizeof(uint32_t) * staged,
                                    hipMemcpyDeviceToDevice, itsStream));
        OC_HIP_CHECK(hipMemcpyAsync(itsTmpOut.get(), itsOut.as<OutRecord>() + N, sizeof(OutRecord) * staged,
                                    hipMemcpyDeviceToDevice, itsStream));
        kPermute<<<gridFor(staged), BLOCK, 0, itsStream>>>(
                static_cast<uint32_t>(staged), N, itsStageNode.as<uint32_t>(), itsNodeStage.as<uint32_t>(),
                itsNodeData.as<uint64_t>(), itsTmpLut.as<uint32_t>(), itsTmpOut.as<OutRecord>(),
                itsLut.as<uint32_t>(), itsOut.as<OutRecord>());
        OC_LAUNCH_CHECK();
    }
}

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
    OC_HIP_CHECK(hipMemcpy(o.denseKey.data(), itsNodeKey.get(), 4 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.count.data(), itsNodeCount.get(), 4 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.pointCount.data(), itsNodePoints.get(), 4 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.dataIdx.data(), itsNodeData.get(), 8 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.parent.data(), itsNodeParent.get(), 4 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.children.data(), itsNodeChild.get(), 32 * m, hipMemcpyDeviceToHost));
    OC_HIP_CHECK(hipMemcpy(o.flags.data(), itsNodeFlags.get(), m, hipMemcpyDeviceToHost));
}

void OctreeBuilder::copyLut(uint32_t* dst, uint64_t count) const
{
    if (count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    OC_HIP_CHECK(hipMemcpy(dst, itsLut.get(), 4 * count, hipMemcpyDeviceToHost));
}

void OctreeBuilder::copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const
{
    if (first + count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    OC_HIP_CHECK(hipMemcpy(dst, itsOut.as<OutRecord>() + first, sizeof(OutRecord) * count, hipMemcpyDeviceToHost));
}

}  // namespace oc
