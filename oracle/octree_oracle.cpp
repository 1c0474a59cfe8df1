// CPU oracle for the LOD octree build path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP builder is compared against.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
// library (libOctreeApi.so) never links, loads or calls it.
//
// It is a multithreaded (std::thread) C++ restatement of the reference's CUDA kernels and host
// drivers, written for clarity rather than speed. It follows the reference's own structure on
// purpose (dense counting grid, per-level propagation, K4-style merge, K5-style distribution,
// post-order per-node subsampling with a dense g^3 grid), so that it is an independent
// formulation of the level-batched, hash-table based GPU design.
//
// Where the reference is racy or undefined, the canonical rules below are implemented
// (the GPU implements the same rules):
//   A2  sparse id = rank in (level ascending from the finest, linear index ascending); root = M-1
//   A3  level-0 flags are initialised explicitly (finished = internal = false, parent = -1)
//   A4  leaf data blocks ordered by (parent level, parent linear index), children in slot order
//   A5  within-leaf order is unconstrained (here: ascending point index); compare sets
//   A6  internal nodes are packed after the N leaf points in DFS post-order
//   A7  implicit winner = minimum original point index in the cell
//   A8  explicit winner = minimum of a keyed 32-bit bijective hash of the point index
//   A9  subsample slot order is unconstrained (here: ascending cell index); compare sets
//   A10 packed colour sums wrap exactly like the reference; amount == 0 decodes to colour 0
//   A12 all child points are processed (the reference's launch bound can skip some)
//   A13 the root is forced finished, so a cloud with N < threshold becomes a single leaf
//   A14 N + I > E is an error (the reference writes out of bounds)
//   A15 the tail of the output buffer beyond N + I is zero
//   A16 intra-cell + weighted, and chunking grids that are not a power of two in [2, 512],
//       are rejected
//
// Parity pins (each checked in tests/): level tables and node statistics (test_octree.cu:51-183),
// the 128^3 cuboid known answers (test_initial_point_counting_sparse.cpp, test_merging_sparse.cpp,
// test_point_distribution_sparse.cpp, test_subsampling_sparse.cpp), the output capacity
// (test_point_cloud_host.cu:46), the hierarchy record layout (tools/potree_to_ply.py:161-181), a
// pure-Python brute force on tiny clouds (tests/bruteforce.py), closed forms for the packed
// colours and invariants (sum of leaf counts = N, membership of every leaf point in its voxel).
// The distance-weighted colour path uses libm expf/sqrt; parity for it is a tolerance (+-1 per
// channel), not bit-exactness.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------------------------------------
// Configuration and small helpers
// ------------------------------------------------------------------------------------------------

struct OracleConfig
{
    uint32_t point_amount;
    uint32_t data_stride;
    uint32_t cloud_type;  // 0 = float32 xyz, 1 = float64 xyz (metadata.cuh:25-29)
    double bb_min[3];
    double bb_max[3];
    double scale[3];
    uint32_t chunking_grid;
    uint32_t merging_threshold;
    float output_factor;
    uint32_t subsampling_grid;
    uint32_t intra;
    uint32_t inter;
    uint32_t replacement;
    uint32_t random;
    uint32_t weighted;
    uint64_t seed;
    uint32_t threads;
};

enum ColorMode
{
    COLOR_NONE,
    COLOR_INTRA,
    COLOR_INTER,
    COLOR_INTER_WEIGHTED
};

#pragma pack(push, 1)
struct OutRecord  // types.cuh:18-24
{
    int32_t x, y, z;
    uint16_t r, g, b;
};
struct HierarchyEntry  // potree_exporter.cuh:28-35
{
    uint8_t type;
    uint8_t bitmask;
    uint32_t points;
    uint64_t byteOffset;
    uint64_t byteSize;
};
#pragma pack(pop)
static_assert(sizeof(OutRecord) == 18, "output record is 18 bytes");
static_assert(sizeof(HierarchyEntry) == 22, "hierarchy entry is 22 bytes");

unsigned threadCount(uint32_t requested)
{
    if (requested)
        return requested;
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? hc : 1;
}

void parallelFor(uint64_t n, unsigned threads, const std::function<void(uint64_t, uint64_t, unsigned)>& fn)
{
    if (n == 0)
        return;
    threads = static_cast<unsigned>(std::min<uint64_t>(threads, n));
    if (threads <= 1)
    {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    uint64_t chunk = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t)
    {
        uint64_t b = t * chunk, e = std::min<uint64_t>(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back(fn, b, e, t);
    }
    for (auto& th : pool)
        th.join();
}

// tools.cu:63-88: only powers of two in [2, 512] give a non-zero depth.
uint32_t octreeDepth(uint32_t grid)
{
    for (uint32_t d = 1; d <= 9; ++d)
        if (grid == (1u << d))
            return d;
    return 0;
}

// octree.cu:13-23: level l (0 = finest) has grid G >> l, (G >> l)^3 nodes, dense offset sum of
// the finer levels.
struct LevelTables
{
    uint32_t depth = 0;
    std::vector<uint32_t> grid, nodes, offset;
    uint32_t dense = 0;
};

LevelTables levelTables(uint32_t chunkingGrid)
{
    LevelTables t;
    t.depth = octreeDepth(chunkingGrid);
    for (uint32_t g = chunkingGrid; g > 0; g >>= 1)
    {
        t.grid.push_back(g);
        t.offset.push_back(t.dense);
        t.nodes.push_back(g * g * g);
        t.dense += g * g * g;
    }
    return t;
}

struct Gridding  // kernel_structs.cuh:18-24
{
    uint32_t gridSize;
    double bbSize;
    double minX, minY, minZ;
    double diag;
};

// kernel_helpers.cuh:11-25 (same expression order; the f32 coordinate is promoted to double)
inline void gridCoords(double px, double py, double pz, const Gridding& g, uint64_t& ix, uint64_t& iy, uint64_t& iz)
{
    double t  = g.bbSize / g.gridSize;
    double uX = (px - g.minX) / t;
    double uY = (py - g.minY) / t;
    double uZ = (pz - g.minZ) / t;
    t         = g.gridSize - 1.0;
    ix        = static_cast<uint64_t>(static_cast<int64_t>(std::fmin(uX, t)));
    iy        = static_cast<uint64_t>(static_cast<int64_t>(std::fmin(uY, t)));
    iz        = static_cast<uint64_t>(static_cast<int64_t>(std::fmin(uZ, t)));
}

inline uint32_t mapPointToGrid(double px, double py, double pz, const Gridding& g)
{
    uint64_t ix, iy, iz;
    gridCoords(px, py, pz, g, ix, iy, iz);
    return static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize);
}

inline void loadPoint(const uint8_t* rec, uint32_t cloudType, double& x, double& y, double& z)
{
    if (cloudType == 0)
    {
        float v[3];
        std::memcpy(v, rec, sizeof(v));
        x = v[0];
        y = v[1];
        z = v[2];
    }
    else
    {
        double v[3];
        std::memcpy(v, rec, sizeof(v));
        x = v[0];
        y = v[1];
        z = v[2];
    }
}

// kernel_helpers.cuh:49-53
inline uint64_t encodeColors(uint16_t r, uint16_t g, uint16_t b)
{
    return (static_cast<uint64_t>(r) << 46) | (static_cast<uint64_t>(g) << 28) | static_cast<uint64_t>(b) << 10 |
           static_cast<uint64_t>(1);
}

// rp_subsampling.cuh:123-129 / fp_subsampling.cuh:95-101; amount == 0 -> 0 (A10)
inline void decodeColors(uint64_t encoded, uint16_t& r, uint16_t& g, uint16_t& b)
{
    auto amount = static_cast<uint16_t>(encoded & 0x3FF);
    if (amount == 0)
    {
        r = g = b = 0;
        return;
    }
    r = static_cast<uint16_t>(((encoded >> 46) & 0xFFFF) / amount);
    g = static_cast<uint16_t>(((encoded >> 28) & 0xFFFF) / amount);
    b = static_cast<uint16_t>(((encoded >> 10) & 0xFFFF) / amount);
}

// Explicit-random winner key (A8). MurmurHash3's 32-bit finalizer (Appleby, public domain) is a
// bijection of its input; keyed with a per-node salt derived by the SplitMix64 finalizer
// (Steele, Lea, Flood 2014) it ranks the points of a node in a reproducible random order with no
// ties. The GPU uses the same two functions.
inline uint64_t splitmix64Mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6BU;
    h ^= h >> 13;
    h *= 0xC2B2AE35U;
    h ^= h >> 16;
    return h;
}

inline uint32_t nodeSalt(uint64_t seed, uint32_t nodeKey)
{
    return static_cast<uint32_t>(splitmix64Mix(seed + 0x9E3779B97F4A7C15ULL * (static_cast<uint64_t>(nodeKey) + 1)) >> 32);
}

inline uint32_t randomRank(uint32_t salt, uint32_t pointIdx)
{
    return fmix32(pointIdx ^ salt);
}

// ------------------------------------------------------------------------------------------------
// The build
// ------------------------------------------------------------------------------------------------

struct Node  // types.cuh:8-16 plus the dense key
{
    uint32_t denseKey   = 0;  // level offset + linear index
    uint32_t level      = 0;
    uint32_t count      = 0;  // dense subtree point count (countingGrid)
    uint32_t pointCount = 0;
    uint32_t parent     = 0xFFFFFFFFu;
    bool isFinished     = false;
    bool isInternal     = false;
    uint64_t dataIdx    = 0;
    int32_t child[8];
};

enum Status
{
    ORACLE_OK                = 0,
    ORACLE_INVALID_CONFIG    = 1,
    ORACLE_POINT_OUTSIDE_BB  = 2,
    ORACLE_OUTPUT_OVERFLOW   = 3,
};

struct Build
{
    OracleConfig cfg{};
    LevelTables lt;
    ColorMode colorMode = COLOR_NONE;
    uint64_t E          = 0;
    uint64_t used       = 0;  // N + I
    std::vector<Node> nodes;
    std::vector<uint32_t> lut;
    std::vector<OutRecord> out;
    std::vector<uint8_t> hierarchy;
    uint32_t exportedNodes    = 0;
    uint64_t exportedPoints   = 0;
    uint32_t internalNodes    = 0;
    uint32_t referenceBoundBinds = 0;  // A12: internal children the reference's launch bound truncates
    uint32_t maxLeafPoints    = 0;
    std::vector<double> nodeMins;     // 3 per node, accumulated top-down (tools.cu:29-61)
};

uint64_t outputCapacity(uint32_t n, float factor)
{
    // point_cloud.cuh:24 : a float32 product, truncated to an integer
    return static_cast<uint64_t>(static_cast<float>(n) * factor);
}

ColorMode resolveColorMode(const OracleConfig& c)
{
    // octree_processor_impl.cu:57-71, rp_subsampling.cu:61-206, fp_subsampling.cu:61-183
    if (c.intra)
        return COLOR_INTRA;
    if (c.inter)
        return c.weighted ? COLOR_INTER_WEIGHTED : COLOR_INTER;
    return COLOR_NONE;
}

int validate(const OracleConfig& c)
{
    if (c.point_amount == 0 || c.cloud_type > 1)
        return ORACLE_INVALID_CONFIG;
    uint32_t coordBytes = c.cloud_type == 0 ? 4 : 8;
    if (c.data_stride < 3 * coordBytes + 3)
        return ORACLE_INVALID_CONFIG;
    if (octreeDepth(c.chunking_grid) == 0)
        return ORACLE_INVALID_CONFIG;
    if (c.subsampling_grid == 0 || c.subsampling_grid > 1024)
        return ORACLE_INVALID_CONFIG;
    if (c.intra && c.weighted)  // A16: the reference dereferences a null averaging grid
        return ORACLE_INVALID_CONFIG;
    if (!(c.bb_max[0] - c.bb_min[0] > 0.0))
        return ORACLE_INVALID_CONFIG;
    uint64_t e = outputCapacity(c.point_amount, c.output_factor);
    if (e < c.point_amount || e > 0xFFFFFFFFull)
        return ORACLE_INVALID_CONFIG;
    return ORACLE_OK;
}

class Builder
{
public:
    Builder(const uint8_t* cloud, const OracleConfig& cfg, Build& b, bool attach = false) : itsCloud(cloud), itsB(b)
    {
        if (attach)
        {
            itsThreads   = 1;
            itsCubicSize = cfg.bb_max[0] - cfg.bb_min[0];
            return;
        }
        itsB.cfg       = cfg;
        itsB.lt        = levelTables(cfg.chunking_grid);
        itsB.colorMode = resolveColorMode(cfg);
        itsB.E         = outputCapacity(cfg.point_amount, cfg.output_factor);
        itsThreads     = threadCount(cfg.threads);
        itsCubicSize   = cfg.bb_max[0] - cfg.bb_min[0];  // metadata.cuh:66-69
    }

    int run()
    {
        int s = countPoints();
        if (s)
            return s;
        propagate();
        createNodes();
        mergeHierarchical();
        distribute();
        s = subsample();
        if (s)
            return s;
        exportHierarchy();
        return ORACLE_OK;
    }

private:
    const uint8_t* itsCloud;
    Build& itsB;
    unsigned itsThreads;
    double itsCubicSize;
    std::vector<uint32_t> itsCount;   // dense pyramid (countingGrid)
    std::vector<int32_t> itsD2S;      // dense -> sparse
    uint64_t itsLeafOffset = 0;

    Gridding chunkGridding() const
    {
        const auto& c = itsB.cfg;
        return Gridding{c.chunking_grid, itsCubicSize, c.bb_min[0], c.bb_min[1], c.bb_min[2], 0.0};
    }

    // point_counting.cuh:30-53 (dense histogram; the sparse ids are assigned canonically later)
    int countPoints()
    {
        const auto& c = itsB.cfg;
        itsCount.assign(itsB.lt.dense, 0);
        Gridding g = chunkGridding();
        uint32_t cells = itsB.lt.nodes[0];
        std::atomic<bool> outside{false};
        uint32_t* cnt = itsCount.data();
        parallelFor(c.point_amount, itsThreads, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i)
            {
                double x, y, z;
                loadPoint(itsCloud + i * c.data_stride, c.cloud_type, x, y, z);
                uint32_t cell = mapPointToGrid(x, y, z, g);
                if (cell >= cells)
                {
                    outside = true;
                    continue;
                }
                __atomic_fetch_add(cnt + cell, 1u, __ATOMIC_RELAXED);
            }
        });
        return outside ? ORACLE_POINT_OUTSIDE_BB : ORACLE_OK;
    }

    // child dense index of slot k (point_count_propagation.cuh:58-70): +x -> 4, +y -> 2, +z -> 1
    uint32_t childDense(uint32_t level, uint32_t x, uint32_t y, uint32_t z, int k) const
    {
        uint32_t lg = itsB.lt.grid[level - 1];
        uint32_t cx = 2 * x + ((k >> 2) & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k & 1);
        return itsB.lt.offset[level - 1] + cx + cy * lg + cz * lg * lg;
    }

    // point_count_propagation.cuh:34-86 (counts only)
    void propagate()
    {
        const auto& lt = itsB.lt;
        for (uint32_t l = 1; l <= lt.depth; ++l)
        {
            uint32_t gs = lt.grid[l];
            parallelFor(lt.nodes[l], itsThreads, [&](uint64_t b, uint64_t e, unsigned) {
                for (uint64_t idx = b; idx < e; ++idx)
                {
                    uint32_t i  = static_cast<uint32_t>(idx);
                    uint32_t xy = gs * gs;
                    uint32_t z = i / xy, y = (i - z * xy) / gs, x = (i - z * xy) % gs;  // tools.cu:205-212
                    uint32_t sum = 0;
                    for (int k = 0; k < 8; ++k)
                        sum += itsCount[childDense(l, x, y, z, k)];
                    itsCount[lt.offset[l] + i] = sum;
                }
            });
        }
    }

    // A2: sparse ids by rank in dense order; kernelInitLeafNodes (octree_initialization.cuh:24-52)
    // plus explicit level-0 flags (A3).
    void createNodes()
    {
        const auto& lt = itsB.lt;
        itsD2S.assign(lt.dense, -1);
        uint32_t m = 0;
        for (uint32_t d = 0; d < lt.dense; ++d)
            if (itsCount[d] > 0)
                itsD2S[d] = static_cast<int32_t>(m++);
        itsB.nodes.assign(m, Node{});
        for (uint32_t l = 0; l <= lt.depth; ++l)
        {
            for (uint32_t i = 0; i < lt.nodes[l]; ++i)
            {
                uint32_t d = lt.offset[l] + i;
                if (itsD2S[d] < 0)
                    continue;
                Node& n      = itsB.nodes[itsD2S[d]];
                n.denseKey   = d;
                n.level      = l;
                n.count      = itsCount[d];
                n.pointCount = itsCount[d];
                for (int k = 0; k < 8; ++k)
                    n.child[k] = -1;
            }
        }
    }

    // hierarchical_merging.cuh:35-122 with the canonical leaf block order (A4): parents are visited
    // level by level in linear order and allocate their block from a running offset.
    void mergeHierarchical()
    {
        const auto& lt = itsB.lt;
        auto& nodes    = itsB.nodes;
        uint64_t lutOffset = 0;
        for (uint32_t l = 1; l <= lt.depth; ++l)
        {
            uint32_t gs = lt.grid[l];
            for (uint32_t i = 0; i < lt.nodes[l]; ++i)
            {
                uint32_t d = lt.offset[l] + i;
                if (itsD2S[d] < 0)
                    continue;
                uint32_t sparse = static_cast<uint32_t>(itsD2S[d]);
                uint32_t xy     = gs * gs;
                uint32_t z = i / xy, y = (i - z * xy) / gs, x = (i - z * xy) % gs;
                Node& node      = nodes[sparse];
                uint32_t pc     = itsCount[d];
                bool finished   = pc >= itsB.cfg.merging_threshold;
                node.pointCount = finished ? 0 : pc;
                node.isInternal = finished;
                node.isFinished = finished;
                uint64_t sum    = 0;
                for (int k = 0; k < 8; ++k)
                {
                    uint32_t cd   = childDense(l, x, y, z, k);
                    node.child[k] = itsCount[cd] > 0 ? itsD2S[cd] : -1;
                    if (node.child[k] != -1)
                    {
                        Node& ch      = nodes[node.child[k]];
                        ch.isFinished = finished;
                        ch.parent     = sparse;
                        sum += ch.pointCount;
                    }
                }
                if (finished && sum > 0)
                {
                    uint64_t idx = lutOffset;
                    lutOffset += sum;
                    for (int k = 0; k < 8; ++k)
                    {
                        if (node.child[k] != -1)
                        {
                            nodes[node.child[k]].dataIdx = idx;
                            idx += nodes[node.child[k]].pointCount;
                        }
                    }
                }
            }
        }
        Node& root = nodes.back();
        if (!root.isFinished)  // A13
        {
            root.isFinished = true;
            root.dataIdx    = 0;
            lutOffset       = root.pointCount;
        }
        itsLeafOffset = lutOffset;
    }

    // point_distributing.cuh:31-76; within-leaf order = ascending point index (A5)
    void distribute()
    {
        const auto& c = itsB.cfg;
        auto& nodes   = itsB.nodes;
        itsB.lut.assign(itsB.E, 0);
        itsB.out.assign(itsB.E, OutRecord{0, 0, 0, 0, 0, 0});
        std::vector<uint32_t> leafOf(c.point_amount);
        Gridding g = chunkGridding();
        parallelFor(c.point_amount, itsThreads, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i)
            {
                double x, y, z;
                loadPoint(itsCloud + i * c.data_stride, c.cloud_type, x, y, z);
                uint32_t s = static_cast<uint32_t>(itsD2S[mapPointToGrid(x, y, z, g)]);
                while (!nodes[s].isFinished)
                    s = nodes[s].parent;
                leafOf[i] = s;
            }
        });
        std::vector<uint32_t> fill(nodes.size(), 0);
        double sx = 1.0 / c.scale[0], sy = 1.0 / c.scale[1], sz = 1.0 / c.scale[2];  // octree_chunking.cu:108-112
        for (uint64_t i = 0; i < c.point_amount; ++i)
        {
            uint32_t s   = leafOf[i];
            uint64_t pos = nodes[s].dataIdx + fill[s]++;
            itsB.lut[pos] = static_cast<uint32_t>(i);
            const uint8_t* rec = itsCloud + i * c.data_stride;
            double x, y, z;
            loadPoint(rec, c.cloud_type, x, y, z);
            OutRecord& o = itsB.out[pos];
            o.x          = static_cast<int32_t>(std::floor((x - c.bb_min[0]) * sx));
            o.y          = static_cast<int32_t>(std::floor((y - c.bb_min[1]) * sy));
            o.z          = static_cast<int32_t>(std::floor((z - c.bb_min[2]) * sz));
            const uint8_t* col = rec + 3 * (c.cloud_type == 0 ? 4 : 8);  // point_distributing.cuh:47-49
            o.r          = col[0];
            o.g          = col[1];
            o.b          = col[2];
        }
    }

public:
    // Per-thread dense subsampling workspace (the reference's countingGrid / averagingGrid / rgba,
    // octree_processor_impl.cu:41-71), reset cell by cell after each node.
    struct Workspace
    {
        std::vector<uint64_t> win;    // winner key per cell, ~0 = empty
        std::vector<uint64_t> color;  // packed colour sum per cell
        std::vector<float> rgbw;      // weighted sums per cell
        std::vector<uint32_t> touched;
    };

    struct Input
    {
        uint32_t pointIdx;
        double x, y, z;
        uint16_t r, g, b;
    };

    struct Subsample
    {
        std::vector<uint32_t> lut;
        std::vector<OutRecord> out;
    };

private:
    // Gridding of an internal node (rp_subsampling.cu:16-60 / fp_subsampling.cu:15-58).
    Gridding nodeGridding(uint32_t level, const double nodeMin[3]) const
    {
        double side   = itsCubicSize / std::pow(2.0, static_cast<double>(itsB.lt.depth - level));
        double bbSize = (nodeMin[0] + side) - nodeMin[0];
        double t      = bbSize / itsB.cfg.subsampling_grid;
        return Gridding{itsB.cfg.subsampling_grid, bbSize, nodeMin[0], nodeMin[1], nodeMin[2],
                        std::pow((t * 3) * (t * 3) * 3, 0.5) / 2.0};
    }

    void subsampleNode(uint32_t nodeIdx, const double nodeMin[3], Workspace& ws, std::vector<Subsample>& subs)
    {
        const auto& c = itsB.cfg;
        auto& nodes   = itsB.nodes;
        const Node& node = nodes[nodeIdx];

        // Gather the child points in slot order (the reference's blockIdx.z = slot).
        std::vector<Input> in;
        for (int k = 0; k < 8; ++k)
        {
            int32_t ch = node.child[k];
            if (ch == -1)
                continue;
            const Node& cn = nodes[ch];
            const uint32_t* clut;
            const OutRecord* cout;
            if (cn.isInternal)
            {
                clut = subs[ch].lut.data();
                cout = subs[ch].out.data();
            }
            else
            {
                clut = itsB.lut.data() + cn.dataIdx;
                cout = itsB.out.data() + cn.dataIdx;
            }
            for (uint32_t j = 0; j < cn.pointCount; ++j)
            {
                Input p;
                p.pointIdx = clut[j];
                loadPoint(itsCloud + static_cast<uint64_t>(p.pointIdx) * c.data_stride, c.cloud_type, p.x, p.y, p.z);
                p.r = cout[j].r;
                p.g = cout[j].g;
                p.b = cout[j].b;
                in.push_back(p);
            }
        }
        subsampleInputs(nodeIdx, nodeMin, in, ws, subs[nodeIdx]);
        nodes[nodeIdx].pointCount = static_cast<uint32_t>(subs[nodeIdx].lut.size());
    }

public:
    // One node's subsample from the given inputs (child points in slot order with their colours).
    void subsampleInputs(uint32_t nodeIdx, const double nodeMin[3], const std::vector<Input>& in, Workspace& ws,
                         Subsample& sub)
    {
        const auto& c = itsB.cfg;
        const Node& node  = itsB.nodes[nodeIdx];
        Gridding gr       = nodeGridding(node.level, nodeMin);
        const uint32_t gs = c.subsampling_grid;
        const uint64_t cells = static_cast<uint64_t>(gs) * gs * gs;
        const ColorMode mode = itsB.colorMode;
        const uint32_t salt  = nodeSalt(c.seed, node.denseKey);

        // Evaluate: occupancy + winner (+ intra colour sums). rp_subsample_evaluation.cuh:37-203,
        // fp_subsample_evaluation.cuh:22-110, canonical winners A7/A8.
        std::vector<uint32_t> cellOf(in.size());
        for (size_t j = 0; j < in.size(); ++j)
        {
            const Input& p = in[j];
            uint32_t cell  = mapPointToGrid(p.x, p.y, p.z, gr);
            if (cell >= cells)
            {
                itsOutside.store(true);
                cell = 0;
            }
            cellOf[j]    = cell;
            uint32_t rank = c.random ? randomRank(salt, p.pointIdx) : p.pointIdx;
            uint64_t key  = (static_cast<uint64_t>(rank) << 32) | p.pointIdx;
            if (ws.win[cell] == ~0ull)
                ws.touched.push_back(cell);
            ws.win[cell] = std::min(ws.win[cell], key);
            if (mode == COLOR_INTRA)
                ws.color[cell] += encodeColors(p.r, p.g, p.b);
        }

        // Inter-cell filtering: kernel_inter_cell_filtering.cuh:20-91 and :109-194.
        if (mode == COLOR_INTER || mode == COLOR_INTER_WEIGHTED)
        {
            for (size_t j = 0; j < in.size(); ++j)
            {
                const Input& p = in[j];
                uint64_t ix, iy, iz;
                gridCoords(p.x, p.y, p.z, gr, ix, iy, iz);
                double t         = gr.bbSize / gr.gridSize;
                uint64_t encoded = encodeColors(p.r, p.g, p.b);
                for (int32_t ox = -1; ox <= 1; ox++)
                    for (int32_t oy = -1; oy <= 1; oy++)
                        for (int32_t oz = -1; oz <= 1; oz++)
                        {
                            int32_t nx = static_cast<int32_t>(ix) + ox;
                            int32_t ny = static_cast<int32_t>(iy) + oy;
                            int32_t nz = static_cast<int32_t>(iz) + oz;
                            if (nx < 0 || ny < 0 || nz < 0 || nx >= static_cast<int32_t>(gs) ||
                                ny >= static_cast<int32_t>(gs) || nz >= static_cast<int32_t>(gs))
                                continue;
                            uint32_t v = static_cast<uint32_t>(nx + ny * gs + nz * gs * gs);
                            if (ws.win[v] == ~0ull)
                                continue;
                            if (mode == COLOR_INTER)
                            {
                                ws.color[v] += encoded;
                            }
                            else
                            {
                                double uX = (gr.minX + nx * t + (t * 0.5)) - p.x;
                                double uY = (gr.minY + ny * t + (t * 0.5)) - p.y;
                                double uZ = (gr.minZ + nz * t + (t * 0.5)) - p.z;
                                auto res  = static_cast<float>(std::sqrt(uX * uX + uY * uY + uZ * uZ) / gr.diag);
                                res       = std::exp(-((res * res) * 10.f));
                                float* acc = ws.rgbw.data() + 4 * static_cast<uint64_t>(v);
                                acc[0] += static_cast<float>(p.r) * res;
                                acc[1] += static_cast<float>(p.g) * res;
                                acc[2] += static_cast<float>(p.b) * res;
                                acc[3] += res;
                            }
                        }
            }
        }

        // Select: one point per occupied cell, slot order = ascending cell (A9). The record keeps
        // the winner's global quantised coordinates (rp_subsampling.cuh:116-121).
        std::sort(ws.touched.begin(), ws.touched.end());
        sub.lut.resize(ws.touched.size());
        sub.out.resize(ws.touched.size());
        double sx = 1.0 / c.scale[0], sy = 1.0 / c.scale[1], sz = 1.0 / c.scale[2];
        for (size_t s = 0; s < ws.touched.size(); ++s)
        {
            uint32_t cell     = ws.touched[s];
            uint32_t pointIdx = static_cast<uint32_t>(ws.win[cell] & 0xFFFFFFFFu);
            const uint8_t* rec = itsCloud + static_cast<uint64_t>(pointIdx) * c.data_stride;
            double x, y, z;
            loadPoint(rec, c.cloud_type, x, y, z);
            OutRecord o;
            o.x = static_cast<int32_t>(std::floor((x - c.bb_min[0]) * sx));
            o.y = static_cast<int32_t>(std::floor((y - c.bb_min[1]) * sy));
            o.z = static_cast<int32_t>(std::floor((z - c.bb_min[2]) * sz));
            if (mode == COLOR_NONE)
            {
                const uint8_t* col = rec + 3 * (c.cloud_type == 0 ? 4 : 8);
                o.r = col[0];
                o.g = col[1];
                o.b = col[2];
            }
            else if (mode == COLOR_INTER_WEIGHTED)
            {
                const float* acc = ws.rgbw.data() + 4 * static_cast<uint64_t>(cell);
                o.r              = static_cast<uint16_t>(acc[0] / acc[3]);
                o.g              = static_cast<uint16_t>(acc[1] / acc[3]);
                o.b              = static_cast<uint16_t>(acc[2] / acc[3]);
            }
            else
            {
                decodeColors(ws.color[cell], o.r, o.g, o.b);
            }
            sub.lut[s] = pointIdx;
            sub.out[s] = o;
            // reset the cell (rp_subsampling.cuh:132-134)
            ws.win[cell] = ~0ull;
            if (!ws.color.empty())
                ws.color[cell] = 0;
            if (!ws.rgbw.empty())
                std::fill_n(ws.rgbw.data() + 4 * static_cast<uint64_t>(cell), 4, 0.f);
        }
        ws.touched.clear();
    }

    void allocateWorkspace(Workspace& w) const
    {
        const uint64_t g     = itsB.cfg.subsampling_grid;
        const uint64_t cells = g * g * g;
        w.win.assign(cells, ~0ull);
        if (itsB.colorMode == COLOR_INTRA || itsB.colorMode == COLOR_INTER)
            w.color.assign(cells, 0);
        if (itsB.colorMode == COLOR_INTER_WEIGHTED)
            w.rgbw.assign(4 * cells, 0.f);
    }

private:

    // calculateChildMinBB (tools.cu:29-61), accumulated top-down exactly as the reference's DFS.
    void computeNodeMins(std::vector<double>& mins)
    {
        auto& nodes = itsB.nodes;
        mins.assign(3 * nodes.size(), 0.0);
        uint32_t root = static_cast<uint32_t>(nodes.size() - 1);
        mins[3 * root + 0] = itsB.cfg.bb_min[0];
        mins[3 * root + 1] = itsB.cfg.bb_min[1];
        mins[3 * root + 2] = itsB.cfg.bb_min[2];
        // ids are ordered by level ascending, so descending id order is top-down
        for (uint32_t id = root + 1; id-- > 0;)
        {
            const Node& n = nodes[id];
            if (n.level == 0)
                continue;
            double side      = itsCubicSize / std::pow(2.0, static_cast<double>(itsB.lt.depth - n.level));
            double childSide = side / 2.0;
            for (int k = 0; k < 8; ++k)
            {
                int32_t ch = n.child[k];
                if (ch == -1)
                    continue;
                double* cm = &mins[3 * ch];
                cm[0]      = mins[3 * id + 0];
                cm[1]      = mins[3 * id + 1];
                cm[2]      = mins[3 * id + 2];
                if (k & 4)
                    cm[0] += childSide;
                if (k & 2)
                    cm[1] += childSide;
                if (k & 1)
                    cm[2] += childSide;
            }
        }
    }

    void postOrderPlace(uint32_t id, uint64_t& cursor, std::vector<Subsample>& subs)
    {
        Node& n = itsB.nodes[id];
        for (int k = 0; k < 8; ++k)
            if (n.child[k] != -1)
                postOrderPlace(static_cast<uint32_t>(n.child[k]), cursor, subs);
        if (n.isInternal)
        {
            n.dataIdx = cursor;  // calculateWritingPosition (kernel_helpers.cuh:55-59)
            cursor += n.pointCount;
        }
    }

    int subsample()
    {
        auto& nodes = itsB.nodes;
        std::vector<double>& mins = itsB.nodeMins;
        computeNodeMins(mins);

        // statistics before subsampling (octree.cu:79-132): maxPointsPerNode sizes the reference's
        // subsampling launches; used here only to flag A12.
        uint32_t maxLeaf = 0;
        for (const Node& n : nodes)
            if (n.isFinished && !n.isInternal)
                maxLeaf = std::max(maxLeaf, n.pointCount);
        itsB.maxLeafPoints = maxLeaf;

        std::vector<Subsample> subs(nodes.size());
        std::vector<Workspace> ws(itsThreads);
        for (uint32_t l = 1; l <= itsB.lt.depth; ++l)
        {
            std::vector<uint32_t> todo;
            for (uint32_t id = 0; id < nodes.size(); ++id)
                if (nodes[id].level == l && nodes[id].isInternal)
                    todo.push_back(id);
            if (todo.empty())
                continue;
            itsB.internalNodes += static_cast<uint32_t>(todo.size());
            std::atomic<uint64_t> next{0};
            unsigned nt = static_cast<unsigned>(std::min<uint64_t>(itsThreads, todo.size()));
            parallelFor(nt, nt, [&](uint64_t, uint64_t, unsigned t) {
                Workspace& w = ws[t];
                if (w.win.empty())
                    allocateWorkspace(w);
                for (;;)
                {
                    uint64_t i = next.fetch_add(1);
                    if (i >= todo.size())
                        break;
                    uint32_t id = todo[i];
                    subsampleNode(id, &mins[3 * id], w, subs);
                }
            });
            // A12: the reference covers ceil(8 * maxLeaf / 128) * 128 local points per child.
            uint64_t blocks = (8ull * maxLeaf + 127) / 128;
            uint64_t cover  = blocks < 65535 ? blocks * 128 : 65535ull * ((blocks + 65534) / 65535) * 128;
            for (uint32_t id : todo)
                for (int k = 0; k < 8; ++k)
                    if (nodes[id].child[k] != -1 && nodes[nodes[id].child[k]].pointCount > cover)
                        itsB.referenceBoundBinds++;
        }
        if (itsOutside.load())
            return ORACLE_POINT_OUTSIDE_BB;

        // Place internal nodes after the leaf points in DFS post-order (A6).
        uint64_t cursor = itsLeafOffset;
        postOrderPlace(static_cast<uint32_t>(nodes.size() - 1), cursor, subs);
        itsB.used = cursor;
        if (cursor > itsB.E)
            return ORACLE_OUTPUT_OVERFLOW;  // A14
        for (uint32_t id = 0; id < nodes.size(); ++id)
        {
            if (!nodes[id].isInternal)
                continue;
            std::copy(subs[id].lut.begin(), subs[id].lut.end(), itsB.lut.begin() + nodes[id].dataIdx);
            std::copy(subs[id].out.begin(), subs[id].out.end(), itsB.out.begin() + nodes[id].dataIdx);
        }
        return ORACLE_OK;
    }

    // potree_exporter.cu:62-111: BFS from the root over finished children in slot order.
    void exportHierarchy()
    {
        auto& nodes = itsB.nodes;
        std::deque<uint32_t> toVisit;
        toVisit.push_back(static_cast<uint32_t>(nodes.size() - 1));
        itsB.hierarchy.clear();
        while (!toVisit.empty())
        {
            uint32_t id = toVisit.front();
            toVisit.pop_front();
            const Node& n = nodes[id];
            uint8_t mask  = 0;
            for (int k = 0; k < 8; ++k)
                if (n.child[k] != -1 && nodes[n.child[k]].isFinished)
                    mask |= static_cast<uint8_t>(1 << k);
            HierarchyEntry e{static_cast<uint8_t>(mask == 0 ? 1 : 0), mask, n.pointCount, n.dataIdx * 18,
                             static_cast<uint64_t>(n.pointCount) * 18};
            const uint8_t* p = reinterpret_cast<const uint8_t*>(&e);
            itsB.hierarchy.insert(itsB.hierarchy.end(), p, p + sizeof(e));
            itsB.exportedPoints += n.pointCount;
            itsB.exportedNodes++;
            for (int k = 0; k < 8; ++k)
                if (n.child[k] != -1 && nodes[n.child[k]].isFinished)
                    toVisit.push_back(static_cast<uint32_t>(n.child[k]));
        }
    }

    std::atomic<bool> itsOutside{false};
};

}  // namespace

// ------------------------------------------------------------------------------------------------
// C interface (ctypes)
// ------------------------------------------------------------------------------------------------

extern "C" {

typedef OracleConfig oracle_config;

void* oracle_build(const uint8_t* cloud, const oracle_config* cfg, int* status)
{
    *status = validate(*cfg);
    if (*status)
        return nullptr;
    auto* b = new Build();
    Builder builder(cloud, *cfg, *b);
    *status = builder.run();
    return b;
}

void oracle_free(void* h)
{
    delete static_cast<Build*>(h);
}

uint32_t oracle_node_amount(void* h)
{
    return static_cast<uint32_t>(static_cast<Build*>(h)->nodes.size());
}

// flags: bit0 = finished, bit1 = internal
void oracle_get_nodes(void* h, uint32_t* denseKey, uint32_t* count, uint32_t* pointCount, uint64_t* dataIdx,
                      int32_t* parent, int32_t* children, uint8_t* flags)
{
    auto* b = static_cast<Build*>(h);
    for (size_t i = 0; i < b->nodes.size(); ++i)
    {
        const Node& n = b->nodes[i];
        denseKey[i]   = n.denseKey;
        count[i]      = n.count;
        pointCount[i] = n.pointCount;
        dataIdx[i]    = n.dataIdx;
        parent[i]     = static_cast<int32_t>(n.parent);
        for (int k = 0; k < 8; ++k)
            children[8 * i + k] = n.child[k];
        flags[i] = static_cast<uint8_t>((n.isFinished ? 1 : 0) | (n.isInternal ? 2 : 0));
    }
}

uint64_t oracle_capacity(void* h)
{
    return static_cast<Build*>(h)->E;
}

uint64_t oracle_used_points(void* h)
{
    return static_cast<Build*>(h)->used;
}

void oracle_get_lut(void* h, uint32_t* dst)
{
    auto* b = static_cast<Build*>(h);
    std::copy(b->lut.begin(), b->lut.end(), dst);
}

void oracle_get_output(void* h, uint8_t* dst)
{
    auto* b = static_cast<Build*>(h);
    std::memcpy(dst, b->out.data(), b->out.size() * sizeof(OutRecord));
}

uint64_t oracle_hierarchy_size(void* h)
{
    return static_cast<Build*>(h)->hierarchy.size();
}

void oracle_get_hierarchy(void* h, uint8_t* dst)
{
    auto* b = static_cast<Build*>(h);
    std::memcpy(dst, b->hierarchy.data(), b->hierarchy.size());
}

// [exportedNodes, exportedPoints, internalNodes, referenceBoundBinds, maxLeafPoints]
void oracle_get_summary(void* h, uint64_t* out5)
{
    auto* b = static_cast<Build*>(h);
    out5[0] = b->exportedNodes;
    out5[1] = b->exportedPoints;
    out5[2] = b->internalNodes;
    out5[3] = b->referenceBoundBinds;
    out5[4] = b->maxLeafPoints;
}

// Recomputes the subsample of internal node `nodeId` of a finished build from caller-supplied
// inputs (point index + colour of every child point, children in slot order). Used to check the
// distance-weighted colours level by level with the GPU's own child colours. Returns the number
// of cells; outIdx / outRgb (3 per cell) receive the winners in ascending cell order.
uint32_t oracle_subsample_node(void* h, const uint8_t* cloud, uint32_t nodeId, uint32_t n, const uint32_t* idx,
                               const uint16_t* rgb, uint32_t* outIdx, uint16_t* outRgb)
{
    auto* b = static_cast<Build*>(h);
    Builder builder(cloud, b->cfg, *b, true);
    std::vector<Builder::Input> in(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        in[i].pointIdx = idx[i];
        loadPoint(cloud + static_cast<uint64_t>(idx[i]) * b->cfg.data_stride, b->cfg.cloud_type, in[i].x, in[i].y,
                  in[i].z);
        in[i].r = rgb[3 * i];
        in[i].g = rgb[3 * i + 1];
        in[i].b = rgb[3 * i + 2];
    }
    Builder::Workspace ws;
    builder.allocateWorkspace(ws);
    Builder::Subsample sub;
    builder.subsampleInputs(nodeId, &b->nodeMins[3 * nodeId], in, ws, sub);
    for (size_t k = 0; k < sub.lut.size(); ++k)
    {
        outIdx[k]         = sub.lut[k];
        outRgb[3 * k]     = sub.out[k].r;
        outRgb[3 * k + 1] = sub.out[k].g;
        outRgb[3 * k + 2] = sub.out[k].b;
    }
    return static_cast<uint32_t>(sub.lut.size());
}

// ---- unit-level helpers -------------------------------------------------------------------------

uint32_t oracle_map_point_to_grid(double x, double y, double z, uint32_t gridSize, double bbSize, double minX,
                                  double minY, double minZ)
{
    return mapPointToGrid(x, y, z, Gridding{gridSize, bbSize, minX, minY, minZ, 0.0});
}

uint64_t oracle_encode_colors(uint16_t r, uint16_t g, uint16_t b)
{
    return encodeColors(r, g, b);
}

void oracle_decode_colors(uint64_t encoded, uint16_t* rgb)
{
    decodeColors(encoded, rgb[0], rgb[1], rgb[2]);
}

uint64_t oracle_output_capacity(uint32_t n, float factor)
{
    return outputCapacity(n, factor);
}

uint32_t oracle_random_rank(uint64_t seed, uint32_t nodeKey, uint32_t pointIdx)
{
    return randomRank(nodeSalt(seed, nodeKey), pointIdx);
}

// depth, then (depth + 1) <= 10 entries of grid sizes, node amounts and dense offsets; returns D
// (nothing is written for an unsupported grid, whose depth is 0)
uint32_t oracle_level_tables(uint32_t grid, uint32_t* depth, uint32_t* grids, uint32_t* nodes, uint32_t* offsets)
{
    LevelTables t = levelTables(grid);
    *depth        = t.depth;
    if (t.depth == 0)
        return 0;
    for (size_t l = 0; l < t.grid.size(); ++l)
    {
        grids[l]   = t.grid[l];
        nodes[l]   = t.nodes[l];
        offsets[l] = t.offset[l];
    }
    return t.dense;
}

// Leaf statistics (octree.cu:79-157) over a caller-supplied tree. stats8 receives
// [leafNodeAmount, internalNodeAmount, minPointsPerNode, maxPointsPerNode, maxLeafDepth,
//  mean (float bits), stdev (float bits), pointSum].
void oracle_node_statistics(uint32_t m, const uint32_t* pointCount, const int32_t* children, const uint8_t* flags,
                            uint32_t root, uint64_t* stats8)
{
    uint32_t leaves = 0, internal = 0, minP = 0xFFFFFFFFu, maxP = 0, maxDepth = 0, sum = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{root, 0}};
    std::vector<uint32_t> leafCounts;
    while (!stack.empty())
    {
        auto [id, depth] = stack.back();
        stack.pop_back();
        if (!(flags[id] & 2))
        {
            maxDepth = std::max(maxDepth, depth);
            leaves++;
            sum += pointCount[id];
            minP = std::min(minP, pointCount[id]);
            maxP = std::max(maxP, pointCount[id]);
            leafCounts.push_back(pointCount[id]);
        }
        else
        {
            internal++;
            for (int k = 7; k >= 0; --k)
                if (children[8 * id + k] != -1)
                    stack.push_back({static_cast<uint32_t>(children[8 * id + k]), depth + 1});
        }
    }
    (void)m;
    float mean = static_cast<float>(sum) / leaves;
    float var  = 0.f;
    for (uint32_t pc : leafCounts)
        var += std::pow(static_cast<float>(pc) - mean, 2.f);
    float stdev = std::sqrt(var / leaves);
    uint32_t mb, sb;
    std::memcpy(&mb, &mean, 4);
    std::memcpy(&sb, &stdev, 4);
    stats8[0] = leaves;
    stats8[1] = internal;
    stats8[2] = minP;
    stats8[3] = maxP;
    stats8[4] = maxDepth;
    stats8[5] = mb;
    stats8[6] = sb;
    stats8[7] = sum;
}

}  // extern "C"
