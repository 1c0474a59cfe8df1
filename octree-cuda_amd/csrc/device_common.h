// Device-side helpers shared by the builder's kernels: record decoding, the cell mapping, packed
// colours, quantisation and the explicit-random ranking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "types.h"

namespace oc {

// Unaligned load of the three coordinates of a packed AoS record (the reference reads them through
// a #pragma pack(1) Vector3, metadata.cuh:13-19). gfx950 serves unaligned dword loads directly.
template <typename T>
__device__ __forceinline__ void loadXYZ(const uint8_t* rec, double& x, double& y, double& z)
{
    T v[3];
    __builtin_memcpy(v, rec, sizeof(v));
    x = static_cast<double>(v[0]);
    y = static_cast<double>(v[1]);
    z = static_cast<double>(v[2]);
}

// Cell coordinates of a point in a grid, with the reference's exact expression order
// (kernel_helpers.cuh:11-25): t = bbSize / G; u = (p - min) / t; clamp to G - 1 with fmin;
// truncate to int64, then reinterpret as uint64.
__device__ __forceinline__ void gridCoords(double px, double py, double pz, const Gridding& g, uint64_t& ix,
                                           uint64_t& iy, uint64_t& iz)
{
    double t  = g.bbSize / g.gridSize;
    double uX = (px - g.minX) / t;
    double uY = (py - g.minY) / t;
    double uZ = (pz - g.minZ) / t;
    t         = g.gridSize - 1.0;
    ix        = static_cast<uint64_t>(static_cast<int64_t>(fmin(uX, t)));
    iy        = static_cast<uint64_t>(static_cast<int64_t>(fmin(uY, t)));
    iz        = static_cast<uint64_t>(static_cast<int64_t>(fmin(uZ, t)));
}

__device__ __forceinline__ uint32_t mapPointToGrid(double px, double py, double pz, const Gridding& g)
{
    uint64_t ix, iy, iz;
    gridCoords(px, py, pz, g, ix, iy, iz);
    return static_cast<uint32_t>(ix + iy * g.gridSize + iz * g.gridSize * g.gridSize);
}

// Packed colour (kernel_helpers.cuh:49-53): 18-bit r, g, b fields and a 10-bit count; sums wrap
// exactly like the reference's uint64 atomics.
__device__ __forceinline__ uint64_t encodeColors(uint16_t r, uint16_t g, uint16_t b)
{
    return (static_cast<uint64_t>(r) << 46) | (static_cast<uint64_t>(g) << 28) | (static_cast<uint64_t>(b) << 10) |
           1ull;
}

// rp_subsampling.cuh:123-129; a zero amount (count == 0 mod 1024) decodes to 0 instead of dividing
// by zero.
__device__ __forceinline__ void decodeColors(uint64_t e, uint16_t& r, uint16_t& g, uint16_t& b)
{
    uint32_t amount = static_cast<uint32_t>(e & 0x3FF);
    if (amount == 0)
    {
        r = g = b = 0;
        return;
    }
    r = static_cast<uint16_t>(static_cast<uint32_t>((e >> 46) & 0xFFFF) / amount);
    g = static_cast<uint16_t>(static_cast<uint32_t>((e >> 28) & 0xFFFF) / amount);
    b = static_cast<uint16_t>(static_cast<uint32_t>((e >> 10) & 0xFFFF) / amount);
}

// Explicit random subsampling ranks the points of a node with MurmurHash3's 32-bit finalizer
// (a bijection, so no ties) keyed by a per-node salt from the SplitMix64 finalizer; the winner of
// a cell is the point of minimum rank. Identical on host and device.
__host__ __device__ __forceinline__ uint64_t splitmix64Mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6BU;
    h ^= h >> 13;
    h *= 0xC2B2AE35U;
    h ^= h >> 16;
    return h;
}

__host__ __device__ __forceinline__ uint32_t nodeSalt(uint64_t seed, uint32_t nodeKey)
{
    return static_cast<uint32_t>(
            splitmix64Mix(seed + 0x9E3779B97F4A7C15ULL * (static_cast<uint64_t>(nodeKey) + 1)) >> 32);
}

// Output record of a point (point_distributing.cuh:62-74): int32(floor((p - cubicMin) * (1 / scale)))
// per axis with the reciprocal taken on the host (octree_chunking.cu:108-112); rgb packs r | g << 8
// | b << 16.
__device__ __forceinline__ OutRecord quantizePoint(double x, double y, double z, uint32_t rgb, const Gridding& g,
                                                   double sx, double sy, double sz)
{
    OutRecord o;
    o.x = static_cast<int32_t>(floor((x - g.minX) * sx));
    o.y = static_cast<int32_t>(floor((y - g.minY) * sy));
    o.z = static_cast<int32_t>(floor((z - g.minZ) * sz));
    o.r = static_cast<uint16_t>(rgb & 0xFF);
    o.g = static_cast<uint16_t>((rgb >> 8) & 0xFF);
    o.b = static_cast<uint16_t>((rgb >> 16) & 0xFF);
    return o;
}

__device__ __forceinline__ void storeRecord(OutRecord* dst, const OutRecord& r)
{
    __builtin_memcpy(dst, &r, sizeof(OutRecord));
}

__device__ __forceinline__ OutRecord loadRecord(const OutRecord* src)
{
    OutRecord r;
    __builtin_memcpy(&r, src, sizeof(OutRecord));
    return r;
}

// Wave-level runs: consecutive lanes holding equal keys (inputs arrive spatially sorted, so lanes
// often share a bitmap word or a cell). Every lane must take part; a lane without work passes a key
// of its own. start / end are the first and last lane of the lane's run.
__device__ __forceinline__ uint64_t shflUp64(uint64_t v, int delta)
{
    const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), delta, 64);
    const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), delta, 64);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int lane)
{
    const uint32_t lo = __shfl(static_cast<uint32_t>(v), lane, 64);
    const uint32_t hi = __shfl(static_cast<uint32_t>(v >> 32), lane, 64);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

struct WaveRun
{
    int lane, start, end;
};

__device__ __forceinline__ WaveRun waveRun(uint64_t key)
{
    WaveRun r;
    r.lane                 = static_cast<int>(threadIdx.x & 63);
    const uint64_t prev    = shflUp64(key, 1);
    const bool head        = r.lane == 0 || prev != key;
    const uint64_t heads   = __ballot(head);
    const uint64_t upToMe  = (r.lane == 63) ? ~0ull : ((2ull << r.lane) - 1);
    r.start                = 63 - __clzll(heads & upToMe);
    const uint64_t after   = heads & ~upToMe;
    r.end                  = after ? (__ffsll(static_cast<long long>(after)) - 2) : 63;
    return r;
}

// Inclusive scan with op over the lanes of each run (log-step, limited to the run).
template <typename V, typename Op>
__device__ __forceinline__ V runScan(V v, const WaveRun& r, Op op)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const V n = static_cast<V>(shflUp64(static_cast<uint64_t>(v), o));
        if (r.lane - o >= r.start)
            v = op(v, n);
    }
    return v;
}

// grid-stride helpers
__device__ __forceinline__ uint64_t globalThread()
{
    return static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ uint64_t globalStride()
{
    return static_cast<uint64_t>(gridDim.x) * blockDim.x;
}

}  // namespace oc
