// Device-wide exclusive prefix sum for the builder (sparse ids, leaf blocks, table slices, slots).
//
// Reduce-then-scan in three launches with no inter-workgroup waiting: (1) each workgroup reduces
// its contiguous range of tiles of SCAN_TILE items, (2) one workgroup scans the per-workgroup sums,
// (3) each workgroup rescans its tiles in order from its offset and hands (index, exclusive prefix,
// value) to an output functor.
// A tile is held striped: thread t holds items k * SCAN_THREADS + t (k < SCAN_ITEMS), so each load
// and store instruction of a wave covers 64 consecutive items, and a full tile is read with no
// bounds test, all SCAN_ITEMS loads in flight at once. Its scan is SCAN_ITEMS row scans (row k =
// the SCAN_THREADS items k * SCAN_THREADS ..): a wave64 shuffle scan per row, the wave totals of
// all rows through LDS, one barrier. Inputs and outputs are functors so the flags / values are
// generated and consumed in the same pass as the scan (e.g. occupancy flags straight from the
// counting grid); an output functor is called once per item, in no particular order, and may
// write the item's own input (in-place scans). Inputs of at most SCAN_SINGLE_TILES tiles take one
// launch of a single workgroup instead (small builds make many small scans, and every launch is a
// runtime call).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oc {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS   = 8;
constexpr int SCAN_TILE    = SCAN_THREADS * SCAN_ITEMS;
constexpr int SCAN_WAVES   = SCAN_THREADS / 64;

template <typename T>
__device__ __forceinline__ T waveInclusiveScan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        T n = __shfl_up(v, o, 64);
        if (lane >= o)
            v += n;
    }
    return v;
}

// Exclusive scan of one value per thread over the workgroup; *total gets the workgroup sum.
template <typename T>
__device__ __forceinline__ T blockExclusiveScan(T v, T& total, T* lds)
{
    T inc        = waveInclusiveScan(v);
    const int w  = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63)
        lds[w] = inc;
    __syncthreads();
    T off = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_WAVES; ++i)
    {
        T s = lds[i];
        off += (i < w) ? s : T(0);
        sum += s;
    }
    __syncthreads();
    total = sum;
    return inc - v + off;
}

// The SCAN_ITEMS items of this thread in the tile starting at tileBase (zero past n).
template <typename T, typename In>
__device__ __forceinline__ void loadTile(In in, uint64_t tileBase, uint64_t n, T (&v)[SCAN_ITEMS])
{
    const uint64_t first = tileBase + threadIdx.x;
    if (tileBase + SCAN_TILE <= n)
    {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            v[k] = in(first + static_cast<uint64_t>(k) * SCAN_THREADS);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
        {
            const uint64_t i = first + static_cast<uint64_t>(k) * SCAN_THREADS;
            v[k]             = i < n ? in(i) : T(0);
        }
    }
}

// Exclusive prefix within the tile of each item of this thread; returns the tile total.
// lds holds SCAN_ITEMS * SCAN_WAVES values.
template <typename T>
__device__ __forceinline__ T tileExclusiveScan(const T (&v)[SCAN_ITEMS], T (&pre)[SCAN_ITEMS], T* lds)
{
    const int w = threadIdx.x >> 6;
    T inc[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        inc[k] = waveInclusiveScan(v[k]);
    if ((threadIdx.x & 63) == 63)
    {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            lds[k * SCAN_WAVES + w] = inc[k];
    }
    __syncthreads();
    T carry = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        T row = 0, before = 0;
#pragma unroll
        for (int i = 0; i < SCAN_WAVES; ++i)
        {
            const T s = lds[k * SCAN_WAVES + i];
            before += (i < w) ? s : T(0);
            row += s;
        }
        pre[k] = carry + before + (inc[k] - v[k]);
        carry += row;
    }
    __syncthreads();
    return carry;
}

// out(i, offset + exclusive prefix, value) for each item i of this thread in the tile (items past n
// skipped).
template <typename T, typename Out>
__device__ __forceinline__ void storeTile(Out out, uint64_t tileBase, uint64_t n, const T (&v)[SCAN_ITEMS],
                                          const T (&pre)[SCAN_ITEMS], T offset)
{
    const uint64_t first = tileBase + threadIdx.x;
    if (tileBase + SCAN_TILE <= n)
    {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            out(first + static_cast<uint64_t>(k) * SCAN_THREADS, offset + pre[k], v[k]);
        return;
    }
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        const uint64_t i = first + static_cast<uint64_t>(k) * SCAN_THREADS;
        if (i < n)
            out(i, offset + pre[k], v[k]);
    }
}

// Workgroup b of a grid of G takes the contiguous tiles [b * per, (b + 1) * per), per = ceil(tiles / G).
__device__ __forceinline__ void tileRange(uint64_t tiles, uint64_t& first, uint64_t& last)
{
    const uint64_t per = (tiles + gridDim.x - 1) / gridDim.x;
    first              = blockIdx.x * per;
    last               = first + per < tiles ? first + per : tiles;
}

template <typename T, typename In>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanReduce(In in, uint64_t n, T* blockSums, uint64_t tiles)
{
    __shared__ T lds[SCAN_WAVES];
    uint64_t first, last;
    tileRange(tiles, first, last);
    T s = 0;
    for (uint64_t tile = first; tile < last; ++tile)
    {
        T v[SCAN_ITEMS];
        loadTile<T>(in, tile * SCAN_TILE, n, v);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            s += v[k];
    }
    T total;
    (void)blockExclusiveScan(s, total, lds);
    if (threadIdx.x == 0)
        blockSums[blockIdx.x] = total;
}

// exclusive scan of the per-workgroup sums, in place
template <typename T>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanTiles(T* tileSums, uint64_t tiles, T* total)
{
    __shared__ T lds[SCAN_ITEMS * SCAN_WAVES];
    T carry = 0;
    for (uint64_t base = 0; base < tiles; base += SCAN_TILE)
    {
        T v[SCAN_ITEMS], pre[SCAN_ITEMS];
        loadTile<T>([tileSums](uint64_t i) { return tileSums[i]; }, base, tiles, v);
        const T chunk = tileExclusiveScan(v, pre, lds);
        storeTile<T>([tileSums](uint64_t i, T p, T) { tileSums[i] = p; }, base, tiles, v, pre, carry);
        carry += chunk;
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

template <typename T, typename In, typename Out>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanDown(In in, Out out, uint64_t n, const T* blockOffsets,
                                                          uint64_t tiles)
{
    __shared__ T lds[SCAN_ITEMS * SCAN_WAVES];
    uint64_t first, last;
    tileRange(tiles, first, last);
    T carry = blockOffsets[blockIdx.x];
    for (uint64_t tile = first; tile < last; ++tile)
    {
        T v[SCAN_ITEMS], pre[SCAN_ITEMS];
        loadTile<T>(in, tile * SCAN_TILE, n, v);
        const T chunk = tileExclusiveScan(v, pre, lds);
        storeTile<T>(out, tile * SCAN_TILE, n, v, pre, carry);
        carry += chunk;
    }
}

// Small inputs: one workgroup walks the tiles in order with a running carry. All SCAN_THREADS
// threads of the workgroup call it with the same lds (SCAN_ITEMS * SCAN_WAVES values); the outputs
// are visible to the workgroup after a barrier.
template <typename T, typename In, typename Out>
__device__ __forceinline__ void blockScanAll(In in, Out out, uint64_t n, T* total, T* lds)
{
    T carry = 0;
    for (uint64_t tileBase = 0; tileBase < n; tileBase += SCAN_TILE)
    {
        T v[SCAN_ITEMS], pre[SCAN_ITEMS];
        loadTile<T>(in, tileBase, n, v);
        const T chunk = tileExclusiveScan(v, pre, lds);
        storeTile<T>(out, tileBase, n, v, pre, carry);
        carry += chunk;
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

// the same as one launch
template <typename T, typename In, typename Out>
static __global__ __launch_bounds__(SCAN_THREADS) void kScanSingle(In in, Out out, uint64_t n, T* total)
{
    __shared__ T lds[SCAN_ITEMS * SCAN_WAVES];
    blockScanAll<T>(in, out, n, total, lds);
}

__host__ __device__ inline uint64_t scanTiles(uint64_t n)
{
    return (n + SCAN_TILE - 1) / SCAN_TILE;
}

// Inputs up to this many tiles are scanned by kScanSingle.
constexpr uint64_t SCAN_SINGLE_TILES = 16;

// Exclusive scan of in(0..n-1); out(i, prefix, value) for each i; *total (device) = sum.
// tileScratch must hold scanTiles(n) values (unused for inputs of at most SCAN_SINGLE_TILES tiles).
template <typename T, typename In, typename Out>
void scanExclusive(uint64_t n, In in, Out out, T* tileScratch, T* total, hipStream_t stream)
{
    uint64_t tiles = scanTiles(n);
    if (tiles <= SCAN_SINGLE_TILES)
    {
        kScanSingle<T, In, Out><<<1, SCAN_THREADS, 0, stream>>>(in, out, n, total);
        return;
    }
    unsigned grid = static_cast<unsigned>(tiles < 8192 ? tiles : 8192);
    kScanReduce<T, In><<<grid, SCAN_THREADS, 0, stream>>>(in, n, tileScratch, tiles);
    kScanTiles<T><<<1, SCAN_THREADS, 0, stream>>>(tileScratch, grid, total);
    kScanDown<T, In, Out><<<grid, SCAN_THREADS, 0, stream>>>(in, out, n, tileScratch, tiles);
}

}  // namespace oc
