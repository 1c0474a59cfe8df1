// Bounding box of a device-resident cloud: the caller-side step of the reference CLI
// (calculateRealBB, src/boundingbox.h:10-33: per-axis fmin / fmax of the coordinates widened to
// double) as a two-pass min/max reduction, and calculateCubicBB (:35-52) on the host.
#include "bounds.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device_common.h"
#include "types.h"

namespace oc {

namespace {

constexpr int BB_THREADS = 256;
constexpr int BB_BLOCKS  = 1024;

__device__ __forceinline__ void waveMinMax(double v[6])
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            v[a]     = fmin(v[a], __shfl_xor(v[a], o, 64));
            v[3 + a] = fmax(v[3 + a], __shfl_xor(v[3 + a], o, 64));
        }
    }
}

__device__ __forceinline__ void blockMinMax(double v[6], double* partial)
{
    __shared__ double lds[BB_THREADS / 64][6];
    waveMinMax(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k)
            lds[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int i = 1; i < BB_THREADS / 64; ++i)
            for (int a = 0; a < 3; ++a)
            {
                v[a]     = fmin(v[a], lds[i][a]);
                v[3 + a] = fmax(v[3 + a], lds[i][3 + a]);
            }
        for (int k = 0; k < 6; ++k)
            partial[k] = v[k];
    }
}

template <typename T>
__global__ __launch_bounds__(BB_THREADS) void kBoundsPartial(const uint8_t* __restrict__ cloud, uint32_t n,
                                                              uint32_t stride, double* __restrict__ partials)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = globalThread(); i < n; i += globalStride())
    {
        double x, y, z;
        loadXYZ<T>(cloud + i * stride, x, y, z);
        v[0] = fmin(v[0], x);
        v[1] = fmin(v[1], y);
        v[2] = fmin(v[2], z);
        v[3] = fmax(v[3], x);
        v[4] = fmax(v[4], y);
        v[5] = fmax(v[5], z);
    }
    blockMinMax(v, partials + 6 * blockIdx.x);
}

__global__ __launch_bounds__(BB_THREADS) void kBoundsFinal(const double* __restrict__ partials, uint32_t blocks,
                                                            double* __restrict__ out)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += BB_THREADS)
        for (int a = 0; a < 3; ++a)
        {
            v[a]     = fmin(v[a], partials[6 * b + a]);
            v[3 + a] = fmax(v[3 + a], partials[6 * b + 3 + a]);
        }
    blockMinMax(v, out);
}

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
}

void checkCloud(const uint8_t* deviceCloud, uint32_t n, uint32_t stride, uint8_t cloudType)
{
    if (!deviceCloud || n == 0)
        throw std::runtime_error("the cloud is empty");
    if (cloudType > 1 || stride < (cloudType == 0 ? 12u : 24u))
        throw std::runtime_error("invalid cloud type or stride");
}

uint32_t reductionBlocks(uint32_t n)
{
    return static_cast<uint32_t>(std::min<uint64_t>(BB_BLOCKS, (static_cast<uint64_t>(n) + BB_THREADS - 1) / BB_THREADS));
}

constexpr uint32_t SCRATCH_DOUBLES = 6 * (BB_BLOCKS + 1);
constexpr int MAX_DEVICES = 64;
std::mutex gScratchMutex[MAX_DEVICES];
double* gScratchPtr[MAX_DEVICES] = {};

}  // namespace

// The partial minima and maxima of the allocation-free variant: one entry of 6 doubles per
// workgroup of the first pass (at most BB_BLOCKS) plus the final one, 49,200 bytes per device.
// In host code this name only stands for the array; its device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

void realBoundingBox(const uint8_t* deviceCloud, uint32_t n, uint32_t stride, uint8_t cloudType, int device,
                     double real[6])
{
    checkCloud(deviceCloud, n, stride, cloudType);
    check(hipSetDevice(device), "hipSetDevice");
    hipStream_t stream;
    check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    double* partials = nullptr;
    const uint32_t blocks = reductionBlocks(n);
    hipError_t e          = hipMalloc(&partials, sizeof(double) * 6 * (blocks + 1));
    if (e == hipSuccess)
    {
        if (cloudType == CLOUD_FLOAT_UINT8_T)
            kBoundsPartial<float><<<blocks, BB_THREADS, 0, stream>>>(deviceCloud, n, stride, partials);
        else
            kBoundsPartial<double><<<blocks, BB_THREADS, 0, stream>>>(deviceCloud, n, stride, partials);
        kBoundsFinal<<<1, BB_THREADS, 0, stream>>>(partials, blocks, partials + 6 * blocks);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(real, partials + 6 * blocks, sizeof(double) * 6, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(stream);
        (void)hipFree(partials);
    }
    (void)hipStreamDestroy(stream);
    check(e, "bounding box reduction");
}

void realBoundingBoxNoAlloc(const uint8_t* deviceCloud, uint32_t n, uint32_t stride, uint8_t cloudType, int device,
                            double real[6])
{
    checkCloud(deviceCloud, n, stride, cloudType);
    if (device < 0 || device >= MAX_DEVICES)
        throw std::runtime_error("device " + std::to_string(device) + " is out of range");
    const uint32_t blocks = reductionBlocks(n);
    if (6ull * (blocks + 1) > SCRATCH_DOUBLES)  // the workgroup cap keeps every cloud size within the array
        throw std::runtime_error("the reduction needs " + std::to_string(blocks) + " partials, more than the " +
                                 std::to_string(BB_BLOCKS) + " the static scratch holds");
    // one call per device at a time: every call on a device shares its scratch array
    std::lock_guard<std::mutex> lock(gScratchMutex[device]);
    check(hipSetDevice(device), "hipSetDevice");
    if (!gScratchPtr[device])
    {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress of the bounding-box scratch");
        gScratchPtr[device] = static_cast<double*>(p);
    }
    double* partials = gScratchPtr[device];
    // the null stream: creating a stream would allocate runtime resources
    if (cloudType == CLOUD_FLOAT_UINT8_T)
        kBoundsPartial<float><<<blocks, BB_THREADS, 0, nullptr>>>(deviceCloud, n, stride, partials);
    else
        kBoundsPartial<double><<<blocks, BB_THREADS, 0, nullptr>>>(deviceCloud, n, stride, partials);
    kBoundsFinal<<<1, BB_THREADS, 0, nullptr>>>(partials, blocks, partials + 6 * blocks);
    check(hipGetLastError(), "bounding box reduction launch");
    check(hipMemcpy(real, partials + 6 * blocks, sizeof(double) * 6, hipMemcpyDeviceToHost), "bounding box copy");
}

void cubicBoundingBox(const double real[6], double cubic[6])  // src/boundingbox.h:35-52
{
    const double dimX = real[3] - real[0];
    const double dimY = real[4] - real[1];
    const double dimZ = real[5] - real[2];
    const double side = std::fmax(std::fmax(dimX, dimY), dimZ);
    cubic[0]          = real[0] - ((side - dimX) / 2.0f);
    cubic[1]          = real[1] - ((side - dimY) / 2.0f);
    cubic[2]          = real[2] - ((side - dimZ) / 2.0f);
    cubic[3]          = cubic[0] + side;
    cubic[4]          = cubic[1] + side;
    cubic[5]          = cubic[2] + side;
}

}  // namespace oc
