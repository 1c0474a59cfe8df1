// Page-locked host memory and the device count through the HIP runtime (the library side of the
// seams declared in slot_builder.h).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "slot_builder.h"

namespace oc {

void* pinnedAllocate(size_t bytes)
{
    void* p      = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess)
        throw std::runtime_error("hipHostMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e));
    return p;
}

void pinnedRelease(void* p)
{
    (void)hipHostFree(p);
}

int deviceCount()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess)
        return 0;
    return count;
}

}  // namespace oc
