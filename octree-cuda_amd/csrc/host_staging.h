// Transfers between device memory and pageable host memory through two page-locked staging
// buffers shared by the process: the copy engine fills or empties one while a pool of host
// threads (parallelMemcpy) moves the other, so the transfer runs near the copy engine's rate and
// the page faults of a freshly allocated destination are taken by several threads. Page-locked
// host memory and small transfers go straight through hipMemcpyAsync.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace oc {

struct HostSegment
{
    void* device;
    const void* host;
    uint64_t bytes;
};

// Host to device on `stream`. When staged, returns once every byte has left the staging buffers
// (the stream is synchronised); otherwise the copies are only enqueued.
void uploadSegments(const HostSegment* segments, size_t count, hipStream_t stream);

// Device to host; returns when the bytes are in host memory.
void downloadStaged(void* host, const void* device, uint64_t bytes, hipStream_t stream);

}  // namespace oc
