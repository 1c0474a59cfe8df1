// Host-side copies of large buffers, split over a small pool of threads. One thread copying into
// freshly allocated pageable memory spends most of its time in page faults (about 2 GB/s); several
// threads fault and copy in parallel.
#pragma once

#include <cstddef>

namespace oc {

// memcpy(dst, src, bytes); the calling thread takes one part, pool threads the others. Small
// copies run on the calling thread alone.
void parallelMemcpy(void* dst, const void* src, size_t bytes);

struct CopyPiece
{
    void* dst;
    const void* src;
    size_t bytes;
};

// every piece's memcpy, the pieces' bytes shared out over the threads as one copy
void parallelMemcpy(const CopyPiece* pieces, size_t count);

}  // namespace oc
