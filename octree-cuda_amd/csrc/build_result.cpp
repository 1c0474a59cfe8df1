#include "build_result.h"

#include <cstring>
#include <stdexcept>

#include "pinned_pool.h"

namespace oc {

std::unique_ptr<HostBuildResult> HostBuildResult::capture(const BuildResult& source, PinnedPool* pool)
{
    auto r       = std::unique_ptr<HostBuildResult>(new HostBuildResult());
    r->itsCloud  = source.cloud();
    r->itsProc   = source.processing();
    r->itsDepth  = source.depth();
    r->itsDense  = source.denseNodes();
    r->itsM      = source.sparseNodes();
    r->itsE      = source.capacity();
    r->itsUsed   = source.used();
    r->itsTimings = source.stageTimings();
    source.copyNodes(r->itsNodes);
    const size_t lutBytes = r->itsUsed * sizeof(uint32_t);
    const size_t bytes    = lutBytes + r->itsUsed * sizeof(OutRecord) + 1;
    if (pool)
        r->itsBlock = pool->acquire(bytes);
    else
        r->itsBlock = std::shared_ptr<uint8_t>(new uint8_t[bytes], std::default_delete<uint8_t[]>());
    uint8_t* base = r->itsBlock.get();
    r->itsLut     = reinterpret_cast<const uint32_t*>(base);
    r->itsOut     = base + lutBytes;
    if (r->itsUsed)
    {
        source.copyLut(reinterpret_cast<uint32_t*>(base), r->itsUsed);
        source.copyOutputRange(base + lutBytes, 0, r->itsUsed);
    }
    return r;
}

void HostBuildResult::copyLut(uint32_t* dst, uint64_t count) const
{
    if (count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    const uint64_t head = count < itsUsed ? count : itsUsed;
    if (head)
        std::memcpy(dst, itsLut, head * sizeof(uint32_t));
    if (count > head)
        std::memset(dst + head, 0, (count - head) * sizeof(uint32_t));
}

void HostBuildResult::copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const
{
    if (first + count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    uint64_t copied = 0;
    if (first < itsUsed)
    {
        copied = (first + count <= itsUsed ? count : itsUsed - first);
        std::memcpy(dst, itsOut + first * sizeof(OutRecord), copied * sizeof(OutRecord));
    }
    if (count > copied)
        std::memset(dst + copied * sizeof(OutRecord), 0, (count - copied) * sizeof(OutRecord));
}

}  // namespace oc
