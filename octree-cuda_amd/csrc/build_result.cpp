#include "build_result.h"

#include <cstring>
#include <stdexcept>

namespace oc {

std::unique_ptr<HostBuildResult> HostBuildResult::capture(const BuildResult& source)
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
    r->itsLut.resize(r->itsUsed);
    r->itsOut.resize(r->itsUsed * sizeof(OutRecord));
    if (r->itsUsed)
    {
        source.copyLut(r->itsLut.data(), r->itsUsed);
        source.copyOutputRange(r->itsOut.data(), 0, r->itsUsed);
    }
    return r;
}

void HostBuildResult::copyLut(uint32_t* dst, uint64_t count) const
{
    if (count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    const uint64_t head = count < itsUsed ? count : itsUsed;
    if (head)
        std::memcpy(dst, itsLut.data(), head * sizeof(uint32_t));
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
        std::memcpy(dst, itsOut.data() + first * sizeof(OutRecord), copied * sizeof(OutRecord));
    }
    if (count > copied)
        std::memset(dst + copied * sizeof(OutRecord), 0, (count - copied) * sizeof(OutRecord));
}

}  // namespace oc
