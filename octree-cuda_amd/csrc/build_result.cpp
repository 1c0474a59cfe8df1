#include "build_result.h"

#include <cstring>
#include <stdexcept>
#include <type_traits>

#include "host_copy.h"
#include "pinned_pool.h"

namespace oc {

uint64_t BuildResult::snapshotBytes() const
{
    return SnapshotLayout(sparseNodes(), used()).bytes;
}

void BuildResult::copySnapshot(uint8_t* dst) const
{
    const SnapshotLayout l(sparseNodes(), used());
    HostOctree o;
    copyNodes(o);
    const uint64_t m = sparseNodes();
    std::memcpy(dst + l.dataIdx, o.dataIdx.data(), 8 * m);
    std::memcpy(dst + l.denseKey, o.denseKey.data(), 4 * m);
    std::memcpy(dst + l.count, o.count.data(), 4 * m);
    std::memcpy(dst + l.pointCount, o.pointCount.data(), 4 * m);
    std::memcpy(dst + l.parent, o.parent.data(), 4 * m);
    std::memcpy(dst + l.children, o.children.data(), 32 * m);
    std::memcpy(dst + l.flags, o.flags.data(), m);
    if (used())
    {
        copyLut(reinterpret_cast<uint32_t*>(dst + l.lut), used());
        copyOutputRange(dst + l.out, 0, used());
    }
}

std::unique_ptr<HostBuildResult> HostBuildResult::prepare(const BuildResult& source, PinnedPool* pool,
                                                          uint8_t** block)
{
    auto r        = std::unique_ptr<HostBuildResult>(new HostBuildResult());
    r->itsCloud   = source.cloud();
    r->itsProc    = source.processing();
    r->itsDepth   = source.depth();
    r->itsDense   = source.denseNodes();
    r->itsM       = source.sparseNodes();
    r->itsE       = source.capacity();
    r->itsUsed    = source.used();
    r->itsTimings = source.stageTimings();
    const SnapshotLayout l(r->itsM, r->itsUsed);
    const uint64_t bytes = l.bytes + 1;
    if (pool)
        r->itsBlock = pool->acquire(bytes);
    else
        r->itsBlock = std::shared_ptr<uint8_t>(new uint8_t[bytes], std::default_delete<uint8_t[]>());
    uint8_t* base = r->itsBlock.get();
    r->itsBase    = base;
    r->itsLut     = reinterpret_cast<const uint32_t*>(base + l.lut);
    r->itsOut     = base + l.out;
    *block        = base;
    return r;
}

std::unique_ptr<HostBuildResult> HostBuildResult::inBlock(const CloudInfo& cloud, const ProcessingInfo& proc,
                                                          uint32_t depth, uint32_t denseNodes, uint32_t sparseNodes,
                                                          uint64_t capacity, uint64_t used,
                                                          const std::vector<float>& timings,
                                                          std::shared_ptr<uint8_t> block, uint8_t* base)
{
    auto r        = std::unique_ptr<HostBuildResult>(new HostBuildResult());
    r->itsCloud   = cloud;
    r->itsProc    = proc;
    r->itsDepth   = depth;
    r->itsDense   = denseNodes;
    r->itsM       = sparseNodes;
    r->itsE       = capacity;
    r->itsUsed    = used;
    r->itsTimings = timings;
    r->itsBlock   = std::move(block);
    const SnapshotLayout l(sparseNodes, used);
    r->itsBase = base;
    r->itsLut  = reinterpret_cast<const uint32_t*>(base + l.lut);
    r->itsOut  = base + l.out;
    return r;
}

std::unique_ptr<HostBuildResult> HostBuildResult::capture(const BuildResult& source, PinnedPool* pool)
{
    uint8_t* block = nullptr;
    auto r         = prepare(source, pool, &block);
    source.copySnapshot(block);
    return r;
}

void HostBuildResult::copyNodes(HostOctree& o) const
{
    const SnapshotLayout l(itsM, itsUsed);
    const uint8_t* base = itsBase;
    const uint64_t m    = itsM;
    auto take = [&](auto& v, uint64_t offset, uint64_t n) {
        using V = typename std::decay_t<decltype(v)>::value_type;
        v.resize(n);
        std::memcpy(v.data(), base + offset, n * sizeof(V));
    };
    take(o.dataIdx, l.dataIdx, m);
    take(o.denseKey, l.denseKey, m);
    take(o.count, l.count, m);
    take(o.pointCount, l.pointCount, m);
    take(o.parent, l.parent, m);
    take(o.children, l.children, 8 * m);
    take(o.flags, l.flags, m);
}

void HostBuildResult::copyLut(uint32_t* dst, uint64_t count) const
{
    if (count > itsE)
        throw std::runtime_error("copy beyond the output capacity");
    const uint64_t head = count < itsUsed ? count : itsUsed;
    if (head)
        parallelMemcpy(dst, itsLut, head * sizeof(uint32_t));
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
        parallelMemcpy(dst, itsOut + first * sizeof(OutRecord), copied * sizeof(OutRecord));
    }
    if (count > copied)
        std::memset(dst + copied * sizeof(OutRecord), 0, (count - copied) * sizeof(OutRecord));
}

}  // namespace oc
