// The finished octree as the exports and the C ABI read it: either the builder itself (device
// resident, direct sessions) or a host snapshot of it (builds run by the BuildService, whose
// builders are reused by the next request).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "types.h"

namespace oc {

struct CloudInfo  // PointCloudInfo (metadata.cuh:52-71)
{
    uint32_t pointAmount = 0;
    uint32_t dataStride  = 0;
    double bbMin[3]      = {0, 0, 0};
    double bbMax[3]      = {0, 0, 0};
    double offset[3]     = {0, 0, 0};
    double scale[3]      = {1, 1, 1};
    uint8_t cloudType    = 0;
};

struct ProcessingInfo  // ProcessingInfo (metadata.cuh:75-88)
{
    uint32_t subsamplingGrid  = 128;
    uint32_t chunkingGrid     = 512;
    uint32_t mergingThreshold = 10000;
    bool useIntraCellAvg      = false;
    bool useInterCellAvg      = false;
    bool useWeightingFunction = false;
    bool useReplacementScheme = false;
    bool useRandomSubsampling = false;
    float outputFactor        = 2.2f;
    uint64_t seed             = 0;
};

// Host copy of the sparse nodes (for export and reports).
struct HostOctree
{
    std::vector<uint32_t> denseKey, count, pointCount;
    std::vector<uint64_t> dataIdx;
    std::vector<int32_t> parent, children;
    std::vector<uint8_t> flags;
};

class BuildResult
{
public:
    virtual ~BuildResult() = default;

    virtual const CloudInfo& cloud() const           = 0;
    virtual const ProcessingInfo& processing() const = 0;
    virtual uint32_t depth() const                   = 0;
    virtual uint32_t denseNodes() const              = 0;
    virtual uint32_t sparseNodes() const             = 0;
    virtual uint64_t capacity() const                = 0;  // E
    virtual uint64_t used() const                    = 0;  // N + I
    virtual const std::vector<float>& stageTimings() const = 0;

    virtual void copyNodes(HostOctree& out) const                                     = 0;
    virtual void copyLut(uint32_t* dst, uint64_t count) const                         = 0;
    virtual void copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const = 0;
    void copyOutput(uint8_t* dst, uint64_t count) const { copyOutputRange(dst, 0, count); }

    // Everything a host snapshot holds, packed into one buffer of snapshotBytes() bytes (layout:
    // SnapshotLayout). A device builder issues all of its copies before a single wait.
    virtual void copySnapshot(uint8_t* dst) const;
    uint64_t snapshotBytes() const;
};

// Packing of a snapshot: the node arrays, then the used LUT entries and output records, each
// section starting on a 16-byte boundary (so a device-side packer copies whole 16-byte words).
struct SnapshotLayout
{
    uint64_t dataIdx, denseKey, count, pointCount, parent, children, flags, lut, out, bytes;
    SnapshotLayout(uint32_t m, uint64_t used)
    {
        auto next  = [](uint64_t at, uint64_t size) { return (at + size + 15) & ~15ull; };
        dataIdx    = 0;
        denseKey   = next(dataIdx, 8ull * m);
        count      = next(denseKey, 4ull * m);
        pointCount = next(count, 4ull * m);
        parent     = next(pointCount, 4ull * m);
        children   = next(parent, 4ull * m);
        flags      = next(children, 32ull * m);
        lut        = next(flags, m);
        out        = next(lut, 4 * used);
        bytes      = out + 18 * used;
    }
};

// Snapshot of a finished build in host memory: the nodes and the used part of the LUT and the
// output; records beyond the used part read as zero (the zero tail of octree.bin).
class HostBuildResult final : public BuildResult
{
public:
    // Copies the source's nodes and the used part of its LUT and output; with a pool, the LUT
    // and output land in one page-locked buffer from it.
    static std::unique_ptr<HostBuildResult> capture(const BuildResult& source, class PinnedPool* pool = nullptr);
    // The snapshot's metadata and its (not yet filled) buffer; the caller fills *block with the
    // source's copySnapshot layout before reading the result.
    static std::unique_ptr<HostBuildResult> prepare(const BuildResult& source, class PinnedPool* pool,
                                                    uint8_t** block);
    // A snapshot that lies at `base` inside a shared block (SnapshotLayout(sparseNodes, used) from
    // there), filled by the caller: the results of a batched build share one host block.
    static std::unique_ptr<HostBuildResult> inBlock(const CloudInfo& cloud, const ProcessingInfo& proc,
                                                    uint32_t depth, uint32_t denseNodes, uint32_t sparseNodes,
                                                    uint64_t capacity, uint64_t used,
                                                    const std::vector<float>& timings,
                                                    std::shared_ptr<uint8_t> block, uint8_t* base);

    const CloudInfo& cloud() const override { return itsCloud; }
    const ProcessingInfo& processing() const override { return itsProc; }
    uint32_t depth() const override { return itsDepth; }
    uint32_t denseNodes() const override { return itsDense; }
    uint32_t sparseNodes() const override { return itsM; }
    uint64_t capacity() const override { return itsE; }
    uint64_t used() const override { return itsUsed; }
    const std::vector<float>& stageTimings() const override { return itsTimings; }

    void copyNodes(HostOctree& out) const override;
    void copyLut(uint32_t* dst, uint64_t count) const override;
    void copyOutputRange(uint8_t* dst, uint64_t first, uint64_t count) const override;

private:
    CloudInfo itsCloud;
    ProcessingInfo itsProc;
    uint32_t itsDepth = 0, itsDense = 0, itsM = 0;
    uint64_t itsE = 0, itsUsed = 0;
    std::vector<float> itsTimings;
    std::shared_ptr<uint8_t> itsBlock;  // owner of the memory the snapshot lies in
    const uint8_t* itsBase = nullptr;   // the snapshot (SnapshotLayout) inside itsBlock
    const uint32_t* itsLut = nullptr;
    const uint8_t* itsOut  = nullptr;
};

}  // namespace oc
