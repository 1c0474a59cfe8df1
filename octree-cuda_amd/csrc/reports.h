// Reports: node statistics (octree.cu:79-157), statistics.json (json_exporter.h:14-99), the leaf
// point-count histogram (histogram.cu:39-119) and the memory report.
#pragma once

#include <cstdint>
#include <string>

#include "build_result.h"

namespace oc {

struct OctreeInfo  // metadata.cuh (OctreeInfo)
{
    uint32_t depth              = 0;
    uint32_t maxLeafDepth       = 0;
    uint32_t leafNodeAmount     = 0;
    uint32_t internalNodeAmount = 0;
    float meanPointsPerLeafNode = 0.f;
    float stdevPointsPerLeafNode = 0.f;
    uint32_t minPointsPerNode   = 0;
    uint32_t maxPointsPerNode   = 0;
    uint32_t nodeAmountSparse   = 0;
    uint32_t nodeAmountDense    = 0;
};

OctreeInfo nodeStatistics(const BuildResult& b, const HostOctree& tree);

void writeJsonReport(const std::string& filename, const BuildResult& b, const HostOctree& tree);
void writeHistogram(const std::string& filename, uint32_t binWidth, const BuildResult& b, const HostOctree& tree);
void writeMemoryReport(const std::string& filename);

}  // namespace oc
