#include "reports.h"

#include "builder.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "json_writer.h"

namespace oc {

// updateNodeStatistics / evaluateNodeProperties / calculatePointVarianceInLeafNoes
// (octree.cu:79-157): recursion from the root over the children of internal nodes.
OctreeInfo nodeStatistics(const BuildResult& b, const HostOctree& tree)
{
    OctreeInfo s;
    s.depth            = b.depth();
    s.nodeAmountSparse = static_cast<uint32_t>(tree.flags.size());
    s.nodeAmountDense  = b.denseNodes();
    s.minPointsPerNode = 0xFFFFFFFFu;
    s.maxPointsPerNode = 0;
    if (tree.flags.empty())
        return s;
    uint32_t pointSum = 0;
    std::vector<uint32_t> leafCounts;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{static_cast<uint32_t>(tree.flags.size() - 1), 0}};
    while (!stack.empty())
    {
        auto [id, level] = stack.back();
        stack.pop_back();
        if (!(tree.flags[id] & NODE_INTERNAL))
        {
            s.maxLeafDepth = std::max(s.maxLeafDepth, level);
            s.leafNodeAmount++;
            pointSum += tree.pointCount[id];
            s.minPointsPerNode = std::min(s.minPointsPerNode, tree.pointCount[id]);
            s.maxPointsPerNode = std::max(s.maxPointsPerNode, tree.pointCount[id]);
            leafCounts.push_back(tree.pointCount[id]);
        }
        else
        {
            s.internalNodeAmount++;
            for (int k = 7; k >= 0; --k)
            {
                int32_t c = tree.children[8 * static_cast<uint64_t>(id) + k];
                if (c >= 0)
                    stack.push_back({static_cast<uint32_t>(c), level + 1});
            }
        }
    }
    s.meanPointsPerLeafNode = static_cast<float>(pointSum) / s.leafNodeAmount;
    float var               = 0.f;
    for (uint32_t pc : leafCounts)
        var += std::pow(static_cast<float>(pc) - s.meanPointsPerLeafNode, 2.f);
    s.stdevPointsPerLeafNode = std::sqrt(var / s.leafNodeAmount);
    return s;
}

void writeJsonReport(const std::string& filename, const BuildResult& b, const HostOctree& tree)
{
    const ProcessingInfo& p = b.processing();
    OctreeInfo s            = nodeStatistics(b, tree);
    Json j;
    j["depth"]                       = Json::u(s.depth);
    j["chunking"]["grid"]            = Json::u(p.chunkingGrid);
    j["chunking"]["mergingThreshold"] = Json::u(p.mergingThreshold);
    j["subsampling"]["grid"]         = Json::u(p.subsamplingGrid);
    j["subsampling"]["strategy"] =
            Json::str(p.useRandomSubsampling ? "EXPLICIT RANDOM SUBSAMPLING" : "IMPLICIT RANDOM SUBSAMPLING");
    if (p.useInterCellAvg)
        j["subsampling"]["averaging"] = Json::str("INTER-CELL AVERAGING");
    if (p.useIntraCellAvg)
        j["subsampling"]["averaging"] = Json::str("INTRA-CELL AVERAGING");
    j["resultNodes"]["octreeNodes"]        = Json::u(s.leafNodeAmount + s.internalNodeAmount);
    j["resultNodes"]["leafNodeAmount"]     = Json::u(s.leafNodeAmount);
    j["resultNodes"]["internalNodeAmount"] = Json::u(s.internalNodeAmount);
    j["resultNodes"]["maxLeafDepth"]       = Json::u(s.maxLeafDepth);
    j["overallNodes"]["sparseOctreeNodes"] = Json::u(s.nodeAmountSparse);
    j["overallNodes"]["denseOctreeNodes"]  = Json::u(s.nodeAmountDense);
    j["overallNodes"]["memorySaving"] =
            Json::d((1 - (static_cast<float>(s.nodeAmountSparse) / s.nodeAmountDense)) * 100);
    j["pointDistribution"]["meanPointsPerLeafNode"]  = Json::d(s.meanPointsPerLeafNode);
    j["pointDistribution"]["stdevPointsPerLeafNode"] = Json::d(s.stdevPointsPerLeafNode);
    j["pointDistribution"]["minPointsPerNode"]       = Json::u(s.minPointsPerNode);
    j["pointDistribution"]["maxPointsPerNode"]       = Json::u(s.maxPointsPerNode);

    // device-side stage durations of the last build (HIP events on the session stream)
    const auto& t             = b.stageTimings();
    const char* names[]       = {"pointCounting", "merging", "pointDistribution", "subsampling"};
    float accumulated         = 0.f;
    for (int k = 0; k < 4 && k < static_cast<int>(t.size()); ++k)
    {
        j["timings"]["kernel"][names[k]]["invocations"] = Json::u(1);
        j["timings"]["kernel"][names[k]]["duration"]    = Json::d(t[k]);
        accumulated += t[k];
    }
    j["timings"]["accumulatedKernelTime"] = Json::d(accumulated);
    if (t.size() > 4)
        j["timings"]["process"]["Generating octree"] = Json::d(t[4]);
    j["memory"]["peak"]      = Json::u(DeviceBuffer::peakAllocated());
    j["memory"]["allocated"] = Json::u(DeviceBuffer::totalAllocated());

    std::ofstream f(filename);
    if (!f)
        throw std::runtime_error("cannot write " + filename);
    f << j.dump(4);
}

// histogram.cu:39-119. A bin width of 0 selects 3.5 * stdev / leaves^(1/3) (Scott's rule); the
// reference sizes the bins as (max - min) / width and indexes the maximum one past the end, so
// one extra bin is kept here.
void writeHistogram(const std::string& filename, uint32_t binWidth, const BuildResult& b, const HostOctree& tree)
{
    OctreeInfo s = nodeStatistics(b, tree);
    if (binWidth == 0)
        binWidth = static_cast<uint32_t>(
                std::ceil(3.5f * (s.stdevPointsPerLeafNode / std::pow(static_cast<float>(s.leafNodeAmount), 1.f / 3.f))));
    if (binWidth == 0)
        binWidth = 1;
    uint32_t span      = s.maxPointsPerNode >= s.minPointsPerNode ? s.maxPointsPerNode - s.minPointsPerNode : 0;
    uint32_t binAmount = span / binWidth + 1;
    std::vector<uint32_t> counts(binAmount, 0);
    std::vector<uint32_t> stack{static_cast<uint32_t>(tree.flags.size() - 1)};
    while (!stack.empty())
    {
        uint32_t id = stack.back();
        stack.pop_back();
        if (!(tree.flags[id] & NODE_INTERNAL))
            ++counts[(tree.pointCount[id] - s.minPointsPerNode) / binWidth];
        else
            for (int k = 0; k < 8; ++k)
            {
                int32_t c = tree.children[8 * static_cast<uint64_t>(id) + k];
                if (c >= 0)
                    stack.push_back(static_cast<uint32_t>(c));
            }
    }
    std::string labels = "labels:[", data = "data:[";
    for (uint32_t i = 0; i < binAmount; ++i)
    {
        labels += "'" + std::to_string(s.minPointsPerNode + i * binWidth) + " - " +
                  std::to_string(s.minPointsPerNode + (i + 1) * binWidth) + "'";
        data += std::to_string(counts[i]);
        if (i + 1 < binAmount)
        {
            labels += ",";
            data += ",";
        }
    }
    labels += "]";
    data += "]";
    std::string label = "'Point Distribution: binWidth(" + std::to_string(binWidth) + "), mergingThreshold(" +
                        std::to_string(b.processing().mergingThreshold) + "), points(" +
                        std::to_string(b.cloud().pointAmount) + ")'";
    std::ofstream f(filename);
    if (!f)
        throw std::runtime_error("cannot write " + filename);
    f << "<html>\n    <head>\n        <script src=\"https://cdn.jsdelivr.net/npm/chart.js@2.8.0\"></script>\n"
         "        <style>\n            html, body {\n            margin: 0;\n            height: 100%;\n"
         "            }\n            canvas {\n            width: 100%;\n            height: 100%;\n"
         "            display: block;\n            }\n        </style>\n    </head>\n    <body>\n"
         "        <canvas id=\"myChart\"></canvas>\n        <script>\n"
         "            var ctx = document.getElementById('myChart').getContext('2d');\n"
         "            var chart = new Chart(ctx, {\n                type: 'bar',\n\n                data:{";
    f << labels + ", datasets:[{label: " + label +
                 ", backgroundColor: 'rgb(255, 99, 132)', borderColor: 'rgb(255, 99, 132)', " + data + "}]";
    f << "               }\n            });\n        </script>\n    </body>\n</html>";
}

// The reference writes an HTML chart of every cudaMalloc/cudaFree at process exit
// (memory_tracker.cuh:31-181); this session-level report states the arena's current and peak
// device allocation when the session is destroyed.
void writeMemoryReport(const std::string& filename)
{
    std::ofstream f(filename);
    if (!f)
        throw std::runtime_error("cannot write " + filename);
    double peakMB = DeviceBuffer::peakAllocated() / 1.0e6;
    double nowMB  = DeviceBuffer::totalAllocated() / 1.0e6;
    f << "<html>\n    <head><title>GPU memory report</title></head>\n    <body>\n"
      << "        <h1>GPU memory report</h1>\n"
      << "        <p>Peak device allocation: " << peakMB << " MB</p>\n"
      << "        <p>Allocated at session end: " << nowMB << " MB</p>\n"
      << "    </body>\n</html>\n";
}

}  // namespace oc
