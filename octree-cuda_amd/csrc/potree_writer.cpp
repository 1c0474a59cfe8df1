#include "potree_writer.h"

#include <algorithm>
#include <deque>
#include <fstream>
#include <stdexcept>

#include "json_writer.h"

namespace oc {

std::vector<HierarchyEntry> hierarchyEntries(const HostOctree& tree, uint64_t* exportedPoints)
{
    std::vector<HierarchyEntry> entries;
    uint64_t points = 0;
    const uint32_t m = static_cast<uint32_t>(tree.flags.size());
    if (m == 0)
        return entries;
    std::deque<uint32_t> toVisit{m - 1};
    while (!toVisit.empty())
    {
        uint32_t id = toVisit.front();
        toVisit.pop_front();
        uint8_t mask = 0;
        for (int k = 0; k < 8; ++k)
        {
            int32_t c = tree.children[8 * static_cast<uint64_t>(id) + k];
            if (c >= 0 && (tree.flags[c] & NODE_FINISHED))
                mask |= static_cast<uint8_t>(1u << k);
        }
        HierarchyEntry e;
        e.type       = mask == 0 ? 1 : 0;
        e.bitmask    = mask;
        e.points     = tree.pointCount[id];
        e.byteOffset = tree.dataIdx[id] * sizeof(OutRecord);
        e.byteSize   = static_cast<uint64_t>(tree.pointCount[id]) * sizeof(OutRecord);
        entries.push_back(e);
        points += tree.pointCount[id];
        for (int k = 0; k < 8; ++k)
        {
            int32_t c = tree.children[8 * static_cast<uint64_t>(id) + k];
            if (c >= 0 && (tree.flags[c] & NODE_FINISHED))
                toVisit.push_back(static_cast<uint32_t>(c));
        }
    }
    if (exportedPoints)
        *exportedPoints = points;
    return entries;
}

static void writeMetadata(const std::string& path, const BuildResult& b, uint32_t exportedPoints,
                          uint32_t exportedNodes)
{
    const CloudInfo& c      = b.cloud();
    const ProcessingInfo& p = b.processing();
    const double side       = c.bbMax[0] - c.bbMin[0];
    const double spacing    = side / p.subsamplingGrid;

    Json md;
    md["version"]     = Json::str("2.0");
    md["name"]        = Json::str("GpuPotreeConverter");
    md["description"] = Json::str("Generated with GpuPotreeConverter");
    md["points"]      = Json::u(exportedPoints);
    md["projection"]  = Json::str("");
    Json flags        = Json::array();
    flags.push(Json::str(p.useReplacementScheme ? "REPLACING" : "ADDITIVE"));
    flags.push(Json::str(p.useRandomSubsampling ? "EXPLICIT-RANDOM-SUBSAMPLING" : "IMPLICIT-RANDOM-SUBSAMPLING"));
    if (p.useInterCellAvg)
        flags.push(Json::str("INTER-CELL-COLOR-FILTERING"));
    else if (p.useIntraCellAvg)
        flags.push(Json::str("INTRA-CELL-COLOR-FILTERING"));
    md["flags"]                       = std::move(flags);
    md["hierarchy"]["firstChunkSize"] = Json::u(static_cast<uint64_t>(exportedNodes) * sizeof(HierarchyEntry));
    md["hierarchy"]["stepSize"]       = Json::u(100);
    md["hierarchy"]["depth"]          = Json::u(20);
    Json offset                       = Json::array();
    for (int a = 0; a < 3; ++a)
        offset.push(Json::i(0));
    md["offset"] = std::move(offset);
    Json scale   = Json::array();
    for (int a = 0; a < 3; ++a)
        scale.push(Json::d(c.scale[a]));
    md["scale"]   = std::move(scale);
    md["spacing"] = Json::d(spacing);
    Json bbMin = Json::array(), bbMax = Json::array();
    for (int a = 0; a < 3; ++a)
    {
        bbMin.push(Json::i(0));
        bbMax.push(Json::d(side));
    }
    md["boundingBox"]["min"] = std::move(bbMin);
    md["boundingBox"]["max"] = std::move(bbMax);
    md["encoding"]           = Json::str("default");

    Json attributes = Json::array();
    {
        Json a;
        a["name"]        = Json::str("position");
        a["description"] = Json::str("");
        a["size"]        = Json::u(12);
        a["numElements"] = Json::u(3);
        a["elementSize"] = Json::u(4);
        a["type"]        = Json::str("int32");
        Json mn = Json::array(), mx = Json::array();
        for (int k = 0; k < 3; ++k)
        {
            mn.push(Json::i(0));
            mx.push(Json::d(side));
        }
        a["min"] = std::move(mn);
        a["max"] = std::move(mx);
        attributes.push(std::move(a));
    }
    {
        Json a;
        a["name"]        = Json::str("rgb");
        a["description"] = Json::str("");
        a["size"]        = Json::u(6);
        a["numElements"] = Json::u(3);
        a["elementSize"] = Json::u(2);
        a["type"]        = Json::str("uint16");
        Json mn = Json::array(), mx = Json::array();
        for (int k = 0; k < 3; ++k)
        {
            mn.push(Json::i(0));
            mx.push(Json::i(65536));
        }
        a["min"] = std::move(mn);
        a["max"] = std::move(mx);
        attributes.push(std::move(a));
    }
    md["attributes"] = std::move(attributes);

    std::ofstream f(path);
    if (!f)
        throw std::runtime_error("cannot write " + path);
    f << md.dump(4);
}

void writePotree(const std::string& directory, const BuildResult& b, const HostOctree& tree)
{
    uint64_t exportedPoints = 0;
    std::vector<HierarchyEntry> entries = hierarchyEntries(tree, &exportedPoints);

    {
        std::ofstream h(directory + "/hierarchy.bin", std::ios::binary);
        if (!h)
            throw std::runtime_error("cannot write " + directory + "/hierarchy.bin");
        h.write(reinterpret_cast<const char*>(entries.data()),
                static_cast<std::streamsize>(entries.size() * sizeof(HierarchyEntry)));
    }
    {
        // octree.bin holds all E output records (potree_exporter.cu:56); the records beyond the
        // used N + I are written as zeros.
        std::ofstream o(directory + "/octree.bin", std::ios::binary);
        if (!o)
            throw std::runtime_error("cannot write " + directory + "/octree.bin");
        const uint64_t used  = b.used();
        const uint64_t total = b.capacity();
        const uint64_t chunk = 8ull << 20;  // records per staging copy
        std::vector<uint8_t> buf(std::min<uint64_t>(chunk, std::max<uint64_t>(used, 1)) * sizeof(OutRecord));
        for (uint64_t first = 0; first < used; first += chunk)
        {
            uint64_t n = std::min<uint64_t>(chunk, used - first);
            b.copyOutputRange(buf.data(), first, n);
            o.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(n * sizeof(OutRecord)));
        }
        std::vector<char> zeros(1 << 20, 0);
        uint64_t tailBytes = (total - used) * sizeof(OutRecord);
        while (tailBytes > 0)
        {
            uint64_t n = std::min<uint64_t>(tailBytes, zeros.size());
            o.write(zeros.data(), static_cast<std::streamsize>(n));
            tailBytes -= n;
        }
    }
    writeMetadata(directory + "/metadata.json", b, static_cast<uint32_t>(exportedPoints),
                  static_cast<uint32_t>(entries.size()));
}

}  // namespace oc
