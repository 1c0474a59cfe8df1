// Potree 2.0 export (potree_exporter.cu:35-173): hierarchy.bin, octree.bin, metadata.json.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "build_result.h"

namespace oc {

// BFS from the root over finished children in slot order (potree_exporter.cu:62-111). Returns
// the 22-byte entries; *exportedPoints gets the sum of their point counts.
std::vector<HierarchyEntry> hierarchyEntries(const HostOctree& tree, uint64_t* exportedPoints);

void writePotree(const std::string& directory, const BuildResult& builder, const HostOctree& tree);

}  // namespace oc
