// Plain data shared by host and device code.
#pragma once

#include <cstddef>
#include <cstdint>

namespace oc {

// Cloud record layout (metadata.cuh:25-29)
enum CloudType : uint8_t
{
    CLOUD_FLOAT_UINT8_T  = 0,
    CLOUD_DOUBLE_UINT8_T = 1
};

// Output record written to octree.bin (types.cuh:18-24): int32 position, uint16 colour.
#pragma pack(push, 1)
struct OutRecord
{
    int32_t x, y, z;
    uint16_t r, g, b;
};

// hierarchy.bin entry (potree_exporter.cuh:28-35)
struct HierarchyEntry
{
    uint8_t type;
    uint8_t bitmask;
    uint32_t points;
    uint64_t byteOffset;
    uint64_t byteSize;
};
#pragma pack(pop)
static_assert(sizeof(OutRecord) == 18, "output record is 18 bytes");
static_assert(sizeof(HierarchyEntry) == 22, "hierarchy entry is 22 bytes");

// A cubic grid over a bounding box (kernel_structs.cuh:18-24).
struct Gridding
{
    double minX, minY, minZ;
    double bbSize;
    uint32_t gridSize;
};
static_assert(offsetof(Gridding, minY) == sizeof(double) && offsetof(Gridding, minZ) == 2 * sizeof(double),
              "the cubic minimum is read as three consecutive doubles");

// Per internal node parameters of one subsampling level.
struct NodeGrid
{
    Gridding grid;
    double diag;     // half diagonal of 3x3x3 cells (rp_subsampling.cu:55-60)
    uint32_t salt;   // explicit-random rank salt
    uint32_t id;     // sparse node id
};

// Node flag bits
constexpr uint8_t NODE_FINISHED = 1;
constexpr uint8_t NODE_INTERNAL = 2;

// Colour filter applied by the subsampling (octree_processor_impl.cu:57-71 and the drivers in
// rp_subsampling.cu / fp_subsampling.cu)
enum ColorMode : uint32_t
{
    COLOR_NONE           = 0,
    COLOR_INTRA          = 1,
    COLOR_INTER          = 2,
    COLOR_INTER_WEIGHTED = 3
};

constexpr uint32_t MAX_LEVELS = 10;  // chunking grid 512 -> depth 9 -> levels 0..9

}  // namespace oc
