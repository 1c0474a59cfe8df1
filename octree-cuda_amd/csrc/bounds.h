// Bounding box of a cloud in device memory (the reference CLI's host loop, src/boundingbox.h).
#pragma once

#include <cstdint>

namespace oc {

// Per-axis minimum and maximum of the coordinates: {minX, minY, minZ, maxX, maxY, maxZ}.
void realBoundingBox(const uint8_t* deviceCloud, uint32_t n, uint32_t stride, uint8_t cloudType, int device,
                     double real[6]);

// The same reduction with its partial results in a static device array instead of an allocation:
// no hipMalloc / hipFree and no stream creation (it runs on the null stream). Calls on one device
// are serialised by a host mutex, since they share that device's array.
void realBoundingBoxNoAlloc(const uint8_t* deviceCloud, uint32_t n, uint32_t stride, uint8_t cloudType, int device,
                            double real[6]);

// The cube with side max(dims) centred on the real box (src/boundingbox.h:35-52).
void cubicBoundingBox(const double real[6], double cubic[6]);

}  // namespace oc
