// Internal interface of the spatial point ordering.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbfx {

// out_order[r] = index of the r-th point in Morton (Z-curve) order of the bounding box
hipError_t spatial_order(const double* x, int64_t n, int dim, int32_t* out_order, hipStream_t s);

}  // namespace rbfx
