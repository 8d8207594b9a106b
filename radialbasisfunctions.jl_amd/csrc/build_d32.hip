// Instantiates the stencil build kernel for double factorisation, up to 32-row systems.
#include "build_kernel.h"

namespace rbfx {

hipError_t launch_build_d32(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    return dispatch_dim<double, 32>(P, dim, n_work, s);
}

}  // namespace rbfx
