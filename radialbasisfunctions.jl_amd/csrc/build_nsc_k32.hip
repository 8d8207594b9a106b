// Instantiates the null-space Cholesky build for stencils of up to 32 points and up to
// 10 monomials: 32 lanes per stencil, 1 row(s) per lane, 2 waves per SIMD, one FP64
// refinement step.
#include "build_nsc_kernel.h"

namespace rbfx {

template <int DIM>
static hipError_t nsc_rc_k32(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    return launch_nsc_one<32, 1, 32, 10, DIM, 1, 1, 2>(P, n_work, s);
}

hipError_t launch_nsc_k32(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    switch (dim) {
    case 1: return nsc_rc_k32<1>(P, n_work, s);
    case 2: return nsc_rc_k32<2>(P, n_work, s);
    case 3: return nsc_rc_k32<3>(P, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
