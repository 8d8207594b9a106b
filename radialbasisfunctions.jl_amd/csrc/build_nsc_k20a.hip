// Instantiates the null-space Cholesky build for stencils of up to 20 points and up to
// 6 monomials: 32 lanes per stencil, 1 row(s) per lane, 3 waves per SIMD, one FP64
// refinement step.
#include "build_nsc_kernel.h"

namespace rbfx {

template <int DIM>
static hipError_t nsc_rc_k20a(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    return launch_nsc_one<32, 1, 20, 6, DIM, 1, 1, 3>(P, n_work, s);
}

hipError_t launch_nsc_k20a(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    switch (dim) {
    case 1: return nsc_rc_k20a<1>(P, n_work, s);
    case 2: return nsc_rc_k20a<2>(P, n_work, s);
    case 3: return nsc_rc_k20a<3>(P, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
