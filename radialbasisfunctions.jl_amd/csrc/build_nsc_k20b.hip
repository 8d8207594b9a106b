// Instantiates the null-space Cholesky build for stencils of up to 20 points and up to
// 6 monomials with 8 lanes of 3 rows per stencil: eight stencils per wavefront, 24 row
// slots for 20 rows (the 32-lane mapping leaves 12 idle) and 3-step group sums, at
// 2 waves per SIMD, one FP64 refinement step (a loop of steps would not fit its registers).
#include "build_nsc_kernel.h"

namespace rbfx {

// (in 3-D at most 6 monomials means degree <= 1, 4 monomials: the smaller table keeps
// the 3-D instantiation within its registers)
template <int DIM>
static hipError_t nsc_rc_k20b(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    return launch_nsc_one<8, 3, 20, DIM == 3 ? 4 : 6, DIM, 1, 1, 2>(P, n_work, s);
}

hipError_t launch_nsc_k20b(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    switch (dim) {
    case 1: return nsc_rc_k20b<1>(P, n_work, s);
    case 2: return nsc_rc_k20b<2>(P, n_work, s);
    case 3: return nsc_rc_k20b<3>(P, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
