// Dispatch of the null-space Cholesky build over stencil shape, dimension and
// operator columns (kernel template in build_nsc_kernel.h).
#include "build_nsc.h"

namespace rbfx {

hipError_t launch_nsc_k20a(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);
hipError_t launch_nsc_k20b(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);
hipError_t launch_nsc_k32(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);

// Built shapes (one operator column). Variant 0, 32 lanes per stencil with one row
// each: up to 20 points with up to 6 monomials (3 waves per SIMD), up to 32 points with
// up to 10 monomials (2 waves per SIMD). Variant 1, 8 lanes per stencil with three rows
// each: up to 20 points with up to 6 monomials (2 waves per SIMD). All compile without
// spilling; shapes with more columns or points, where the refinement step's working
// set spilled, are solved by the rook-pivoted kernels.
bool nsc_fits(int dim, int kmax, int m, int n_cols, int variant)
{
    if (dim < 1 || dim > 3 || n_cols != 1 || kmax < 1 || m < 0 || m >= kmax) return false;
    if (variant == 0) return kmax <= 32 && m <= 10;
    if (variant == 1) return kmax <= 20 && m <= (dim == 3 ? 4 : 6);
    return false;
}

hipError_t launch_build_nsc(const BuildDev& P, int dim, int kmax, int m, int variant, int64_t n_work, hipStream_t s)
{
    if (!nsc_fits(dim, kmax, m, P.n_cols, variant)) return hipErrorInvalidValue;
    // uniform builds only: the kernel takes every stencil's k and m from P.k and P.n_mono
    if (P.adj_ptr || P.deg_map || P.k != kmax || P.n_mono != m) return hipErrorInvalidValue;
    if (variant == 1) return launch_nsc_k20b(P, dim, n_work, s);
    if (kmax <= 20 && m <= 6) return launch_nsc_k20a(P, dim, n_work, s);
    return launch_nsc_k32(P, dim, n_work, s);
}

}  // namespace rbfx
