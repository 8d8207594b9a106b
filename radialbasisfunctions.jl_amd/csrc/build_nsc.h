// Null-space Cholesky build: interface between the C ABI layer and the kernels of
// build_nsc*.hip.
#pragma once
#include "build.h"

namespace rbfx {

// Sign s with s * Z^T Phi Z positive definite on the null space Z of P^T (P the
// polynomial block), or 0 when the basis / degree pairing gives no such guarantee.
// PHS r^p with odd p is conditionally positive definite of order ceil(p/2) with sign
// (-1)^ceil(p/2): r (order 1, needs degree >= 0), r^3 (order 2, degree >= 1), r^5
// (order 3, degree >= 2), r^7 (order 4, degree >= 3); IMQ and Gaussian are positive
// definite for any degree.
__host__ __device__ inline int nsc_sign(int basis_kind, int poly_deg)
{
    switch (basis_kind) {
    case RBFX_PHS1: return poly_deg >= 0 ? -1 : 0;
    case RBFX_PHS3: return poly_deg >= 1 ? 1 : 0;
    case RBFX_PHS5: return poly_deg >= 2 ? -1 : 0;
    case RBFX_PHS7: return poly_deg >= 3 ? 1 : 0;
    case RBFX_IMQ:
    case RBFX_GAUSSIAN: return 1;
    default: return 0;
    }
}

// Whether a null-space kernel of the given variant is built for stencils of at most
// kmax points, m monomials (the largest over the launch; m < kmax) and n_cols operator
// columns in dim dimensions. Variant 0 (32 lanes per stencil) is built for one operator
// column, up to 32 points and 10 monomials; variant 1 (8 lanes of three rows per
// stencil) for one column, up to 20 points and 6 monomials.
bool nsc_fits(int dim, int kmax, int m, int n_cols, int variant);

// FP64 null-space build of the work (P.list / P.order / all stencils) of a uniform build
// (no adj_ptr, no degree map: every stencil has P.k points and P.n_mono monomials, with
// kmax = P.k and m = P.n_mono within nsc_fits). Stencils whose
// factorisation breaks down, whose refinement does not settle, or that the kernel does
// not serve (Hermite stencils are never passed here) are appended to P.fail_list for
// the rook-pivoted kernels; every other stencil gets status P.ok_status.
hipError_t launch_build_nsc(const BuildDev& P, int dim, int kmax, int m, int variant, int64_t n_work, hipStream_t s);

}  // namespace rbfx
