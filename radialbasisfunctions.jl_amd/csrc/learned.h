// Internal interface of the learned weight path (surrogate inference + acceptance).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "build.h"
#include "mlp.h"

namespace rbfx {

// Canonical frame of a 2-D stencil: offsets xi_i = (x_i - x_e)/h, rotated so the first
// stencil point with |xi| > 0 lies on +x, then reflected (y -> -y) when the y offsets
// sum to a negative value. The Laplacian weights of the PHS + polynomial system are
// invariant under both maps, so the network sees one representative per orbit.
template <typename Real>
__host__ __device__ inline void canonical_rotation(Real x1, Real y1, Real& c, Real& s)
{
    const Real r = sqrt(x1 * x1 + y1 * y1);
    if (r > (Real)0) {
        c = x1 / r;
        s = y1 / r;
    } else {
        c = (Real)1;
        s = (Real)0;
    }
}

struct LearnedDev {
    int64_t n_eval;
    const double* data;
    const double* eval;
    const int32_t* adj_idx;  // n_eval * k (fixed stencil size)
    const int32_t* order;    // optional processing order of the stencils
    int k;
    int basis_kind;
    double eps;
    int n_mono;
    int mono_e[kMaxMono][3];
    TermDev terms[RBFX_MAX_TERMS];
    int n_terms;
    double op_scale_order;  // derivative order of the operator (weights scale as h^-order)
    // network
    const __bf16* wbf16;
    const float* params;  // biases are read from the flat FP32 array
    const double* out_mean;
    const double* out_std;
    double accept_tol;
    // outputs
    int32_t* out_col;
    double* out_val;
    int32_t* status;
    int32_t* reject_list;
    int32_t* reject_count;
    double* out_rho;  // optional per-stencil relative residual
    int64_t n_blocks;  // workgroups with work (set by the launcher; the grid is padded)
};

hipError_t launch_learned_build(const LearnedDev& P, hipStream_t s);

// canonical offsets (n_eval * k * 2 FP64), network inputs (n_eval * kInPad BF16: the
// 2k offsets' BF16 high parts, then their BF16 remainders, then zeros), the stencil
// radius h and the canonical reference distance r1
hipError_t launch_stencil_features(const double* data, const double* eval, int64_t n_eval, int k,
                                   const int32_t* adj_idx, double* out_xi, __bf16* out_x, double* out_h,
                                   double* out_r1, hipStream_t s);

}  // namespace rbfx
