// Internal interface between the C ABI layer and the build kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbfx.h"

namespace rbfx {

struct TermDev {
    int col, kind, d1, d2;
    double coeff;
};

// Device-side copy of rbfx_build_args plus work-list plumbing.
struct BuildDev {
    int64_t n_data, n_eval;
    const double* data;
    const double* eval;
    const int64_t* adj_ptr;
    const int32_t* adj_idx;
    int k;
    int basis_kind;
    double eps;
    int n_cols, n_terms;
    TermDev terms[RBFX_MAX_TERMS];
    int poly_deg;
    int n_mono;                 // binomial(dim + deg, dim), 0 for deg = -1
    int mono_e[RBFX_MAX_N][3];  // exponents of each monomial (graded order)
    const uint8_t* is_boundary;
    const int32_t* bnd_index;
    const double* bc_alpha;
    const double* bc_beta;
    const double* normals;
    const int64_t* out_ptr;
    int32_t* out_col;
    double* out_val;
    int64_t out_nnz;
    double* out_lambda;  // optional: n_cols * n_eval * n_mono multipliers
    int32_t* status;
    // optional work list (stencil ids) whose length lives in device memory
    const int32_t* list;
    const int32_t* list_len;
    // FP32 path: stencils that must be re-solved in FP64 are appended here
    int32_t* fail_list;
    int32_t* fail_count;
    int refine_steps;
    double accept_tol;
    int ok_status;  // status written for a stencil solved by this launch
};

// largest supported polynomial degree (any dimension), and the exponent table
constexpr int kMaxPolyDeg = 4;
int fill_mono_table(int dim, int deg, int (*e)[3]);

hipError_t launch_build(const BuildDev& P, int dim, int nmax, bool fp32, int64_t n_work, hipStream_t s);

}  // namespace rbfx
