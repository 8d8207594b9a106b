// Internal interface between the C ABI layer and the build kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbfx.h"

namespace rbfx {

// Workgroups are dealt to the 8 XCDs round-robin. With a grid padded to a multiple of
// 8, this maps them so each XCD walks one contiguous eighth of the work (in spatial
// order, its L2 then serves the coordinate gathers of neighbouring stencils).
// Returns -1 for the padding blocks.
__device__ __forceinline__ int64_t xcd_contiguous_block(int64_t nb_real)
{
    const int64_t nb = gridDim.x;
    const int64_t b = blockIdx.x;
    const int64_t lb = (b & 7) * (nb >> 3) + (b >> 3);
    return lb < nb_real ? lb : -1;
}
__host__ __device__ inline int64_t pad_to_xcds(int64_t nb) { return (nb + 7) / 8 * 8; }

// largest supported polynomial degree (any dimension) and the longest monomial
// table (3-D, degree 4)
constexpr int kMaxPolyDeg = 4;
constexpr int kMaxMono = 35;

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
    int n_mono;               // binomial(dim + deg, dim), 0 for deg = -1
    int mono_e[kMaxMono][3];  // exponents of each monomial (graded order: the table of a
                              // lower degree is a prefix)
    const int8_t* deg_map;    // optional per-stencil degree (-1..poly_deg)
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
    // optional processing order of all n_eval stencils (used without a work list)
    const int32_t* order;
    // optional work list (stencil ids) whose length lives in device memory
    const int32_t* list;
    const int32_t* list_len;
    // FP32 path: stencils that must be re-solved in FP64 are appended here
    int32_t* fail_list;
    int32_t* fail_count;
    int refine_steps;
    double accept_tol;
    int ok_status;  // status written for a stencil solved by this launch
    int64_t n_blocks;  // workgroups with work (set by the launcher; the grid is padded)
    // optional: the offsets x_j - x_e of each work position's k points, [position][k][dim],
    // written by launch_pack_offsets for this launch's work order (fixed k, no Hermite
    // rows); the column indices are then already written too
    const double* packed;
};

int fill_mono_table(int dim, int deg, int (*e)[3]);

// number of monomials of total degree <= deg in dim variables (0 for deg = -1)
__host__ __device__ inline int n_mono_of(int dim, int deg)
{
    if (deg < 0) return 0;
    int num = 1, den = 1;
    for (int i = 1; i <= dim; ++i) {
        num *= deg + i;
        den *= i;
    }
    return num / den;
}

// monomial count of stencil e: its own degree when a degree map is given
__device__ __forceinline__ int stencil_n_mono(const BuildDev& P, int64_t e, int dim)
{
    if (!P.deg_map) return P.n_mono;
    const int d = P.deg_map[e];
    return (d >= -1 && d <= P.poly_deg) ? n_mono_of(dim, d) : -1;  // -1: invalid degree
}

// FP64 builds of stencils of up to RBFX_MAX_N rows with the system in LDS: the full
// matrix (k_build_wide) or its packed lower triangle (k_build_sym, half the LDS);
// rows_cap is the largest n in the launch
hipError_t launch_build_wide(const BuildDev& P, int dim, int rows_cap, int64_t n_work, hipStream_t s);
bool wide_fits(int rows_cap, int n_cols);
hipError_t launch_build_sym(const BuildDev& P, int dim, int rows_cap, int64_t n_work, hipStream_t s);
bool sym_fits(int rows_cap, int n_cols);

hipError_t launch_build(const BuildDev& P, int dim, int nmax, bool fp32, int64_t n_work, hipStream_t s);
// the offsets and column indices of every work position of P (its list, or order, or
// identity), for a build launch with P.packed = out; n_work bounds the positions (the
// list's device length is read by the kernel)
hipError_t launch_pack_offsets(const BuildDev& P, int dim, int64_t n_work, double* out, hipStream_t s);

}  // namespace rbfx
