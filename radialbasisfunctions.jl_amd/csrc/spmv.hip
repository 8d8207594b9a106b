// Multi-field CSR SpMV for gfx950: Y_v = alpha * W_v X + beta * Y_v.
//
// Replaces the apply path of RadialBasisOperator: _eval_op(op, x) = op.weights * x
// and mul!(y, op, x, alpha, beta) (src/operators/operators.jl:217-218, 529-542),
// which the reference runs as a serial SparseArrays CSC product. Here the weights
// are CSR (one row per stencil, so a row is a contiguous k-run of column indices and
// values) and several value arrays that share one stencil set (gradient components,
// Laplacian) are applied in one pass, so the index array is read once.
//
// Two mappings, both HBM-bound:
//   - many fields (F >= 8): one lane per (row, field); the F lanes of a row read the
//     same index/value (one cache line serves the group) and gather X rows
//     (point-major X makes each gather one contiguous F*8-byte segment);
//   - few fields: G lanes per row split the row's nonzeros, each keeps F partial
//     sums per value array, then a butterfly reduction inside the group.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spmv.h"

namespace rbfx {

// Workgroups are dealt to the 8 XCDs round-robin; remap so each XCD walks one
// contiguous range of (spatially ordered) rows and its own L2 serves their gathers.
// The grid is padded to a multiple of 8; blocks mapped past the end return -1.
__device__ __forceinline__ int64_t xcd_block(int64_t nb_real)
{
    const int64_t nb = gridDim.x;  // multiple of 8
    const int64_t b = blockIdx.x;
    const int64_t lb = (b & 7) * (nb >> 3) + (b >> 3);
    return lb < nb_real ? lb : -1;
}

__host__ inline int64_t pad8(int64_t nb) { return (nb + 7) / 8 * 8; }

template <int NV>
__global__ __launch_bounds__(256) void k_spmv_fields(const SpmvDev P, int G /* pow2 >= F */, int64_t nb_real)
{
    const int lane = threadIdx.x;
    const int f = lane % G;
    const int rows_per_block = blockDim.x / G;
    const int64_t lb = xcd_block(nb_real);
    if (lb < 0) return;
    const int64_t r = lb * rows_per_block + lane / G;
    if (r >= P.n_rows || f >= P.n_fields) return;
    const int64_t row = P.row_order ? (int64_t)P.row_order[r] : r;
    const int64_t b = P.row_ptr ? P.row_ptr[row] : row * (int64_t)P.k;
    const int64_t e = P.row_ptr ? P.row_ptr[row + 1] : b + P.k;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    const double* xf = P.x + (int64_t)f * P.x_ld_field;
    for (int64_t q = b; q < e; ++q) {
        const int64_t c = __builtin_nontemporal_load(P.col_idx + q);
        const double xv = xf[c * P.x_ld_point];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += __builtin_nontemporal_load(P.vals + (int64_t)v * P.nnz + q) * xv;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double* yp = P.y + (int64_t)v * P.y_ld_val + row * P.y_ld_point + (int64_t)f * P.y_ld_field;
        const double out = P.alpha * acc[v];
        *yp = (P.beta == 0.0) ? out : out + P.beta * *yp;
    }
}

template <int NV, int G, int FMAX>
__global__ __launch_bounds__(256) void k_spmv_rows(const SpmvDev P, int64_t nb_real)
{
    const int sub = threadIdx.x % G;
    const int64_t lb = xcd_block(nb_real);
    if (lb < 0) return;  // whole block: the butterfly below stays inside live blocks
    const int64_t r = (lb * blockDim.x + threadIdx.x) / G;
    const bool live = r < P.n_rows;
    const int64_t row = live ? (P.row_order ? (int64_t)P.row_order[r] : r) : 0;
    double acc[NV][FMAX];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int f = 0; f < FMAX; ++f) acc[v][f] = 0.0;
    if (live) {
        const int64_t b = P.row_ptr ? P.row_ptr[row] : row * (int64_t)P.k;
        const int64_t e = P.row_ptr ? P.row_ptr[row + 1] : b + P.k;
        for (int64_t q = b + sub; q < e; q += G) {
            const int64_t c = __builtin_nontemporal_load(P.col_idx + q);
            double xv[FMAX];
#pragma unroll
            for (int f = 0; f < FMAX; ++f) xv[f] = (f < P.n_fields) ? P.x[c * P.x_ld_point + (int64_t)f * P.x_ld_field] : 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double w = __builtin_nontemporal_load(P.vals + (int64_t)v * P.nnz + q);
#pragma unroll
                for (int f = 0; f < FMAX; ++f) acc[v][f] += w * xv[f];
            }
        }
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1)
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int f = 0; f < FMAX; ++f) acc[v][f] += __shfl_xor(acc[v][f], m, G);
    if (live && sub == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int f = 0; f < FMAX; ++f) {
                if (f < P.n_fields) {
                    double* yp = P.y + (int64_t)v * P.y_ld_val + row * P.y_ld_point + (int64_t)f * P.y_ld_field;
                    const double out = P.alpha * acc[v][f];
                    *yp = (P.beta == 0.0) ? out : out + P.beta * *yp;
                }
            }
    }
}

template <int NV>
static hipError_t spmv_nv(const SpmvDev& P, double avg_row, hipStream_t s)
{
    const int TB = 256;
    if (P.n_fields >= 8) {
        int G = 8;
        while (G < P.n_fields) G <<= 1;
        if (G > 64) return hipErrorInvalidValue;
        const int rows_per_block = TB / G;
        const int64_t grid = (P.n_rows + rows_per_block - 1) / rows_per_block;
        hipLaunchKernelGGL(k_spmv_fields<NV>, dim3((unsigned)pad8(grid)), dim3(TB), 0, s, P, G, grid);
        return hipGetLastError();
    }
    const int64_t threads_g4 = P.n_rows * 4, threads_g8 = P.n_rows * 8, threads_g16 = P.n_rows * 16;
#define RBFX_SPMV_LAUNCH(GG, FF, TH)                                                                   \
    {                                                                                                  \
        const int64_t nbr = ((TH) + TB - 1) / TB;                                                      \
        hipLaunchKernelGGL((k_spmv_rows<NV, GG, FF>), dim3((unsigned)pad8(nbr)), dim3(TB), 0, s, P, nbr); \
        return hipGetLastError();                                                                      \
    }
    if (P.n_fields <= 1) {
        if (avg_row < 12) { RBFX_SPMV_LAUNCH(4, 1, threads_g4) }
        if (avg_row < 40) { RBFX_SPMV_LAUNCH(8, 1, threads_g8) }
        RBFX_SPMV_LAUNCH(16, 1, threads_g16)
    }
    if (P.n_fields <= 4) {
        if (avg_row < 12) { RBFX_SPMV_LAUNCH(4, 4, threads_g4) }
        if (avg_row < 40) { RBFX_SPMV_LAUNCH(8, 4, threads_g8) }
        RBFX_SPMV_LAUNCH(16, 4, threads_g16)
    }
    if (avg_row < 40) { RBFX_SPMV_LAUNCH(8, 8, threads_g8) }
    RBFX_SPMV_LAUNCH(16, 8, threads_g16)
#undef RBFX_SPMV_LAUNCH
}

hipError_t launch_spmv(const SpmvDev& P, hipStream_t s)
{
    const double avg_row = P.n_rows > 0 ? (double)P.nnz / (double)P.n_rows : 0.0;
    switch (P.n_valarrays) {
    case 1: return spmv_nv<1>(P, avg_row, s);
    case 2: return spmv_nv<2>(P, avg_row, s);
    case 3: return spmv_nv<3>(P, avg_row, s);
    case 4: return spmv_nv<4>(P, avg_row, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
