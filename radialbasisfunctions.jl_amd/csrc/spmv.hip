// Multi-field CSR SpMV for gfx950: Y_v = alpha * W_v X + beta * Y_v.
//
// Replaces the apply path of RadialBasisOperator: _eval_op(op, x) = op.weights * x
// and mul!(y, op, x, alpha, beta) (src/operators/operators.jl:217-218, 529-542),
// which the reference runs as a serial SparseArrays CSC product. Here the weights
// are CSR (one row per stencil, so a row is a contiguous k-run of column indices and
// values) and several value arrays that share one stencil set (gradient components,
// Laplacian) are applied in one pass, so the index array is read once.
//
// Mappings, all HBM-bound:
//   - many fields (F >= 8): one lane per (row, field); the F lanes of a row read the
//     same index/value (one cache line serves the group) and gather X rows
//     (point-major X makes each gather one contiguous F*8-byte segment);
//   - few fields: G lanes per row split the row's nonzeros, each keeps F partial
//     sums per value array, then a reduction inside the group; rows of one length are
//     read with vector loads straight from HBM (k_spmv_vec), CSR rows of varying
//     length are staged through LDS (k_spmv_staged).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spmv.h"

namespace rbfx {

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

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

// Tiles of a launch: the grid (a multiple of 8 workgroups) is dealt to the 8 XCDs
// round-robin; workgroup b takes tiles of XCD (b % 8)'s contiguous eighth, every
// (grid / 8)-th one. With one workgroup per tile (the grid used here; fewer, longer-
// lived workgroups measured slower, since staging and compute then serialise) each
// workgroup runs one tile.
struct TileRange {
    int64_t first, end, stride;
};
__device__ __forceinline__ TileRange xcd_tile_range(int64_t n_tiles)
{
    const int64_t per_x = gridDim.x >> 3;
    const int64_t x = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int64_t tpx = (n_tiles + 7) >> 3;
    const int64_t lo = x * tpx;
    return {lo + j, lo + tpx < n_tiles ? lo + tpx : n_tiles, per_x};
}

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
    const int64_t srow = P.rows_permuted ? r : row;  // where the row's entries are stored
    const int64_t b = P.row_ptr ? P.row_ptr[srow] : srow * (int64_t)P.k;
    const int64_t e = P.row_ptr ? P.row_ptr[srow + 1] : b + P.k;
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
    const int64_t srow = P.rows_permuted ? r : row;  // where the row's entries are stored
    double acc[NV][FMAX];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int f = 0; f < FMAX; ++f) acc[v][f] = 0.0;
    if (live) {
        const int64_t b = P.row_ptr ? P.row_ptr[srow] : srow * (int64_t)P.k;
        const int64_t e = P.row_ptr ? P.row_ptr[srow + 1] : b + P.k;
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

// Staged mapping, for rows stored back to back in processing order: a workgroup owns
// R = 256/G consecutive rows, whose entries [b0, b1) it first copies into LDS with
// consecutive lanes on consecutive entries, so every load instruction of the index and
// value streams reads whole cache lines exactly once; then G lanes per row gather X
// (U independent gathers in flight per lane) and reduce as in k_spmv_rows.
// G lanes per row: gather X (U independent gathers in flight per lane), accumulate
// the row's products, butterfly-reduce inside the group and write Y. Entries are read
// from `col` / `val` (value array v at val + v * vstride) at offsets relative to b0.
// (IDX: the type of entry offsets, 32-bit inside a staged tile; UNITX: one field with
// unit point stride, so a gather is x[c])
template <int NV, int G, int FMAX, int U, bool UNITX, typename IDX, typename ColPtr, typename ValPtr>
__device__ __forceinline__ void rows_gather_reduce(const SpmvDev& P, int64_t r0, int64_t r1, int64_t b0, ColPtr col,
                                                   ValPtr val, int64_t vstride)
{
    const int sub = threadIdx.x % G;
    const int64_t r = r0 + threadIdx.x / G;
    const bool live = r < r1;
    const int64_t row = live ? (P.row_order ? (int64_t)P.row_order[r] : r) : 0;
    double acc[NV][FMAX];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int f = 0; f < FMAX; ++f) acc[v][f] = 0.0;
    if (live) {
        const IDX qb = (IDX)((P.row_ptr ? P.row_ptr[r] : r * (int64_t)P.k) - b0);
        const IDX qe = (IDX)((P.row_ptr ? P.row_ptr[r + 1] : (r + 1) * (int64_t)P.k) - b0);
        for (IDX q0 = qb + sub; q0 < qe; q0 += U * G) {
            double xv[U][FMAX], w[U][NV];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const IDX q = q0 + u * G;
                const bool ok = q < qe;
                const int32_t c = ok ? col[q] : 0;
                if constexpr (UNITX) {
                    xv[u][0] = ok ? P.x[c] : 0.0;
                } else {
#pragma unroll
                    for (int f = 0; f < FMAX; ++f)
                        xv[u][f] = (ok && f < P.n_fields) ? P.x[(int64_t)c * P.x_ld_point + (int64_t)f * P.x_ld_field]
                                                          : 0.0;
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) w[u][v] = ok ? val[v * vstride + q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int f = 0; f < FMAX; ++f) acc[v][f] += w[u][v] * xv[u][f];
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

// Staged mapping, for rows stored back to back in processing order: a workgroup owns
// R = 256/G consecutive rows, whose entries [b0, b1) it first copies into LDS with
// consecutive lanes on consecutive entries, so every load instruction of the index and
// value streams reads whole cache lines exactly once; then G lanes per row gather X.
// A workgroup whose entries exceed the LDS capacity (a longest-row bound that was too
// small) reads them from HBM instead.
template <int NV, int G, int FMAX, int U, bool UNITX>
__global__ __launch_bounds__(256) void k_spmv_staged(const SpmvDev P, int64_t nb_real, int cap)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    double* s_val = reinterpret_cast<double*>(s_raw);                        // NV x cap
    int32_t* s_col = reinterpret_cast<int32_t*>(s_val + (int64_t)NV * cap);  // cap
    constexpr int R = 256 / G;
    const int64_t lb = xcd_block(nb_real);
    if (lb < 0) return;  // whole block
    const int64_t r0 = lb * R;
    const int64_t r1 = r0 + R < P.n_rows ? r0 + R : P.n_rows;
    const int64_t b0 = P.row_ptr ? P.row_ptr[r0] : r0 * (int64_t)P.k;
    const int64_t b1 = P.row_ptr ? P.row_ptr[r1] : r1 * (int64_t)P.k;
    const int64_t n_e = b1 - b0;
    if (n_e > cap) {
        rows_gather_reduce<NV, G, FMAX, U, UNITX, int64_t>(P, r0, r1, 0, P.col_idx, P.vals, P.nnz);
        return;
    }
    // 16-byte loads (2 values / 4 indices per lane) where the run and both copies are
    // 16-byte aligned, else one element per lane
    bool vec = (n_e & 3) == 0 && (cap & 1) == 0 && ((uintptr_t)(P.col_idx + b0) & 15) == 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) vec = vec && ((uintptr_t)(P.vals + (int64_t)v * P.nnz + b0) & 15) == 0;
    if (vec) {
        const int n4 = (int)(n_e >> 2), n2 = (int)(n_e >> 1);
        const i32x4* c4 = reinterpret_cast<const i32x4*>(P.col_idx + b0);
        i32x4* sc4 = reinterpret_cast<i32x4*>(s_col);
        for (int i = threadIdx.x; i < n4; i += 256) sc4[i] = __builtin_nontemporal_load(c4 + i);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const f64x2* v2 = reinterpret_cast<const f64x2*>(P.vals + (int64_t)v * P.nnz + b0);
            f64x2* sv2 = reinterpret_cast<f64x2*>(s_val + (int64_t)v * cap);
            for (int i = threadIdx.x; i < n2; i += 256) sv2[i] = __builtin_nontemporal_load(v2 + i);
        }
    } else {
        for (int i = threadIdx.x; i < n_e; i += 256) {
            s_col[i] = __builtin_nontemporal_load(P.col_idx + b0 + i);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                s_val[v * cap + i] = __builtin_nontemporal_load(P.vals + (int64_t)v * P.nnz + b0 + i);
        }
    }
    __syncthreads();
    rows_gather_reduce<NV, G, FMAX, U, UNITX, int>(P, r0, r1, b0, (const int32_t*)s_col, (const double*)s_val,
                                                   (int64_t)cap);
}

// Many fields (F >= 8): LPR lanes per row, each owning FPL consecutive fields (FPL = 2:
// 16-byte gathers and stores, when X and Y hold a point's fields contiguously), so a
// row's gather of one entry reads its F values from one or two cache lines. Each lane
// walks its row's entries with U independent gathers in flight and needs no reduction.
template <int NV, int LPR, int FPL, int U, typename ColPtr, typename ValPtr>
__device__ __forceinline__ void fields_rows(const SpmvDev& P, int64_t r0, int64_t r1, int64_t b0, ColPtr col,
                                            ValPtr val, int64_t vstride)
{
    const int fl = threadIdx.x % LPR;
    const int f0 = fl * FPL;
    const int64_t r = r0 + threadIdx.x / LPR;
    if (r >= r1 || f0 >= P.n_fields) return;
    const int64_t row = P.row_order ? (int64_t)P.row_order[r] : r;
    const int64_t qb = (P.row_ptr ? P.row_ptr[r] : r * (int64_t)P.k) - b0;
    const int64_t qe = (P.row_ptr ? P.row_ptr[r + 1] : (r + 1) * (int64_t)P.k) - b0;
    const double* xf = P.x + (int64_t)f0 * P.x_ld_field;
    double acc[NV][FPL];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int j = 0; j < FPL; ++j) acc[v][j] = 0.0;
    for (int64_t q0 = qb; q0 < qe; q0 += U) {
        double xv[U][FPL], w[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = q0 + u;
            const bool ok = q < qe;
            const int64_t c = ok ? (int64_t)col[q] : 0;
            if constexpr (FPL == 2) {
                const f64x2 t = *reinterpret_cast<const f64x2*>(xf + c * P.x_ld_point);
                xv[u][0] = ok ? t.x : 0.0;
                xv[u][1] = ok ? t.y : 0.0;
            } else {
                xv[u][0] = ok ? xf[c * P.x_ld_point] : 0.0;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) w[u][v] = ok ? val[v * vstride + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < FPL; ++j) acc[v][j] += w[u][v] * xv[u][j];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double* yp = P.y + (int64_t)v * P.y_ld_val + row * P.y_ld_point + (int64_t)f0 * P.y_ld_field;
        if constexpr (FPL == 2) {
            f64x2 out = {P.alpha * acc[v][0], P.alpha * acc[v][1]};
            if (P.beta != 0.0) {
                const f64x2 old = *reinterpret_cast<const f64x2*>(yp);
                out.x += P.beta * old.x;
                out.y += P.beta * old.y;
            }
            *reinterpret_cast<f64x2*>(yp) = out;
        } else {
            const double out = P.alpha * acc[v][0];
            *yp = (P.beta == 0.0) ? out : out + P.beta * *yp;
        }
    }
}

// The workgroup's rows (256 / LPR, stored back to back) first have their index and
// value runs copied into LDS, as in k_spmv_staged; a workgroup whose entries exceed
// the LDS it was given (a longest-row bound that was too small) reads HBM instead.
template <int NV, int LPR, int FPL, int U>
__global__ __launch_bounds__(256) void k_spmv_fields_staged(const SpmvDev P, int64_t nb_real, int cap)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    double* s_val = reinterpret_cast<double*>(s_raw);                        // NV x cap
    int32_t* s_col = reinterpret_cast<int32_t*>(s_val + (int64_t)NV * cap);  // cap
    constexpr int R = 256 / LPR;
    const TileRange tr = xcd_tile_range(nb_real);
    for (int64_t lb = tr.first; lb < tr.end; lb += tr.stride) {  // uniform over the workgroup
        const int64_t r0 = lb * R;
        const int64_t r1 = r0 + R < P.n_rows ? r0 + R : P.n_rows;
        const int64_t b0 = P.row_ptr ? P.row_ptr[r0] : r0 * (int64_t)P.k;
        const int64_t b1 = P.row_ptr ? P.row_ptr[r1] : r1 * (int64_t)P.k;
        const int64_t n_e = b1 - b0;
        if (n_e > cap) {
            fields_rows<NV, LPR, FPL, U>(P, r0, r1, 0, P.col_idx, P.vals, P.nnz);
            continue;  // LDS untouched in this tile
        }
        for (int i = threadIdx.x; i < n_e; i += 256) {
            s_col[i] = __builtin_nontemporal_load(P.col_idx + b0 + i);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                s_val[v * cap + i] = __builtin_nontemporal_load(P.vals + (int64_t)v * P.nnz + b0 + i);
        }
        __syncthreads();
        fields_rows<NV, LPR, FPL, U>(P, r0, r1, b0, (const int32_t*)s_col, (const double*)s_val, (int64_t)cap);
        __syncthreads();  // the next tile overwrites the LDS copy
    }
}

// Source-tile mapping (many fields, rows stored in processing order with a source-tile
// plan): a workgroup owns one tile of consecutive rows. It copies into LDS, with
// coalesced loads, the tile's distinct column ids, its entries (local column positions
// and values) and then the X rows of those columns (every 16-byte piece a thread loads
// issued before any is stored, so up to kTilePieces loads per thread are in flight); then
// LPR lanes per row, FPL fields each, walk the row's entries entirely from LDS. A column
// that several rows of the tile share (a source every ~60 targets of a regrid use)
// crosses the cache hierarchy once per tile instead of once per entry. The entry order
// of every row is the other mappings' own, so the sums are bit-identical to theirs.
template <int NV, int LPR, int FPL, int U>
__device__ __forceinline__ void tiled_rows(const SpmvDev& P, int64_t r0, int64_t r1, int64_t b0, const double* s_x,
                                           const double* s_w, const uint16_t* s_l, int cap_e)
{
    const int F = P.n_fields;
    const int fl = threadIdx.x % LPR;
    const int f0 = fl * FPL;
    const int64_t r = r0 + threadIdx.x / LPR;
    if (r >= r1 || f0 >= F) return;
    const int64_t row = P.row_order ? (int64_t)P.row_order[r] : r;
    const int qb = (int)((P.row_ptr ? P.row_ptr[r] : r * (int64_t)P.k) - b0);
    const int qe = (int)((P.row_ptr ? P.row_ptr[r + 1] : (r + 1) * (int64_t)P.k) - b0);
    double acc[NV][FPL];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int j = 0; j < FPL; ++j) acc[v][j] = 0.0;
    for (int q0 = qb; q0 < qe; q0 += U) {
        double xv[U][FPL], w[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u;
            const bool ok = q < qe;
            const int l = ok ? (int)s_l[q] : 0;
            if constexpr (FPL == 2) {
                const f64x2 t = *reinterpret_cast<const f64x2*>(s_x + l * F + f0);
                xv[u][0] = ok ? t.x : 0.0;
                xv[u][1] = ok ? t.y : 0.0;
            } else {
                xv[u][0] = ok ? s_x[l * F + f0] : 0.0;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) w[u][v] = ok ? s_w[v * cap_e + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < FPL; ++j) acc[v][j] += w[u][v] * xv[u][j];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double* yp = P.y + (int64_t)v * P.y_ld_val + row * P.y_ld_point + (int64_t)f0 * P.y_ld_field;
        if constexpr (FPL == 2) {
            f64x2 out = {P.alpha * acc[v][0], P.alpha * acc[v][1]};
            if (P.beta != 0.0) {
                const f64x2 old = *reinterpret_cast<const f64x2*>(yp);
                out.x += P.beta * old.x;
                out.y += P.beta * old.y;
            }
            *reinterpret_cast<f64x2*>(yp) = out;
        } else {
            const double out = P.alpha * acc[v][0];
            *yp = (P.beta == 0.0) ? out : out + P.beta * *yp;
        }
    }
}

// X bytes a source-tile workgroup stages, and the 16-byte pieces of it per thread
constexpr int kTileXBytes = 32768;
constexpr int kTilePieces = kTileXBytes / (16 * 256);

// LDS of a source-tile workgroup: X rows (cap_src x F doubles), values (NV x cap_e),
// local column positions (cap_e) and column ids (cap_src)
static size_t tiled_lds(int nv, int n_fields, int cap_src, int cap_e)
{
    return (size_t)cap_src * n_fields * 8 + (size_t)nv * cap_e * 8 + (size_t)((cap_e * 2 + 15) / 16 * 16) +
           (size_t)cap_src * 4;
}

// (FPL = 2 needs fields_pair: X rows 16-byte aligned with the fields contiguous; a tile
// whose distinct columns or entries exceed the LDS it was given gathers from X directly)
// (one value array: with several, the kernel spills at any occupancy, and those applies
// keep the other mappings. 3 waves per SIMD: three workgroups per CU, which their LDS
// allows; left to itself the compiler spends 256 VGPRs on the unrolled copy and runs one.)
template <int NV, int LPR, int FPL, int U>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(3))) void
k_spmv_fields_tiled(
    const SpmvDev P, int64_t n_tiles, int cap_src, int cap_e)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    const int F = P.n_fields;
    double* s_x = reinterpret_cast<double*>(s_raw);                                            // cap_src x F
    double* s_w = s_x + (int64_t)cap_src * F;                                                  // NV x cap_e
    uint16_t* s_l = reinterpret_cast<uint16_t*>(s_w + (int64_t)NV * cap_e);                     // cap_e
    int32_t* s_src = reinterpret_cast<int32_t*>(reinterpret_cast<unsigned char*>(s_l) + (cap_e * 2 + 15) / 16 * 16);
    constexpr int RP = 256 / LPR;  // rows per pass of the workgroup
    const TileRange tr = xcd_tile_range(n_tiles);
    for (int64_t t = tr.first; t < tr.end; t += tr.stride) {  // uniform over the workgroup
        const int64_t r0 = t * P.tile_rows;
        const int64_t r1 = r0 + P.tile_rows < P.n_rows ? r0 + P.tile_rows : P.n_rows;
        const int64_t b0 = P.row_ptr ? P.row_ptr[r0] : r0 * (int64_t)P.k;
        const int n_e = (int)((P.row_ptr ? P.row_ptr[r1] : r1 * (int64_t)P.k) - b0);
        const int ns = P.tile_src_n[t];
        if (ns > cap_src || n_e > cap_e) {
            for (int64_t ra = r0; ra < r1; ra += RP)
                fields_rows<NV, LPR, FPL, U>(P, ra, ra + RP < r1 ? ra + RP : r1, 0, P.col_idx, P.vals, P.nnz);
            continue;  // LDS untouched in this tile
        }
        const int32_t* src = P.tile_src + P.tile_src_off[t];
        for (int i = threadIdx.x; i < ns; i += 256) s_src[i] = src[i];
        for (int i = threadIdx.x; i < n_e; i += 256) {
            s_l[i] = P.tile_col[b0 + i];
#pragma unroll
            for (int v = 0; v < NV; ++v) s_w[v * cap_e + i] = __builtin_nontemporal_load(P.vals + (int64_t)v * P.nnz + b0 + i);
        }
        __syncthreads();
        const int FP = FPL == 2 ? F >> 1 : F;  // pieces per X row (16 or 8 bytes each)
        const int n_p = ns * FP;               // <= kTilePieces * 256 (cap_src is sized so)
        if constexpr (FPL == 2) {
            f64x2 tmp[kTilePieces];
#pragma unroll
            for (int p = 0; p < kTilePieces; ++p) {
                const int i = threadIdx.x + p * 256;
                if (i < n_p) {
                    const int si = i / FP, f2 = i - si * FP;
                    tmp[p] = *reinterpret_cast<const f64x2*>(P.x + (int64_t)s_src[si] * P.x_ld_point + 2 * f2);
                }
            }
#pragma unroll
            for (int p = 0; p < kTilePieces; ++p) {
                const int i = threadIdx.x + p * 256;
                if (i < n_p) {
                    const int si = i / FP, f2 = i - si * FP;
                    *reinterpret_cast<f64x2*>(s_x + si * F + 2 * f2) = tmp[p];
                }
            }
        } else {
            for (int i = threadIdx.x; i < n_p; i += 256) {
                const int si = i / F, f = i - si * F;
                s_x[si * F + f] = P.x[(int64_t)s_src[si] * P.x_ld_point + (int64_t)f * P.x_ld_field];
            }
        }
        __syncthreads();
        for (int64_t ra = r0; ra < r1; ra += RP)
            tiled_rows<NV, LPR, FPL, U>(P, ra, ra + RP < r1 ? ra + RP : r1, b0, s_x, s_w, s_l, cap_e);
        __syncthreads();  // the next tile overwrites the LDS copies
    }
}

// One workgroup per tile: the tile's (column, entry) pairs sorted by column in LDS
// (bitonic, NE a power of two >= the tile's entries; padding sorts last), the first of
// every run of equal columns flagged, an inclusive scan of the flags giving each entry
// its column's position among the tile's distinct columns.
template <int NE>
__global__ __launch_bounds__(256) void k_source_tiles(int64_t n_rows, int k, const int64_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col, int tile_rows, int phase,
                                                      int32_t* __restrict__ tile_n,
                                                      const int64_t* __restrict__ tile_off,
                                                      int32_t* __restrict__ tile_src, uint16_t* __restrict__ tile_col)
{
    __shared__ int32_t s_key[NE];
    __shared__ uint16_t s_id[NE];
    __shared__ int32_t s_inc[NE];
    __shared__ int32_t s_part[256];
    constexpr int PER = NE / 256;
    const int64_t t = blockIdx.x;
    const int64_t r0 = t * tile_rows;
    const int64_t r1 = r0 + tile_rows < n_rows ? r0 + tile_rows : n_rows;
    const int64_t b0 = row_ptr ? row_ptr[r0] : r0 * (int64_t)k;
    const int64_t b1 = row_ptr ? row_ptr[r1] : r1 * (int64_t)k;
    const int n_e = (int)(b1 - b0);  // <= NE: the host checks the longest tile
    for (int i = threadIdx.x; i < NE; i += 256) {
        s_key[i] = i < n_e ? col[b0 + i] : INT32_MAX;
        s_id[i] = (uint16_t)i;
    }
    __syncthreads();
    for (int size = 2; size <= NE; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < NE / 2; i += 256) {
                const int lo = 2 * stride * (i / stride) + (i % stride);
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const int32_t a = s_key[lo], b = s_key[hi];
                if ((a > b) == up) {
                    s_key[lo] = b;
                    s_key[hi] = a;
                    const uint16_t ia = s_id[lo];
                    s_id[lo] = s_id[hi];
                    s_id[hi] = ia;
                }
            }
            __syncthreads();
        }
    }
    // flags (first of a run of equal valid columns), scanned: each thread PER consecutive
    int run = 0;
    const int i0 = threadIdx.x * PER;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = i0 + j;
        run += (i < n_e && (i == 0 || s_key[i] != s_key[i - 1])) ? 1 : 0;
        s_inc[i] = run;
    }
    s_part[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int j = 0; j < 256; ++j) {
            const int v = s_part[j];
            s_part[j] = acc;
            acc += v;
        }
    }
    __syncthreads();
    const int base = s_part[threadIdx.x];
#pragma unroll
    for (int j = 0; j < PER; ++j) s_inc[i0 + j] += base;
    __syncthreads();
    if (phase == 0) {
        if (threadIdx.x == 0) tile_n[t] = n_e > 0 ? s_inc[n_e - 1] : 0;
        return;
    }
    const int64_t off = tile_off[t];
    for (int i = threadIdx.x; i < n_e; i += 256) {
        const int lid = s_inc[i] - 1;
        tile_col[b0 + s_id[i]] = (uint16_t)lid;
        if (i == 0 || s_key[i] != s_key[i - 1]) tile_src[off + lid] = s_key[i];
    }
}

hipError_t launch_source_tiles(int64_t n_rows, int k, const int64_t* row_ptr, const int32_t* col, int tile_rows,
                               int phase, int32_t* tile_n, const int64_t* tile_off, int32_t* tile_src,
                               uint16_t* tile_col, hipStream_t s)
{
    const int64_t n_tiles = (n_rows + tile_rows - 1) / tile_rows;
    if (n_tiles <= 0) return hipSuccess;
    const int64_t most = row_ptr ? (int64_t)kTileEntries : (int64_t)tile_rows * k;  // (row_ptr: checked by the caller)
    if (most <= 2048)
        hipLaunchKernelGGL(k_source_tiles<2048>, dim3((unsigned)n_tiles), dim3(256), 0, s, n_rows, k, row_ptr, col,
                           tile_rows, phase, tile_n, tile_off, tile_src, tile_col);
    else
        hipLaunchKernelGGL(k_source_tiles<kTileEntries>, dim3((unsigned)n_tiles), dim3(256), 0, s, n_rows, k, row_ptr,
                           col, tile_rows, phase, tile_n, tile_off, tile_src, tile_col);
    return hipGetLastError();
}

// LDS bytes of one staged workgroup
static int64_t staged_lds(int nv, int G, int64_t max_row) { return (256 / G) * max_row * (8 * (int64_t)nv + 4); }
constexpr int64_t kStagedLdsMax = 48 * 1024;

template <int NV, int FMAX, bool UNITX>
static hipError_t spmv_staged(const SpmvDev& P, hipStream_t s)
{
    const int64_t L = P.max_row;
#define RBFX_STAGED(GG, UU)                                                                                      \
    {                                                                                                            \
        const int64_t nbr = (P.n_rows + 256 / (GG) - 1) / (256 / (GG));                                          \
        const int cap = (int)((256 / (GG)) * L);                                                                 \
        hipLaunchKernelGGL((k_spmv_staged<NV, GG, FMAX, UU, UNITX>), dim3((unsigned)pad8(nbr)), dim3(256),   \
                           (unsigned)staged_lds(NV, GG, L), s, P, nbr, cap);                                     \
        return hipGetLastError();                                                                                \
    }
    if (L <= 16) RBFX_STAGED(4, 4)
    if (L <= 20) RBFX_STAGED(4, 5)
    if (L <= 32) RBFX_STAGED(8, 4)
    if (L <= 40) RBFX_STAGED(8, 5)
    RBFX_STAGED(16, 4)
#undef RBFX_STAGED
}

// lanes per row of the staged mapping for the longest row L (as chosen in spmv_staged)
static int staged_g(int64_t L) { return L <= 20 ? 4 : (L <= 40 ? 8 : 16); }

// Vector mapping, for rows of one length k (no row_ptr) and few fields: G = ceil(k / E)
// lanes per row, each owning E consecutive entries, which it reads with one 4E-byte
// index load and E/2 16-byte value loads per value array (E = 4 or 2; E = 1: scalar
// loads). A wavefront holds U units of floor(64 / G) consecutive rows (U = 2 for
// k = 20, else 1), so its loads cover one contiguous run of rows and read every byte
// of it once, without an LDS round trip; each lane has the loads of its U rows in
// flight together. The G partial sums of a row are added by the row's first lane.
typedef int i32x2 __attribute__((ext_vector_type(2)));

template <int NV, int E, int FMAX, bool UNITX, bool PERM, int GS, int U>
__global__ __launch_bounds__(256) void k_spmv_vec(const SpmvDev P, int G_rt, int64_t nb_real)
{
    const int G = GS > 0 ? GS : G_rt;  // lanes per row (a constant for k = 20)
    const int64_t lb = xcd_block(nb_real);
    if (lb < 0) return;
    const int lane = threadIdx.x & 63;
    const int rw = 64 / G;  // rows per wavefront and unit
    const int rl = lane / G, sub = lane - rl * G;
    // a wavefront takes U units of rw consecutive rows: unit u holds rows r0 + u * rw + [0, rw)
    const int64_t r0 = (lb * 4 + (threadIdx.x >> 6)) * (int64_t)(U * rw);
    const int k = P.k;
    const int64_t last = P.n_rows - 1;
    // rows past the end (and the lanes past the wavefront's last row) read row n_rows - 1,
    // so every load below is in bounds and issued unconditionally; they write nothing
    int64_t r[U];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t ru = r0 + u * rw + rl;
        live[u] = rl < rw && ru <= last;
        r[u] = live[u] ? ru : last;
    }
    const int j0 = sub * E;
    int c[U][E];
    double w[U][NV][E];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        // PERM: the entries are stored in processing order, so their addresses do not
        // wait for the row_order read (done after the gathers are issued)
        const int64_t srow = PERM ? r[u] : (P.row_order ? (int64_t)P.row_order[r[u]] : r[u]);
        const int64_t q = srow * (int64_t)k + j0;
        if constexpr (E == 4) {
            const i32x4 c4 = *reinterpret_cast<const i32x4*>(P.col_idx + q);
            c[u][0] = c4.x; c[u][1] = c4.y; c[u][2] = c4.z; c[u][3] = c4.w;
        } else if constexpr (E == 2) {
            const i32x2 c2 = *reinterpret_cast<const i32x2*>(P.col_idx + q);
            c[u][0] = c2.x; c[u][1] = c2.y;
        } else {
            c[u][0] = P.col_idx[q];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* vp = P.vals + (int64_t)v * P.nnz + q;
            if constexpr (E >= 2) {
#pragma unroll
                for (int h = 0; h < E / 2; ++h) {
                    const f64x2 t = *reinterpret_cast<const f64x2*>(vp + 2 * h);
                    w[u][v][2 * h] = t.x;
                    w[u][v][2 * h + 1] = t.y;
                }
            } else {
                w[u][v][0] = vp[0];
            }
        }
    }
    // E divides k (E = 1, or k a multiple of 2 or 4), so every lane's E entries lie in
    // its row: the gathers of all U units are issued unconditionally, all in flight at once
    double xv[U][E][FMAX];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if constexpr (UNITX) {
                xv[u][e][0] = P.x[c[u][e]];
            } else {
#pragma unroll
                for (int f = 0; f < FMAX; ++f)
                    xv[u][e][f] = f < P.n_fields ? P.x[(int64_t)c[u][e] * P.x_ld_point + (int64_t)f * P.x_ld_field] : 0.0;
            }
        }
    double acc[U][NV][FMAX];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int f = 0; f < FMAX; ++f) {
                double t = 0.0;
#pragma unroll
                for (int e = 0; e < E; ++e) t += w[u][v][e] * xv[u][e][f];
                acc[u][v][f] = t;
            }
    // segmented tree sum over the row's G lanes: at distance o = 1, 2, 4, ... each lane
    // adds the partial of the lane o places on when that lane is in the same row, so
    // the row's first lane ends with the whole sum (lanes past the last row of the
    // wavefront read wrapped lanes; their sums are not used)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        if (o >= G) break;  // uniform
        const bool in = sub + o < G;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int f = 0; f < FMAX; ++f) {
                    const double t = __shfl_down(acc[u][v][f], o, 64);
                    acc[u][v][f] += in ? t : 0.0;
                }
    }
    if (sub != 0) return;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!live[u]) continue;
        const int64_t row = P.row_order ? (int64_t)P.row_order[r[u]] : r[u];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int f = 0; f < FMAX; ++f) {
                if (f < P.n_fields) {
                    double* yp = P.y + (int64_t)v * P.y_ld_val + row * P.y_ld_point + (int64_t)f * P.y_ld_field;
                    const double out = P.alpha * acc[u][v][f];
                    *yp = (P.beta == 0.0) ? out : out + P.beta * *yp;
                }
            }
    }
}

// entries per lane of the vector mapping (0: the mapping does not serve this apply):
// rows of one length whose runs keep the index and value loads aligned
int vec_entries(const SpmvDev& P, int nv)
{
    if (P.row_ptr || P.k < 1 || P.n_fields >= 8) return 0;
    bool a16 = ((uintptr_t)P.col_idx & 15) == 0, a8 = ((uintptr_t)P.col_idx & 7) == 0;
    bool v16 = true;
    for (int v = 0; v < nv; ++v) v16 = v16 && ((uintptr_t)(P.vals + (int64_t)v * P.nnz) & 15) == 0;
    int e = 1;
    if (P.k % 4 == 0 && a16 && v16) e = 4;
    else if (P.k % 2 == 0 && a8 && v16) e = 2;
    return (P.k + e - 1) / e <= 64 ? e : 0;
}

#ifndef RBFX_VEC_UNITS
#define RBFX_VEC_UNITS 2
#endif
// units of rows per wavefront in the vector mapping: each lane holds the loads of U
// rows in flight at once
constexpr int kVecUnits = RBFX_VEC_UNITS;

template <int NV, int FMAX, bool UNITX, bool PERM>
static hipError_t spmv_vec_perm(const SpmvDev& P, int e, hipStream_t s)
{
    if (P.n_rows <= 0) return hipSuccess;  // the kernel reads row n_rows - 1
    const int G = (P.k + e - 1) / e;
    const int U = (e == 4 && G == 5) ? kVecUnits : 1;
    const int64_t rows_per_block = 4 * U * (64 / G);
    const int64_t nbr = (P.n_rows + rows_per_block - 1) / rows_per_block;
    const dim3 grid((unsigned)pad8(nbr));
    if (e == 4 && G == 5)
        hipLaunchKernelGGL((k_spmv_vec<NV, 4, FMAX, UNITX, PERM, 5, kVecUnits>), grid, dim3(256), 0, s, P, G, nbr);
    else if (e == 4)
        hipLaunchKernelGGL((k_spmv_vec<NV, 4, FMAX, UNITX, PERM, 0, 1>), grid, dim3(256), 0, s, P, G, nbr);
    else if (e == 2)
        hipLaunchKernelGGL((k_spmv_vec<NV, 2, FMAX, UNITX, PERM, 0, 1>), grid, dim3(256), 0, s, P, G, nbr);
    else
        hipLaunchKernelGGL((k_spmv_vec<NV, 1, FMAX, UNITX, PERM, 0, 1>), grid, dim3(256), 0, s, P, G, nbr);
    return hipGetLastError();
}

template <int NV, int FMAX, bool UNITX>
static hipError_t spmv_vec(const SpmvDev& P, int e, hipStream_t s)
{
    return (P.rows_permuted || !P.row_order) ? spmv_vec_perm<NV, FMAX, UNITX, true>(P, e, s)
                                             : spmv_vec_perm<NV, FMAX, UNITX, false>(P, e, s);
}

// many-field layout: two fields per lane when X and Y hold each point's fields
// contiguously and 16-byte aligned (and F is even); lanes per row a power of 2 >= 4
static bool fields_pair(const SpmvDev& P)
{
    return (P.n_fields % 2) == 0 && P.x_ld_field == 1 && (P.x_ld_point % 2) == 0 && ((uintptr_t)P.x & 15) == 0 &&
           P.y_ld_field == 1 && (P.y_ld_point % 2) == 0 && (P.y_ld_val % 2) == 0 && ((uintptr_t)P.y & 15) == 0;
}
static int fields_lpr(const SpmvDev& P)
{
    const int fpl = fields_pair(P) ? 2 : 1;
    int lpr = 4;
    while (lpr * fpl < P.n_fields) lpr <<= 1;
    return lpr;
}

bool spmv_staged_fits(const SpmvDev& P, int nv)
{
    if ((P.row_order && !P.rows_permuted) || P.max_row <= 0 || P.nnz <= 0 || P.n_fields > 64) return false;
    const int lanes = P.n_fields >= 8 ? fields_lpr(P) : staged_g(P.max_row);
    return (256 / lanes) * P.max_row * (8 * (int64_t)nv + 4) <= kStagedLdsMax;
}

template <int NV>
static hipError_t spmv_nv(const SpmvDev& P, double avg_row, hipStream_t s)
{
    const int TB = 256;
    // vector mapping: rows of one length, few fields (the automatic choice for them)
    const int ve = vec_entries(P, NV);
    if (P.mapping == 3 && !ve) return hipErrorInvalidValue;
    if (ve && (P.mapping == 0 || P.mapping == 3)) {
        if (P.n_fields <= 1) return P.x_ld_point == 1 ? spmv_vec<NV, 1, true>(P, ve, s) : spmv_vec<NV, 1, false>(P, ve, s);
        if (P.n_fields <= 4) return spmv_vec<NV, 4, false>(P, ve, s);
        return spmv_vec<NV, 8, false>(P, ve, s);
    }
    // staged mapping: rows stored in processing order, a known longest row, LDS in budget
    const bool staged_ok = P.n_fields < 8 && spmv_staged_fits(P, NV);
    if (P.mapping == 2 && P.n_fields < 8 && !staged_ok) return hipErrorInvalidValue;
    if (staged_ok && P.mapping != 1) {
        if (P.n_fields <= 1)
            return P.x_ld_point == 1 ? spmv_staged<NV, 1, true>(P, s) : spmv_staged<NV, 1, false>(P, s);
        if (P.n_fields <= 4) return spmv_staged<NV, 4, false>(P, s);
        return spmv_staged<NV, 8, false>(P, s);
    }
    if (P.mapping == 4 && !(NV == 1 && P.tile_col && P.n_fields >= 8)) return hipErrorInvalidValue;
    if (P.n_fields >= 8) {
        int G = 8;
        while (G < P.n_fields) G <<= 1;
        if (G > 64) return hipErrorInvalidValue;
        const bool pair = fields_pair(P);
        const int lpr = fields_lpr(P);
        // source tiles, when asked for: measured on the 16-field 8M-row regrid at 2.84 ms
        // against 1.73 ms for the staged mapping below (each tile's staging latency and
        // three LDS reads per entry and lane outweigh the 5.7x fewer X gathers), so the
        // automatic choice stays the staged one
        if constexpr (NV == 1) if (P.tile_col && P.mapping == 4) {
            const int64_t n_tiles = (P.n_rows + P.tile_rows - 1) / P.tile_rows;
            const int cap = kTileXBytes / (8 * P.n_fields);
            const int cap_e = tile_entries_lds(NV);
            const size_t lds = tiled_lds(NV, P.n_fields, cap, cap_e);
#define RBFX_FTILED(LL, FF)                                                                                        \
    {                                                                                                              \
        hipLaunchKernelGGL((k_spmv_fields_tiled<NV, LL, FF, 4>), dim3((unsigned)pad8(n_tiles)), dim3(256), lds, s, P, \
                           n_tiles, cap, cap_e);                                                                   \
        return hipGetLastError();                                                                                  \
    }
            if (pair) {
                if (lpr == 4) RBFX_FTILED(4, 2)
                if (lpr == 8) RBFX_FTILED(8, 2)
                if (lpr == 16) RBFX_FTILED(16, 2)
                RBFX_FTILED(32, 2)
            }
            if (lpr == 8) RBFX_FTILED(8, 1)
            if (lpr == 16) RBFX_FTILED(16, 1)
            if (lpr == 32) RBFX_FTILED(32, 1)
            RBFX_FTILED(64, 1)
#undef RBFX_FTILED
        }
        const int64_t cap = (256 / lpr) * P.max_row;
        const bool fstaged = spmv_staged_fits(P, NV);
        if (P.mapping == 2 && !fstaged) return hipErrorInvalidValue;
        if (fstaged && P.mapping != 1) {
            const int64_t nbr = (P.n_rows + 256 / lpr - 1) / (256 / lpr);
            const size_t lds = (size_t)cap * (8 * NV + 4);
            // 6 gathers in flight per lane: 16-field k = 30 regrid 1.78 ms, against 1.83 ms
            // for 10, 1.81 for 5, 1.86 for 4, 1.94 for 15
#define RBFX_FSTAGED(LL, FF)                                                                                      \
    {                                                                                                             \
        hipLaunchKernelGGL((k_spmv_fields_staged<NV, LL, FF, 6>), dim3((unsigned)pad8(nbr)), dim3(256), lds, s, P, \
                           nbr, (int)cap);                                                                        \
        return hipGetLastError();                                                                                 \
    }
            if (pair) {
                if (lpr == 4) RBFX_FSTAGED(4, 2)
                if (lpr == 8) RBFX_FSTAGED(8, 2)
                if (lpr == 16) RBFX_FSTAGED(16, 2)
                RBFX_FSTAGED(32, 2)
            }
            if (lpr == 8) RBFX_FSTAGED(8, 1)
            if (lpr == 16) RBFX_FSTAGED(16, 1)
            if (lpr == 32) RBFX_FSTAGED(32, 1)
            RBFX_FSTAGED(64, 1)
#undef RBFX_FSTAGED
        }
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

// one thread per (processing row, entry): streaming reads of the source row, writes in order
__global__ __launch_bounds__(256) void k_permute_rows(int64_t n_rows, int k, const int32_t* __restrict__ order,
                                                      const int32_t* __restrict__ col, const double* __restrict__ vals,
                                                      int64_t nnz, int nv, int32_t* __restrict__ out_col,
                                                      double* __restrict__ out_vals)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_rows * k) return;
    const int64_t r = i / k, j = i % k;
    const int64_t src = (int64_t)order[r] * k + j;
    out_col[i] = col[src];
    for (int v = 0; v < nv; ++v) out_vals[(int64_t)v * nnz + i] = vals[(int64_t)v * nnz + src];
}

hipError_t launch_permute_rows(int64_t n_rows, int k, const int32_t* row_order, const int32_t* col, const double* vals,
                               int64_t nnz, int n_valarrays, int32_t* out_col, double* out_vals, hipStream_t s)
{
    const int64_t n = n_rows * k;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_permute_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_rows, k, row_order, col,
                       vals, nnz, n_valarrays, out_col, out_vals);
    return hipGetLastError();
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
