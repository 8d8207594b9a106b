// Fused RBF-FD stencil assemble-and-solve for gfx950.
//
// Replaces the per-stencil hot loop of build_weights_kernel
// (src/solve/execution.jl:186-265): _build_collocation_matrix! (assembly.jl:95-115),
// _build_rhs! (assembly.jl:127-185), the Hermite dispatch (assembly.jl:240-467) and
// _solve_system! = bunchkaufman!(Symmetric(A,:U), true) (assembly.jl:220-222).
//
// Mapping: one 64-lane wavefront per stencil. The saddle-point system
//     [ 0   P^T ] [lambda]   [L P (x_e)    ]
//     [ P   Phi ] [  w   ] = [L phi(x_e, .)]
// (polynomial rows first) is assembled in a centred, radius-scaled frame directly
// into registers: lane L owns row L, r[j] holds A[L][j] (static j), and the operator
// columns ride along as extra register columns r[NMAX + c] (augmented system).
//
// Factorisation: symmetric-indefinite LDL^T with rook pivoting (the pivot tests of
// Ashcraft, Grimes & Lewis 1998, as LAPACK ?sytrf_rook uses). Rows never move
// between lanes; a "query" of column j broadcasts row j through LDS (by symmetry it
// is column j) and takes a cross-lane argmax. The pivot row(s) go to LDS, where they
// stay as the rows of U = D L^T, and every still-active lane applies the rank-1 or
// rank-2 update to its register row. The step loop is a runtime loop, so registers
// are only ever indexed statically. Back substitution is right-looking: when x_p is
// known every lane that pivoted earlier reads its own U-row entry for column p.
//
// FP32 mode keeps an FP64 copy of the row and of the right-hand sides: the FP32
// solution is refined with FP64 residuals b - A x and FP32 corrections that replay
// the stored multipliers; stencils whose correction does not settle are queued for
// the FP64 kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "build.h"
#include "device_math.h"

// waves per SIMD the register allocation targets (64-row systems hold twice the
// registers of 32-row ones)
template <typename T, int NMAX>
struct BuildOccupancy {
    static constexpr int waves = (NMAX <= 32) ? 2 : 1;
};

namespace rbfx {

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ double readlane(double v, int lane)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float readlane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ int readlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& idx)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const T ov = __shfl_xor(v, m, 64);
        const int oi = __shfl_xor(idx, m, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

// ---------------------------------------------------------------- monomials
// x^p for a runtime exponent 0 <= p <= kMaxPolyDeg, by repeated multiplication
__device__ __forceinline__ double ipow(double x, int p)
{
    double v = 1.0;
#pragma unroll
    for (int q = 0; q < kMaxPolyDeg; ++q)
        if (q < p) v *= x;
    return v;
}

// operator term `kind` applied to the monomial x^e
template <int DIM>
__device__ __forceinline__ double mono_op(const int (&e)[3], int kind, int d1, int d2, const double (&x)[DIM])
{
    if (kind == RBFX_OP_MIXED && d1 == d2) kind = RBFX_OP_D2;
    if (kind == RBFX_OP_LAP) {
        double acc = 0.0;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
            if (e[dd] >= 2) {
                double t = (double)(e[dd] * (e[dd] - 1));
#pragma unroll
                for (int d = 0; d < DIM; ++d) t *= ipow(x[d], d == dd ? e[d] - 2 : e[d]);
                acc += t;
            }
        }
        return acc;
    }
    double coef = 1.0, val = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int o = 0;
        if (kind == RBFX_OP_D1 && d == d1) o = 1;
        if (kind == RBFX_OP_D2 && d == d1) o = 2;
        if (kind == RBFX_OP_MIXED && (d == d1 || d == d2)) o = 1;
        if (o > e[d]) return 0.0;
        if (o == 1) coef *= (double)e[d];
        if (o == 2) coef *= (double)(e[d] * (e[d] - 1));
        val *= ipow(x[d], e[d] - o);
    }
    return coef * val;
}

template <int DIM>
__device__ __forceinline__ double mono_normal(const int (&e)[3], const double (&n)[DIM], const double (&x)[DIM])
{
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) acc += n[d] * mono_op<DIM>(e, RBFX_OP_D1, d, 0, x);
    return acc;
}

// term scale h^-order in the radius-scaled frame
__device__ __forceinline__ double term_scale(const TermDev& t, double hinv)
{
    return t.coeff * (t.kind == RBFX_OP_VALUE ? 1.0 : (t.kind == RBFX_OP_D1 ? hinv : hinv * hinv));
}

// ------------------------------------------------------- back substitution
// Solves U x = y given the pivot sequence (s_piv, s_blk), the U rows in s_U and
// each lane's pivot slot. y: this lane's forward-eliminated RHS entry.
template <typename T, int NMAX>
__device__ __forceinline__ T back_substitute(const T (*s_U)[NMAX], const int* s_piv, const int* s_blk, const int n,
                                             const int myslot, const T diag, T y)
{
    const T dinv = (T)1 / diag;
    T acc = (T)0, x = (T)0;
    const int L = lane_id();
    const int ms = myslot < NMAX ? myslot : 0;
    for (int s = n - 1; s >= 0; --s) {
        const int blk = uniform(s_blk[s]);
        if (blk == 1) {
            const int p = uniform(s_piv[s]);
            const T xp = readlane((y - acc) * dinv, p);
            if (L == p) x = xp;
            if (myslot < s) acc += s_U[ms][p] * xp;
        } else if (blk == 0 && s > 0) {
            // second slot of a 2x2 block (s-1, s)
            const int p = uniform(s_piv[s - 1]), q = uniform(s_piv[s]);
            const T a = readlane(diag, p), bq = s_U[s - 1][q], c = readlane(diag, q);
            const T y0 = readlane(y - acc, p), y1 = readlane(y - acc, q);
            const T idet = (T)1 / (a * c - bq * bq);
            const T x0 = (y0 * c - y1 * bq) * idet, x1 = (y1 * a - y0 * bq) * idet;
            if (L == p) x = x0;
            if (L == q) x = x1;
            if (myslot < s - 1) acc += s_U[ms][p] * x0 + s_U[ms][q] * x1;
            --s;
        }
    }
    return x;
}

// forward replay of the stored multipliers on a new RHS (FP32 refinement)
template <typename T>
__device__ __forceinline__ T forward_replay(const T (*s_Lm)[64], const int* s_piv, const int* s_blk, const int n,
                                            const int myslot, T y)
{
    const int L = lane_id();
    for (int s = 0; s < n; ++s) {
        const int blk = uniform(s_blk[s]);
        if (blk == 1) {
            const T yp = readlane(y, uniform(s_piv[s]));
            if (myslot > s) y -= s_Lm[s][L] * yp;
        } else if (blk == 2) {
            const T y0 = readlane(y, uniform(s_piv[s])), y1 = readlane(y, uniform(s_piv[s + 1]));
            if (myslot > s + 1) y -= s_Lm[s][L] * y0 + s_Lm[s + 1][L] * y1;
            ++s;
        }
    }
    return y;
}

// ------------------------------------------------------------------ kernel
template <typename T, int NMAX, int DIM>
__global__ __launch_bounds__(64, (BuildOccupancy<T, NMAX>::waves)) void k_build(const BuildDev P)
{
    constexpr bool REFINE = sizeof(T) == 4;
    constexpr int RC = RBFX_MAX_COLS;
    constexpr int PT = NMAX * NMAX / 4;  // >= kk * M for any split kk + M <= NMAX

    __shared__ double s_xi[NMAX][DIM];
    __shared__ double s_ba[NMAX], s_bb[NMAX], s_bn[NMAX][DIM];
    __shared__ int s_bt[NMAX];
    __shared__ int s_piv[NMAX], s_blk[NMAX];
    __shared__ T s_col[NMAX];
    __shared__ double s_x[64];
    // U rows (factorisation) alias the P table (assembly)
    constexpr int UB = NMAX * NMAX * (int)sizeof(T), PB = PT * 8;
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[UB > PB ? UB : PB];
    T(*s_U)[NMAX] = reinterpret_cast<T(*)[NMAX]>(s_raw);
    double* s_P = reinterpret_cast<double*>(s_raw);
    __shared__ T s_Lm[REFINE ? NMAX : 1][64];
    __shared__ double s_b64[REFINE ? RC : 1][64];
    __shared__ int s_mono[RBFX_MAX_N][3];
    __shared__ TermDev s_terms[RBFX_MAX_TERMS];

    const int L = lane_id();
    const int n_cols = P.n_cols;
    const int M = P.n_mono;
    const int n_terms = P.n_terms;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    // operator and monomial tables to LDS once per block (keeps them out of SGPRs)
    {
        const int* src = &P.mono_e[0][0];
        for (int i = L; i < RBFX_MAX_N * 3; i += 64) (&s_mono[0][0])[i] = src[i];
        if (L < RBFX_MAX_TERMS) s_terms[L] = P.terms[L];
    }
    __syncthreads();

    {
        const int64_t w = blockIdx.x;
        if (w >= n_work) return;
        const int64_t e = P.list ? (int64_t)P.list[w] : w;
        const int64_t s0 = P.adj_ptr ? P.adj_ptr[e] : e * (int64_t)P.k;
        const int kk = P.adj_ptr ? (int)(P.adj_ptr[e + 1] - s0) : P.k;
        const int64_t o0 = P.out_ptr ? P.out_ptr[e] : e * (int64_t)P.k;
        const int n = M + kk;
        if (kk < 1 || n > NMAX) {
            if (L == 0 && P.status) P.status[e] = RBFX_ST_TOO_LARGE;
            return;
        }

        // ---- this lane's stencil point (point rows live at lanes M..M+kk-1)
        const int pi = L - M;
        const bool is_pt = pi >= 0 && pi < kk;
        int g = 0;
        double xe[DIM], D[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xe[d] = P.eval[e * DIM + d];
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) D[d] = 0.0;
        if (is_pt) {
            g = P.adj_idx[s0 + pi];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                D[d] = P.data[(int64_t)g * DIM + d] - xe[d];
                r2 += D[d] * D[d];
            }
        }

        // ---- stencil classification (types.jl:208-223)
        bool hermite = false;
        int bt = 0;  // 0 interior, 1 Dirichlet, 2 Neumann/Robin
        double ba = 0.0, bb = 0.0, bn[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) bn[d] = 0.0;
        if (P.is_boundary) {
            const bool ib = is_pt && P.is_boundary[g];
            hermite = __ballot(ib) != 0ull;
            if (hermite && e < P.n_data && P.is_boundary[e]) {
                const int bi = P.bnd_index[e];
                if (P.bc_alpha[bi] == 1.0 && P.bc_beta[bi] == 0.0) {
                    // identity row (fill_dirichlet_entry!, execution.jl:297-303)
                    if (L < n_cols) P.out_val[(int64_t)L * P.out_nnz + o0] = 1.0;
                    if (L == 0) {
                        if (P.out_col) P.out_col[o0] = (int32_t)e;
                        if (P.status) P.status[e] = RBFX_ST_DIRICHLET;
                    }
                    return;
                }
            }
            if (hermite && ib) {
                const int bi = P.bnd_index[g];
                ba = P.bc_alpha[bi];
                bb = P.bc_beta[bi];
#pragma unroll
                for (int d = 0; d < DIM; ++d) bn[d] = P.normals[(int64_t)bi * DIM + d];
                bt = (ba == 1.0 && bb == 0.0) ? 1 : 2;
            }
        }

        // ---- radius-scaled local frame
        double h = sqrt(wave_max(r2));
        if (!(h > 0.0)) h = 1.0;
        const double hinv = 1.0 / h;
        double xi[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = D[d] / h;
        bb = bb / h;  // beta acts through a first derivative
        if (is_pt) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) s_xi[pi][d] = xi[d];
            if (hermite) {
                s_bt[pi] = bt;
                s_ba[pi] = ba;
                s_bb[pi] = bb;
#pragma unroll
                for (int d = 0; d < DIM; ++d) s_bn[pi][d] = bn[d];
            }
        }
        int eval_local = -1;
        if (hermite) {
            bool eq = is_pt;
#pragma unroll
            for (int d = 0; d < DIM; ++d) eq = eq && (D[d] == 0.0);
            const unsigned long long m = __ballot(eq);
            if (!m) {
                if (L == 0 && P.status) P.status[e] = RBFX_ST_EVAL_NOT_FOUND;
                __syncthreads();
                return;
            }
            eval_local = uniform(__builtin_ctzll(m)) - M;
        }
        __syncthreads();

        BasisDev B;
        B.kind = P.basis_kind;
        B.eps = (B.kind == RBFX_IMQ || B.kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;

        // ---- assembly (FP64) into registers
        T r[NMAX + RC];
        double a64[REFINE ? NMAX : 1];
        double diag64 = 0.0;
#pragma unroll
        for (int j = 0; j < NMAX + RC; ++j) r[j] = (T)0;
#pragma unroll
        for (int j = 0; j < (REFINE ? NMAX : 1); ++j) a64[j] = 0.0;
        if (is_pt) {
#pragma unroll
            for (int a = 0; a < NMAX; ++a) {
                if (a < M) {
                    double pv = mono_op<DIM>(s_mono[a], RBFX_OP_VALUE, 0, 0, xi);
                    if (hermite && bt == 2) pv = ba * pv + bb * mono_normal<DIM>(s_mono[a], bn, xi);
                    r[a] = (T)pv;
                    if constexpr (REFINE) a64[a] = pv;
                    s_P[pi * M + a] = pv;
                }
            }
#pragma unroll
            for (int j = 0; j < NMAX; ++j) {
                if (j >= M && j < n) {
                    const int jj = j - M;
                    double Dj[DIM], q2 = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        Dj[d] = xi[d] - s_xi[jj][d];
                        q2 += Dj[d] * Dj[d];
                    }
                    double v = phi_r2(B, q2);
                    if (hermite) {
                        const int tj = s_bt[jj];
                        if (bt == 2 || tj == 2) {
                            double nj[DIM];
#pragma unroll
                            for (int d = 0; d < DIM; ++d) nj[d] = s_bn[jj][d];
                            double Di, Djn, D2;
                            phi_grad_hess_dirs<DIM>(B, Dj, q2, bn, nj, Di, Djn, D2);
                            const double aj = s_ba[jj], bj = s_bb[jj];
                            if (bt != 2) v = aj * v - bj * Djn;
                            else if (tj != 2) v = ba * v + bb * Di;
                            else v = ba * aj * v - ba * bj * Djn + bb * aj * Di + bb * bj * D2;
                        }
                    }
                    if (jj == pi) diag64 = v;
                    r[j] = (T)v;
                    if constexpr (REFINE) a64[j] = v;
                }
            }
        }
        __syncthreads();
        if (L < M) {
#pragma unroll
            for (int j = 0; j < NMAX; ++j) {
                if (j >= M && j < n) {
                    const double v = s_P[(j - M) * M + L];
                    r[j] = (T)v;
                    if constexpr (REFINE) a64[j] = v;
                }
            }
        }

        // ---- right-hand sides (FP64), augmented as register columns NMAX..NMAX+RC-1
        {
            double bcol[RC];
#pragma unroll
            for (int c = 0; c < RC; ++c) bcol[c] = 0.0;
            const bool f7 = hermite && s_bt[eval_local] == 2;
            if (is_pt) {
                double De[DIM], q2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    De[d] = -xi[d];
                    q2 += De[d] * De[d];
                }
                for (int t = 0; t < n_terms; ++t) {
                    const TermDev tm = s_terms[t];
                    double v = phi_term<DIM>(B, tm.kind, tm.d1, tm.d2, De, q2);
                    if (hermite && bt == 2) v = ba * v + bb * phi_term_normal<DIM>(B, tm.kind, tm.d1, De, q2, bn);
                    const double sc = term_scale(tm, hinv);
#pragma unroll
                    for (int c = 0; c < RC; ++c)
                        if (tm.col == c) bcol[c] += sc * v;
                }
            } else if (L < M) {
                double z[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) z[d] = 0.0;
                const int(&ea)[3] = s_mono[L < RBFX_MAX_N ? L : 0];
                if (f7) {
                    // the eval node's boundary operator applied to P (assembly.jl:452-467)
                    const double eav = s_ba[eval_local], ebv = s_bb[eval_local];
                    double en[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) en[d] = s_bn[eval_local][d];
                    const double v = eav * mono_op<DIM>(ea, RBFX_OP_VALUE, 0, 0, z) + ebv * mono_normal<DIM>(ea, en, z);
#pragma unroll
                    for (int c = 0; c < RC; ++c) bcol[c] = v;
                } else {
                    for (int t = 0; t < n_terms; ++t) {
                        const TermDev tm = s_terms[t];
                        const double v = term_scale(tm, hinv) * mono_op<DIM>(ea, tm.kind, tm.d1, tm.d2, z);
#pragma unroll
                        for (int c = 0; c < RC; ++c)
                            if (tm.col == c) bcol[c] += v;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                r[NMAX + c] = (T)bcol[c];
                if constexpr (REFINE) s_b64[c][L] = bcol[c];
            }
        }
        __syncthreads();

        // ---- factorisation with rook pivoting (runtime step loop)
        T diag = (T)diag64;
        bool act = L < n;
        int myslot = 1 << 20;
        unsigned long long amask = __ballot(act);
        const T alpha = (T)0.6403882032022076;  // (1 + sqrt(17)) / 8
        bool failed = false;
        int slot = 0;
        while (amask != 0ull) {
            const int c0 = uniform(__builtin_ctzll(amask));
            // query column c0 (row c0 by symmetry)
            if (L == c0) {
#pragma unroll
                for (int j = 0; j < NMAX; ++j) s_col[j] = r[j];
            }
            __syncthreads();
            T v = (act && L != c0) ? (T)fabs(s_col[L & (NMAX - 1)]) : (T)-1;
            int vi = L;
            wave_argmax(v, vi);
            __syncthreads();
            T colmax = v > (T)0 ? v : (T)0;
            int imax = uniform(vi);
            const T absakk = (T)fabs(readlane(diag, c0));
            if (!(absakk > (T)0 || colmax > (T)0)) { failed = true; break; }
            int kstep = 1, p = c0, kp = c0;
            if (absakk < alpha * colmax) {
                for (int it = 0; it < 2 * NMAX; ++it) {
                    if (L == imax) {
#pragma unroll
                        for (int j = 0; j < NMAX; ++j) s_col[j] = r[j];
                    }
                    __syncthreads();
                    T rv = (act && L != imax) ? (T)fabs(s_col[L & (NMAX - 1)]) : (T)-1;
                    int ri = L;
                    wave_argmax(rv, ri);
                    __syncthreads();
                    const T rowmax = rv > (T)0 ? rv : (T)0;
                    const int jmax = uniform(ri);
                    const T aii = (T)fabs(readlane(diag, imax));
                    if (!(aii < alpha * rowmax)) { kp = imax; break; }
                    if (p == jmax || rowmax <= colmax || it == 2 * NMAX - 1) { kp = imax; kstep = 2; break; }
                    p = imax;
                    colmax = rowmax;
                    imax = jmax;
                }
            }
            if (kstep == 1) {
                const int pv = kp;
                if (L == pv) {
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) s_U[slot][j] = r[j];
                }
                if (L == 0) { s_piv[slot] = pv; s_blk[slot] = 1; }
                __syncthreads();
                const T d = readlane(diag, pv);
                if (d == (T)0) { failed = true; break; }
                const T dinv = (T)1 / d;
                T rb[RC];
#pragma unroll
                for (int c = 0; c < RC; ++c) rb[c] = (c < n_cols) ? readlane(r[NMAX + c], pv) : (T)0;
                if (act && L != pv) {
                    const T a = s_U[slot][L];
                    const T l = a * dinv;
                    diag -= l * a;
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) r[j] -= l * s_U[slot][j];
#pragma unroll
                    for (int c = 0; c < RC; ++c)
                        if (c < n_cols) r[NMAX + c] -= l * rb[c];
                    if constexpr (REFINE) s_Lm[slot][L] = l;
                }
                if (L == pv) { act = false; myslot = slot; }
                slot += 1;
            } else {
                const int pp = p, qq = kp;
                if (L == pp) {
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) s_U[slot][j] = r[j];
                }
                if (L == qq) {
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) s_U[slot + 1][j] = r[j];
                }
                if (L == 0) {
                    s_piv[slot] = pp; s_blk[slot] = 2;
                    s_piv[slot + 1] = qq; s_blk[slot + 1] = 0;
                }
                __syncthreads();
                const T a = readlane(diag, pp), bq = s_U[slot][qq], c = readlane(diag, qq);
                const T det = a * c - bq * bq;
                if (det == (T)0) { failed = true; break; }
                const T idet = (T)1 / det;
                T rb0[RC], rb1[RC];
#pragma unroll
                for (int cc = 0; cc < RC; ++cc) {
                    rb0[cc] = (cc < n_cols) ? readlane(r[NMAX + cc], pp) : (T)0;
                    rb1[cc] = (cc < n_cols) ? readlane(r[NMAX + cc], qq) : (T)0;
                }
                if (act && L != pp && L != qq) {
                    const T u0 = s_U[slot][L], u1 = s_U[slot + 1][L];
                    const T l0 = (u0 * c - u1 * bq) * idet;
                    const T l1 = (u1 * a - u0 * bq) * idet;
                    diag -= l0 * u0 + l1 * u1;
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) r[j] -= l0 * s_U[slot][j] + l1 * s_U[slot + 1][j];
#pragma unroll
                    for (int cc = 0; cc < RC; ++cc)
                        if (cc < n_cols) r[NMAX + cc] -= l0 * rb0[cc] + l1 * rb1[cc];
                    if constexpr (REFINE) {
                        s_Lm[slot][L] = l0;
                        s_Lm[slot + 1][L] = l1;
                    }
                }
                if (L == pp) { act = false; myslot = slot; }
                if (L == qq) { act = false; myslot = slot + 1; }
                slot += 2;
            }
            amask = __ballot(act);
            __syncthreads();
        }

        int st = failed ? RBFX_ST_SINGULAR : P.ok_status;
        if (!failed) {
            bool converged = true;
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                if (c < n_cols) {
                    const T x = back_substitute<T, NMAX>(s_U, s_piv, s_blk, n, myslot, diag, r[NMAX + c]);
                    if constexpr (REFINE) {
                        // FP64 residual b - A x, FP32 correction on the same factors
                        double x64 = (double)x;
                        double rel = 1.0;
                        for (int it = 0; it < P.refine_steps; ++it) {
                            s_x[L] = x64;
                            __syncthreads();
                            double res = s_b64[c][L];
#pragma unroll
                            for (int j = 0; j < NMAX; ++j) res -= a64[j] * s_x[j];
                            __syncthreads();
                            const T y = forward_replay<T>(s_Lm, s_piv, s_blk, n, myslot, (T)res);
                            const T dx = back_substitute<T, NMAX>(s_U, s_piv, s_blk, n, myslot, diag, y);
                            x64 += (double)dx;
                            const double num = wave_max(L < n ? fabs((double)dx) : 0.0);
                            const double den = wave_max(L < n ? fabs(x64) : 0.0);
                            rel = den > 0.0 ? num / den : num;
                            if (rel <= 0.25 * P.accept_tol) break;
                        }
                        if (!(rel <= P.accept_tol)) converged = false;
                        if (is_pt) P.out_val[(int64_t)c * P.out_nnz + o0 + pi] = x64;
                    } else {
                        if (is_pt) P.out_val[(int64_t)c * P.out_nnz + o0 + pi] = (double)x;
                    }
                }
            }
            if constexpr (REFINE) {
                st = converged ? RBFX_ST_REFINED : RBFX_ST_FALLBACK64;
                if (!converged && L == 0 && P.fail_list) {
                    const int slot2 = atomicAdd(P.fail_count, 1);
                    P.fail_list[slot2] = (int32_t)e;
                }
            }
        } else {
            if (REFINE && P.fail_list) {
                // FP32 pivot breakdown: retry in FP64 before calling the system singular
                st = RBFX_ST_FALLBACK64;
                if (L == 0) {
                    const int slot2 = atomicAdd(P.fail_count, 1);
                    P.fail_list[slot2] = (int32_t)e;
                }
            }
            for (int c = 0; c < n_cols; ++c)
                if (is_pt) P.out_val[(int64_t)c * P.out_nnz + o0 + pi] = __builtin_nan("");
        }
        if (is_pt && P.out_col) P.out_col[o0 + pi] = g;
        if (L == 0 && P.status) P.status[e] = st;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- dispatch
template <typename T, int NMAX, int DIM>
inline hipError_t launch_one(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    if (n_work <= 0) return hipSuccess;
    // one wavefront per stencil; a work-list launch sizes the grid by its upper bound
    const int64_t grid = n_work;
    hipLaunchKernelGGL((k_build<T, NMAX, DIM>), dim3((unsigned)grid), dim3(64), 0, s, P);
    return hipGetLastError();
}

template <typename T, int NMAX>
inline hipError_t dispatch_dim(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    switch (dim) {
    case 1: return launch_one<T, NMAX, 1>(P, n_work, s);
    case 2: return launch_one<T, NMAX, 2>(P, n_work, s);
    case 3: return launch_one<T, NMAX, 3>(P, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
