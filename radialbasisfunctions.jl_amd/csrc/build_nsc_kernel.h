// Null-space Cholesky stencil build for gfx950 (kernel template).
//
// Replaces, like k_build (build_kernel.h), the per-stencil loop of build_weights_kernel
// (src/solve/execution.jl:186-265): _build_collocation_matrix! (assembly.jl:95-115),
// _build_rhs! (assembly.jl:127-185) and _solve_system! (assembly.jl:220-222), for
// interior stencils whose saddle-point system
//     [ Phi  P ] [ w      ]   [ L phi (x_e, .) ]
//     [ P^T  0 ] [ lambda ] = [ L P (x_e)      ]
// is definite on the null space of P^T (nsc_sign). No pivot search: with the
// Householder factorisation P = Q [R; 0], w = Q [w1; u] where R^T w1 = L P and
//     s (Q2^T Phi Q2) u = s Q2^T (L phi - Phi Q1 w1),   Q = [Q1 Q2],
// and s (Q2^T Phi Q2) is symmetric positive definite, factorised as L D L^T without
// pivoting. One FP64 refinement step on the original system (residual of the
// assembled system, correction through the same factors) follows; a stencil whose
// LDL^T meets a non-positive pivot, or whose correction does not settle, is handed to
// the rook-pivoted kernels, which remain the drop-in for every system.
//
// Mapping: a group of GS lanes per stencil (64/GS stencils per wavefront); lane li owns
// the point rows li + s*GS, s < SL. Every loop over rows, columns and monomials is
// unrolled to its compile-time bound (KMAX points, MMAX monomials) with the actual k
// and m as predicates, so each register is addressed statically. Vectors a step needs
// whole (a reflector, a pivot row, a solution) pass through the group's LDS; scalar
// sums over the group are DPP / swizzle butterflies whose every lane ends with the
// same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "build_kernel.h"
#include "build_nsc.h"

namespace rbfx {

// ---------------------------------------------------------------- group collectives
__device__ __forceinline__ double swz_xor16(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(unsigned)u, 0x401F);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(unsigned)(u >> 32), 0x401F);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over aligned groups of GS lanes. Every step adds two partial sums that the
// partner lane adds in the other order, so all lanes of a group hold identical bits:
// quad_perm [1,0,3,2] and [2,3,0,1] (lane xor 1, 2), row_half_mirror (i <-> 7-i over
// quad sums), row_mirror (i <-> 15-i over 8-lane sums), swizzle xor 16, xor 32.
template <int GS>
__device__ __forceinline__ double gsum(double v)
{
    static_assert(GS == 8 || GS == 16 || GS == 32 || GS == 64, "group of 8..64 lanes");
    v = v + dpp<0xB1>(v);
    v = v + dpp<0x4E>(v);
    v = v + dpp<0x141>(v);
    if constexpr (GS >= 16) v = v + dpp<0x140>(v);
    if constexpr (GS >= 32) v = v + swz_xor16(v);
    if constexpr (GS == 64) v = v + __shfl_xor(v, 32, 64);
    return v;
}

template <int GS>
__device__ __forceinline__ double gmax(double v)
{
    v = fmax(v, dpp<0xB1>(v));
    v = fmax(v, dpp<0x4E>(v));
    v = fmax(v, dpp<0x141>(v));
    if constexpr (GS >= 16) v = fmax(v, dpp<0x140>(v));
    if constexpr (GS >= 32) v = fmax(v, swz_xor16(v));
    if constexpr (GS == 64) v = fmax(v, __shfl_xor(v, 32, 64));
    return v;
}

// Square root, reciprocal square root and reciprocal in about one ulp: the hardware
// estimates (v_rsq_f64, v_rcp_f64) refined by Newton steps. Their rounding differs from
// the correctly rounded operations by at most an ulp or so, a relative perturbation of
// the assembled system of order 1e-16, inside the conditioning bound the build meets.
// nsc_sqrt(0) = 0 (the PHS diagonal: the estimate is taken at max(x, 1e-300), so
// x = 0 gives s = 0 through every step); arguments are finite and not negative.
__device__ __forceinline__ double nsc_rsq(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}
__device__ __forceinline__ double nsc_sqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(fmax(x, 1e-300));
    double s = x * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    return fma(fma(-s, s, x), h, s);
}
__device__ __forceinline__ double nsc_rcp(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    y = fma(y, fma(-d, y, 1.0), y);
    y = fma(y, fma(-d, y, 1.0), y);
    return y;
}

// phi(r^2) with the basis kind a compile-time constant (the assembly and residual
// loops are instantiated once per kind, so no kind switch sits inside them)
template <int KIND>
__device__ __forceinline__ double phi_kind(double eps, double r2)
{
    if constexpr (KIND == RBFX_PHS1) return nsc_sqrt(r2);
    else if constexpr (KIND == RBFX_PHS3) return r2 * nsc_sqrt(r2);
    else if constexpr (KIND == RBFX_PHS5) return (r2 * r2) * nsc_sqrt(r2);
    else if constexpr (KIND == RBFX_PHS7) return (r2 * r2 * r2) * nsc_sqrt(r2);
    else if constexpr (KIND == RBFX_IMQ) return nsc_rsq((eps * eps) * r2 + 1.0);
    else return exp(-(eps * eps) * r2);
}

// ---------------------------------------------------------------- per-group LDS
template <int GS, int SL, int KMAX, int MMAX, int DIM, int RC>
struct NscLds {
    static constexpr int NR = GS * SL;        // row-index space of the group
    double xi[NR][DIM];                       // scaled offsets of the stencil points
    double vec[2][NR];                        // broadcast vectors of a step
    double rhs[RC][NR];                       // right-hand side of the point rows
    double res[RC][NR];                       // residual of a refinement step (point rows)
    double wsol[RC][NR];                      // solution w, per operator column
    double wt[RC][NR];                        // the last solve's Q^T w
    double piv[KMAX + 2];                     // pivot row of an LDL^T step / a sent value
    double R[MMAX][MMAX];                     // triangular factor of P
    double tau[MMAX];                         // reflector scales
    double g[RC][MMAX];                       // polynomial right-hand side
    double gres[RC][MMAX];                    // residual of a refinement step (polynomial rows)
    double w1[MMAX];                          // R^-T g of the column being solved
    double btm[RC][MMAX];                     // the last solve's (Q^T rhs)[:M]
    double lam[RC][MMAX];                     // Lagrange multipliers
    double dlam[RC][MMAX];                    // their correction in a refinement step
};

// ------------------------------------------------------------------ kernel
// W: waves per SIMD the register allocation targets (every instantiation is checked
// to compile without spilling vector registers)
template <int GS, int SL, int KMAX, int MMAX, int DIM, int RC, int NREF, int W>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(W))) void
k_nsc(const BuildDev P)
{
    static_assert(64 % GS == 0 && KMAX <= GS * SL && MMAX <= GS * SL, "rows fit the group");
    constexpr int G = 64 / GS;
    constexpr int NR = GS * SL;
    using Lds = NscLds<GS, SL, KMAX, MMAX, DIM, RC>;
    __shared__ Lds s_st[G];
    __shared__ int s_mono[kMaxMono][3];
    __shared__ TermDev s_terms[RBFX_MAX_TERMS];

    const int L = lane_id();
    const int g = L / GS, li = L % GS;
    Lds& S = s_st[g];
    const int n_cols = P.n_cols, n_terms = P.n_terms;
    {
        const int* src = &P.mono_e[0][0];
        for (int i = L; i < kMaxMono * 3; i += 64) (&s_mono[0][0])[i] = src[i];
        if (L < RBFX_MAX_TERMS) s_terms[L] = P.terms[L];
    }
    const int64_t lb = xcd_contiguous_block(P.n_blocks);
    if (lb < 0) return;  // padding block: the whole workgroup leaves
    // the group's tables start at zero: entries beyond a stencil's own size (rows of R
    // past its monomials, say) are then zeros, never stale values
    for (int i = li; i < (int)(sizeof(Lds) / sizeof(double)); i += GS) reinterpret_cast<double*>(&S)[i] = 0.0;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    const int64_t wi = lb * G + g;
    // (long-lived per-lane flags are ints: as bools they would be held as lane masks in
    // scalar registers, which run out and spill)
    int alive = wi < n_work;
    int64_t e = alive ? (P.list ? (int64_t)P.list[wi] : (P.order ? (int64_t)P.order[wi] : wi)) : 0;
    if (e < 0 || e >= P.n_eval) {
        alive = 0;
        e = 0;
    }
    // Uniform builds only (fixed k, one polynomial degree; the host checks it, and that
    // k and m fit the instantiation): the stencil shape is a kernel argument, so every
    // condition on it is a scalar branch. A group without work (past the end of the
    // launch, or handed back) computes on zeros, harmlessly, and writes nothing.
    const int kk = P.k;
    const int M = P.n_mono;
    const int m_wave = M;
    int handoff = 0;
    int fail = 0;
    if (P.adj_ptr || P.deg_map) {  // (the launcher refuses such builds; nothing is read)
        handoff = alive;
        alive = 0;
    }
    const int64_t o0 = alive ? (P.out_ptr ? P.out_ptr[e] : e * (int64_t)kk) : 0;
    const int64_t s0 = e * (int64_t)kk;

    // ---- this lane's points (rows li + s*GS)
    int is_pt[SL];
    double xi[SL][DIM];
    int gidx[SL];
    double xe[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xe[d] = alive ? P.eval[e * DIM + d] : 0.0;
    double r2max = 0.0;
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        const int row = li + s * GS;
        is_pt[s] = alive && row < kk;  // (for the gather; the arithmetic uses row < kk)
        gidx[s] = 0;
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[s][d] = 0.0;
        if (is_pt[s]) {
            if (P.packed) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[s][d] = P.packed[(wi * (int64_t)P.k + row) * DIM + d];
            } else {
                gidx[s] = P.adj_idx[s0 + row];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[s][d] = P.data[(int64_t)gidx[s] * DIM + d] - xe[d];
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) r2 += xi[s][d] * xi[s][d];
        }
        r2max = fmax(r2max, r2);
    }
    // Hermite stencils (a boundary point among the neighbours, or a boundary eval node)
    // are the rook kernels' work: hand them back
    if (P.is_boundary) {
        bool ib = alive && e < P.n_data && P.is_boundary[e];
#pragma unroll
        for (int s = 0; s < SL; ++s) ib = ib || (is_pt[s] && P.is_boundary[gidx[s]]);
        if (alive && group_bits<GS>(__ballot(ib)) != 0ull) {
            handoff = 1;
            alive = 0;
        }
    }
    // point rows of the arithmetic: the same lanes in every group
#pragma unroll
    for (int s = 0; s < SL; ++s) is_pt[s] = li + s * GS < kk;
    double h = sqrt(gmax<GS>(r2max));
    if (!(h > 0.0)) h = 1.0;
    const double hinv = 1.0 / h;
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        const int row = li + s * GS;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xi[s][d] = xi[s][d] / h;
            S.xi[row][d] = is_pt[s] ? xi[s][d] : 0.0;
        }
    }
    const int kind = P.basis_kind;
    const double eps_h = (kind == RBFX_IMQ || kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;
    const int sgn = nsc_sign(kind, P.poly_deg);

    // ---- right-hand sides of the point rows (FP64, term scale h^-order folded in)
    double b0[SL][RC];
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int c = 0; c < RC; ++c) b0[s][c] = 0.0;
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        if (!is_pt[s]) continue;
        double De[DIM], q2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            De[d] = -xi[s][d];
            q2 += De[d] * De[d];
        }
#define RBFX_NSC_RHS(BK)                                                                 \
    {                                                                                    \
        BasisDev Bk;                                                                     \
        Bk.kind = (BK);                                                                  \
        Bk.eps = eps_h;                                                                  \
        for (int t = 0; t < n_terms; ++t) {                                              \
            const TermDev tm = s_terms[t];                                               \
            const double v = term_scale(tm, hinv) * phi_term<DIM>(Bk, tm.kind, tm.d1, tm.d2, De, q2); \
            _Pragma("unroll") for (int c = 0; c < RC; ++c) if (tm.col == c) b0[s][c] += v; \
        }                                                                                \
    }
        switch (kind) {
        case RBFX_PHS1: RBFX_NSC_RHS(RBFX_PHS1) break;
        case RBFX_PHS3: RBFX_NSC_RHS(RBFX_PHS3) break;
        case RBFX_PHS5: RBFX_NSC_RHS(RBFX_PHS5) break;
        case RBFX_PHS7: RBFX_NSC_RHS(RBFX_PHS7) break;
        case RBFX_IMQ: RBFX_NSC_RHS(RBFX_IMQ) break;
        default: RBFX_NSC_RHS(RBFX_GAUSSIAN) break;
        }
#undef RBFX_NSC_RHS
    }
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int c = 0; c < RC; ++c) S.rhs[c][li + s * GS] = b0[s][c];
    // polynomial right-hand side L P(x_e) (the centre is the origin of the frame), one
    // monomial per owning lane; zero beyond the stencil's monomials
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        const int a = li + s * GS;
        if (a < MMAX) {
            double ga[RC];
#pragma unroll
            for (int c = 0; c < RC; ++c) ga[c] = 0.0;
            if (a < M) {
                const double z[DIM] = {};
                for (int t = 0; t < n_terms; ++t) {
                    const TermDev tm = s_terms[t];
                    const double v = term_scale(tm, hinv) * mono_op<DIM>(s_mono[a], tm.kind, tm.d1, tm.d2, z);
#pragma unroll
                    for (int c = 0; c < RC; ++c)
                        if (tm.col == c) ga[c] += v;
                }
            }
#pragma unroll
            for (int c = 0; c < RC; ++c) S.g[c][a] = ga[c];
        }
    }

    // ---- monomials of the own points (graded order, the table's: the exponents are
    // compile-time constants here, so no per-lane selects)
    double p[SL][MMAX];
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int a = 0; a < MMAX; ++a)
            p[s][a] = (is_pt[s] && a < M) ? mono_graded<DIM>(a, xi[s]) : 0.0;
    __syncthreads();

    // ---- collocation rows Phi[row][j] = phi(|xi_row - xi_j|). Rows and columns past the
    // stencil (k <= j < KMAX) hold phi of zero offsets: finite values that never reach a
    // stencil entry, since every reflector is zero there and the LDL^T and the solves
    // stop at row and column k.
    double phi[SL][KMAX];
#define RBFX_NSC_ASSEMBLE(BK)                                                            \
    _Pragma("unroll") for (int j = 0; j < KMAX; ++j)                                     \
    {                                                                                    \
        double xj[DIM];                                                                  \
        _Pragma("unroll") for (int d = 0; d < DIM; ++d) xj[d] = S.xi[j][d];              \
        _Pragma("unroll") for (int s = 0; s < SL; ++s)                                   \
        {                                                                                \
            double q2 = 0.0;                                                             \
            _Pragma("unroll") for (int d = 0; d < DIM; ++d)                              \
            {                                                                            \
                const double t_ = xi[s][d] - xj[d];                                      \
                q2 += t_ * t_;                                                           \
            }                                                                            \
            phi[s][j] = phi_kind<BK>(eps_h, q2);                                         \
        }                                                                                \
    }
    switch (kind) {
    case RBFX_PHS1: RBFX_NSC_ASSEMBLE(RBFX_PHS1) break;
    case RBFX_PHS3: RBFX_NSC_ASSEMBLE(RBFX_PHS3) break;
    case RBFX_PHS5: RBFX_NSC_ASSEMBLE(RBFX_PHS5) break;
    case RBFX_PHS7: RBFX_NSC_ASSEMBLE(RBFX_PHS7) break;
    case RBFX_IMQ: RBFX_NSC_ASSEMBLE(RBFX_IMQ) break;
    default: RBFX_NSC_ASSEMBLE(RBFX_GAUSSIAN) break;
    }

    // ---- Householder QR of P, applied to Phi from both sides: Phi <- H Phi H
    // (H = I - tau v v^T, v zero above row c). Each lane keeps the reflectors' entries
    // of its own rows (vr, in the registers of the columns of P they replace): the
    // solves apply Q and Q^T with them and a sum over the group.
    double vr[SL][MMAX];
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int a = 0; a < MMAX; ++a) vr[s][a] = 0.0;
#pragma unroll
    for (int c = 0; c < MMAX; ++c) {
        if (c < m_wave) {
            const bool act = c < M;
            double a1 = 0.0, a2 = 0.0;
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                const int row = li + s * GS;
                const double x = row >= c ? p[s][c] : 0.0;
                a1 += x * x;
                a2 += row == c ? x : 0.0;
            }
            const double nrm2 = gsum<GS>(a1), xc = gsum<GS>(a2);
            const double sq = sqrt(nrm2);
            const double alpha = xc >= 0.0 ? -sq : sq;
            const double den = nrm2 - alpha * xc;  // |v|^2 / 2
            if (act && !(den > 0.0 && den < 1e300)) fail = 1;  // P rank-deficient
            const double tau = (act && den > 0.0) ? nsc_rcp(den) : 0.0;
            double v[SL];
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                const int row = li + s * GS;
                v[s] = !act ? 0.0 : (row > c ? p[s][c] : (row == c ? xc - alpha : 0.0));
                S.vec[0][row] = v[s];
                vr[s][c] = v[s];
            }
            if (li == c % GS) {
                S.R[c][c] = act ? alpha : 1.0;
                S.tau[c] = tau;
            }
            __syncthreads();
            double y[SL];
#pragma unroll
            for (int s = 0; s < SL; ++s) y[s] = 0.0;
#pragma unroll
            for (int j = 0; j < KMAX; ++j) {
                if (j >= c) {
                    const double vj = S.vec[0][j];
#pragma unroll
                    for (int s = 0; s < SL; ++s) y[s] = fma(phi[s][j], vj, y[s]);
                }
            }
            double beta_l = 0.0;
#pragma unroll
            for (int s = 0; s < SL; ++s) beta_l = fma(v[s], y[s], beta_l);
            const double beta = gsum<GS>(beta_l);
            // the later columns of P
#pragma unroll
            for (int c2 = 1; c2 < MMAX; ++c2) {
                if (c2 > c && c2 < m_wave) {
                    double loc = 0.0;
#pragma unroll
                    for (int s = 0; s < SL; ++s) loc = fma(v[s], p[s][c2], loc);
                    const double sp = tau * gsum<GS>(loc);
#pragma unroll
                    for (int s = 0; s < SL; ++s) p[s][c2] = fma(-sp, v[s], p[s][c2]);
                }
            }
            // row c of R is final once column c is eliminated
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                if (act && li + s * GS == c) {
#pragma unroll
                    for (int c2 = 1; c2 < MMAX; ++c2)
                        if (c2 > c) S.R[c][c2] = p[s][c2];
                }
            }
            double z[SL];
            const double hb = 0.5 * tau * tau * beta;
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                z[s] = fma(tau, y[s], -hb * v[s]);
                S.vec[1][li + s * GS] = z[s];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KMAX; ++j) {
                const double zj = S.vec[1][j];
                if (j >= c) {
                    const double vj = S.vec[0][j];
#pragma unroll
                    for (int s = 0; s < SL; ++s) phi[s][j] = fma(-z[s], vj, fma(-v[s], zj, phi[s][j]));
                } else {
#pragma unroll
                    for (int s = 0; s < SL; ++s) phi[s][j] = fma(-v[s], zj, phi[s][j]);
                }
            }
            __syncthreads();
        }
    }

    // ---- LDL^T (unit L, no pivoting) of the null-space block s * Phi~[M:, M:]. The
    // block is negated in place when s = -1; L overwrites the strictly lower part of a
    // row (row i keeps l_ij in column j < i), and d the diagonal (kept in dreg).
    if (sgn < 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const bool rowin = li + s * GS >= M;
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (rowin && j >= M) phi[s][j] = -phi[s][j];
        }
    }
    double dreg[SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) dreg[s] = 1.0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        const bool step = j >= M && j < kk;
        if (!step) continue;
        const int js = j / GS, jl = j % GS;
        if (step && li == jl) {
#pragma unroll
            for (int l = 0; l < KMAX; ++l)
                if (l >= j) S.piv[l] = phi[js][l];
        }
        __syncthreads();
        const double d = S.piv[j];
        const bool bad = step && !(d > 0.0 && d < 1e300);
        if (bad) fail = 1;
        const double dinv = step && !bad ? nsc_rcp(d) : 0.0;
        if (li == jl) dreg[js] = step && !bad ? d : 1.0;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int row = li + s * GS;
            const bool upd = step && !bad && row > j && row < kk;
            const double lij = upd ? S.piv[row < KMAX ? row : KMAX - 1] * dinv : 0.0;
            if (upd) phi[s][j] = lij;
#pragma unroll
            for (int l = 1; l < KMAX; ++l)
                if (l > j) phi[s][l] = fma(-lij, S.piv[l], phi[s][l]);
        }
        __syncthreads();
    }
    // (a group that failed fails as a whole: every lane saw the same values)

    // ---- solve with the factors, one operator column at a time: rhs on the point rows
    // and the polynomial RHS gp (S.g[cc], or S.gres[cc] in a refinement step); leaves w
    // in S.wsol[cc] (add: += to it), Q^T w in S.wt[cc] and (Q^T rhs)[:M] in S.btm[cc];
    // returns this lane's largest |w|
    auto solve_col = [&](const int cc, const double (&rhs)[SL], const bool add, const double* gp)
                         __attribute__((always_inline)) {
        double bt[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) bt[s] = rhs[s];
        // bt = Q^T rhs = H_{M-1} ... H_0 rhs
#pragma unroll
        for (int c = 0; c < MMAX; ++c) {
            if (c < m_wave) {
                double vc[SL], loc = 0.0;
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    vc[s] = vr[s][c];
                    loc = fma(vc[s], bt[s], loc);
                }
                const double sp = S.tau[c] * gsum<GS>(loc);
#pragma unroll
                for (int s = 0; s < SL; ++s) bt[s] = fma(-sp, vc[s], bt[s]);
            }
        }
        // w1 = R^-T g: forward substitution, entry a finalised by its owner lane
        double acc[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) acc[s] = 0.0;
#pragma unroll
        for (int a = 0; a < MMAX; ++a) {
            const int as = a / GS, al = a % GS;
            if (a < m_wave) {
                if (li == al) {
                    S.w1[a] = a < M ? (gp[a] - acc[as]) * nsc_rcp(S.R[a][a]) : 0.0;
                    S.btm[cc][a] = bt[as];
                }
                __syncthreads();
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    const int row = li + s * GS;
                    if (row > a && row < MMAX) acc[s] = fma(S.R[a][row], S.w1[a], acc[s]);
                }
            } else if (li == al) {
                S.w1[a] = 0.0;
                S.btm[cc][a] = 0.0;
            }
        }
        __syncthreads();
        // t = s (bt - Phi~[:, :M] w1) on the null-space rows
        double t[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            double v = bt[s];
#pragma unroll
            for (int a = 0; a < MMAX; ++a)
                if (a < m_wave) v = fma(-phi[s][a], S.w1[a], v);
            t[s] = sgn < 0 ? -v : v;
        }
        // forward substitution L y = t: row j, once final, goes to the rows below it
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const bool step = j >= M && j < kk;
            if (!step) continue;
            const int js = j / GS, jl = j % GS;
            if (step && li == jl) S.piv[KMAX] = t[js];
            __syncthreads();
            const double tj = S.piv[KMAX];
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                const int row = li + s * GS;
                if (step && row > j && row < kk) t[s] = fma(-phi[s][j], tj, t[s]);
            }
            __syncthreads();
        }
        // D z = y, then L^T u = z: row i, once final, sends l_ij u_i to every row j < i
#pragma unroll
        for (int s = 0; s < SL; ++s) t[s] = t[s] * nsc_rcp(dreg[s]);
#pragma unroll
        for (int i = KMAX - 1; i >= 1; --i) {
            const bool step = i > M && i < kk;
            if (!step) continue;
            const int is = i / GS, il = i % GS;
            if (step && li == il) {
#pragma unroll
                for (int j = 0; j < KMAX - 1; ++j)
                    if (j < i) S.vec[0][j] = phi[is][j] * t[is];
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                const int row = li + s * GS;
                if (step && row >= M && row < i) t[s] -= S.vec[0][row];
            }
            __syncthreads();
        }
        // wt = [w1; u] (to LDS for the multipliers), then w = Q wt = H_0 ... H_{M-1} wt
        double w[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int row = li + s * GS;
            const bool pt = row < kk;
            const double w1v = S.w1[row < MMAX ? row : 0];
            w[s] = !pt ? 0.0 : (row < M ? w1v : t[s]);
            S.wt[cc][row] = w[s];
        }
#pragma unroll
        for (int c = MMAX - 1; c >= 0; --c) {
            if (c < m_wave) {
                double vc[SL], loc = 0.0;
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    vc[s] = vr[s][c];
                    loc = fma(vc[s], w[s], loc);
                }
                const double sp = S.tau[c] * gsum<GS>(loc);
#pragma unroll
                for (int s = 0; s < SL; ++s) w[s] = fma(-sp, vc[s], w[s]);
            }
        }
        double wmax = 0.0;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int row = li + s * GS;
            S.wsol[cc][row] = add ? S.wsol[cc][row] + w[s] : w[s];
            wmax = fmax(wmax, fabs(w[s]));
        }
        __syncthreads();
        return wmax;
    };

    // multipliers of the last solve of column cc: R lambda = btm - Phi~[:M, :] wt (rows
    // below M of Phi~ are transformed collocation rows, untouched by the LDL^T); into
    // out[cc] (S.lam or S.dlam)
    auto multipliers = [&](const int cc, double (&out)[RC][MMAX]) __attribute__((always_inline)) {
        double q[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int row = li + s * GS;
            double v = S.btm[cc][row < MMAX ? row : 0];
            if (row < MMAX) {
#pragma unroll
                for (int j = 0; j < KMAX; ++j) v = fma(-phi[s][j], S.wt[cc][j], v);
            }
            q[s] = v;
        }
        double acc[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) acc[s] = 0.0;
#pragma unroll
        for (int a = MMAX - 1; a >= 0; --a) {
            const int as = a / GS, al = a % GS;
            if (a < m_wave) {
                if (li == al) out[cc][a] = a < M ? (q[as] - acc[as]) * nsc_rcp(S.R[a][a]) : 0.0;
                __syncthreads();
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    const int row = li + s * GS;
                    if (row < a) acc[s] = fma(S.R[row][a], out[cc][a], acc[s]);
                }
            } else if (li == al) {
                out[cc][a] = 0.0;
            }
        }
        __syncthreads();
    };

    // (a runtime loop over the columns: the column's vectors pass through LDS, so the
    // registers do not grow with the number of columns)
#pragma unroll 1
    for (int cc = 0; cc < n_cols; ++cc) {
        double rhs[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) rhs[s] = S.rhs[cc][li + s * GS];
        solve_col(cc, rhs, false, S.g[cc]);
        if (NREF > 0 || P.out_lambda) multipliers(cc, S.lam);
    }

    if constexpr (NREF > 0) {
        // FP64 refinement on the assembled (untransformed) system: the residual at
        // (w, lambda), a correction through the same factors, w += dw, lambda += dlambda.
        // Up to NREF steps, until the correction of every group of the wavefront is
        // below 1e-7 of its solution; a group whose correction does not settle goes to the
        // rook-pivoted kernel. Each step contracts the error by about the relative
        // accuracy of the factors. Where that is poor, for nearly flat kernels (an IMQ
        // with eps h of a few hundredths: eigenvalues of the null-space block from 1e-5
        // down to 1e-14 against |Phi| ~ k), it stays near 1 and more steps do not help,
        // so every built kernel takes one step (NREF = 1).
        int settled = 0;
#pragma unroll 1
        for (int it = 0; it < NREF; ++it) {
            // (the residual evaluates the monomials again from a copy of the coordinates the
            // compiler must treat as new: reusing the first evaluation would hold them in
            // registers through the whole factorisation)
            double xr[SL][DIM];
#pragma unroll
            for (int s = 0; s < SL; ++s)
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    xr[s][d] = xi[s][d];
                    asm volatile("" : "+v"(xr[s][d]));
                }
            double r[SL][RC];
#pragma unroll
            for (int s = 0; s < SL; ++s) {
#pragma unroll
                for (int c = 0; c < RC; ++c) r[s][c] = S.rhs[c][li + s * GS];
#pragma unroll
                for (int a = 0; a < MMAX; ++a) {
                    if (a < m_wave) {
                        const double pa = (is_pt[s] && a < M) ? mono_graded<DIM>(a, xr[s]) : 0.0;
#pragma unroll
                        for (int c = 0; c < RC; ++c) r[s][c] = fma(-pa, S.lam[c][a], r[s][c]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
    // (the compiler barrier at the top keeps the solution's LDS loads inside each basis
    // case: loads common to all cases would otherwise be hoisted above the switch and held)
#define RBFX_NSC_RESID(BK)                                                               \
    asm volatile("" ::: "memory");                                                       \
    _Pragma("unroll") for (int j = 0; j < KMAX; ++j)                                     \
    {                                                                                    \
        double xj[DIM];                                                                  \
        _Pragma("unroll") for (int d = 0; d < DIM; ++d) xj[d] = S.xi[j][d];              \
        _Pragma("unroll") for (int s = 0; s < SL; ++s)                                   \
        {                                                                                \
            double q2 = 0.0;                                                             \
            _Pragma("unroll") for (int d = 0; d < DIM; ++d)                              \
            {                                                                            \
                const double t_ = xi[s][d] - xj[d];                                      \
                q2 += t_ * t_;                                                           \
            }                                                                            \
            const double v_ = phi_kind<BK>(eps_h, q2);                                   \
            _Pragma("unroll") for (int c = 0; c < RC; ++c) r[s][c] = fma(-v_, S.wsol[c][j], r[s][c]); \
        }                                                                                \
        __builtin_amdgcn_sched_barrier(0);                                               \
        }
            switch (kind) {
            case RBFX_PHS1: RBFX_NSC_RESID(RBFX_PHS1) break;
            case RBFX_PHS3: RBFX_NSC_RESID(RBFX_PHS3) break;
            case RBFX_PHS5: RBFX_NSC_RESID(RBFX_PHS5) break;
            case RBFX_PHS7: RBFX_NSC_RESID(RBFX_PHS7) break;
            case RBFX_IMQ: RBFX_NSC_RESID(RBFX_IMQ) break;
            default: RBFX_NSC_RESID(RBFX_GAUSSIAN) break;
            }
#undef RBFX_NSC_RESID
            // polynomial rows: g - P^T w, into S.gres
#pragma unroll
            for (int a = 0; a < MMAX; ++a) {
                if (a < m_wave) {
                    double pa[SL];
#pragma unroll
                    for (int s = 0; s < SL; ++s)
                        pa[s] = (is_pt[s] && a < M) ? mono_graded<DIM>(a, xr[s]) : 0.0;
#pragma unroll
                    for (int c = 0; c < RC; ++c) {
                        if (c < n_cols) {
                            double loc = 0.0;
#pragma unroll
                            for (int s = 0; s < SL; ++s) loc = fma(pa[s], S.wsol[c][li + s * GS], loc);
                            const double ptw = gsum<GS>(loc);
                            if (li == a % GS) S.gres[c][a] = a < M ? S.g[c][a] - ptw : 0.0;
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int s = 0; s < SL; ++s)
#pragma unroll
                for (int c = 0; c < RC; ++c) S.res[c][li + s * GS] = r[s][c];
            __syncthreads();
            double dmax = 0.0, wmax = 0.0;
#pragma unroll 1
            for (int cc = 0; cc < n_cols; ++cc) {
                double rhs[SL];
#pragma unroll
                for (int s = 0; s < SL; ++s) rhs[s] = S.res[cc][li + s * GS];
                dmax = fmax(dmax, solve_col(cc, rhs, true, S.gres[cc]));
                // (lambda enters the next step's residual; after a single step it is only
                // an output)
                if (NREF > 1 || P.out_lambda) {
                    multipliers(cc, S.dlam);
#pragma unroll
                    for (int s = 0; s < SL; ++s) {
                        const int a = li + s * GS;
                        if (a < MMAX) S.lam[cc][a] += S.dlam[cc][a];
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int s = 0; s < SL; ++s) wmax = fmax(wmax, fabs(S.wsol[cc][li + s * GS]));
            }
            dmax = gmax<GS>(dmax);
            wmax = gmax<GS>(wmax);
            settled = dmax <= 1e-7 * wmax;
            // (groups without work count as settled; the wavefront leaves together)
            if constexpr (NREF > 1)
                if (__ballot(alive && !settled) == 0ull) break;
        }
        if (!settled) fail = 1;
    }

    // ---- outputs
    if (alive && !fail) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int row = li + s * GS;
            if (is_pt[s]) {
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (c < n_cols) P.out_val[(int64_t)c * P.out_nnz + o0 + row] = S.wsol[c][row];
                if (P.out_col && !P.packed) P.out_col[o0 + row] = gidx[s];
            }
            if (P.out_lambda && row < M) {
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (c < n_cols)
                        P.out_lambda[((int64_t)c * P.n_eval + e) * P.n_mono + row] = S.lam[c][row];
            }
        }
        if (li == 0 && P.status) P.status[e] = P.ok_status;
    }
    // hand-offs and failures: queued for the rook-pivoted kernels (one lane per group)
    if ((handoff || (alive && fail)) && li == 0 && P.fail_list) {
        const int q = atomicAdd(P.fail_count, 1);
        P.fail_list[q] = (int32_t)e;
    }
}

// NREF: refinement steps at most (a loop of steps keeps every factor register live
// around it: the 8-lane mapping then spills, the 32-lane ones lose a wave per SIMD)
template <int GS, int SL, int KMAX, int MMAX, int DIM, int RC, int NREF, int W>
inline hipError_t launch_nsc_one(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    if (n_work <= 0) return hipSuccess;
    constexpr int G = 64 / GS;
    BuildDev Q = P;
    Q.n_blocks = (n_work + G - 1) / G;
    hipLaunchKernelGGL((k_nsc<GS, SL, KMAX, MMAX, DIM, RC, NREF, W>), dim3((unsigned)pad_to_xcds(Q.n_blocks)), dim3(64),
                       0, s, Q);
    return hipGetLastError();
}

}  // namespace rbfx
