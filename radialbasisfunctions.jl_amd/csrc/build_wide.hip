// Fused build for stencil systems of 65..RBFX_MAX_N rows (k up to 60 with a degree-3
// polynomial gives 70): one wavefront per stencil, the augmented system [A | b] in LDS
// (row-major: the n columns, padding to a multiple of eight, then the RC right-hand sides;
// odd leading dimension). Lane li owns row m0 + li, and for systems of more than 64 rows
// (m0 = n - 64 > 0) also row li < m0: the second rows are the first ones, the polynomial
// rows, which the pivot order freezes within the first few steps, so the second pass of
// the trailing update is skipped for the rest of the factorisation.
//
// The arithmetic is that of k_build (build_kernel.h): centred, radius-scaled
// saddle-point system assembled in FP64, rook-pivoted LDL^T (pivot tests of Ashcraft,
// Grimes & Lewis) with rows that never move, the trailing update applied to the
// still-active rows, and a back substitution over the rows frozen at their pivot step.
// Only the storage differs: the rows live in LDS instead of registers, so a pivot-row
// read is an LDS load and the pivot search reads the column entries of the lane's own
// rows. The active columns are kept in a prefix of every row, and the trailing update
// runs over that prefix eight columns per LDS round trip.
#include "build_kernel.h"

namespace rbfx {

static_assert(RBFX_MAX_N <= 128, "at most two rows per lane");

// maximum over the 64 lanes, as a scalar: rotations within each row of 16 lanes, then the
// four row maxima read from lanes 0, 16, 32 and 48 (no LDS permute)
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t k)
{
    k = max(k, (uint32_t)dpp_i<0x121>((int)k));
    k = max(k, (uint32_t)dpp_i<0x122>((int)k));
    k = max(k, (uint32_t)dpp_i<0x124>((int)k));
    k = max(k, (uint32_t)dpp_i<0x128>((int)k));
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)k, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)k, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)k, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)k, 48);
    return max(max(a, b), max(c, d));
}

__device__ __forceinline__ uint32_t row_key(double v, int row)
{
    const uint32_t hi = (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);
    return (hi & ~127u) | (uint32_t)(127 - row);
}

// TWO: rows per lane (1 for systems of up to 64 rows, 2 up to 128)
template <int DIM, int RC, int TWO>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64))) void k_build_wide(const BuildDev P, int rows_cap,
                                                                                  int lda)
{
    constexpr int ROWS = 64 * TWO;
    // the system (rows_cap x lda), then the point tables of the assembly, sized by rows_cap
    // so that small systems leave LDS for more wavefronts per CU: the scaled points, and
    // for builds with boundary data the normals, boundary coefficients and types
    extern __shared__ __attribute__((aligned(16))) double s_A[];
    double(*s_xi)[DIM] = reinterpret_cast<double(*)[DIM]>(s_A + (int64_t)rows_cap * lda);
    double(*s_bn)[DIM] = s_xi + rows_cap;
    double* s_ba = reinterpret_cast<double*>(s_bn + rows_cap);
    double* s_bb = s_ba + rows_cap;
    int* s_bt = reinterpret_cast<int*>(s_bb + rows_cap);
    const bool bnd_tables = P.is_boundary != nullptr;
    __shared__ int s_piv[ROWS], s_blk[ROWS];
    __shared__ int s_mono[kMaxMono][3];
    __shared__ TermDev s_terms[RBFX_MAX_TERMS];

    const int L = threadIdx.x;
    const int n_cols = P.n_cols, n_terms = P.n_terms;
    for (int i = L; i < kMaxMono * 3; i += 64) (&s_mono[0][0])[i] = (&P.mono_e[0][0])[i];
    if (L < RBFX_MAX_TERMS) s_terms[L] = P.terms[L];

    // one stencil per workgroup: every exit below is uniform over the workgroup
    const int64_t lb = xcd_contiguous_block(P.n_blocks);
    if (lb < 0) return;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    if (lb >= n_work) return;
    const int64_t e = P.list ? (int64_t)P.list[lb] : (P.order ? (int64_t)P.order[lb] : lb);
    if (e < 0 || e >= P.n_eval) return;
    const int64_t s0 = P.adj_ptr ? P.adj_ptr[e] : e * (int64_t)P.k;
    const int kk = P.adj_ptr ? (int)(P.adj_ptr[e + 1] - s0) : P.k;
    const int64_t o0 = P.out_ptr ? P.out_ptr[e] : e * (int64_t)P.k;
    const int M = stencil_n_mono(P, e, DIM);
    if (M < 0) {
        if (L == 0 && P.status) P.status[e] = RBFX_ST_INVALID;
        return;
    }
    if (kk < 1 || M + kk > rows_cap) {
        if (L == 0 && P.status) P.status[e] = RBFX_ST_TOO_LARGE;
        return;
    }
    const int n = M + kk;
    const int rhs0 = (n + 7) & ~7;  // position of the first right-hand side in a row
    // row ownership: slot 0 of lane L holds row m0 + L, slot 1 (TWO == 2) row L < m0
    const int m0 = TWO == 2 ? max(n - 64, 0) : 0;
    auto rowof = [&](int t) -> int { return (TWO == 2 && t == 1) ? L : L + m0; };
    auto owns = [&](int t) -> bool { return (TWO == 2 && t == 1) ? L < m0 : L + m0 < n; };
    // the owner of row r (uniform): its lane and slot
    auto lane_of = [&](int r) -> int { return r < m0 ? r : r - m0; };
    auto slot_t = [&](int r) -> int { return (TWO == 2 && r < m0) ? 1 : 0; };

    // ---- the stencil points this lane loads: q = L (and q = L + 64)
    double xe[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xe[d] = P.eval[e * DIM + d];
    int gidx[TWO] = {};
    double D[TWO][DIM];
    double r2max = 0.0;
    bool ib[TWO] = {};
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int q = L + 64 * t;
#pragma unroll
        for (int d = 0; d < DIM; ++d) D[t][d] = 0.0;
        if (q < kk) {
            gidx[t] = P.adj_idx[s0 + q];
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                D[t][d] = P.data[(int64_t)gidx[t] * DIM + d] - xe[d];
                r2 += D[t][d] * D[t][d];
            }
            r2max = fmax(r2max, r2);
            ib[t] = P.is_boundary && P.is_boundary[gidx[t]];
        }
    }
    // stencil classification (types.jl:208-223)
    const bool hermite = P.is_boundary && (__ballot(ib[0] || ib[TWO - 1]) != 0ull);
    // a Dirichlet eval node is an identity row whatever its neighbours are (the row
    // length count_nonzeros gives it, execution.jl:39-69)
    if (P.is_boundary && e < P.n_data && P.is_boundary[e]) {
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
    double h = sqrt(group_max<64>(r2max));
    if (!(h > 0.0)) h = 1.0;
    const double hinv = 1.0 / h;
    unsigned long long at_eval[TWO];
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int q = L + 64 * t;
        bool eq = false;
        if (q < kk) {
            int bt = 0;
            double ba = 0.0, bb = 0.0;
            if (hermite && ib[t]) {
                const int bi = P.bnd_index[gidx[t]];
                ba = P.bc_alpha[bi];
                bb = P.bc_beta[bi];
#pragma unroll
                for (int d = 0; d < DIM; ++d) s_bn[q][d] = P.normals[(int64_t)bi * DIM + d];
                bt = (ba == 1.0 && bb == 0.0) ? 1 : 2;
            }
            eq = hermite;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                s_xi[q][d] = D[t][d] / h;
                eq = eq && (D[t][d] == 0.0);
            }
            if (bnd_tables) {
                s_bt[q] = bt;
                s_ba[q] = ba;
                s_bb[q] = bb / h;  // beta acts through a first derivative
            }
        }
        at_eval[t] = __ballot(eq);
    }
    int eval_local = 0;
    if (hermite) {
        const unsigned long long ae1 = TWO == 2 ? at_eval[TWO - 1] : 0ull;
        if (!at_eval[0] && !ae1) {
            if (L == 0 && P.status) P.status[e] = RBFX_ST_EVAL_NOT_FOUND;
            return;
        }
        eval_local = at_eval[0] ? __builtin_ctzll(at_eval[0]) : 64 + __builtin_ctzll(ae1);
    }
    __syncthreads();

    BasisDev B;
    B.kind = P.basis_kind;
    B.eps = (B.kind == RBFX_IMQ || B.kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;

    // ---- assembly (FP64): the lane's rows; columns 0..n-1, then the RHS at rhs0..rhs0+RC-1
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int i = rowof(t);
        if (!owns(t)) continue;
        double* Ai = s_A + (int64_t)i * lda;
        const bool is_poly = i < M;
        const int pi = i - M;
        double xi[DIM], bn[DIM];
        int bt = 0;
        double ba = 0.0, bb = 0.0;
        if (!is_poly) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xi[d] = s_xi[pi][d];
                bn[d] = hermite ? s_bn[pi][d] : 0.0;
            }
            if (hermite) {
                bt = s_bt[pi];
                ba = s_ba[pi];
                bb = s_bb[pi];
            }
        }
        // columns 0..M-1. A point row holds its monomials; the polynomial rows' entries at
        // the point columns are the same numbers (symmetry of the saddle-point matrix), so
        // each point lane also writes its value into column i of polynomial row a, and the
        // polynomial lanes only zero their corner block.
        for (int a = 0; a < M; ++a) {
            if (is_poly) {
                Ai[a] = 0.0;
            } else {
                double v = mono_op<DIM>(s_mono[a], RBFX_OP_VALUE, 0, 0, xi);
                if (hermite && bt == 2) v = ba * v + bb * mono_normal<DIM>(s_mono[a], bn, xi);
                Ai[a] = v;
                s_A[(int64_t)a * lda + i] = v;
            }
        }
        // columns M..n-1 of the point rows: phi between the points (instantiated once per
        // basis kind; without boundary data four independent entries per iteration)
        auto point_cols = [&](auto KC) __attribute__((always_inline)) {
            if (is_poly) return;
            if (!hermite) {
#pragma unroll 4
                for (int jj = 0; jj < kk; ++jj) {
                    double q2 = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        const double t_ = xi[d] - s_xi[jj][d];
                        q2 += t_ * t_;
                    }
                    Ai[M + jj] = phi_r2_k<decltype(KC)::value>(B.eps, q2);
                }
                return;
            }
            for (int jj = 0; jj < kk; ++jj) {
                double xj[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xj[d] = s_xi[jj][d];
                const int tj = s_bt[jj];
                double Dj[DIM], q2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    Dj[d] = xi[d] - xj[d];
                    q2 += Dj[d] * Dj[d];
                }
                double v = phi_r2_k<decltype(KC)::value>(B.eps, q2);
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
                Ai[M + jj] = v;
            }
        };
        switch (B.kind) {
        case RBFX_PHS1: point_cols(std::integral_constant<int, RBFX_PHS1>{}); break;
        case RBFX_PHS3: point_cols(std::integral_constant<int, RBFX_PHS3>{}); break;
        case RBFX_PHS5: point_cols(std::integral_constant<int, RBFX_PHS5>{}); break;
        case RBFX_PHS7: point_cols(std::integral_constant<int, RBFX_PHS7>{}); break;
        case RBFX_IMQ: point_cols(std::integral_constant<int, RBFX_IMQ>{}); break;
        default: point_cols(std::integral_constant<int, RBFX_GAUSSIAN>{}); break;
        }
        // right-hand sides
        double bcol[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) bcol[c] = 0.0;
        if (!is_poly) {
            double De[DIM], q2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                De[d] = -xi[d];
                q2 += De[d] * De[d];
            }
            for (int tt = 0; tt < n_terms; ++tt) {
                const TermDev tm = s_terms[tt];
                double v = phi_term<DIM>(B, tm.kind, tm.d1, tm.d2, De, q2);
                if (hermite && bt == 2) v = ba * v + bb * phi_term_normal<DIM>(B, tm.kind, tm.d1, De, q2, bn);
                const double sc = term_scale(tm, hinv);
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (tm.col == c) bcol[c] += sc * v;
            }
        } else {
            double z[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) z[d] = 0.0;
            const int(&ea)[3] = s_mono[i];
            if (hermite && s_bt[eval_local] == 2) {
                // the eval node's boundary operator applied to P (assembly.jl:452-467)
                const double eav = s_ba[eval_local], ebv = s_bb[eval_local];
                double en[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) en[d] = s_bn[eval_local][d];
                const double v = eav * mono_op<DIM>(ea, RBFX_OP_VALUE, 0, 0, z) + ebv * mono_normal<DIM>(ea, en, z);
#pragma unroll
                for (int c = 0; c < RC; ++c) bcol[c] = v;
            } else {
                for (int tt = 0; tt < n_terms; ++tt) {
                    const TermDev tm = s_terms[tt];
                    const double v = term_scale(tm, hinv) * mono_op<DIM>(ea, tm.kind, tm.d1, tm.d2, z);
#pragma unroll
                    for (int c = 0; c < RC; ++c)
                        if (tm.col == c) bcol[c] += v;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) Ai[rhs0 + c] = bcol[c];
    }
    __syncthreads();

    // ---- factorisation with rook pivoting; rows never move, pivoted rows freeze
    const double alpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8
    bool act[TWO];
    int myslot[TWO];
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        act[t] = owns(t);
        myslot[t] = 1 << 20;
    }
    int slot = 0;
    bool failed = false;
    auto A = [&](int i, int j) -> double { return s_A[(int64_t)i * lda + j]; };
    // Column c of the active rows goes to this lane's registers (v[t] = A(rowof(t), c)),
    // and the largest |A(r, c)| over the active rows r != excl comes back as a packed
    // (magnitude, row) key. The pivot tests then take single entries of that column from
    // their owning lanes (colval) instead of another LDS round trip.
    //
    // Column positions: the columns of the active rows sit in positions 0..n_act-1 of
    // every row (the right-hand sides at rhs0.., past the padding). Freezing a pivot swaps
    // its column with the last active position in every row, frozen rows included, so
    // the trailing update only touches the active prefix, and an entry (r, c) is always
    // at position pos(c) of row r. pos[t] is the position of column rowof(t).
    int pos[TWO];
#pragma unroll
    for (int t = 0; t < TWO; ++t) pos[t] = owns(t) ? rowof(t) : 1 << 20;
    int n_act = n;
    auto posof = [&](int c) -> int {
        const int cc = __builtin_amdgcn_readfirstlane(c);
        return slot_t(cc) ? rdlane(pos[TWO - 1], lane_of(cc)) : rdlane(pos[0], lane_of(cc));
    };
    double v[TWO];
    auto colkey = [&](int c, int excl) -> uint32_t {
        uint32_t k = 0u;
        const int pc = posof(c);
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int r = rowof(t);
            v[t] = act[t] ? A(r, pc) : 0.0;
            if (act[t] && r != excl) k = max(k, row_key(fabs(v[t]), r));
        }
        return wave_max_u32(k);
    };
    auto colval = [&](int r) -> double {
        const int rr = __builtin_amdgcn_readfirstlane(r);
        return slot_t(rr) ? rdlane(v[TWO - 1], lane_of(rr)) : rdlane(v[0], lane_of(rr));
    };
    for (;;) {
        const unsigned long long b0 = __ballot(act[0]), b1 = TWO == 2 ? __ballot(act[TWO - 1]) : 0ull;
        if (!b0 && !b1) break;
        // the lowest active row: slot 1 holds rows 0..m0-1, slot 0 rows m0..n-1
        const int c0 = b1 ? (int)__builtin_ctzll(b1) : m0 + (int)__builtin_ctzll(b0);
        // column c0 = row c0 by symmetry of the active block
        const uint32_t ck = colkey(c0, c0);
        int imax = ck ? 127 - (int)(ck & 127u) : c0;
        const double akk = colval(c0);
        double colmax = ck ? fabs(colval(imax)) : 0.0;
        const double absakk = fabs(akk);
        int kstep = 1, p = c0, kp = c0;
        // A(p, p), A(kp, kp) and, for a 2x2 pivot, A(p, kp) as the search ends
        double app = akk, akp = akk, apk = akk;
        bool need_rook = absakk < alpha * colmax;
        if (!(absakk > 0.0 || colmax > 0.0)) {
            failed = true;  // zero column: singular
            break;
        }
        for (int it = 0; it < 2 * ROWS && need_rook; ++it) {
            const uint32_t rk = colkey(imax, imax);
            const int ri = rk ? 127 - (int)(rk & 127u) : imax;
            const double rowmax = rk ? fabs(colval(ri)) : 0.0;
            const double aim = colval(imax);
            const double aii = fabs(aim);
            if (!(aii < alpha * rowmax)) {
                kp = imax;
                akp = aim;
                need_rook = false;
            } else if (p == ri || rowmax <= colmax || it == 2 * ROWS - 1) {
                kp = imax;
                kstep = 2;
                akp = aim;
                apk = colval(p);  // column imax holds A(p, imax)
                need_rook = false;
            } else {
                p = imax;
                app = aim;
                colmax = rowmax;
                imax = ri;
            }
        }
        // (the same on every lane; read from lane 0 so the compiler keeps the pivot
        // bookkeeping and the column masks below in scalar registers)
        kstep = __builtin_amdgcn_readfirstlane(kstep);
        const int p0 = __builtin_amdgcn_readfirstlane((kstep == 1) ? kp : p);
        const int p1 = __builtin_amdgcn_readfirstlane(kp);  // equals p0 for a 1x1 pivot
        // A(p0, p0), A(p1, p1), A(p0, p1)
        const double d0 = kstep == 1 ? akp : app, d1 = akp, bq = kstep == 1 ? akp : apk;
        const double det = d0 * d1 - bq * bq;
        if ((kstep == 1 && d0 == 0.0) || (kstep == 2 && det == 0.0)) {
            failed = true;
            break;
        }
        const double idet = 1.0 / det;
        if (L == 0) {
            s_piv[slot] = p0;
            s_blk[slot] = kstep;
            if (kstep == 2) {
                s_piv[slot + 1] = p1;
                s_blk[slot + 1] = 0;
            }
        }
        // freeze: the pivot columns move to the end of the active prefix in every row
        auto swap_to = [&](int c, int last) {
            const int pc = posof(c);
            if (pc == last) return;
#pragma unroll
            for (int t = 0; t < TWO; ++t) {
                if (!owns(t)) continue;
                const int r = rowof(t);
                double* Ar = s_A + (int64_t)r * lda;
                const double x0 = Ar[pc], x1 = Ar[last];
                Ar[pc] = x1;
                Ar[last] = x0;
                if (pos[t] == last) pos[t] = pc;
                else if (r == c) pos[t] = last;
            }
        };
        swap_to(p0, n_act - 1);
        if (kstep == 2) swap_to(p1, n_act - 2);
        n_act -= kstep;
        const double* U0 = s_A + (int64_t)p0 * lda;
        const double* U1 = s_A + (int64_t)p1 * lda;
        // The active prefix (rounded up to eight: the padding before rhs0 and the frozen
        // columns past the prefix are never read) and the right-hand sides are updated
        // eight columns at a time: uniform column offsets become immediate offsets of
        // paired LDS loads, all loads of a chunk are issued before its stores, and per
        // element the FMAs run in the same order as a column-by-column update.
        const int na8 = (n_act + 7) & ~7;
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int r = rowof(t);
            if (!act[t] || r == p0 || r == p1) continue;
            const double u0 = U0[pos[t]];
            double l0, l1 = 0.0;
            if (kstep == 1) {
                l0 = u0 / d0;
            } else {
                const double u1 = U1[pos[t]];
                l0 = (u0 * d1 - u1 * bq) * idet;
                l1 = (u1 * d0 - u0 * bq) * idet;
            }
            const double ml0 = -l0, ml1 = -l1;
            double* Ar = s_A + (int64_t)r * lda;
            // a 1x1 and a 2x2 pivot each have their own loop, so a chunk's loads stay together
            auto chunk1 = [&](auto W, int j0) __attribute__((always_inline)) {
                constexpr int CW = decltype(W)::value;
                double a[CW], w0[CW];
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    a[q] = Ar[j0 + q];
                    w0[q] = U0[j0 + q];
                }
#pragma unroll
                for (int q = 0; q < CW; ++q) Ar[j0 + q] = fma(ml0, w0[q], a[q]);
            };
            auto chunk2 = [&](auto W, int j0) __attribute__((always_inline)) {
                constexpr int CW = decltype(W)::value;
                double a[CW], w0[CW], w1[CW];
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    a[q] = Ar[j0 + q];
                    w0[q] = U0[j0 + q];
                    w1[q] = U1[j0 + q];
                }
#pragma unroll
                for (int q = 0; q < CW; ++q) Ar[j0 + q] = fma(ml1, w1[q], fma(ml0, w0[q], a[q]));
            };
            using C8 = std::integral_constant<int, 8>;
            using CR = std::integral_constant<int, RC>;
            if (kstep == 1) {
                for (int j0 = 0; j0 < na8; j0 += 8) chunk1(C8{}, j0);
                chunk1(CR{}, rhs0);
            } else {
                for (int j0 = 0; j0 < na8; j0 += 8) chunk2(C8{}, j0);
                chunk2(CR{}, rhs0);
            }
        }
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int r = rowof(t);
            if (act[t] && r == p0) {
                act[t] = false;
                myslot[t] = slot;
            }
            if (act[t] && kstep == 2 && r == p1) {
                act[t] = false;
                myslot[t] = slot + 1;
            }
        }
        slot += kstep;
        __syncthreads();
    }
    __syncthreads();

    // ---- back substitution over the frozen rows, last pivot first
    double x[TWO][RC], acc[TWO][RC];
#pragma unroll
    for (int t = 0; t < TWO; ++t)
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            x[t][c] = 0.0;
            acc[t][c] = 0.0;
        }
    if (!failed) {
        const int nslots = slot;
        // the right-hand sides (final once every row has frozen) and the pivot list go to
        // registers: each slot then costs one round of LDS loads (the pivot block and the
        // lane's entries in the pivot columns), issued together
        double y[TWO][RC];
        int pv[TWO], bk[TWO];
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
#pragma unroll
            for (int c = 0; c < RC; ++c) y[t][c] = owns(t) ? A(rowof(t), rhs0 + c) : 0.0;
            // slot sl of the pivot list sits in lane sl % 64, register sl / 64
            const int sl = L + 64 * t;
            pv[t] = sl < nslots ? s_piv[sl] : 0;
            bk[t] = sl < nslots ? s_blk[sl] : -1;
        }
        auto slot_of = [&](const int(&v)[TWO], int sl) -> int {
            return (TWO == 2 && sl >= 64) ? rdlane(v[TWO - 1], sl & 63) : rdlane(v[0], sl & 63);
        };
        for (int s = nslots - 1; s >= 0; --s) {
            const int blk = slot_of(bk, s);
            const int bprev = s > 0 ? slot_of(bk, s - 1) : -1;
            const bool one = blk == 1, two = (blk == 0) && (bprev == 2);
            if (!one && !two) continue;  // first slot of a 2x2 block: solved with the second
            const int p = slot_of(pv, one ? s : s - 1);
            const int q = two ? slot_of(pv, s) : p;
            const int pl = lane_of(p), pt = slot_t(p), ql = lane_of(q), qt = slot_t(q);
            const int pp = posof(p), pq = posof(q);
            const double dp = A(p, pp), dq = A(q, pq), bq = A(p, pq);
            double arp[TWO], arq[TWO];
#pragma unroll
            for (int t = 0; t < TWO; ++t) {
                const int r = rowof(t);
                arp[t] = owns(t) ? A(r, pp) : 0.0;
                arq[t] = (two && owns(t)) ? A(r, pq) : 0.0;
            }
            // (one division per pivot: a 1x1 pivot needs only 1/dp, a 2x2 only 1/det)
            double idet = 0.0, dinv = 0.0;
            if (two) idet = 1.0 / (dp * dq - bq * bq);
            else dinv = 1.0 / dp;
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                if (c >= n_cols) continue;
                // y - acc of rows p and q, from their owning lanes
                const double rp0 = y[0][c] - acc[0][c];
                const double rp1 = TWO == 2 ? y[TWO - 1][c] - acc[TWO - 1][c] : 0.0;
                const double yp = pt ? rdlane(rp1, pl) : rdlane(rp0, pl);
                const double yq = qt ? rdlane(rp1, ql) : rdlane(rp0, ql);
                const double x0 = two ? (yp * dq - yq * bq) * idet : yp * dinv;
                const double x1 = two ? (yq * dp - yp * bq) * idet : 0.0;
#pragma unroll
                for (int t = 0; t < TWO; ++t) {
                    const int r = rowof(t);
                    if (!owns(t)) continue;
                    if (r == p) x[t][c] = x0;
                    if (two && r == q) x[t][c] = x1;
                    const bool upd = one ? (myslot[t] < s) : (myslot[t] < s - 1);
                    if (upd) acc[t][c] += two ? arp[t] * x0 + arq[t] * x1 : arp[t] * x0;
                }
            }
        }
    }

    // ---- outputs: weights of the point rows, multipliers of the polynomial rows
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int i = rowof(t);
        if (!owns(t)) continue;
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            if (c >= n_cols) continue;
            const double v = failed ? __builtin_nan("") : x[t][c];
            if (i >= M) P.out_val[(int64_t)c * P.out_nnz + o0 + (i - M)] = v;
            else if (P.out_lambda) P.out_lambda[((int64_t)c * P.n_eval + e) * P.n_mono + i] = v;
        }
        // row i holds stencil point i - M
        if (i >= M && P.out_col) P.out_col[o0 + (i - M)] = P.adj_idx[s0 + (i - M)];
    }
    if (L == 0 && P.status) P.status[e] = failed ? RBFX_ST_SINGULAR : P.ok_status;
}

// ---------------------------------------------------------------- packed symmetric variant
//
// k_build_sym: the same build with half the LDS, for more wavefronts per CU. The
// saddle-point matrix is symmetric, and so is every trailing block of its elimination,
// so only the lower triangle is kept, by position: lane li owns position li (and
// li + 64), row j holding its entries at positions 0..j. Pivoting is a symmetric
// permutation: the pivot row and column move together to the end of the active prefix
// (rows and columns of two positions swap), so the active block is always positions
// 0..n_act-1, its pivot rows are the frozen rows just past it (contiguous, read as
// broadcasts by the update), and a frozen row keeps its entries at the positions of
// the rows frozen after it, which the back substitution reads. The original row index
// travels with each row (the pivot search starts from the active row of the lowest
// original index, as in k_build_wide; ties in the search go to the lower position).
// The stencil points take positions 0..k-1 and the monomials k..k+m-1, so for systems
// of more than 64 rows the second lane slot holds monomial rows, whose assembly is a
// zero fill, and which the first steps freeze (every step freezes the top active
// positions): the second update pass is skipped from then on.
//
// Row starts: row j holds positions 0..j padded to a multiple of eight (the update's
// eight-column chunks stay inside the row), plus the fewest entries (0..31) that put the
// starts of each 16 consecutive rows in 16 different 8-byte bank pairs modulo 16 and of
// each 32 consecutive rows in 32 different ones modulo 32: a column read by the lanes of
// a wavefront is then free of bank conflicts whether the LDS serves 16 or 32 lanes of
// 8-byte reads per cycle.
constexpr int kSymRows = 128;
struct SymRowStarts {
    int rb[kSymRows + 1];
};
constexpr SymRowStarts make_sym_row_starts()
{
    SymRowStarts s{};
    bool u32[kSymRows / 32 + 1][32] = {};
    bool u16[kSymRows / 16 + 1][16] = {};
    s.rb[0] = 0;
    u32[0][0] = true;
    u16[0][0] = true;
    for (int j = 1; j <= kSymRows; ++j) {
        const int base = s.rb[j - 1] + ((j + 7) & ~7);  // row j - 1 holds j entries
        int p = 0;
        while (p < 31 && (u32[j / 32][(base + p) & 31] || u16[j / 16][(base + p) & 15])) ++p;
        s.rb[j] = base + p;
        u32[j / 32][(base + p) & 31] = true;
        u16[j / 16][(base + p) & 15] = true;
    }
    return s;
}
__constant__ SymRowStarts kSymRb = make_sym_row_starts();
constexpr SymRowStarts kSymRbHost = make_sym_row_starts();
static_assert(RBFX_MAX_N <= kSymRows, "positions fit the row-start table");
constexpr bool sym_rows_spread()
{
    for (int g = 0; g < kSymRows; g += 16) {
        bool seen[16] = {};
        for (int j = g; j < g + 16; ++j) {
            if (seen[kSymRbHost.rb[j] & 15]) return false;
            seen[kSymRbHost.rb[j] & 15] = true;
        }
    }
    for (int g = 0; g < kSymRows; g += 32) {
        bool seen[32] = {};
        for (int j = g; j < g + 32; ++j) {
            if (seen[kSymRbHost.rb[j] & 31]) return false;
            seen[kSymRbHost.rb[j] & 31] = true;
        }
    }
    return true;
}
static_assert(sym_rows_spread(), "row starts spread over the LDS banks");

// HERM: the build may carry boundary data. Launches without any (P.is_boundary null)
// whose LDS is small enough use the instantiation without the Hermite paths, whose
// smaller register allocation leaves room for more wavefronts per SIMD
template <int DIM, int RC, int TWO, bool HERM>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64))) void k_build_sym(const BuildDev P, int rows_cap,
                                                                                 int tri)
{
    constexpr int ROWS = 64 * TWO;
    static_assert(HERM || TWO == 1, "the instantiation without Hermite paths serves up to 64 rows");
    // the packed triangle (tri doubles) is the whole dynamic LDS: the right-hand sides,
    // the stencil points and their boundary data stay in the registers of the lanes that
    // own them and are read by the others with lane reads
    extern __shared__ __attribute__((aligned(16))) double s_T[];
    // (the static LDS counts toward the wavefronts per CU: the block list is bytes, and
    // the monomial exponents and operator terms are read from the kernel arguments)
    __shared__ int8_t s_blk[ROWS];

    const int L = threadIdx.x;
    const int n_cols = P.n_cols, n_terms = P.n_terms;
    auto mono_exps = [&](int a, int(&ex)[3]) {
        ex[0] = P.mono_e[a][0];
        ex[1] = P.mono_e[a][1];
        ex[2] = P.mono_e[a][2];
    };

    // one stencil per workgroup: every exit below is uniform over the workgroup
    const int64_t lb = xcd_contiguous_block(P.n_blocks);
    if (lb < 0) return;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    if (lb >= n_work) return;
    const int64_t e = P.list ? (int64_t)P.list[lb] : (P.order ? (int64_t)P.order[lb] : lb);
    if (e < 0 || e >= P.n_eval) return;
    const int64_t s0 = P.adj_ptr ? P.adj_ptr[e] : e * (int64_t)P.k;
    const int kk = P.adj_ptr ? (int)(P.adj_ptr[e + 1] - s0) : P.k;
    const int64_t o0 = P.out_ptr ? P.out_ptr[e] : e * (int64_t)P.k;
    const int M = stencil_n_mono(P, e, DIM);
    if (M < 0) {
        if (L == 0 && P.status) P.status[e] = RBFX_ST_INVALID;
        return;
    }
    if (kk < 1 || M + kk > rows_cap) {
        if (L == 0 && P.status) P.status[e] = RBFX_ST_TOO_LARGE;
        return;
    }
    const int n = M + kk;

    // ---- the stencil points this lane loads: q = L (and q = L + 64)
    double xe[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xe[d] = P.eval[e * DIM + d];
    int gidx[TWO] = {};
    double D[TWO][DIM];
    double r2max = 0.0;
    bool ib[TWO] = {};
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int q = L + 64 * t;
#pragma unroll
        for (int d = 0; d < DIM; ++d) D[t][d] = 0.0;
        if (q < kk) {
            gidx[t] = P.adj_idx[s0 + q];
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                D[t][d] = P.data[(int64_t)gidx[t] * DIM + d] - xe[d];
                r2 += D[t][d] * D[t][d];
            }
            r2max = fmax(r2max, r2);
            ib[t] = HERM && P.is_boundary && P.is_boundary[gidx[t]];
        }
    }
    // stencil classification (types.jl:208-223)
    const bool hermite = HERM && P.is_boundary && (__ballot(ib[0] || ib[TWO - 1]) != 0ull);
    // a Dirichlet eval node is an identity row whatever its neighbours are (the row
    // length count_nonzeros gives it, execution.jl:39-69)
    if (HERM && P.is_boundary && e < P.n_data && P.is_boundary[e]) {
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
    double h = sqrt(group_max<64>(r2max));
    if (!(h > 0.0)) h = 1.0;
    const double hinv = 1.0 / h;
    unsigned long long at_eval[TWO];
    // this lane's points q = L + 64 t: the scaled coordinates and, for stencils with
    // boundary points, the boundary type (1 Dirichlet, 2 Neumann/Robin), the coefficients
    // (beta scaled by 1/h: it acts through a first derivative) and the normal
    double xo[TWO][DIM], bno[TWO][DIM], bao[TWO], bbo[TWO];
    int bto[TWO];
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int q = L + 64 * t;
        bool eq = false;
        bao[t] = 0.0;
        bbo[t] = 0.0;
        bto[t] = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xo[t][d] = 0.0;
            bno[t][d] = 0.0;
        }
        if (q < kk) {
            if (hermite && ib[t]) {
                const int bi = P.bnd_index[gidx[t]];
                const double ba = P.bc_alpha[bi], bb = P.bc_beta[bi];
#pragma unroll
                for (int d = 0; d < DIM; ++d) bno[t][d] = P.normals[(int64_t)bi * DIM + d];
                bto[t] = (ba == 1.0 && bb == 0.0) ? 1 : 2;
                bao[t] = ba;
                bbo[t] = bb / h;
            }
            eq = hermite;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xo[t][d] = D[t][d] / h;
                eq = eq && (D[t][d] == 0.0);
            }
        }
        at_eval[t] = __ballot(eq);
    }
    // point q's data (q uniform), from the lane that loaded it
    auto pq_sel = [&](int q) -> int { return __builtin_amdgcn_readfirstlane(q); };
    auto px = [&](int q, int d) -> double {
        const int c = pq_sel(q);
        return (TWO == 2 && c >= 64) ? rdlane(xo[TWO - 1][d], c & 63) : rdlane(xo[0][d], c & 63);
    };
    auto pbn = [&](int q, int d) -> double {
        const int c = pq_sel(q);
        return (TWO == 2 && c >= 64) ? rdlane(bno[TWO - 1][d], c & 63) : rdlane(bno[0][d], c & 63);
    };
    auto pba = [&](int q) -> double {
        const int c = pq_sel(q);
        return (TWO == 2 && c >= 64) ? rdlane(bao[TWO - 1], c & 63) : rdlane(bao[0], c & 63);
    };
    auto pbb = [&](int q) -> double {
        const int c = pq_sel(q);
        return (TWO == 2 && c >= 64) ? rdlane(bbo[TWO - 1], c & 63) : rdlane(bbo[0], c & 63);
    };
    auto pbt = [&](int q) -> int {
        const int c = pq_sel(q);
        return (TWO == 2 && c >= 64) ? rdlane(bto[TWO - 1], c & 63) : rdlane(bto[0], c & 63);
    };
    int eval_local = 0;
    if (hermite) {
        const unsigned long long ae1 = TWO == 2 ? at_eval[TWO - 1] : 0ull;
        if (!at_eval[0] && !ae1) {
            if (L == 0 && P.status) P.status[e] = RBFX_ST_EVAL_NOT_FOUND;
            return;
        }
        eval_local = at_eval[0] ? __builtin_ctzll(at_eval[0]) : 64 + __builtin_ctzll(ae1);
    }
    __syncthreads();

    BasisDev B;
    B.kind = P.basis_kind;
    B.eps = (B.kind == RBFX_IMQ || B.kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;

    // positions: point q at q, monomial a at kk + a; orig: the row index of the
    // saddle-point system (monomials 0..M-1, then the points)
    int rbj[TWO], orig[TWO];
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int j = L + 64 * t;
        rbj[t] = kSymRb.rb[j];
        orig[t] = j < kk ? M + j : j - kk;
    }
    // start of row c (c uniform): from the lane that owns position c
    auto rb_of = [&](int c) -> int {
        const int cc = __builtin_amdgcn_readfirstlane(c);
        return (TWO == 2 && cc >= 64) ? rdlane(rbj[TWO - 1], cc & 63) : rdlane(rbj[0], cc & 63);
    };

    // ---- assembly (FP64) of the lower triangle and the right-hand sides (registers: lane
    // L holds those of its positions)
    double bcol[TWO][RC];
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
#pragma unroll
        for (int c = 0; c < RC; ++c) bcol[t][c] = 0.0;
        const int j = L + 64 * t;
        if (j >= n) continue;
        double* Tj = s_T + rbj[t];
        const bool is_poly = j >= kk;
        if (is_poly) {
            // monomial row a = j - kk: its entries at the monomial positions kk..j are zero
            // (its point entries are written by the point lanes below)
            for (int c = kk; c <= j; ++c) Tj[c] = 0.0;
        }
        double xi[DIM], bn[DIM];
        int bt = 0;
        double ba = 0.0, bb = 0.0;
        if (!is_poly) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xi[d] = xo[t][d];
                bn[d] = bno[t][d];
            }
            if (hermite) {
                bt = bto[t];
                ba = bao[t];
                bb = bbo[t];
            }
            // the point's monomial values: column j of monomial row a (position kk + a)
            for (int a = 0; a < M; ++a) {
                int ea[3];
                mono_exps(a, ea);
                double v = mono_op<DIM>(ea, RBFX_OP_VALUE, 0, 0, xi);
                if (hermite && bt == 2) v = ba * v + bb * mono_normal<DIM>(ea, bn, xi);
                s_T[rb_of(kk + a) + j] = v;
            }
        }
        // phi between point j and the points jj <= j (positions jj): instantiated once per
        // basis kind; without boundary data four independent entries per iteration
        auto point_cols = [&](auto KC) __attribute__((always_inline)) {
            if (is_poly) return;
            if (!hermite) {
#pragma unroll 4
                for (int jj = 0; jj < kk; ++jj) {
                    double q2 = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        const double t_ = xi[d] - px(jj, d);
                        q2 += t_ * t_;
                    }
                    const double v = phi_r2_k<decltype(KC)::value>(B.eps, q2);
                    if (jj <= j) Tj[jj] = v;
                }
                return;
            }
            for (int jj = 0; jj <= j; ++jj) {
                double xj[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xj[d] = px(jj, d);
                const int tj = pbt(jj);
                double Dj[DIM], q2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    Dj[d] = xi[d] - xj[d];
                    q2 += Dj[d] * Dj[d];
                }
                double v = phi_r2_k<decltype(KC)::value>(B.eps, q2);
                if (bt == 2 || tj == 2) {
                    double nj[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) nj[d] = pbn(jj, d);
                    double Di, Djn, D2;
                    phi_grad_hess_dirs<DIM>(B, Dj, q2, bn, nj, Di, Djn, D2);
                    const double aj = pba(jj), bj = pbb(jj);
                    if (bt != 2) v = aj * v - bj * Djn;
                    else if (tj != 2) v = ba * v + bb * Di;
                    else v = ba * aj * v - ba * bj * Djn + bb * aj * Di + bb * bj * D2;
                }
                Tj[jj] = v;
            }
        };
        switch (B.kind) {
        case RBFX_PHS1: point_cols(std::integral_constant<int, RBFX_PHS1>{}); break;
        case RBFX_PHS3: point_cols(std::integral_constant<int, RBFX_PHS3>{}); break;
        case RBFX_PHS5: point_cols(std::integral_constant<int, RBFX_PHS5>{}); break;
        case RBFX_PHS7: point_cols(std::integral_constant<int, RBFX_PHS7>{}); break;
        case RBFX_IMQ: point_cols(std::integral_constant<int, RBFX_IMQ>{}); break;
        default: point_cols(std::integral_constant<int, RBFX_GAUSSIAN>{}); break;
        }
        // right-hand sides (kept in registers until the whole assembly is done)
        if (!is_poly) {
            double De[DIM], q2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                De[d] = -xi[d];
                q2 += De[d] * De[d];
            }
            for (int tt = 0; tt < n_terms; ++tt) {
                const TermDev tm = P.terms[tt];
                double v = phi_term<DIM>(B, tm.kind, tm.d1, tm.d2, De, q2);
                if (hermite && bt == 2) v = ba * v + bb * phi_term_normal<DIM>(B, tm.kind, tm.d1, De, q2, bn);
                const double sc = term_scale(tm, hinv);
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (tm.col == c) bcol[t][c] += sc * v;
            }
        } else {
            double z[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) z[d] = 0.0;
            int ea[3];
            mono_exps(j - kk, ea);
            if (hermite && pbt(eval_local) == 2) {
                // the eval node's boundary operator applied to P (assembly.jl:452-467)
                const double eav = pba(eval_local), ebv = pbb(eval_local);
                double en[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) en[d] = pbn(eval_local, d);
                const double v = eav * mono_op<DIM>(ea, RBFX_OP_VALUE, 0, 0, z) + ebv * mono_normal<DIM>(ea, en, z);
#pragma unroll
                for (int c = 0; c < RC; ++c) bcol[t][c] = v;
            } else {
                for (int tt = 0; tt < n_terms; ++tt) {
                    const TermDev tm = P.terms[tt];
                    const double v = term_scale(tm, hinv) * mono_op<DIM>(ea, tm.kind, tm.d1, tm.d2, z);
#pragma unroll
                    for (int c = 0; c < RC; ++c)
                        if (tm.col == c) bcol[t][c] += v;
                }
            }
        }
    }
    __syncthreads();

    // ---- factorisation with rook pivoting and symmetric moves of the pivots
    const double alpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8
    int n_act = n;
    int slot = 0;
    bool failed = false;
    // Column c (uniform) of the active block to this lane's registers (v[t] = A(j, c)
    // for its positions j < n_act: row j's entry if c <= j, else row c's entry j), and
    // the largest |A(j, c)| over active j != excl as a packed (magnitude, position) key.
    double v[TWO];
    auto colkey = [&](int c, int excl) -> uint32_t {
        const int cc = __builtin_amdgcn_readfirstlane(c);
        const int rbc = rb_of(cc);
        uint32_t k = 0u;
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int j = L + 64 * t;
            const bool a_ = j < n_act;
            v[t] = a_ ? s_T[j >= cc ? rbj[t] + cc : rbc + j] : 0.0;
            if (a_ && j != excl) k = max(k, row_key(fabs(v[t]), j));
        }
        return wave_max_u32(k);
    };
    auto colval = [&](int r) -> double {
        const int rr = __builtin_amdgcn_readfirstlane(r);
        return (TWO == 2 && rr >= 64) ? rdlane(v[TWO - 1], rr & 63) : rdlane(v[0], rr & 63);
    };
    // symmetric swap of positions a and b (both active): the rows and the columns, for
    // every row of the system (frozen rows keep their entries at the active positions);
    // each entry of the lower triangle that moves is moved by exactly one lane
    auto sym_swap = [&](int a, int b) {
        const int lo = __builtin_amdgcn_readfirstlane(min(a, b)), hi = __builtin_amdgcn_readfirstlane(max(a, b));
        if (lo == hi) return;
        const int rlo = rb_of(lo), rhi = rb_of(hi);
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int k = L + 64 * t;
            if (k >= n || k == hi) continue;
            int x0, x1;
            if (k < lo) {
                x0 = rlo + k;  // (lo, k) <-> (hi, k)
                x1 = rhi + k;
            } else if (k == lo) {
                x0 = rlo + lo;  // the diagonal entries
                x1 = rhi + hi;
            } else if (k < hi) {
                x0 = rbj[t] + lo;  // (k, lo) <-> (hi, k)
                x1 = rhi + k;
            } else {
                x0 = rbj[t] + lo;  // (k, lo) <-> (k, hi)
                x1 = rbj[t] + hi;
            }
            const double y0 = s_T[x0], y1 = s_T[x1];
            s_T[x0] = y1;
            s_T[x1] = y0;
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const double ylo = (TWO == 2 && lo >= 64) ? rdlane(bcol[TWO - 1][c], lo & 63) : rdlane(bcol[0][c], lo & 63);
            const double yhi = (TWO == 2 && hi >= 64) ? rdlane(bcol[TWO - 1][c], hi & 63) : rdlane(bcol[0][c], hi & 63);
#pragma unroll
            for (int t = 0; t < TWO; ++t) {
                const int k = L + 64 * t;
                if (k == lo) bcol[t][c] = yhi;
                else if (k == hi) bcol[t][c] = ylo;
            }
        }
        const int olo = (TWO == 2 && lo >= 64) ? rdlane(orig[TWO - 1], lo & 63) : rdlane(orig[0], lo & 63);
        const int ohi = (TWO == 2 && hi >= 64) ? rdlane(orig[TWO - 1], hi & 63) : rdlane(orig[0], hi & 63);
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int k = L + 64 * t;
            if (k == lo) orig[t] = ohi;
            else if (k == hi) orig[t] = olo;
        }
    };
    while (n_act > 0) {
        // the active row of the lowest original index starts the search
        uint32_t ko = 0u;
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int j = L + 64 * t;
            if (j < n_act) ko = max(ko, ((uint32_t)(255 - orig[t]) << 8) | (uint32_t)j);
        }
        const int c0 = (int)(wave_max_u32(ko) & 255u);
        const uint32_t ck = colkey(c0, c0);
        int imax = ck ? 127 - (int)(ck & 127u) : c0;
        const double akk = colval(c0);
        double colmax = ck ? fabs(colval(imax)) : 0.0;
        const double absakk = fabs(akk);
        int kstep = 1, p = c0, kp = c0;
        // A(p, p), A(kp, kp) and, for a 2x2 pivot, A(p, kp) as the search ends
        double app = akk, akp = akk, apk = akk;
        bool need_rook = absakk < alpha * colmax;
        if (!(absakk > 0.0 || colmax > 0.0)) {
            failed = true;  // zero column: singular
            break;
        }
        for (int it = 0; it < 2 * ROWS && need_rook; ++it) {
            const uint32_t rk = colkey(imax, imax);
            const int ri = rk ? 127 - (int)(rk & 127u) : imax;
            const double rowmax = rk ? fabs(colval(ri)) : 0.0;
            const double aim = colval(imax);
            const double aii = fabs(aim);
            if (!(aii < alpha * rowmax)) {
                kp = imax;
                akp = aim;
                need_rook = false;
            } else if (p == ri || rowmax <= colmax || it == 2 * ROWS - 1) {
                kp = imax;
                kstep = 2;
                akp = aim;
                apk = colval(p);  // column imax holds A(p, imax)
                need_rook = false;
            } else {
                p = imax;
                app = aim;
                colmax = rowmax;
                imax = ri;
            }
        }
        kstep = __builtin_amdgcn_readfirstlane(kstep);
        const int p0 = __builtin_amdgcn_readfirstlane((kstep == 1) ? kp : p);
        const int p1 = __builtin_amdgcn_readfirstlane(kp);  // equals p0 for a 1x1 pivot
        // A(p0, p0), A(p1, p1), A(p0, p1)
        const double d0 = kstep == 1 ? akp : app, d1 = akp, bq = kstep == 1 ? akp : apk;
        const double det = d0 * d1 - bq * bq;
        if ((kstep == 1 && d0 == 0.0) || (kstep == 2 && det == 0.0)) {
            failed = true;
            break;
        }
        const double idet = 1.0 / det;
        if (L == 0) {
            s_blk[slot] = (int8_t)kstep;
            if (kstep == 2) s_blk[slot + 1] = 0;
        }
        // p0 to position n_act - 1, p1 (2x2) to n_act - 2
        sym_swap(p0, n_act - 1);
        if (kstep == 2) sym_swap(p1 == n_act - 1 ? p0 : p1, n_act - 2);
        n_act -= kstep;
        // the pivot rows: p0 at Q0, p1 at Q1 = n_act (2x2)
        const int Q0 = n_act + kstep - 1;
        const double* U0 = s_T + rb_of(Q0);
        const double* U1 = s_T + rb_of(n_act);
        // the pivot rows' right-hand sides, from their lanes
        double v0[RC], v1[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            v0[c] = (TWO == 2 && Q0 >= 64) ? rdlane(bcol[TWO - 1][c], Q0 & 63) : rdlane(bcol[0][c], Q0 & 63);
            v1[c] = (TWO == 2 && n_act >= 64) ? rdlane(bcol[TWO - 1][c], n_act & 63) : rdlane(bcol[0][c], n_act & 63);
        }
        // The trailing update of the active rows j < n_act over their columns 0..j, eight
        // at a time (the chunk past the diagonal lands in the row's padding), and of their
        // right-hand sides; per element the FMAs run in the same order as in k_build_wide.
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int j = L + 64 * t;
            if (j >= n_act) continue;
            const double u0 = U0[j];
            double l0, l1 = 0.0;
            if (kstep == 1) {
                l0 = u0 / d0;
            } else {
                const double u1 = U1[j];
                l0 = (u0 * d1 - u1 * bq) * idet;
                l1 = (u1 * d0 - u0 * bq) * idet;
            }
            const double ml0 = -l0, ml1 = -l1;
            double* Tj = s_T + rbj[t];
            auto chunk1 = [&](auto W, double* Ar, const double* W0, int j0) __attribute__((always_inline)) {
                constexpr int CW = decltype(W)::value;
                double a[CW], w0[CW];
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    a[q] = Ar[j0 + q];
                    w0[q] = W0[j0 + q];
                }
#pragma unroll
                for (int q = 0; q < CW; ++q) Ar[j0 + q] = fma(ml0, w0[q], a[q]);
            };
            auto chunk2 = [&](auto W, double* Ar, const double* W0, const double* W1, int j0)
                              __attribute__((always_inline)) {
                constexpr int CW = decltype(W)::value;
                double a[CW], w0[CW], w1[CW];
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    a[q] = Ar[j0 + q];
                    w0[q] = W0[j0 + q];
                    w1[q] = W1[j0 + q];
                }
#pragma unroll
                for (int q = 0; q < CW; ++q) Ar[j0 + q] = fma(ml1, w1[q], fma(ml0, w0[q], a[q]));
            };
            using C8 = std::integral_constant<int, 8>;
            if (kstep == 1) {
                for (int j0 = 0; j0 < n_act; j0 += 8)
                    if (j0 <= j) chunk1(C8{}, Tj, U0, j0);
#pragma unroll
                for (int c = 0; c < RC; ++c) bcol[t][c] = fma(ml0, v0[c], bcol[t][c]);
            } else {
                for (int j0 = 0; j0 < n_act; j0 += 8)
                    if (j0 <= j) chunk2(C8{}, Tj, U0, U1, j0);
#pragma unroll
                for (int c = 0; c < RC; ++c) bcol[t][c] = fma(ml1, v1[c], fma(ml0, v0[c], bcol[t][c]));
            }
        }
        slot += kstep;
        __syncthreads();
    }
    __syncthreads();

    // ---- back substitution, last pivot (position 0) first: slot s froze into position
    // n - 1 - s, so the rows frozen before a pivot are the positions above it
    double x[TWO][RC], acc[TWO][RC];
#pragma unroll
    for (int t = 0; t < TWO; ++t)
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            x[t][c] = 0.0;
            acc[t][c] = 0.0;
        }
    if (!failed) {
        const int nslots = slot;
        double y[TWO][RC];
        int bk[TWO];
#pragma unroll
        for (int t = 0; t < TWO; ++t) {
            const int j = L + 64 * t;
#pragma unroll
            for (int c = 0; c < RC; ++c) y[t][c] = j < n ? bcol[t][c] : 0.0;
            // slot sl of the block list sits in lane sl % 64, register sl / 64
            bk[t] = j < nslots ? (int)s_blk[j] : -1;
        }
        auto slot_of = [&](const int(&w)[TWO], int sl) -> int {
            return (TWO == 2 && sl >= 64) ? rdlane(w[TWO - 1], sl & 63) : rdlane(w[0], sl & 63);
        };
        for (int s = nslots - 1; s >= 0; --s) {
            const int blk = slot_of(bk, s);
            const int bprev = s > 0 ? slot_of(bk, s - 1) : -1;
            const bool one = blk == 1, two = (blk == 0) && (bprev == 2);
            if (!one && !two) continue;  // first slot of a 2x2 block: solved with the second
            // the block's pivots: pp (p0, slot s - 1 of a 2x2) and pq (p1, slot s)
            const int pq = n - 1 - s;
            const int pp = two ? pq + 1 : pq;
            const int rbp = rb_of(pp), rbq = rb_of(pq);
            const double dp = s_T[rbp + pp], dq = s_T[rbq + pq], bq = s_T[rbp + pq];
            double arp[TWO], arq[TWO];
#pragma unroll
            for (int t = 0; t < TWO; ++t) {
                const int j = L + 64 * t;
                const bool above = j > pp && j < n;
                arp[t] = above ? s_T[rbj[t] + pp] : 0.0;
                arq[t] = (two && above) ? s_T[rbj[t] + pq] : 0.0;
            }
            // (one division per pivot: a 1x1 pivot needs only 1/dp, a 2x2 only 1/det)
            double idet = 0.0, dinv = 0.0;
            if (two) idet = 1.0 / (dp * dq - bq * bq);
            else dinv = 1.0 / dp;
            const int pl = pp & 63, pt = pp >> 6, ql = pq & 63, qt = pq >> 6;
            // (the 1x1 and 2x2 solves are separate paths: a 1x1 pivot reads one right-hand
            // side and updates with one column)
            if (!two) {
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    if (c >= n_cols) continue;
                    const double rp0 = y[0][c] - acc[0][c];
                    const double rp1 = TWO == 2 ? y[TWO - 1][c] - acc[TWO - 1][c] : 0.0;
                    const double x0 = (pt ? rdlane(rp1, pl) : rdlane(rp0, pl)) * dinv;
#pragma unroll
                    for (int t = 0; t < TWO; ++t) {
                        const int j = L + 64 * t;
                        if (j == pp) x[t][c] = x0;
                        if (j > pp && j < n) acc[t][c] += arp[t] * x0;
                    }
                }
                continue;
            }
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                if (c >= n_cols) continue;
                const double rp0 = y[0][c] - acc[0][c];
                const double rp1 = TWO == 2 ? y[TWO - 1][c] - acc[TWO - 1][c] : 0.0;
                const double yp = pt ? rdlane(rp1, pl) : rdlane(rp0, pl);
                const double yq = qt ? rdlane(rp1, ql) : rdlane(rp0, ql);
                const double x0 = (yp * dq - yq * bq) * idet;
                const double x1 = (yq * dp - yp * bq) * idet;
#pragma unroll
                for (int t = 0; t < TWO; ++t) {
                    const int j = L + 64 * t;
                    if (j == pp) x[t][c] = x0;
                    if (j == pq) x[t][c] = x1;
                    if (j > pp && j < n) acc[t][c] += arp[t] * x0 + arq[t] * x1;
                }
            }
        }
    }

    // ---- outputs: weights of the point rows, multipliers of the monomial rows
#pragma unroll
    for (int t = 0; t < TWO; ++t) {
        const int j = L + 64 * t;
        if (j >= n) continue;
        const int i = orig[t];
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            if (c >= n_cols) continue;
            const double v_ = failed ? __builtin_nan("") : x[t][c];
            if (i >= M) P.out_val[(int64_t)c * P.out_nnz + o0 + (i - M)] = v_;
            else if (P.out_lambda) P.out_lambda[((int64_t)c * P.n_eval + e) * P.n_mono + i] = v_;
        }
        if (i >= M && P.out_col) P.out_col[o0 + (i - M)] = P.adj_idx[s0 + (i - M)];
    }
    if (L == 0 && P.status) P.status[e] = failed ? RBFX_ST_SINGULAR : P.ok_status;
}

static int rc_class(int n_cols) { return n_cols <= 1 ? 1 : (n_cols <= 4 ? 4 : 8); }
// odd: the columns rounded up to a multiple of eight, then the right-hand sides
static int wide_lda(int rows_cap, int n_cols) { return (((rows_cap + 7) & ~7) + rc_class(n_cols)) | 1; }
// dynamic LDS: the system, then per row the scaled point and, for builds with boundary
// data, the normal, two boundary coefficients and a boundary type
static size_t wide_lds(int rows_cap, int n_cols, int dim, bool bnd)
{
    return (size_t)rows_cap * wide_lda(rows_cap, n_cols) * 8 + (size_t)rows_cap * dim * 8 +
           (bnd ? (size_t)rows_cap * (dim + 2) * 8 + (size_t)rows_cap * 4 : 0);
}

// the system plus the kernel's static tables within 64 KB of LDS
constexpr int64_t kWideStaticLds = 2 * 1024;
bool wide_fits(int rows_cap, int n_cols)
{
    return rows_cap >= 1 && rows_cap <= RBFX_MAX_N &&
           (int64_t)wide_lds(rows_cap, n_cols, 3, true) + kWideStaticLds <= 65536;
}

template <int DIM, int RC>
static hipError_t launch_wide_rc(const BuildDev& P, int rows_cap, int64_t n_work, hipStream_t s)
{
    BuildDev Q = P;
    Q.n_blocks = n_work;
    const int lda = wide_lda(rows_cap, P.n_cols);
    const size_t lds = wide_lds(rows_cap, P.n_cols, DIM, P.is_boundary != nullptr);
    if (rows_cap <= 64)
        hipLaunchKernelGGL((k_build_wide<DIM, RC, 1>), dim3((unsigned)pad_to_xcds(n_work)), dim3(64), lds, s, Q,
                           rows_cap, lda);
    else
        hipLaunchKernelGGL((k_build_wide<DIM, RC, 2>), dim3((unsigned)pad_to_xcds(n_work)), dim3(64), lds, s, Q,
                           rows_cap, lda);
    return hipGetLastError();
}

template <int DIM>
static hipError_t launch_wide_dim(const BuildDev& P, int rows_cap, int64_t n_work, hipStream_t s)
{
    switch (rc_class(P.n_cols)) {
    case 1: return launch_wide_rc<DIM, 1>(P, rows_cap, n_work, s);
    case 4: return launch_wide_rc<DIM, 4>(P, rows_cap, n_work, s);
    default: return launch_wide_rc<DIM, 8>(P, rows_cap, n_work, s);
    }
}

hipError_t launch_build_wide(const BuildDev& P, int dim, int rows_cap, int64_t n_work, hipStream_t s)
{
    if (n_work <= 0) return hipSuccess;
    if (!wide_fits(rows_cap, P.n_cols)) return hipErrorInvalidValue;
    switch (dim) {
    case 1: return launch_wide_dim<1>(P, rows_cap, n_work, s);
    case 2: return launch_wide_dim<2>(P, rows_cap, n_work, s);
    case 3: return launch_wide_dim<3>(P, rows_cap, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

// dynamic LDS of k_build_sym: the packed triangle
static int sym_tri(int rows_cap) { return kSymRbHost.rb[rows_cap]; }
static size_t sym_lds(int rows_cap, int n_cols)
{
    (void)n_cols;  // (the right-hand sides live in registers)
    return (size_t)sym_tri(rows_cap) * 8;
}

bool sym_fits(int rows_cap, int n_cols)
{
    return rows_cap >= 1 && rows_cap <= RBFX_MAX_N &&
           (int64_t)sym_lds(rows_cap, n_cols) + kWideStaticLds <= 65536;
}

template <int DIM, int RC>
static hipError_t launch_sym_rc(const BuildDev& P, int rows_cap, int64_t n_work, hipStream_t s)
{
    BuildDev Q = P;
    Q.n_blocks = n_work;
    const int tri = sym_tri(rows_cap);
    // the instantiation without Hermite paths (fewer VGPRs) where the LDS leaves room for
    // more than the 12 wavefronts per CU of the other's registers
    const size_t lds = sym_lds(rows_cap, P.n_cols);
    const bool lean = !P.is_boundary && rows_cap <= 64 && 160 * 1024 / (lds + kWideStaticLds) > 12;
    if (lean)
        hipLaunchKernelGGL((k_build_sym<DIM, RC, 1, false>), dim3((unsigned)pad_to_xcds(n_work)), dim3(64), lds, s, Q,
                           rows_cap, tri);
    else if (rows_cap <= 64)
        hipLaunchKernelGGL((k_build_sym<DIM, RC, 1, true>), dim3((unsigned)pad_to_xcds(n_work)), dim3(64), lds, s, Q,
                           rows_cap, tri);
    else
        hipLaunchKernelGGL((k_build_sym<DIM, RC, 2, true>), dim3((unsigned)pad_to_xcds(n_work)), dim3(64), lds, s, Q,
                           rows_cap, tri);
    return hipGetLastError();
}

template <int DIM>
static hipError_t launch_sym_dim(const BuildDev& P, int rows_cap, int64_t n_work, hipStream_t s)
{
    switch (rc_class(P.n_cols)) {
    case 1: return launch_sym_rc<DIM, 1>(P, rows_cap, n_work, s);
    case 4: return launch_sym_rc<DIM, 4>(P, rows_cap, n_work, s);
    default: return launch_sym_rc<DIM, 8>(P, rows_cap, n_work, s);
    }
}

hipError_t launch_build_sym(const BuildDev& P, int dim, int rows_cap, int64_t n_work, hipStream_t s)
{
    if (n_work <= 0) return hipSuccess;
    if (!sym_fits(rows_cap, P.n_cols)) return hipErrorInvalidValue;
    switch (dim) {
    case 1: return launch_sym_dim<1>(P, rows_cap, n_work, s);
    case 2: return launch_sym_dim<2>(P, rows_cap, n_work, s);
    case 3: return launch_sym_dim<3>(P, rows_cap, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
