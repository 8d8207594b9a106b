// Fused RBF-FD stencil assemble-and-solve for gfx950 (kernel template).
//
// Replaces the per-stencil hot loop of build_weights_kernel
// (src/solve/execution.jl:186-265): _build_collocation_matrix! (assembly.jl:95-115),
// _build_rhs! (assembly.jl:127-185), the Hermite dispatch (assembly.jl:240-467) and
// _solve_system! = bunchkaufman!(Symmetric(A,:U), true) (assembly.jl:220-222).
//
// Mapping: a group of NMAX lanes per stencil (two stencils per wavefront when
// NMAX = 32, one when NMAX = 64). The saddle-point system
//     [ 0   P^T ] [lambda]   [L P (x_e)    ]
//     [ P   Phi ] [  w   ] = [L phi(x_e, .)]
// (polynomial rows first) is assembled in a centred, radius-scaled frame directly
// into registers: lane li of a group owns row li, r[j] holds A[li][j], and the
// operator columns ride along as register columns r[NMAX + c] (augmented system).
//
// The matrix part of a row sits in 1024-bit vector registers. Where an index is
// uniform over a group (the assembly column, a queried column, a pivot), the entry is
// read or written by indexed register moves (s_set_gpr_idx_on) with a scalar index; a
// wave holding two groups does this once per group and selects by lane.
//
// Factorisation: symmetric-indefinite LDL^T with rook pivoting (the pivot tests of
// Ashcraft, Grimes & Lewis 1998, as LAPACK ?sytrf_rook uses). Rows never move between
// lanes. Column queries read each lane's own register (A is symmetric), a group
// argmax over packed 32-bit (magnitude, lane) keys (DPP rotations inside 16-lane rows,
// one cross-row exchange, one unsigned max per level) picks the largest,
// and the pivot row(s) go through LDS as the broadcast source of the rank-1/rank-2
// update. A lane's registers freeze when its row pivots, so they keep its row of U
// (U = D L^T, plus the eliminated right-hand sides) for the back substitution, which
// is right-looking: once x_p is known every lane that pivoted earlier subtracts its
// own U entry for column p.
//
// FP32 mode keeps an FP64 copy of the row and of the right-hand sides: the FP32
// solution is refined with FP64 residuals b - A x and FP32 corrections that replay the
// multipliers (kept per lane, indexed by pivot row); stencils whose correction does
// not settle are queued for the FP64 kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "build.h"
#include "device_math.h"
#include "mono.h"

namespace rbfx {

// ---------------------------------------------------------------- lane helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdlane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Every control used here (quad_perm, row_ror, row_mirror, row_half_mirror) reads a
// lane of the same row of 16, so all source lanes are valid and, with full row and
// bank masks, every lane is written: the instruction needs no old value to merge into
// (which would cost a register initialisation per use).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    static_assert(CTRL <= 0xFF || (CTRL >= 0x121 && CTRL <= 0x12F) || CTRL == 0x140 || CTRL == 0x141,
                  "a control whose every source lane is inside the row");
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(unsigned)u), hi = dpp_i<CTRL>((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

// argmax key of a non-negative magnitude: the value's leading bits (sign dropped, 8 or
// 11 exponent bits, then mantissa bits) with the low 6 bits replaced by 63 - lane,
// so an unsigned max picks the largest value and, among values equal to ~2^-14
// relative, the lowest lane. 0 marks "no candidate". The exact magnitude is read
// back from the chosen lane; only the choice among near-ties is coarsened.
__device__ __forceinline__ uint32_t argmax_key(double v, int li)
{
    const uint32_t hi = (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);
    return (hi & ~63u) | (uint32_t)(63 - li);
}
__device__ __forceinline__ uint32_t argmax_key(float v, int li)
{
    return ((uint32_t)__float_as_int(v) & ~63u) | (uint32_t)(63 - li);
}

// all-reduce of unsigned keys (max) over aligned groups of NMAX lanes
template <int NMAX>
__device__ __forceinline__ uint32_t group_max_u32(uint32_t k)
{
    k = max(k, (uint32_t)dpp_i<0x121>((int)k));
    k = max(k, (uint32_t)dpp_i<0x122>((int)k));
    k = max(k, (uint32_t)dpp_i<0x124>((int)k));
    k = max(k, (uint32_t)dpp_i<0x128>((int)k));
#pragma unroll
    for (int m = 16; m < NMAX; m <<= 1) k = max(k, (uint32_t)__shfl_xor((int)k, m, 64));
    return k;
}

template <int NMAX>
__device__ __forceinline__ double group_max(double v)
{
    v = fmax(v, dpp<0x121>(v));
    v = fmax(v, dpp<0x122>(v));
    v = fmax(v, dpp<0x124>(v));
    v = fmax(v, dpp<0x128>(v));
#pragma unroll
    for (int m = 16; m < NMAX; m <<= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

// value of lane `src` of this lane's group; src must be uniform over each group
template <int NMAX, typename T>
__device__ __forceinline__ T group_bcast(T v, int src)
{
    if constexpr (NMAX == 64) {
        return rdlane(v, __builtin_amdgcn_readfirstlane(src));
    } else {
        const int a = rdlane(src, 0), b = rdlane(src, 32);
        const T va = rdlane(v, a), vb = rdlane(v, 32 + b);
        return lane_id() >= 32 ? vb : va;
    }
}

// bits of a wave ballot that belong to this lane's group
template <int NMAX>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long m)
{
    if constexpr (NMAX == 64) return m;
    const int g = lane_id() / NMAX;
    return (m >> (g * NMAX)) & ((1ull << NMAX) - 1ull);
}

// ------------------------------------------------ register rows
// A lane's row: NMAX matrix entries held in 1024-bit vector registers (16 doubles or
// 32 floats each), then RC right-hand-side columns in plain registers. Static indices
// (unrolled loops) address single registers through get/set. An entry at an index that
// is uniform over a group (the assembly column, a queried column, a pivot) is an indexed
// register move (s_set_gpr_idx_on + v_mov) in each vector and a select of the vector,
// one fixed branch-free sequence per group index (group_row_at, RBFX_ROW_SET,
// RBFX_GROUP_ROW_PUT). Variable-index reads receive the vectors by value (RBFX_ROWV):
// a read through a reference is compiled as a load at a variable offset, which keeps
// the row in scratch memory; variable-index writes are macros in the kernel body for
// the same reason. (Plain registers selected by a switch on the index, a scalar branch
// tree per access, measured slower: C4 FP64 73.0 against 67.5 ms at two waves.)

template <typename T, int NMAX, int RC>
struct VecRow {
    static constexpr int C = 128 / (int)sizeof(T);
    static constexpr int NC = NMAX / C;
    static_assert(NMAX % C == 0 && NC >= 1 && NC <= 4, "the matrix part fills one to four vector registers");
    typedef T elem;
    typedef T vec __attribute__((ext_vector_type(C)));
    // separate members (not an array), so that every access is a whole-vector access
    // of one variable and the row never leaves registers
    vec v0, v1, v2, v3;
    T t[RC > 0 ? RC : 1];

    __device__ __forceinline__ void zero()
    {
        v0 = (vec)(T)0;
        if constexpr (NC > 1) v1 = (vec)(T)0;
        if constexpr (NC > 2) v2 = (vec)(T)0;
        if constexpr (NC > 3) v3 = (vec)(T)0;
#pragma unroll
        for (int c = 0; c < (RC > 0 ? RC : 1); ++c) t[c] = (T)0;
    }
    // entry q of vector c
    static __device__ __forceinline__ T vget(vec v0, vec v1, vec v2, vec v3, unsigned c, unsigned q)
    {
        T x = v0[q];
        if constexpr (NC > 1) {
            const T y = v1[q];
            x = c == 1u ? y : x;
        }
        if constexpr (NC > 2) {
            const T y = v2[q];
            x = c == 2u ? y : x;
        }
        if constexpr (NC > 3) {
            const T y = v3[q];
            x = c == 3u ? y : x;
        }
        return x;
    }
    // compile-time index j in [0, NMAX + RC) (after unrolling). The right-hand-side
    // registers are only ever addressed with constant subscripts, so that no access
    // with a variable offset into the row exists before the loops are unrolled.
    __device__ __forceinline__ T get(int j) const
    {
        if (j < NMAX) return vget(v0, v1, v2, v3, (unsigned)j / C, (unsigned)j % C);
        T x = (T)0;
#pragma unroll
        for (int c = 0; c < RC; ++c) x = (j - NMAX == c) ? t[c] : x;
        return x;
    }
    __device__ __forceinline__ void set(int j, T x)
    {
        if (j < NMAX) {
            const unsigned c = (unsigned)j / C, q = (unsigned)j % C;
            v0[q] = c == 0u ? x : v0[q];
            if constexpr (NC > 1) v1[q] = c == 1u ? x : v1[q];
            if constexpr (NC > 2) v2[q] = c == 2u ? x : v2[q];
            if constexpr (NC > 3) v3[q] = c == 3u ? x : v3[q];
        } else {
#pragma unroll
            for (int c = 0; c < RC; ++c) t[c] = (j - NMAX == c) ? x : t[c];
        }
    }
    __device__ __forceinline__ vec a0() const { return v0; }
    __device__ __forceinline__ vec a1() const { return v1; }
    __device__ __forceinline__ vec a2() const { return v2; }
    __device__ __forceinline__ vec a3() const { return v3; }

    template <int GN>
    static __device__ __forceinline__ T group_at(vec v0, vec v1, vec v2, vec v3, int idx)
    {
        if constexpr (GN == 64) {
            const unsigned j = (unsigned)__builtin_amdgcn_readfirstlane(idx);
            return vget(v0, v1, v2, v3, j / C, j & (C - 1));
        } else {
            const unsigned a = (unsigned)rdlane(idx, 0), b = (unsigned)rdlane(idx, 32);
            const T va = vget(v0, v1, v2, v3, a / C, a & (C - 1));
            const T vb = vget(v0, v1, v2, v3, b / C, b & (C - 1));
            return lane_id() >= 32 ? vb : va;
        }
    }
};

// arguments of a variable-index read of a row
#define RBFX_ROWV(row) (row).a0(), (row).a1(), (row).a2(), (row).a3()

// row entry idx, idx uniform over each group of NMAX lanes
template <int NMAX, typename Row, typename A0, typename A1, typename A2, typename A3>
__device__ __forceinline__ typename Row::elem group_row_at(A0 a0, A1 a1, A2 a2, A3 a3, int idx)
{
    return Row::template group_at<NMAX>(a0, a1, a2, a3, idx);
}

// row entry j (wave-uniform, in [0, NMAX)) = x on the lanes where pred holds. A macro,
// not a function: a variable-index write compiled inside a callee becomes a store
// through a pointer at a variable offset, which after inlining pins the row in scratch
// memory; written in the kernel body it stays an indexed register move.
#define RBFX_ROW_PUT(row, j, x, pred)                                                          \
    do {                                                                                       \
        using Row_ = std::remove_reference_t<decltype(row)>;                                   \
        const unsigned j_ = (unsigned)(j);                                                     \
        const auto x_ = (x);                                                                   \
        const bool p_ = (pred);                                                                \
        const unsigned q_ = j_ & (unsigned)(Row_::C - 1), c_ = j_ / (unsigned)Row_::C;         \
        (row).v0[q_] = (p_ && c_ == 0u) ? x_ : (row).v0[q_];                                   \
        if constexpr (Row_::NC > 1) (row).v1[q_] = (p_ && c_ == 1u) ? x_ : (row).v1[q_];       \
        if constexpr (Row_::NC > 2) (row).v2[q_] = (p_ && c_ == 2u) ? x_ : (row).v2[q_];       \
        if constexpr (Row_::NC > 3) (row).v3[q_] = (p_ && c_ == 3u) ? x_ : (row).v3[q_];       \
    } while (0)

// row entry j (wave-uniform, in [0, NMAX)) = x on every lane
#define RBFX_ROW_SET(row, j, x) RBFX_ROW_PUT(row, j, x, true)

// row[idx] = x on the lanes where pred holds; idx uniform over each group of NMAX lanes
#define RBFX_GROUP_ROW_PUT(NMAX_, row, idx, x, pred)                                           \
    do {                                                                                       \
        if constexpr ((NMAX_) == 64) {                                                         \
            RBFX_ROW_PUT(row, __builtin_amdgcn_readfirstlane(idx), x, pred);                   \
        } else {                                                                               \
            const int ga_ = rdlane((int)(idx), 0), gb_ = rdlane((int)(idx), 32);               \
            RBFX_ROW_PUT(row, ga_, x, (pred) && lane_id() < 32);                               \
            RBFX_ROW_PUT(row, gb_, x, (pred) && lane_id() >= 32);                              \
        }                                                                                      \
    } while (0)

// term scale h^-order in the radius-scaled frame
__device__ __forceinline__ double term_scale(const TermDev& t, double hinv)
{
    return t.coeff * (t.kind == RBFX_OP_VALUE ? 1.0 : (t.kind == RBFX_OP_D1 ? hinv : hinv * hinv));
}

// per-stencil tables in LDS
template <typename T, int NMAX, int DIM, int UW>
struct StencilLds {
    // monomial values at the stencil points, pm[pi * M + a] (two stencils per wavefront:
    // k * M <= 16 * 16 when k + M <= 32)
    static constexpr int kPm = NMAX == 32 ? 256 : 1;
    double pm[kPm];
    double xi[NMAX][DIM];  // scaled offsets of the stencil points
    double ba[NMAX], bb[NMAX], bn[NMAX][DIM];
    T prow[2][2][UW];  // candidate / pivot rows: two slots, double-buffered by step parity
    int bt[NMAX];      // 0 interior, 1 Dirichlet, 2 Neumann/Robin (Hermite stencils)
    int piv[NMAX], blk[NMAX];
};

// ------------------------------------------------------- back substitution
// Solves U x = y (RC right-hand sides at once) for this lane's group. Row p of U is
// the frozen register row of the lane that pivoted at slot s (s_piv[s] = p), with the
// block structure in s_blk (1: 1x1, 2: first slot of a 2x2 block, 0: its second slot).
template <int NMAX, int RC, typename Row, typename A0, typename A1, typename A2, typename A3,
          typename T = typename Row::elem>
__device__ __forceinline__ void back_substitute(A0 r0, A1 r1, A2 r2, A3 r3, const int* s_piv, const int* s_blk, int nslots,
                                                int nslots_wave, int myslot, T diag, int n_cols, const T (&y)[RC],
                                                T (&x)[RC])
{
    const int li = lane_id() % NMAX;
    const T dinv = (T)1 / diag;
    T acc[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) {
        acc[c] = (T)0;
        x[c] = (T)0;
    }
    for (int s = nslots_wave - 1; s >= 0; --s) {
        const int blk = s < nslots ? s_blk[s] : -1;
        const int bprev = (s < nslots && s > 0) ? s_blk[s - 1] : -1;
        const bool one = blk == 1, two = (blk == 0) && (bprev == 2);
        const int p = one ? s_piv[s] : (two ? s_piv[s - 1] : 0);
        const int q = two ? s_piv[s] : 0;
        // the second row of a 2x2 block only when one of the wave's groups has one here
        const bool two_any = __ballot(two) != 0ull;
        const T up = group_row_at<NMAX, Row>(r0, r1, r2, r3, p);
        T uq = (T)0, dp = (T)1, dq = (T)1, bq = (T)0, idet = (T)1;
        if (two_any) {
            dp = group_bcast<NMAX>(diag, p);
            uq = group_row_at<NMAX, Row>(r0, r1, r2, r3, q);
            dq = group_bcast<NMAX>(diag, q);
            bq = group_bcast<NMAX>(uq, p);  // U[p][q]
            idet = (T)1 / (dp * dq - bq * bq);
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            if (c < n_cols) {
                const T res = y[c] - acc[c];
                const T xp = group_bcast<NMAX>(res * dinv, p);
                const T y0 = two_any ? group_bcast<NMAX>(res, p) : (T)0;
                const T y1 = two_any ? group_bcast<NMAX>(res, q) : (T)0;
                const T x0 = two ? (y0 * dq - y1 * bq) * idet : xp;
                const T x1 = two ? (y1 * dp - y0 * bq) * idet : (T)0;
                x[c] = ((one || two) && li == p) ? x0 : ((two && li == q) ? x1 : x[c]);
                const bool upd = one ? (myslot < s) : (two && myslot < s - 1);
                const T dacc = two ? up * x0 + uq * x1 : up * x0;
                acc[c] += upd ? dacc : (T)0;
            }
        }
    }
}

// forward replay of the multipliers (indexed by pivot row) on a new right-hand side
template <int NMAX, typename Row, typename A0, typename A1, typename A2, typename A3,
          typename T = typename Row::elem>
__device__ __forceinline__ T forward_replay(A0 m0, A1 m1, A2 m2, A3 m3, const int* s_piv, const int* s_blk, int nslots,
                                            int nslots_wave, int myslot, T y)
{
    for (int s = 0; s < nslots_wave; ++s) {
        const int blk = s < nslots ? s_blk[s] : -1;
        const int p = blk >= 1 ? s_piv[s] : 0;
        const int q = blk == 2 ? s_piv[s + 1] : 0;
        const bool two_any = __ballot(blk == 2) != 0ull;
        const T y0 = group_bcast<NMAX>(y, p), y1 = two_any ? group_bcast<NMAX>(y, q) : (T)0;
        const T l0 = group_row_at<NMAX, Row>(m0, m1, m2, m3, p);
        const T l1 = two_any ? group_row_at<NMAX, Row>(m0, m1, m2, m3, q) : (T)0;
        if (blk == 1 && myslot > s) y -= l0 * y0;
        if (blk == 2 && myslot > s + 1) y -= l0 * y0 + l1 * y1;
    }
    return y;
}

// r += a * u over a broadcast LDS row, in chunks of 8 columns: the scheduling
// barriers keep at most one chunk of the row in flight, so the register row is
// never spilled to make room for prefetched operands
template <int NMAX, int RC, typename Row, typename T>
__device__ __forceinline__ void axpy_row(Row& r, T a, const T* u, int n_cols)
{
#pragma unroll
    for (int j0 = 0; j0 < NMAX; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8; ++j) r.set(j, fma(a, u[j], r.get(j)));
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
        if (c < n_cols) r.set(NMAX + c, fma(a, u[NMAX + c], r.get(NMAX + c)));
}

// Waves per SIMD the register allocation targets. The FP64 two-stencil kernels (up to
// 4 right-hand sides) fit 3 waves (168 VGPRs) without spilling in 1-D and 2-D; in 3-D
// they would spill at 3 waves and run at 2. No build kernel may spill to scratch memory:
// with a spilling three-wave kernel (52-108 bytes per lane), repeated identical builds
// of 1-D stencils (11-row systems, short-lived wavefronts) differed from run to run in
// single stencils, up to wrong and singular results, on the wavefronts that were the
// third on their compute unit; the spill-free kernels are bit-identical in every
// repetition. The one-stencil kernels (64-row register rows) need more than 256
// registers: at one wave per SIMD the allocator keeps the overflow in accumulation
// registers (AGPRs), so they too use no scratch memory.
template <typename T, int NMAX, int DIM, int RC>
struct BuildOccupancy {
    static constexpr int waves = NMAX == 64 ? 1 : ((sizeof(T) == 8 && NMAX <= 32 && RC <= 4 && DIM <= 2) ? 3 : 2);
};

// ------------------------------------------------------------------ kernel
template <typename T, int NMAX, int DIM, int RC>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64),
                          amdgpu_waves_per_eu(BuildOccupancy<T, NMAX, DIM, RC>::waves))) void
k_build(const BuildDev P)
{
    constexpr bool REFINE = sizeof(T) == 4;
    constexpr int G = 64 / NMAX;   // stencils per wavefront
    constexpr int UW = NMAX + RC;  // register row: matrix entries + right-hand sides
    constexpr int NL = REFINE ? NMAX : 1;
    // two groups per wavefront select registers by a per-group (not wave-uniform) index:
    // candidate rows then pass through LDS instead of register switches
    constexpr bool SLOT_ROWS = G > 1;

    __shared__ StencilLds<T, NMAX, DIM, UW> s_st[G];
    __shared__ double s_x[G][NL];
    __shared__ int s_mono[kMaxMono][3];
    __shared__ TermDev s_terms[RBFX_MAX_TERMS];

    const int L = lane_id();
    const int g = L / NMAX, li = L % NMAX;
    StencilLds<T, NMAX, DIM, UW>& S = s_st[g];
    const int n_cols = P.n_cols;
    const int n_terms = P.n_terms;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    {
        const int* src = &P.mono_e[0][0];
        for (int i = L; i < kMaxMono * 3; i += 64) (&s_mono[0][0])[i] = src[i];
        if (L < RBFX_MAX_TERMS) s_terms[L] = P.terms[L];
    }

    const int64_t lb = xcd_contiguous_block(P.n_blocks);
    if (lb < 0) return;  // padding block: the whole workgroup leaves
    // the row slots start at zero: a group that has finished (or never started) applies
    // its slot rows with zero multipliers, which must leave its registers unchanged
    for (int j = li; j < 2 * 2 * UW; j += NMAX) (&S.prow[0][0][0])[j] = (T)0;
    const int64_t wi = lb * G + g;
    bool alive = wi < n_work;
    int64_t e = alive ? (P.list ? (int64_t)P.list[wi] : (P.order ? (int64_t)P.order[wi] : wi)) : 0;
    if (e < 0 || e >= P.n_eval) {  // an order entry out of range: nothing to solve
        alive = false;
        e = 0;
    }
    const int64_t s0 = alive ? (P.adj_ptr ? P.adj_ptr[e] : e * (int64_t)P.k) : 0;
    const int kk = alive ? (P.adj_ptr ? (int)(P.adj_ptr[e + 1] - s0) : P.k) : 0;
    const int64_t o0 = alive ? (P.out_ptr ? P.out_ptr[e] : e * (int64_t)P.k) : 0;
    const int m_e = alive ? stencil_n_mono(P, e, DIM) : P.n_mono;
    const int M = m_e < 0 ? 0 : m_e;  // monomials of this stencil (its degree)
    if (alive && m_e < 0) {
        if (li == 0 && P.status) P.status[e] = RBFX_ST_INVALID;
        alive = false;
    }
    if (alive && (kk < 1 || M + kk > NMAX)) {
        if (li == 0 && P.status) P.status[e] = RBFX_ST_TOO_LARGE;
        alive = false;
    }

    // ---- this lane's stencil point (point rows live at lanes M..M+kk-1 of the group)
    const int pi = li - M;
    bool is_pt = alive && pi >= 0 && pi < kk;
    int gidx = 0;
    double xe[DIM], D[DIM];
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xe[d] = alive ? P.eval[e * DIM + d] : 0.0;
        D[d] = 0.0;
    }
    if (is_pt) {
        if (P.packed) {  // offsets packed by position: no wait on the list / adjacency reads
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                D[d] = P.packed[(wi * (int64_t)P.k + pi) * DIM + d];
                r2 += D[d] * D[d];
            }
        } else {
            gidx = P.adj_idx[s0 + pi];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                D[d] = P.data[(int64_t)gidx * DIM + d] - xe[d];
                r2 += D[d] * D[d];
            }
        }
    }

    // ---- stencil classification (types.jl:208-223)
    bool hermite = false;
    int bt = 0;
    double ba = 0.0, bb = 0.0, bn[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) bn[d] = 0.0;
    if (P.is_boundary) {
        const bool ib = is_pt && P.is_boundary[gidx];
        hermite = group_bits<NMAX>(__ballot(ib)) != 0ull;
        // a Dirichlet eval node is an identity row whatever its neighbours are, the rule
        // count_nonzeros sizes its row by (execution.jl:39-69), so the row never holds
        // more than the one entry it was given
        if (alive && e < P.n_data && P.is_boundary[e]) {
            const int bi = P.bnd_index[e];
            if (P.bc_alpha[bi] == 1.0 && P.bc_beta[bi] == 0.0) {
                // identity row (fill_dirichlet_entry!, execution.jl:297-303)
                if (li < n_cols) P.out_val[(int64_t)li * P.out_nnz + o0] = 1.0;
                if (li == 0) {
                    if (P.out_col) P.out_col[o0] = (int32_t)e;
                    if (P.status) P.status[e] = RBFX_ST_DIRICHLET;
                }
                alive = false;
                is_pt = false;
            }
        }
        if (hermite && ib && alive) {
            const int bi = P.bnd_index[gidx];
            ba = P.bc_alpha[bi];
            bb = P.bc_beta[bi];
#pragma unroll
            for (int d = 0; d < DIM; ++d) bn[d] = P.normals[(int64_t)bi * DIM + d];
            bt = (ba == 1.0 && bb == 0.0) ? 1 : 2;
        }
    }

    // ---- radius-scaled local frame
    double h = sqrt(group_max<NMAX>(is_pt ? r2 : 0.0));
    if (!(h > 0.0)) h = 1.0;
    const double hinv = 1.0 / h;
    double xi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xi[d] = D[d] / h;
    bb = bb / h;  // beta acts through a first derivative
    if (is_pt) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) S.xi[pi][d] = xi[d];
        S.bt[pi] = bt;
        if (hermite) {
            S.ba[pi] = ba;
            S.bb[pi] = bb;
#pragma unroll
            for (int d = 0; d < DIM; ++d) S.bn[pi][d] = bn[d];
        }
    }
    int eval_local = 0;
    {
        bool eq = is_pt && hermite;
#pragma unroll
        for (int d = 0; d < DIM; ++d) eq = eq && (D[d] == 0.0);
        const unsigned long long m = group_bits<NMAX>(__ballot(eq));
        if (alive && hermite) {
            if (!m) {
                if (li == 0 && P.status) P.status[e] = RBFX_ST_EVAL_NOT_FOUND;
                alive = false;
                is_pt = false;
            } else {
                eval_local = __builtin_ctzll(m) - M;
            }
        }
    }
    // monomials of the own point, once: the assembly reads them for the polynomial
    // columns of this row and the polynomial rows of this column
    constexpr bool PM_TABLE = NMAX == 32;
    if constexpr (PM_TABLE) {
        if (is_pt)
            for (int a = 0; a < M; ++a) S.pm[pi * M + a] = mono_op<DIM>(s_mono[a], RBFX_OP_VALUE, 0, 0, xi);
    }
    const int n = alive ? M + kk : 0;
    int n_wave = n;
#pragma unroll
    for (int m = NMAX; m < 64; m <<= 1) n_wave = max(n_wave, __shfl_xor(n_wave, m, 64));
    n_wave = __builtin_amdgcn_readfirstlane(n_wave);
    __syncthreads();

    BasisDev B;
    B.kind = P.basis_kind;
    B.eps = (B.kind == RBFX_IMQ || B.kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;

    // ---- assembly (FP64): column j of every row in step j (j uniform over the wave)
    using RowT = VecRow<T, NMAX, RC>;
    using MulRow = VecRow<T, NMAX, 0>;
    RowT r;
    VecRow<double, NMAX, 0> a64;  // FP64 copy of the matrix row (FP32 mode only)
    double diag64 = 0.0;
    r.zero();
    if constexpr (REFINE) a64.zero();
    const bool is_poly = alive && li < M;
    for (int j = 0; j < n_wave; ++j) {
        // the shape parameter re-enters the loop as an opaque value, so the constants
        // derived from it for every basis kind are not hoisted out of the loop (they
        // would stay live across it and spill)
        BasisDev Bj = B;
        asm volatile("" : "+v"(Bj.eps));
        double v = 0.0;
        if (j < M) {
            // point row: monomial j at the own point (Neumann/Robin rows: alpha P + beta dP/dn)
            if (is_pt) {
                if constexpr (PM_TABLE) v = S.pm[pi * M + j];
                else v = mono_op<DIM>(s_mono[j], RBFX_OP_VALUE, 0, 0, xi);
                if (hermite && bt == 2) v = ba * v + bb * mono_normal<DIM>(s_mono[j], bn, xi);
            }
        } else {
            const int jj = j - M;
            if (jj < kk) {
                double xj[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xj[d] = S.xi[jj][d];
                const int tj = S.bt[jj];
                if (is_poly) {
                    // polynomial row li: monomial li at point jj
                    if constexpr (PM_TABLE) v = S.pm[jj * M + li];
                    else v = mono_op<DIM>(s_mono[li], RBFX_OP_VALUE, 0, 0, xj);
                    if (hermite && tj == 2) {
                        double nj[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) nj[d] = S.bn[jj][d];
                        v = S.ba[jj] * v + S.bb[jj] * mono_normal<DIM>(s_mono[li], nj, xj);
                    }
                } else if (is_pt) {
                    double Dj[DIM], q2 = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        Dj[d] = xi[d] - xj[d];
                        q2 += Dj[d] * Dj[d];
                    }
                    v = phi_r2(Bj, q2);
                    if (hermite && (bt == 2 || tj == 2)) {
                        double nj[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) nj[d] = S.bn[jj][d];
                        double Di, Djn, D2;
                        phi_grad_hess_dirs<DIM>(Bj, Dj, q2, bn, nj, Di, Djn, D2);
                        const double aj = S.ba[jj], bj = S.bb[jj];
                        if (bt != 2) v = aj * v - bj * Djn;
                        else if (tj != 2) v = ba * v + bb * Di;
                        else v = ba * aj * v - ba * bj * Djn + bb * aj * Di + bb * bj * D2;
                    }
                    if (jj == pi) diag64 = v;
                }
            }
        }
        RBFX_ROW_SET(r, j, (T)v);
        if constexpr (REFINE) RBFX_ROW_SET(a64, j, v);
    }

    // ---- right-hand sides (FP64), augmented as register columns NMAX..NMAX+RC-1
    double b64[REFINE ? RC : 1];
    {
        double bcol[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) bcol[c] = 0.0;
        const bool f7 = alive && hermite && S.bt[eval_local] == 2;
        if (is_pt) {
            double De[DIM], q2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                De[d] = -xi[d];
                q2 += De[d] * De[d];
            }
            // one term loop per basis kind: with the kind a constant, only its own
            // derivative formulas are compiled into the loop (with the kind a runtime
            // value, the compiler hoists the loop-invariant parts of every kind's
            // formulas out of the loop and spills them)
#define RBFX_RHS_TERMS(BK)                                                                              \
    {                                                                                                   \
        BasisDev Bk;                                                                                    \
        Bk.kind = (BK);                                                                                 \
        Bk.eps = B.eps;                                                                                 \
        for (int t = 0; t < n_terms; ++t) {                                                             \
            const TermDev tm = s_terms[t];                                                              \
            double v = phi_term<DIM>(Bk, tm.kind, tm.d1, tm.d2, De, q2);                                \
            if (hermite && bt == 2) v = ba * v + bb * phi_term_normal<DIM>(Bk, tm.kind, tm.d1, De, q2, bn); \
            const double sc = term_scale(tm, hinv);                                                     \
            for (int c = 0; c < RC; ++c)                                                                \
                if (tm.col == c) bcol[c] += sc * v;                                                     \
        }                                                                                               \
    }
            switch (B.kind) {
            case RBFX_PHS1: RBFX_RHS_TERMS(RBFX_PHS1) break;
            case RBFX_PHS3: RBFX_RHS_TERMS(RBFX_PHS3) break;
            case RBFX_PHS5: RBFX_RHS_TERMS(RBFX_PHS5) break;
            case RBFX_PHS7: RBFX_RHS_TERMS(RBFX_PHS7) break;
            case RBFX_IMQ: RBFX_RHS_TERMS(RBFX_IMQ) break;
            default: RBFX_RHS_TERMS(RBFX_GAUSSIAN) break;
            }
#undef RBFX_RHS_TERMS
        } else if (is_poly) {
            double z[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) z[d] = 0.0;
            const int(&ea)[3] = s_mono[li];
            if (f7) {
                // the eval node's boundary operator applied to P (assembly.jl:452-467)
                const double eav = S.ba[eval_local], ebv = S.bb[eval_local];
                double en[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) en[d] = S.bn[eval_local][d];
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
            r.set(NMAX + c, (T)bcol[c]);
            if constexpr (REFINE) b64[c] = bcol[c];
        }
    }

    // ---- factorisation with rook pivoting (runtime step loop; every lane takes part)
    const T alpha = (T)0.6403882032022076;  // (1 + sqrt(17)) / 8
    T diag = (T)diag64;
    MulRow Lm;  // multipliers by pivot row (FP32 mode only)
    if constexpr (REFINE) Lm.zero();
    bool act = alive && li < n;
    int myslot = 1 << 20;
    int slot = 0;
    bool failed = false;
    for (int step = 0;; ++step) {
        const unsigned long long gact = group_bits<NMAX>(__ballot(act));
        if (__ballot(gact != 0ull) == 0ull) break;  // every group of the wave is done
        const bool busy = gact != 0ull;             // this lane's group still factorises
        const int c0 = busy ? (int)__builtin_ctzll(gact) : 0;
        const int buf = step & 1;
        int kstep = 1, p = c0, kp = c0;
        int sp = 0, sk = 0;  // LDS slots of rows p and kp
        bool need_rook;
        if constexpr (SLOT_ROWS) {
            // Candidate rows pass through LDS (two slots per group): row c0 first, which is
            // column c0 of the active block by symmetry, then the rook candidates. The pivot
            // row(s) are then already in their slots for the update.
            if (busy && li == c0) {
#pragma unroll
                for (int j = 0; j < UW; ++j) S.prow[buf][0][j] = r.get(j);
            }
            __syncthreads();
            const T cv = (T)fabs(S.prow[buf][0][li]);
            const uint32_t ck = group_max_u32<NMAX>((act && li != c0) ? argmax_key(cv, li) : 0u);
            int imax = ck ? 63 - (int)(ck & 63u) : c0;
            T colmax = ck ? (T)fabs(S.prow[buf][0][imax]) : (T)0;
            const T absakk = (T)fabs(S.prow[buf][0][c0]);
            need_rook = busy && absakk < alpha * colmax;
            if (busy && !(absakk > (T)0 || colmax > (T)0)) {
                failed = true;  // zero column: singular
                need_rook = false;
            }
            // rook iterations: every group takes part in the collective steps
            for (int it = 0; it < 2 * NMAX; ++it) {
                if (__ballot(need_rook) == 0ull) break;
                const int si = 1 - sp;  // row imax into the slot not holding row p
                if (need_rook && li == imax) {
#pragma unroll
                    for (int j = 0; j < UW; ++j) S.prow[buf][si][j] = r.get(j);
                }
                __syncthreads();
                const T rv = need_rook ? (T)fabs(S.prow[buf][si][li]) : (T)0;
                const uint32_t rk = group_max_u32<NMAX>((need_rook && act && li != imax) ? argmax_key(rv, li) : 0u);
                const int ri = rk ? 63 - (int)(rk & 63u) : imax;
                if (need_rook) {
                    const T rowmax = rk ? (T)fabs(S.prow[buf][si][ri]) : (T)0;
                    const T aii = (T)fabs(S.prow[buf][si][imax]);
                    if (!(aii < alpha * rowmax)) {
                        kp = imax;
                        sk = si;
                        need_rook = false;
                    } else if (p == ri || rowmax <= colmax || it == 2 * NMAX - 1) {
                        kp = imax;
                        sk = si;
                        kstep = 2;
                        need_rook = false;
                    } else {
                        p = imax;
                        sp = si;
                        colmax = rowmax;
                        imax = ri;
                    }
                }
            }
        } else {
            // one group per wavefront: the selection index is wave-uniform, so a column
            // entry is a scalar-indexed register read and a value a single lane read
            const T cv = (T)fabs(group_row_at<NMAX, RowT>(RBFX_ROWV(r), c0));  // column c0 = row c0 by symmetry
            const uint32_t ck = group_max_u32<NMAX>((act && li != c0) ? argmax_key(cv, li) : 0u);
            int imax = ck ? 63 - (int)(ck & 63u) : c0;
            T colmax = ck ? group_bcast<NMAX>(cv, imax) : (T)0;
            const T absakk = (T)fabs(group_bcast<NMAX>(diag, c0));
            need_rook = busy && absakk < alpha * colmax;
            if (busy && !(absakk > (T)0 || colmax > (T)0)) {
                failed = true;  // zero column: singular
                need_rook = false;
            }
            for (int it = 0; it < 2 * NMAX; ++it) {
                if (__ballot(need_rook) == 0ull) break;
                const T rv = (T)fabs(group_row_at<NMAX, RowT>(RBFX_ROWV(r), imax));
                const uint32_t rk = group_max_u32<NMAX>((need_rook && act && li != imax) ? argmax_key(rv, li) : 0u);
                const int ri = rk ? 63 - (int)(rk & 63u) : imax;
                const T rowmax = rk ? group_bcast<NMAX>(rv, ri) : (T)0;
                const T aii = (T)fabs(group_bcast<NMAX>(diag, imax));
                if (need_rook) {
                    if (!(aii < alpha * rowmax)) {
                        kp = imax;
                        need_rook = false;
                    } else if (p == ri || rowmax <= colmax || it == 2 * NMAX - 1) {
                        kp = imax;
                        kstep = 2;
                        need_rook = false;
                    } else {
                        p = imax;
                        colmax = rowmax;
                        imax = ri;
                    }
                }
            }
            // the pivot row(s) to LDS, the broadcast source of the update: row p0 in slot 0,
            // row p1 in slot 1 (which repeats row p0 for a 1x1 pivot)
            const bool two = __ballot(busy && !failed && kstep == 2) != 0ull;
            const int p0r = (kstep == 1) ? kp : p;
            if (busy && !failed && li == p0r) {
#pragma unroll
                for (int j = 0; j < UW; ++j) S.prow[buf][0][j] = r.get(j);
            }
            if (two && busy && !failed && li == kp) {
#pragma unroll
                for (int j = 0; j < UW; ++j) S.prow[buf][1][j] = r.get(j);
            }
            __syncthreads();
            sp = 0;
            sk = two ? 1 : 0;
        }
        // a singular group stops here (failed is uniform over the group)
        const bool go = busy && !failed;
        const int p0 = (kstep == 1) ? kp : p;
        const int p1 = kp;                      // equals p0 for a 1x1 pivot
        const T* U0 = S.prow[buf][(kstep == 1 && SLOT_ROWS) ? sk : sp];
        const T* U1 = S.prow[buf][sk];          // equals U0 for a 1x1 pivot
        // a wave with a 2x2 pivot in either group runs the two-row update for both;
        // a 1x1 group then applies its pivot row twice, the second time with a zero multiplier
        const bool two_any = __ballot(go && kstep == 2) != 0ull;
        if (go && li == 0) {
            S.piv[slot] = p0;
            S.blk[slot] = kstep;
            if (kstep == 2) {
                S.piv[slot + 1] = p1;
                S.blk[slot + 1] = 0;
            }
        }
        const T d0 = U0[p0];
        const T d1 = U1[p1];
        const T u0 = U0[li];
        T l0, l1 = (T)0;
        if (kstep == 1) {
            if (go && d0 == (T)0) failed = true;
            l0 = u0 / d0;
        } else {
            const T bq = U0[p1];
            const T det = d0 * d1 - bq * bq;
            if (go && det == (T)0) failed = true;
            const T u1 = U1[li];
            const T idet = (T)1 / det;
            l0 = (u0 * d1 - u1 * bq) * idet;
            l1 = (u1 * d0 - u0 * bq) * idet;
        }
        // rows still to be eliminated; every other lane applies zero multipliers, which
        // leaves its (finite) registers unchanged
        const bool mine = go && !failed && act && li != p0 && li != p1;
        if (!mine) {
            l0 = (T)0;
            l1 = (T)0;
        }
        // in-place fused multiply-adds keep the register row the only live copy
        const T ml0 = -l0, ml1 = -l1;
        diag = fma(ml0, u0, diag);
        axpy_row<NMAX, RC>(r, ml0, U0, n_cols);
        if (two_any) {
            diag = fma(ml1, U1[li], diag);
            axpy_row<NMAX, RC>(r, ml1, U1, n_cols);
        }
        if constexpr (REFINE) {
            RBFX_GROUP_ROW_PUT(NMAX, Lm, p0, l0, mine);
            if (two_any) RBFX_GROUP_ROW_PUT(NMAX, Lm, p1, l1, mine && kstep == 2);
        }
        if (go && !failed) {
            if (li == p0) {
                act = false;
                myslot = slot;
            }
            if (kstep == 2 && li == p1) {
                act = false;
                myslot = slot + 1;
            }
            slot += kstep;
        }
        if (failed) act = false;
    }
    __syncthreads();

    // a singular group has left all its rows inactive and reports NaN weights
    const bool gfail = group_bits<NMAX>(__ballot(failed)) != 0ull;
    int st = gfail ? RBFX_ST_SINGULAR : P.ok_status;
    const int nslots = gfail ? 0 : slot;
    int nslots_wave = nslots;
#pragma unroll
    for (int m = NMAX; m < 64; m <<= 1) nslots_wave = max(nslots_wave, __shfl_xor(nslots_wave, m, 64));
    nslots_wave = __builtin_amdgcn_readfirstlane(nslots_wave);

    T y[RC], x[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) y[c] = r.get(NMAX + c);
    back_substitute<NMAX, RC, RowT>(RBFX_ROWV(r), S.piv, S.blk, nslots, nslots_wave, myslot, diag, n_cols, y, x);
    if constexpr (!REFINE) {
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            if (c < n_cols && is_pt)
                P.out_val[(int64_t)c * P.out_nnz + o0 + pi] = gfail ? __builtin_nan("") : (double)x[c];
            if (c < n_cols && is_poly && P.out_lambda)
                P.out_lambda[((int64_t)c * P.n_eval + e) * P.n_mono + li] = gfail ? __builtin_nan("") : (double)x[c];
        }
    } else {
        // FP64 residual b - A x, FP32 correction on the same factors, per column
        bool converged = !gfail;
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            if (c < n_cols) {
                double x64 = (double)x[c];
                double rel = 1.0, rel_prev = 2.0;
                for (int it = 0; it < P.refine_steps; ++it) {
                    s_x[g][li] = x64;
                    __syncthreads();
                    double res = b64[c];
#pragma unroll
                    for (int j = 0; j < NMAX; ++j) res -= a64.get(j) * s_x[g][j];
                    __syncthreads();
                    T yy[1] = {forward_replay<NMAX, MulRow>(RBFX_ROWV(Lm), S.piv, S.blk, nslots, nslots_wave, myslot, (T)res)};
                    T dx[1];
                    back_substitute<NMAX, 1, RowT>(RBFX_ROWV(r), S.piv, S.blk, nslots, nslots_wave, myslot, diag, 1, yy, dx);
                    x64 += (double)dx[0];
                    const double num = group_max<NMAX>(li < n ? fabs((double)dx[0]) : 0.0);
                    const double den = group_max<NMAX>(li < n ? fabs(x64) : 0.0);
                    rel = den > 0.0 ? num / den : num;
                    // a group stops when converged, or when the correction does not
                    // contract (the FP32 factors are too inaccurate for this system:
                    // it goes to the FP64 re-solve without spending the remaining steps)
                    const bool stalled = !(rel < 0.5 * rel_prev) || !(rel < 1.0);
                    rel_prev = rel;
                    if (__ballot(!(rel <= 0.25 * P.accept_tol) && !stalled) == 0ull) break;
                }
                if (!(rel <= P.accept_tol)) converged = false;
                if (is_pt) P.out_val[(int64_t)c * P.out_nnz + o0 + pi] = gfail ? __builtin_nan("") : x64;
                if (is_poly && P.out_lambda)
                    P.out_lambda[((int64_t)c * P.n_eval + e) * P.n_mono + li] = gfail ? __builtin_nan("") : x64;
            }
        }
        if (alive) {
            st = converged ? RBFX_ST_REFINED : RBFX_ST_FALLBACK64;
            if (st == RBFX_ST_FALLBACK64 && li == 0 && P.fail_list) {
                // refinement did not settle or FP32 pivots broke down: re-solve in FP64
                const int s2 = atomicAdd(P.fail_count, 1);
                P.fail_list[s2] = (int32_t)e;
            }
        }
    }
    if (is_pt && P.out_col && !P.packed) P.out_col[o0 + pi] = gidx;
    if (alive && li == 0 && P.status) P.status[e] = st;
}

template <typename T, int NMAX, int DIM, int RC>
inline hipError_t launch_one(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    if (n_work <= 0) return hipSuccess;
    constexpr int G = 64 / NMAX;
    // one group of NMAX lanes per stencil; a work-list launch sizes the grid by its upper bound
    BuildDev Q = P;
    Q.n_blocks = (n_work + G - 1) / G;
    hipLaunchKernelGGL((k_build<T, NMAX, DIM, RC>), dim3((unsigned)pad_to_xcds(Q.n_blocks)), dim3(64), 0, s, Q);
    return hipGetLastError();
}

template <typename T, int NMAX, int DIM>
inline hipError_t dispatch_rc(const BuildDev& P, int64_t n_work, hipStream_t s)
{
    if (P.n_cols <= 1) return launch_one<T, NMAX, DIM, 1>(P, n_work, s);
    if (P.n_cols <= 4) return launch_one<T, NMAX, DIM, 4>(P, n_work, s);
    return launch_one<T, NMAX, DIM, 8>(P, n_work, s);
}

template <typename T, int NMAX>
inline hipError_t dispatch_dim(const BuildDev& P, int dim, int64_t n_work, hipStream_t s)
{
    switch (dim) {
    case 1: return dispatch_rc<T, NMAX, 1>(P, n_work, s);
    case 2: return dispatch_rc<T, NMAX, 2>(P, n_work, s);
    case 3: return dispatch_rc<T, NMAX, 3>(P, n_work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
