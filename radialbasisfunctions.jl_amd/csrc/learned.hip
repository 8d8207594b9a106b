// Learned weight path for gfx950: surrogate inference fused with FP64 acceptance.
//
// No reference counterpart. Each 256-thread workgroup takes 64 stencils:
//   A. gather the k neighbours, scale by the stencil radius h, rotate/reflect into the
//      canonical frame (canonical_rotation), write the network inputs (each offset as
//      a BF16 high part plus a BF16 remainder) to LDS;
//   B. evaluate the MLP (128 -> 256 -> 256 -> 256 -> 32, GELU) on the matrix cores
//      (v_mfma_f32_16x16x32_bf16, activations in LDS, weights streamed from L2);
//   C. in FP64: de-standardise the outputs (the network predicts w~ r1^2, r1 the
//      distance of the canonical reference neighbour) to w~, project w~ onto exact
//      polynomial reproduction (P^T w~ = L P, minimal-norm correction through the
//      Gram matrix G = P^T P), refit lambda by least squares, and form the collocation
//      residual s - P lambda with s = L phi - Phi w~ (O(k^2) against the O(k^3) solve);
//   D. accept when max|residual| <= tol * max|L phi|: write w = w~ h^-order and the
//      column indices; otherwise queue the stencil for the solver.
// Four lanes (a DPP quad) share a stencil in phases A, C and D; rows i = q, q+4, ...
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "learned.h"
#include "mono.h"

namespace rbfx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64;           // stencils per workgroup
constexpr int kActLd = kHid + 8;  // BF16 row stride of the activation tile (16-B aligned rows)

__device__ __forceinline__ double quad_sum(double v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}
__device__ __forceinline__ double quad_max(double v)
{
    v = fmax(v, __shfl_xor(v, 1, 64));
    v = fmax(v, __shfl_xor(v, 2, 64));
    return v;
}

// monomial a (graded exponents e) at a 2-D point
__device__ __forceinline__ double mono2(const int (&e)[3], double x, double y)
{
    double v = 1.0;
#pragma unroll
    for (int p = 0; p < kMaxPolyDeg; ++p) {
        if (p < e[0]) v *= x;
        if (p < e[1]) v *= y;
    }
    return v;
}

// acc[i][j] += act[16 i + .][k] * Wt[ncol0 + 16 j + .][k] over k < KL, for the 64 rows
template <int KL, int NT>
__device__ __forceinline__ void mlp_tiles(const __bf16 (*act)[kActLd], const __bf16* __restrict__ Wt, int ncol0,
                                          f32x4 (&acc)[4][NT], int lane)
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int k0 = 0; k0 < KL; k0 += 32) {
        bf16x8 a[4], b[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j)
            b[j] = *reinterpret_cast<const bf16x8*>(Wt + (int64_t)(ncol0 + j * 16 + fr) * KL + k0 + fk);
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&act[i * 16 + fr][k0 + fk]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// one hidden layer: act[:, :N] = gelu(act[:, :KL] Wt^T + b), in place (all 4 waves)
template <int KL>
__device__ __forceinline__ void hidden_layer(__bf16 (*act)[kActLd], const __bf16* __restrict__ Wt,
                                             const float* __restrict__ bias, int w, int lane)
{
    f32x4 acc[4][4];
    mlp_tiles<KL, 4>(act, Wt, w * 64, acc, lane);
    __syncthreads();  // every wave has read its inputs
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = w * 64 + j * 16 + (lane & 15);
            const float b = bias[col];
#pragma unroll
            for (int q = 0; q < 4; ++q) act[i * 16 + (lane >> 4) * 4 + q][col] = (__bf16)gelu_tanh(acc[i][j][q] + b);
        }
    __syncthreads();
}

template <int K, int MM>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(2))) void k_learned(
    const LearnedDev P)
{
    constexpr int R = (K + 3) / 4;  // rows per quad lane
    __shared__ __attribute__((aligned(16))) __bf16 s_act[kBM][kActLd];
    __shared__ double s_xi[kBM][K][2];
    __shared__ float s_out[kBM][kOutPad + 1];
    __shared__ double s_w[kBM][K];
    __shared__ int s_mono[MM][3];
    __shared__ TermDev s_terms[RBFX_MAX_TERMS];

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int sl = t >> 2, q = t & 3;
    const int64_t lb = xcd_contiguous_block(P.n_blocks);
    if (lb < 0) return;  // padding block: the whole workgroup leaves
    const int64_t wi = lb * kBM + sl;
    int64_t e = (P.order && wi < P.n_eval) ? (int64_t)P.order[wi] : wi;
    const bool live = wi < P.n_eval && e >= 0 && e < P.n_eval;
    if (!live) e = 0;
    if (t < MM * 3) (&s_mono[0][0])[t] = (&P.mono_e[0][0])[t];
    if (t < RBFX_MAX_TERMS) s_terms[t] = P.terms[t];

    // ---- A. radius-scaled offsets and the canonical frame
    double xe = 0.0, ye = 0.0;
    if (live) {
        xe = P.eval[2 * e];
        ye = P.eval[2 * e + 1];
    }
    double dx[R], dy[R];
    int gi[R];
    double h2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        dx[r] = 0.0;
        dy[r] = 0.0;
        gi[r] = 0;
        if (live && i < K) {
            gi[r] = P.adj_idx[e * K + i];
            dx[r] = P.data[2 * (int64_t)gi[r]] - xe;
            dy[r] = P.data[2 * (int64_t)gi[r] + 1] - ye;
            h2 = fmax(h2, dx[r] * dx[r] + dy[r] * dy[r]);
        }
    }
    h2 = quad_max(h2);
    double h = sqrt(h2);
    if (!(h > 0.0)) h = 1.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        dx[r] = dx[r] / h;
        dy[r] = dy[r] / h;
        if (i < K) {
            s_xi[sl][i][0] = dx[r];
            s_xi[sl][i][1] = dy[r];
        }
    }
    __syncthreads();
    double c = 1.0, s = 0.0, r1 = 1.0;
    for (int i = 0; i < K; ++i) {
        const double x = s_xi[sl][i][0], y = s_xi[sl][i][1];
        if (x != 0.0 || y != 0.0) {
            canonical_rotation(x, y, c, s);
            r1 = sqrt(x * x + y * y);
            break;
        }
    }
    double sy = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double x = c * dx[r] + s * dy[r];
        const double y = -s * dx[r] + c * dy[r];
        dx[r] = x;
        dy[r] = y;
        if (q + 4 * r < K) sy += y;
    }
    sy = quad_sum(sy);
    const double fl = sy < 0.0 ? -1.0 : 1.0;
    __syncthreads();  // the scan above read s_xi
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        dy[r] *= fl;
        if (i < K) {
            s_xi[sl][i][0] = dx[r];
            s_xi[sl][i][1] = dy[r];
            // BF16 high part, then the BF16 rounding of the remainder
            const __bf16 hx = (__bf16)(float)dx[r], hy = (__bf16)(float)dy[r];
            s_act[sl][2 * i] = hx;
            s_act[sl][2 * i + 1] = hy;
            s_act[sl][2 * K + 2 * i] = (__bf16)(float)(dx[r] - (double)(float)hx);
            s_act[sl][2 * K + 2 * i + 1] = (__bf16)(float)(dy[r] - (double)(float)hy);
        }
    }
    for (int col = 4 * K + q; col < kInPad; col += 4) s_act[sl][col] = (__bf16)0.0f;
    __syncthreads();

    // ---- B. network
    hidden_layer<kInPad>(s_act, P.wbf16 + mlp_wb_off(0), P.params + mlp_b_off(0), w, lane);
    hidden_layer<kHid>(s_act, P.wbf16 + mlp_wb_off(1), P.params + mlp_b_off(1), w, lane);
    hidden_layer<kHid>(s_act, P.wbf16 + mlp_wb_off(2), P.params + mlp_b_off(2), w, lane);
    if (w < kOutPad / 16) {
        f32x4 acc[4][1];
        mlp_tiles<kHid, 1>(s_act, P.wbf16 + mlp_wb_off(3), w * 16, acc, lane);
        const int col = w * 16 + (lane & 15);
        const float b = P.params[mlp_b_off(3) + col];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) s_out[i * 16 + (lane >> 4) * 4 + qq][col] = acc[i][0][qq] + b;
    }
    __syncthreads();

    // ---- C. FP64 projection, multiplier refit and collocation residual
    BasisDev B;
    B.kind = P.basis_kind;
    B.eps = (B.kind == RBFX_IMQ || B.kind == RBFX_GAUSSIAN) ? P.eps * h : P.eps;
    double wt[R];
    double ptw[MM], G[MM][MM];
    double pr[R][MM];  // the monomials at this lane's rows, evaluated once
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int a = 0; a < MM; ++a) pr[r][a] = mono2(s_mono[a], dx[r], dy[r]);
#pragma unroll
    for (int a = 0; a < MM; ++a) {
        ptw[a] = 0.0;
#pragma unroll
        for (int b = 0; b < MM; ++b) G[a][b] = 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        wt[r] = 0.0;
        if (i < K) {
            // the network predicts w~ r1^2 (r1: the canonical reference distance)
            wt[r] = ((double)s_out[sl][i] * P.out_std[i] + P.out_mean[i]) / (r1 * r1);
#pragma unroll
            for (int a = 0; a < MM; ++a) {
                ptw[a] += pr[r][a] * wt[r];
#pragma unroll
                for (int b = 0; b <= a; ++b) G[a][b] += pr[r][a] * pr[r][b];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < MM; ++a) {
        ptw[a] = quad_sum(ptw[a]);
#pragma unroll
        for (int b = 0; b <= a; ++b) G[a][b] = quad_sum(G[a][b]);
    }
    // Cholesky G = L L^T in place (lower triangle)
    bool spd = true;
#pragma unroll
    for (int a = 0; a < MM; ++a) {
        double d = G[a][a];
#pragma unroll
        for (int b = 0; b < a; ++b) d -= G[a][b] * G[a][b];
        spd = spd && d > 0.0;
        d = sqrt(fmax(d, 1e-300));
        G[a][a] = d;
#pragma unroll
        for (int c2 = a + 1; c2 < MM; ++c2) {
            double v = G[c2][a];
#pragma unroll
            for (int b = 0; b < a; ++b) v -= G[c2][b] * G[a][b];
            G[c2][a] = v / d;
        }
    }
    // L P at the centre (the operator's action on each monomial, scaled frame)
    double rhs[MM];
    {
        const double z[2] = {0.0, 0.0};
#pragma unroll
        for (int a = 0; a < MM; ++a) {
            double lp = 0.0;
            for (int tt = 0; tt < P.n_terms; ++tt) {
                const TermDev tm = s_terms[tt];
                lp += tm.coeff * mono_op<2>(s_mono[a], tm.kind, tm.d1, tm.d2, z);
            }
            rhs[a] = lp - ptw[a];
        }
    }
    auto chol_solve = [&](double (&v)[MM]) {
#pragma unroll
        for (int a = 0; a < MM; ++a) {
#pragma unroll
            for (int b = 0; b < a; ++b) v[a] -= G[a][b] * v[b];
            v[a] /= G[a][a];
        }
#pragma unroll
        for (int a = MM - 1; a >= 0; --a) {
#pragma unroll
            for (int b = a + 1; b < MM; ++b) v[a] -= G[b][a] * v[b];
            v[a] /= G[a][a];
        }
    };
    chol_solve(rhs);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        double corr = 0.0;
#pragma unroll
        for (int a = 0; a < MM; ++a) corr += pr[r][a] * rhs[a];
        wt[r] += corr;
        if (i < K) s_w[sl][i] = wt[r];
    }
    __syncthreads();
    // s_i = L phi(x_e, x_i) - sum_j phi(|xi_i - xi_j|) w~_j
    double sres[R], lmax = 0.0;
    double pts[MM];
#pragma unroll
    for (int a = 0; a < MM; ++a) pts[a] = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = q + 4 * r;
        sres[r] = 0.0;
        if (i < K) {
            const double De[2] = {-dx[r], -dy[r]};
            const double q2 = dx[r] * dx[r] + dy[r] * dy[r];
            double lphi = 0.0;
            for (int tt = 0; tt < P.n_terms; ++tt) {
                const TermDev tm = s_terms[tt];
                lphi += tm.coeff * phi_term<2>(B, tm.kind, tm.d1, tm.d2, De, q2);
            }
            lmax = fmax(lmax, fabs(lphi));
            double acc = 0.0;
            if (B.kind == RBFX_PHS3) {  // uniform: the surrogate's own basis, without the basis switch
#pragma unroll 4
                for (int j = 0; j < K; ++j) {
                    const double ex = dx[r] - s_xi[sl][j][0], ey = dy[r] - s_xi[sl][j][1];
                    const double q2 = ex * ex + ey * ey;
                    acc += q2 * sqrt(q2) * s_w[sl][j];
                }
            } else {
#pragma unroll 4
                for (int j = 0; j < K; ++j) {
                    const double ex = dx[r] - s_xi[sl][j][0], ey = dy[r] - s_xi[sl][j][1];
                    acc += phi_r2(B, ex * ex + ey * ey) * s_w[sl][j];
                }
            }
            sres[r] = lphi - acc;
#pragma unroll
            for (int a = 0; a < MM; ++a) pts[a] += pr[r][a] * sres[r];
        }
    }
#pragma unroll
    for (int a = 0; a < MM; ++a) pts[a] = quad_sum(pts[a]);
    chol_solve(pts);  // lambda
    double rmax = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (q + 4 * r < K) {
            double v = sres[r];
#pragma unroll
            for (int a = 0; a < MM; ++a) v -= pr[r][a] * pts[a];
            rmax = fmax(rmax, fabs(v));
        }
    }
    rmax = quad_max(rmax);
    lmax = quad_max(lmax);
    const double rho = lmax > 0.0 ? rmax / lmax : rmax;

    // ---- D. accept or queue for the solver
    if (!live) return;
    const bool ok = spd && rho <= P.accept_tol;
    if (P.out_rho && q == 0) P.out_rho[e] = rho;
    if (ok) {
        // w = w~ h^-order (the scaled frame absorbs the radius)
        double sc = 1.0;
        for (int o = 0; o < (int)P.op_scale_order; ++o) sc /= h;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = q + 4 * r;
            if (i < K) {
                P.out_val[e * K + i] = wt[r] * sc;
                if (P.out_col) P.out_col[e * K + i] = gi[r];
            }
        }
        if (q == 0 && P.status) P.status[e] = RBFX_ST_LEARNED;
    } else if (q == 0) {
        const int slot = atomicAdd(P.reject_count, 1);
        P.reject_list[slot] = (int32_t)e;
    }
}

hipError_t launch_learned_build(const LearnedDev& P, hipStream_t s)
{
    if (P.n_eval <= 0) return hipSuccess;
    if (P.k != 20 || P.n_mono != 6) return hipErrorInvalidValue;
    LearnedDev Q = P;
    Q.n_blocks = (P.n_eval + kBM - 1) / kBM;
    hipLaunchKernelGGL((k_learned<20, 6>), dim3((unsigned)pad_to_xcds(Q.n_blocks)), dim3(256), 0, s, Q);
    return hipGetLastError();
}

// ---- canonical features for training data: one thread per stencil, the same
// arithmetic as phase A above (quad partial sums of the y offsets included)
__global__ __launch_bounds__(256) void k_features(const double* data, const double* eval, int64_t n_eval, int k,
                                                  const int32_t* adj, double* out_xi, __bf16* out_x, double* out_h,
                                                  double* out_r1)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n_eval) return;
    const double xe = eval[2 * e], ye = eval[2 * e + 1];
    double hq[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < k; ++i) {
        const int64_t j = adj[e * k + i];
        const double dx = data[2 * j] - xe, dy = data[2 * j + 1] - ye;
        hq[i & 3] = fmax(hq[i & 3], dx * dx + dy * dy);
    }
    double h = sqrt(fmax(fmax(hq[0], hq[1]), fmax(hq[2], hq[3])));
    if (!(h > 0.0)) h = 1.0;
    double c = 1.0, s = 0.0, r1 = 1.0;
    for (int i = 0; i < k; ++i) {
        const int64_t j = adj[e * k + i];
        const double x = (data[2 * j] - xe) / h, y = (data[2 * j + 1] - ye) / h;
        if (x != 0.0 || y != 0.0) {
            canonical_rotation(x, y, c, s);
            r1 = sqrt(x * x + y * y);
            break;
        }
    }
    double sq[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < k; ++i) {
        const int64_t j = adj[e * k + i];
        const double x = (data[2 * j] - xe) / h, y = (data[2 * j + 1] - ye) / h;
        sq[i & 3] += -s * x + c * y;
    }
    const double fl = ((sq[0] + sq[1]) + (sq[2] + sq[3])) < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < k; ++i) {
        const int64_t j = adj[e * k + i];
        const double x = (data[2 * j] - xe) / h, y = (data[2 * j + 1] - ye) / h;
        const double xr = c * x + s * y, yr = fl * (-s * x + c * y);
        if (out_xi) {
            out_xi[(e * k + i) * 2] = xr;
            out_xi[(e * k + i) * 2 + 1] = yr;
        }
        if (out_x && 4 * k <= kInPad) {
            const __bf16 hx = (__bf16)(float)xr, hy = (__bf16)(float)yr;
            out_x[e * kInPad + 2 * i] = hx;
            out_x[e * kInPad + 2 * i + 1] = hy;
            out_x[e * kInPad + 2 * k + 2 * i] = (__bf16)(float)(xr - (double)(float)hx);
            out_x[e * kInPad + 2 * k + 2 * i + 1] = (__bf16)(float)(yr - (double)(float)hy);
        }
    }
    if (out_x)
        for (int col = 4 * k; col < kInPad; ++col) out_x[e * kInPad + col] = (__bf16)0.0f;
    if (out_h) out_h[e] = h;
    if (out_r1) out_r1[e] = r1;
}

hipError_t launch_stencil_features(const double* data, const double* eval, int64_t n_eval, int k,
                                   const int32_t* adj_idx, double* out_xi, __bf16* out_x, double* out_h,
                                   double* out_r1, hipStream_t s)
{
    if (n_eval <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((n_eval + 255) / 256)), dim3(256), 0, s, data, eval, n_eval, k,
                       adj_idx, out_xi, out_x, out_h, out_r1);
    return hipGetLastError();
}

}  // namespace rbfx
