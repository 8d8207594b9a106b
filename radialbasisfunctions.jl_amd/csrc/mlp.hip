// Training kernels of the stencil-weight MLP on the gfx950 matrix cores.
//
// One generic BF16 GEMM, C[m][n] = sum_k A(m,k) B(n,k), with element strides on both
// operands so the same kernel serves the forward product (H Wt^T), the input
// gradient (dZ Wt) and the weight gradient (dZ^T H), plus fused epilogues (bias,
// GELU, GELU backward, split-K accumulation). A 64x64 output tile per 256-thread
// workgroup; each wavefront owns a 32x32 quadrant as 2x2 v_mfma_f32_16x16x32_bf16
// tiles. Operand tiles pass through LDS with the reduction index contiguous, so a
// lane's fragment (8 consecutive k of one row or column) is a single 16-byte read.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"

namespace rbfx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64, kBK = 32, kLdsK = kBK + 8;  // +8: rows 80 B apart (16-B aligned, fewer bank conflicts)

__global__ __launch_bounds__(256) void k_gemm(const GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[kTile][kLdsK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[kTile][kLdsK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int kbeg = blockIdx.z * g.k_split;
    const int kend = min(g.K, kbeg + g.k_split);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // each thread stages 8 consecutive k of one row of each operand tile
    const int lr = t >> 2, lk = (t & 3) * 8;
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        {
            const int m = m0 + lr;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = k0 + lk + q;
                As[lr][lk + q] = (m < g.M && k < kend) ? g.A[(int64_t)m * g.sAm + (int64_t)k * g.sAk] : (__bf16)0.0f;
            }
            const int n = n0 + lr;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = k0 + lk + q;
                Bs[lr][lk + q] = (n < g.N && k < kend) ? g.B[(int64_t)n * g.sBn + (int64_t)k * g.sBk] : (__bf16)0.0f;
            }
        }
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + fr][fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + fr][fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // D layout of a 16x16 tile: lane holds column (lane % 16), rows 4*(lane / 16) + q
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][q];
                switch (g.epi) {
                case EPI_F32:
                    if (g.accumulate) atomicAdd(&g.C[(int64_t)m * g.ldc + n], v);
                    else g.C[(int64_t)m * g.ldc + n] = v;
                    break;
                case EPI_BIAS:
                    g.C[(int64_t)m * g.ldc + n] = v + g.bias[n];
                    break;
                case EPI_BIAS_GELU: {
                    const float z = v + g.bias[n];
                    g.C[(int64_t)m * g.ldc + n] = z;
                    g.H[(int64_t)m * g.ldh + n] = (__bf16)gelu_tanh(z);
                    break;
                }
                default: {  // EPI_GELU_BWD
                    const float z = g.Z[(int64_t)m * g.ldz + n];
                    g.H[(int64_t)m * g.ldh + n] = (__bf16)(v * gelu_tanh_grad(z));
                    break;
                }
                }
            }
}

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
    GemmArgs a = g;
    if (a.k_split <= 0 || !a.accumulate) a.k_split = a.K;
    a.k_split = (a.k_split + kBK - 1) / kBK * kBK;
    const int splits = (a.K + a.k_split - 1) / a.k_split;
    dim3 grid((a.N + kTile - 1) / kTile, (a.M + kTile - 1) / kTile, splits);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// loss = mean_m w_m mean_n (Y-T)^2 (w_m = 1 without row weights);
// dY = 2 w_m (Y - T) / (M n_out) as BF16 (zero in the padding columns)
__global__ __launch_bounds__(256) void k_mse_grad(int M, int n_out, const float* Y, int64_t ldy, const float* T,
                                                  int64_t ldt, const float* row_weight, __bf16* dY, int64_t ld_dy,
                                                  float* loss)
{
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)M * ld_dy;
    float l = 0.f;
    if (i < total) {
        const int64_t m = i / ld_dy;
        const int n = (int)(i % ld_dy);
        const float wm = row_weight ? row_weight[m] : 1.0f;
        float d = 0.f;
        if (n < n_out) {
            d = Y[m * ldy + n] - T[m * ldt + n];
            l = wm * d * d;
        }
        dY[i] = (__bf16)(2.0f * wm * d / ((float)M * (float)n_out));
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / ((float)M * (float)n_out));
}

hipError_t launch_mse_grad(int M, int n_out, const float* Y, int64_t ldy, const float* T, int64_t ldt,
                           const float* row_weight, __bf16* dY, int64_t ld_dy, float* loss, hipStream_t s)
{
    const int64_t total = (int64_t)M * ld_dy;
    hipLaunchKernelGGL(k_mse_grad, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, M, n_out, Y, ldy, T, ldt,
                       row_weight, dY, ld_dy, loss);
    return hipGetLastError();
}

// out[n] = sum_m X[m][n] (FP32 accumulation of BF16 values); one thread per column chunk
__global__ __launch_bounds__(256) void k_colsum_bf16(int M, int N, const __bf16* X, int64_t ldx, float* out)
{
    const int n = blockIdx.x * 64 + (threadIdx.x & 63);
    const int part = threadIdx.x >> 6;
    __shared__ float red[4][64];
    float acc = 0.f;
    if (n < N)
        for (int m = part; m < M; m += 4) acc += (float)X[(int64_t)m * ldx + n];
    red[part][threadIdx.x & 63] = acc;
    __syncthreads();
    if (part == 0 && n < N) out[n] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

hipError_t launch_colsum_bf16(int M, int N, const __bf16* X, int64_t ldx, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum_bf16, dim3((unsigned)((N + 63) / 64)), dim3(256), 0, s, M, N, X, ldx, out);
    return hipGetLastError();
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ __launch_bounds__(256) void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t launch_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1,
                       c2);
    return hipGetLastError();
}

// BF16 copy of the Wt blocks of the flat FP32 parameter array
__global__ __launch_bounds__(256) void k_params_to_bf16(const float* p, __bf16* wb)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= mlp_n_wbf16()) return;
    int l = 0;
    while (l + 1 < kMlpLayers && i >= mlp_wb_off(l + 1)) ++l;
    wb[i] = (__bf16)p[mlp_w_off(l) + (i - mlp_wb_off(l))];
}

hipError_t launch_params_to_bf16(const float* p, __bf16* wb, hipStream_t s)
{
    hipLaunchKernelGGL(k_params_to_bf16, dim3((unsigned)((mlp_n_wbf16() + 255) / 256)), dim3(256), 0, s, p, wb);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_f32(float* p, int64_t n, float v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace rbfx
