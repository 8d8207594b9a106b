// Internal interface of the stencil-weight MLP (training and inference kernels).
//
// Network: in (2k offsets, each split into a BF16 high part and a BF16 remainder, so
// 4k inputs zero-padded to kInPad) -> kHid -> kHid -> kHid -> out (k weights + m
// multipliers, zero-padded to kOutPad); GELU (tanh form) on the hidden layers; BF16
// operands, FP32 accumulation on the matrix cores. The split keeps ~16 mantissa bits
// of the offsets, which the 1/r^2 sensitivity of the weights needs.
//
// Parameters (FP32 master copy, one flat array): for each layer l, Wt_l [N_l x K_l]
// (output-major, the reduction index contiguous) followed by b_l [N_l]. The BF16
// copy used by the matrix cores holds the Wt_l blocks only, back to back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbfx {

constexpr int kMlpLayers = 4;
constexpr int kInPad = 128;
constexpr int kHid = 256;
constexpr int kOutPad = 32;

__host__ __device__ constexpr int mlp_k(int l) { return l == 0 ? kInPad : kHid; }
__host__ __device__ constexpr int mlp_n(int l) { return l == kMlpLayers - 1 ? kOutPad : kHid; }
// offset of Wt_l in the flat FP32 parameter array (b_l follows it)
__host__ __device__ constexpr int64_t mlp_w_off(int l)
{
    int64_t o = 0;
    for (int i = 0; i < l; ++i) o += (int64_t)mlp_n(i) * mlp_k(i) + mlp_n(i);
    return o;
}
__host__ __device__ constexpr int64_t mlp_b_off(int l) { return mlp_w_off(l) + (int64_t)mlp_n(l) * mlp_k(l); }
__host__ __device__ constexpr int64_t mlp_n_params() { return mlp_w_off(kMlpLayers); }
// offset of Wt_l in the BF16 copy
__host__ __device__ constexpr int64_t mlp_wb_off(int l)
{
    int64_t o = 0;
    for (int i = 0; i < l; ++i) o += (int64_t)mlp_n(i) * mlp_k(i);
    return o;
}
__host__ __device__ constexpr int64_t mlp_n_wbf16() { return mlp_wb_off(kMlpLayers); }

// C[m][n] = sum_k A(m,k) B(n,k) with element strides, BF16 operands, FP32 accumulation
enum GemmEpi {
    EPI_F32 = 0,        // C (FP32) = acc            (accumulate: C += acc, atomically)
    EPI_BIAS = 1,       // C (FP32) = acc + bias[n]
    EPI_BIAS_GELU = 2,  // C (FP32) = acc + bias[n]; H (BF16) = gelu(C)
    EPI_GELU_BWD = 3,   // H (BF16) = acc * gelu'(Z[m][n])
};

struct GemmArgs {
    int M, N, K;
    const __bf16* A;
    int64_t sAm, sAk;
    const __bf16* B;
    int64_t sBn, sBk;
    float* C;
    int64_t ldc;
    const float* bias;
    __bf16* H;
    int64_t ldh;
    const float* Z;
    int64_t ldz;
    int epi;
    int accumulate;
    int k_split;  // reduction length per grid.z slice (multiple of 32)
};

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s);

// training step pieces
hipError_t launch_mse_grad(int M, int n_out, const float* Y, int64_t ldy, const float* T, int64_t ldt,
                           const float* row_weight, __bf16* dY, int64_t ld_dy, float* loss, hipStream_t s);
hipError_t launch_colsum_bf16(int M, int N, const __bf16* X, int64_t ldx, float* out, hipStream_t s);
hipError_t launch_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s);
hipError_t launch_params_to_bf16(const float* p, __bf16* wb, hipStream_t s);
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);

// tanh(u) = 1 - 2 / (exp(2u) + 1): one hardware exp and one reciprocal. exp(2u)
// overflows to +inf for large u (giving 1) and underflows to 0 for very negative u
// (giving -1); near u = 0 the absolute error is a few float ulps, far below the BF16
// rounding the activations then take.
__device__ __forceinline__ float tanh_exp(float u)
{
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * u) + 1.0f);
}
__device__ __forceinline__ float gelu_tanh(float z)
{
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    const float t = tanh_exp(c * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_tanh_grad(float z)
{
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z);
    const float t = tanh_exp(u);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

}  // namespace rbfx
