// Spatial (Morton / Z-order) ordering of points for gfx950.
//
// No reference counterpart beyond reorder_points! (src/utils.jl:45-55), which sorts
// points on the host. Here the order is used without moving any user data: the
// apply walks its rows in this order, so the X gathers of consecutive rows hit the
// same cache lines. Keys: each coordinate quantised to 64/dim bits on the bounding
// box, bits interleaved; hipCUB radix sort of (key, index) pairs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "order.h"

namespace rbfx {

__global__ __launch_bounds__(256) void k_bbox_blocks(const double* __restrict__ x, int64_t n, int dim,
                                                     double* __restrict__ part)
{
    __shared__ double smn[3][256], smx[3][256];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
#pragma unroll
        for (int d = 0; d < 3; ++d)  // constant indexes keep mn/mx in registers
            if (d < dim) {
                const double v = x[i * dim + d];
                mn[d] = fmin(mn[d], v);
                mx[d] = fmax(mx[d], v);
            }
    for (int d = 0; d < 3; ++d) {
        smn[d][threadIdx.x] = mn[d];
        smx[d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                smn[d][threadIdx.x] = fmin(smn[d][threadIdx.x], smn[d][threadIdx.x + w]);
                smx[d][threadIdx.x] = fmax(smx[d][threadIdx.x], smx[d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            part[blockIdx.x * 6 + d] = smn[d][0];
            part[blockIdx.x * 6 + 3 + d] = smx[d][0];
        }
}

// spread the low `bits` bits of v so consecutive bits land `stride` apart
__device__ __forceinline__ uint64_t spread_bits(uint64_t v, int bits, int stride)
{
    uint64_t r = 0;
    for (int b = 0; b < bits; ++b) r |= ((v >> b) & 1ull) << (b * stride);
    return r;
}

__global__ __launch_bounds__(256) void k_morton(const double* __restrict__ x, int64_t n, int dim, double lo0,
                                                double lo1, double lo2, double sc0, double sc1, double sc2,
                                                uint64_t* __restrict__ key, int32_t* __restrict__ idx)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= n) return;
    const int bits = 63 / dim;
    const double lim = (double)((1ull << bits) - 1ull);
    const double lo[3] = {lo0, lo1, lo2}, sc[3] = {sc0, sc1, sc2};
    uint64_t k = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d < dim) {
            double q = (x[i * dim + d] - lo[d]) * sc[d];
            q = q < 0.0 ? 0.0 : (q > lim ? lim : q);
            k |= spread_bits((uint64_t)q, bits, dim) << d;
        }
    key[i] = k;
    idx[i] = (int32_t)i;
}

hipError_t spatial_order(const double* x, int64_t n, int dim, int32_t* out_order, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int nb = (int)std::min<int64_t>((n + 255) / 256, 1024);
    double* part = nullptr;
    uint64_t *key = nullptr, *key2 = nullptr;
    int32_t* idx = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipMalloc(&part, sizeof(double) * 6 * nb);
    if (e == hipSuccess) e = hipMalloc(&key, sizeof(uint64_t) * n);
    if (e == hipSuccess) e = hipMalloc(&key2, sizeof(uint64_t) * n);
    if (e == hipSuccess) e = hipMalloc(&idx, sizeof(int32_t) * n);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bbox_blocks, dim3(nb), dim3(256), 0, s, x, n, dim, part);
        e = hipGetLastError();
    }
    double lo[3] = {0, 0, 0}, sc[3] = {0, 0, 0};
    if (e == hipSuccess) {
        double* hp = new double[6 * nb];
        e = hipMemcpyAsync(hp, part, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) {
            const int bits = 63 / dim;
            for (int d = 0; d < dim; ++d) {
                double mn = 1e300, mx = -1e300;
                for (int b = 0; b < nb; ++b) {
                    mn = std::min(mn, hp[b * 6 + d]);
                    mx = std::max(mx, hp[b * 6 + 3 + d]);
                }
                lo[d] = mn;
                const double ext = mx - mn;
                sc[d] = ext > 0 ? (double)((1ull << bits) - 1ull) / ext : 0.0;
            }
        }
        delete[] hp;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_morton, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, dim, lo[0], lo[1],
                           lo[2], sc[0], sc[1], sc[2], key, idx);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2, idx, out_order, (int)n, 0, 64, s);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key2, idx, out_order, (int)n, 0, 64, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (tmp) (void)hipFree(tmp);
    if (idx) (void)hipFree(idx);
    if (key2) (void)hipFree(key2);
    if (key) (void)hipFree(key);
    if (part) (void)hipFree(part);
    return e;
}

}  // namespace rbfx
