// Exact grid-hashed k nearest neighbours for gfx950.
//
// Replaces find_neighbors (src/utils.jl:12-18: KDTree + knn(tree, eval, k, true)).
// The result is the set of the k data points with the smallest squared Euclidean
// distance to each query, ascending, computed in FP64 in dimension order without
// fused multiply-adds (the same rounding a KD-tree query sees); equal distances
// are ordered by lower index.
//
// Data points are binned into a uniform grid (counting sort: count, scan,
// scatter), queries are binned into the same grid so that a wavefront works on
// spatially adjacent queries, and every query scans rings of cells outward until
// no unvisited cell can hold a point closer than its current k-th neighbour.
// The top-k list lives in registers (static-index insertion network).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "knn.h"
#include "knn_grid.h"

namespace rbfx {

struct Grid {
    double lo[3];
    double inv_s;  // 1 / cell size
    double s;
    int n[3];      // cells per dimension
    int dim;
};

__device__ __forceinline__ int cell_coord(double x, double lo, double inv_s, int n)
{
    int c = (int)floor((x - lo) * inv_s);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

template <int DIM>
__device__ __forceinline__ int64_t cell_of(const double* p, const Grid& G)
{
    int64_t id = 0;
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) id = id * G.n[d] + cell_coord(p[d], G.lo[d], G.inv_s, G.n[d]);
    return id;
}

// ---------------------------------------------------------------- bbox
__global__ void k_bbox_partial(const double* __restrict__ x, int64_t n, int dim, double* out)
{
    __shared__ double smin[3][256], smax[3][256];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < dim; ++d) {
            const double v = x[i * dim + d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    for (int d = 0; d < 3; ++d) { smin[d][threadIdx.x] = mn[d]; smax[d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + w]);
                smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            out[blockIdx.x * 6 + d] = smin[d][0];
            out[blockIdx.x * 6 + 3 + d] = smax[d][0];
        }
}

// ---------------------------------------------------------------- binning
// plain fill kernel rather than a memset node, so the sequence stays capturable
__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <int DIM>
__global__ void k_cell_count(const double* __restrict__ x, int64_t n, Grid G, int32_t* __restrict__ cell,
                             int32_t* __restrict__ count)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = (int32_t)cell_of<DIM>(x + i * DIM, G);
        cell[i] = c;
        atomicAdd(&count[c], 1);
    }
}

template <int DIM>
__global__ void k_scatter(const double* __restrict__ x, int64_t n, const int32_t* __restrict__ cell,
                          int32_t* __restrict__ cursor, int32_t* __restrict__ sorted_idx,
                          double* __restrict__ sorted_x)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t slot = atomicAdd(&cursor[cell[i]], 1);
        sorted_idx[slot] = (int32_t)i;
        if (sorted_x)
#pragma unroll
            for (int d = 0; d < DIM; ++d) sorted_x[(int64_t)slot * DIM + d] = x[i * DIM + d];
    }
}

// ---------------------------------------------------------------- query
// (d2, idx) lexicographic order
__device__ __forceinline__ bool lt(double a, int ia, double b, int ib) { return a < b || (a == b && ia < ib); }

template <int DIM>
__device__ __forceinline__ double sqdist(const double* a, const double (&q)[DIM])
{
    // no contraction: same rounding as sum((x .- y).^2) evaluated term by term. The
    // operators must be written here, inside the pragma's scope: arithmetic helpers
    // defined elsewhere keep their own contraction setting and were fused into FMAs.
#pragma clang fp contract(off)
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double t = a[d] - q[d];
        const double t2 = t * t;
        s = s + t2;
    }
    return s;
}

template <int DIM, int KCAP>
__global__ __launch_bounds__(128) void k_query(const double* __restrict__ sx, const int32_t* __restrict__ sidx,
                                               const int32_t* __restrict__ cell_start,
                                               const double* __restrict__ query,
                                               const int32_t* __restrict__ qorder, int64_t n_query, Grid G,
                                               int kmax, const int64_t* __restrict__ q_ptr,
                                               int32_t* __restrict__ out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n_query) return;
    const int64_t qi = qorder ? qorder[t] : t;
    double q[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[d] = query[qi * DIM + d];
    int c[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) c[d] = cell_coord(q[d], G.lo[d], G.inv_s, G.n[d]);

    double bd[KCAP];
    int bi[KCAP];
    // slots [0, KCAP-kmax) hold sentinels below every real distance, so the
    // insertion threshold is always the last slot
#pragma unroll
    for (int s = 0; s < KCAP; ++s) {
        bd[s] = (s < KCAP - kmax) ? -1.0 : __builtin_huge_val();
        bi[s] = (s < KCAP - kmax) ? -1 : 0x7fffffff;
    }
    int filled = 0;
    int maxn = G.n[0];
    for (int d = 1; d < DIM; ++d) maxn = max(maxn, G.n[d]);

    for (int rho = 0;; ++rho) {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) { lo[d] = c[d] - rho; hi[d] = c[d] + rho; }
        // iterate the shell of cells at Chebyshev distance rho
        const int z0 = (DIM == 3) ? max(lo[2], 0) : 0, z1 = (DIM == 3) ? min(hi[2], G.n[2] - 1) : 0;
        for (int cz = z0; cz <= z1; ++cz) {
            const bool zedge = (DIM == 3) && (cz == lo[2] || cz == hi[2]);
            const int y0 = (DIM >= 2) ? max(lo[1], 0) : 0, y1 = (DIM >= 2) ? min(hi[1], G.n[1] - 1) : 0;
            for (int cy = y0; cy <= y1; ++cy) {
                const bool yedge = (DIM >= 2) && (cy == lo[1] || cy == hi[1]);
                const bool full_row = zedge || yedge || (DIM == 1 && rho == 0);
                const int x0 = max(lo[0], 0), x1 = min(hi[0], G.n[0] - 1);
                const int step = full_row ? 1 : (hi[0] - lo[0] > 0 ? hi[0] - lo[0] : 1);
                for (int cx = full_row ? x0 : lo[0]; cx <= (full_row ? x1 : hi[0]); cx += step) {
                    if (cx < 0 || cx >= G.n[0]) continue;
                    const int64_t cid = ((int64_t)cz * G.n[1] + cy) * G.n[0] + cx;
                    const int32_t b = cell_start[cid], e = cell_start[cid + 1];
                    for (int32_t s = b; s < e; ++s) {
                        const double d2 = sqdist<DIM>(sx + (int64_t)s * DIM, q);
                        const int id = sidx[s];
                        if (!lt(d2, id, bd[KCAP - 1], bi[KCAP - 1])) continue;
                        // insertion position p = number of entries preceding the candidate
                        // (the list is sorted, so they are slots [0, p)); then shift right
                        int p = 0;
#pragma unroll
                        for (int u = 0; u < KCAP - 1; ++u) p += lt(bd[u], bi[u], d2, id) ? 1 : 0;
#pragma unroll
                        for (int u = KCAP - 1; u >= 1; --u) {
                            bd[u] = u > p ? bd[u - 1] : (u == p ? d2 : bd[u]);
                            bi[u] = u > p ? bi[u - 1] : (u == p ? id : bi[u]);
                        }
                        if (p == 0) { bd[0] = d2; bi[0] = id; }
                        ++filled;
                    }
                }
            }
        }
        // lower bound on the distance of any point outside the visited box
        double bound = 1e300;
        bool covers_all = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double blo = G.lo[d] + (double)(c[d] - rho) * G.s;
            const double bhi = G.lo[d] + (double)(c[d] + rho + 1) * G.s;
            if (c[d] - rho > 0) bound = fmin(bound, q[d] - blo);
            if (c[d] + rho + 1 < G.n[d]) bound = fmin(bound, bhi - q[d]);
            if (c[d] - rho > 0 || c[d] + rho + 1 < G.n[d]) covers_all = false;
        }
        if (covers_all || rho > maxn) break;
        if (filled >= kmax && bound > 0.0 && bd[KCAP - 1] < bound * bound * (1.0 - 1e-9)) break;
    }

    const int64_t base = q_ptr ? q_ptr[qi] : qi * (int64_t)kmax;
    const int kq = q_ptr ? (int)(q_ptr[qi + 1] - q_ptr[qi]) : kmax;
#pragma unroll
    for (int s = 0; s < KCAP; ++s) {
        const int o = s - (KCAP - kmax);
        if (o >= 0 && o < kq) out[base + o] = bi[s];
    }
}

// ---------------------------------------------------------------- driver
template <int DIM>
static hipError_t knn_dim(KnnWorkspace& ws, const double* data, int64_t n_data, const double* query,
                          int64_t n_query, int k, const int64_t* q_ptr, int32_t* out, hipStream_t st,
                          const double* bb /* host bbox: lo[3], hi[3] */)
{
    Grid G;
    G.dim = DIM;
    double ext[3] = {0, 0, 0};
    for (int d = 0; d < DIM; ++d) ext[d] = bb[3 + d] - bb[d];
    double s = 1.0;
    const int64_t ncell = knn_grid_size(DIM, ext, n_data, &s, G.n);
    G.s = s;
    G.inv_s = 1.0 / s;
    for (int d = 0; d < 3; ++d) G.lo[d] = (d < DIM) ? bb[d] : 0.0;

    hipError_t err;
    if ((err = ws.reserve(ncell, n_data, n_query, DIM)) != hipSuccess) return err;
    const int TB = 256;
    const int gridN = (int)std::min<int64_t>((n_data + TB - 1) / TB, 65535);
    const int gridQ = (int)std::min<int64_t>((n_query + TB - 1) / TB, 65535);

    // data binning
    const int gridC = (int)std::min<int64_t>((ncell + 1 + TB - 1) / TB, 65535);
    hipLaunchKernelGGL(k_fill_i32, dim3(gridC), dim3(TB), 0, st, ws.count, ncell + 1, 0);
    hipLaunchKernelGGL(k_cell_count<DIM>, dim3(gridN), dim3(TB), 0, st, data, n_data, G, ws.cell, ws.count);
    size_t tmp_bytes = ws.scan_tmp_bytes;
    if ((err = hipcub::DeviceScan::ExclusiveSum(ws.scan_tmp, tmp_bytes, ws.count, ws.start, (int)(ncell + 1), st)) !=
        hipSuccess)
        return err;
    if ((err = hipMemcpyAsync(ws.cursor, ws.start, sizeof(int32_t) * (ncell + 1), hipMemcpyDeviceToDevice, st)) !=
        hipSuccess)
        return err;
    hipLaunchKernelGGL(k_scatter<DIM>, dim3(gridN), dim3(TB), 0, st, data, n_data, ws.cell, ws.cursor,
                       ws.sidx, ws.sx);
    // query binning (spatially coherent wavefronts)
    hipLaunchKernelGGL(k_fill_i32, dim3(gridC), dim3(TB), 0, st, ws.qcount, ncell + 1, 0);
    hipLaunchKernelGGL(k_cell_count<DIM>, dim3(gridQ), dim3(TB), 0, st, query, n_query, G, ws.qcell, ws.qcount);
    tmp_bytes = ws.scan_tmp_bytes;
    if ((err = hipcub::DeviceScan::ExclusiveSum(ws.scan_tmp, tmp_bytes, ws.qcount, ws.qstart, (int)(ncell + 1),
                                                st)) != hipSuccess)
        return err;
    hipLaunchKernelGGL(k_scatter<DIM>, dim3(gridQ), dim3(TB), 0, st, query, n_query, ws.qcell, ws.qstart,
                       ws.qorder, (double*)nullptr);

    const int QB = 128;
    const dim3 gq((unsigned)((n_query + QB - 1) / QB));
    // the insertion network costs KCAP - 1 steps per accepted candidate, so KCAP
    // follows k closely in the common range
    if (k <= 16)
        hipLaunchKernelGGL((k_query<DIM, 16>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else if (k <= 20)
        hipLaunchKernelGGL((k_query<DIM, 20>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else if (k <= 24)
        hipLaunchKernelGGL((k_query<DIM, 24>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else if (k <= 32)
        hipLaunchKernelGGL((k_query<DIM, 32>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else if (k <= 40)
        hipLaunchKernelGGL((k_query<DIM, 40>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else if (k <= 48)
        hipLaunchKernelGGL((k_query<DIM, 48>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    else
        hipLaunchKernelGGL((k_query<DIM, 64>), gq, dim3(QB), 0, st, ws.sx, ws.sidx, ws.start, query, ws.qorder,
                           n_query, G, k, q_ptr, out);
    return hipGetLastError();
}

hipError_t KnnWorkspace::reserve(int64_t ncell, int64_t n_data, int64_t n_query, int dim)
{
    auto grow = [](void** p, size_t& cap, size_t need) -> hipError_t {
        if (need <= cap) return hipSuccess;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(p, need);
        if (e == hipSuccess) cap = need;
        return e;
    };
    hipError_t e;
    const size_t nc = (size_t)(ncell + 1);
    if ((e = grow((void**)&count, cap_count, nc * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&start, cap_start, nc * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&cursor, cap_cursor, nc * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&qcount, cap_qcount, nc * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&qstart, cap_qstart, nc * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&cell, cap_cell, (size_t)n_data * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&sidx, cap_sidx, (size_t)n_data * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&sx, cap_sx, (size_t)n_data * dim * 8)) != hipSuccess) return e;
    if ((e = grow((void**)&qcell, cap_qcell, (size_t)n_query * 4)) != hipSuccess) return e;
    if ((e = grow((void**)&qorder, cap_qorder, (size_t)n_query * 4)) != hipSuccess) return e;
    size_t need = 0;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, count, start, (int)nc)) != hipSuccess) return e;
    if ((e = grow(&scan_tmp, scan_tmp_bytes, need + 256)) != hipSuccess) return e;
    if ((e = grow((void**)&bbox, cap_bbox, 6 * 8 * 1024)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t run_knn(KnnWorkspace& ws, const double* data, int64_t n_data, const double* query, int64_t n_query,
                   int dim, int k, const int64_t* q_ptr, int32_t* out, hipStream_t st)
{
    hipError_t e;
    if ((e = ws.reserve(1, n_data, n_query, dim)) != hipSuccess) return e;
    // bounding box of data and queries together (queries outside the data box
    // still bin into boundary cells, so only the data box matters for grid size)
    const int nb = 1024;
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, data, n_data, dim, ws.bbox);
    static thread_local double hb[6 * 1024];
    if ((e = hipMemcpyAsync(hb, ws.bbox, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    double bb[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) {
            bb[d] = fmin(bb[d], hb[b * 6 + d]);
            bb[3 + d] = fmax(bb[3 + d], hb[b * 6 + 3 + d]);
        }
    switch (dim) {
    case 1: return knn_dim<1>(ws, data, n_data, query, n_query, k, q_ptr, out, st, bb);
    case 2: return knn_dim<2>(ws, data, n_data, query, n_query, k, q_ptr, out, st, bb);
    case 3: return knn_dim<3>(ws, data, n_data, query, n_query, k, q_ptr, out, st, bb);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rbfx
