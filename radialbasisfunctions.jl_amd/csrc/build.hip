// Dispatch of the stencil build kernel over factorisation precision and system size,
// and the host-side monomial exponent table the kernel reads.
#include "build.h"

namespace rbfx {

hipError_t launch_build_d32(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);
hipError_t launch_build_d64(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);
hipError_t launch_build_f32(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);
hipError_t launch_build_f64(const BuildDev& P, int dim, int64_t n_work, hipStream_t s);

// monomials of total degree <= deg in graded order (any order spans the same space;
// the stencil weights do not depend on it)
int fill_mono_table(int dim, int deg, int (*e)[3])
{
    int c = 0;
    for (int tot = 0; tot <= deg; ++tot) {
        if (dim == 1) {
            e[c][0] = tot; e[c][1] = 0; e[c][2] = 0; ++c;
        } else if (dim == 2) {
            for (int a = tot; a >= 0; --a) { e[c][0] = a; e[c][1] = tot - a; e[c][2] = 0; ++c; }
        } else {
            for (int a = tot; a >= 0; --a)
                for (int b = tot - a; b >= 0; --b) { e[c][0] = a; e[c][1] = b; e[c][2] = tot - a - b; ++c; }
        }
    }
    return c;
}

// One thread per (work position, neighbour): the dependent reads of the gather (work
// list or order -> adjacency -> coordinates) run here at full occupancy, and the build
// kernel then reads each stencil's offsets with one independent, coalesced load per
// lane. The subtraction is the build kernel's own, so the offsets are bit-identical.
template <int DIM>
__global__ __launch_bounds__(256) void k_pack_offsets(const BuildDev P, int64_t n_bound, double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int k = P.k;
    const int64_t wi = t / k;
    const int pi = (int)(t - wi * k);
    if (wi >= n_bound) return;
    const int64_t n_work = P.list ? (int64_t)*P.list_len : P.n_eval;
    if (wi >= n_work) return;
    const int64_t e = P.list ? (int64_t)P.list[wi] : (P.order ? (int64_t)P.order[wi] : wi);
    if (e < 0 || e >= P.n_eval) return;  // the build kernel reports nothing for it either
    const int32_t gidx = P.adj_idx[e * k + pi];
    double* o = out + (wi * k + pi) * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) o[d] = P.data[(int64_t)gidx * DIM + d] - P.eval[e * DIM + d];
    if (P.out_col) P.out_col[e * k + pi] = gidx;
}

hipError_t launch_pack_offsets(const BuildDev& P, int dim, int64_t n_work, double* out, hipStream_t s)
{
    const int64_t n = n_work * (int64_t)P.k;
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (dim == 1) hipLaunchKernelGGL(k_pack_offsets<1>, grid, dim3(256), 0, s, P, n_work, out);
    else if (dim == 2) hipLaunchKernelGGL(k_pack_offsets<2>, grid, dim3(256), 0, s, P, n_work, out);
    else hipLaunchKernelGGL(k_pack_offsets<3>, grid, dim3(256), 0, s, P, n_work, out);
    return hipGetLastError();
}

hipError_t launch_build(const BuildDev& P, int dim, int nmax, bool fp32, int64_t n_work, hipStream_t s)
{
    if (fp32) return nmax <= 32 ? launch_build_f32(P, dim, n_work, s) : launch_build_f64(P, dim, n_work, s);
    return nmax <= 32 ? launch_build_d32(P, dim, n_work, s) : launch_build_d64(P, dim, n_work, s);
}

}  // namespace rbfx
