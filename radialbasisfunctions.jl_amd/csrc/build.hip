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

hipError_t launch_build(const BuildDev& P, int dim, int nmax, bool fp32, int64_t n_work, hipStream_t s)
{
    if (fp32) return nmax <= 32 ? launch_build_f32(P, dim, n_work, s) : launch_build_f64(P, dim, n_work, s);
    return nmax <= 32 ? launch_build_d32(P, dim, n_work, s) : launch_build_d64(P, dim, n_work, s);
}

}  // namespace rbfx
