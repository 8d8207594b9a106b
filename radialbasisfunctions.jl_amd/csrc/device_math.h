// Device-side basis and monomial math for the RBF-FD weight build (gfx950).
//
// Every stencil is assembled in a centred, radius-scaled frame: xi = (x - x_e)/h with
// h the stencil radius. Polyharmonic splines are homogeneous, so phi(h r) = h^p phi(r)
// and the factor h^p cancels into the Lagrange multipliers; IMQ and Gaussian keep
// phi exactly by using eps*h. A term of derivative order o picks up h^-o in the RHS.
// The formulas restate the reference's closed forms:
//   src/basis/polyharmonic_spline.jl:142-467 (PHS1/3/5/7, including the 3-D Laplacian
//   constants 2/r, 12r, 30r^3, 56r^5 used in every dimension, and the Hermite normal
//   forms), src/basis/inverse_multiquadric.jl:33-83, src/basis/gaussian.jl:122-163,
//   src/basis/basis.jl:87-92 (mixed partials), src/RadialBasisFunctions.jl:111-114
//   (avoid_inf = 1e-16).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbfx.h"

namespace rbfx {

constexpr double kAvoidInf = 1.0e-16;

struct BasisDev {
    int kind;
    double eps;  // already multiplied by h for IMQ / Gaussian
};

__device__ __forceinline__ bool is_phs(int kind)
{
    return kind == RBFX_PHS1 || kind == RBFX_PHS3 || kind == RBFX_PHS5 || kind == RBFX_PHS7;
}

// phi as a function of r^2
__device__ __forceinline__ double phi_r2(const BasisDev& b, double r2)
{
    switch (b.kind) {
    case RBFX_PHS1: return sqrt(r2);
    case RBFX_PHS3: return r2 * sqrt(r2);
    case RBFX_PHS5: return (r2 * r2) * sqrt(r2);
    case RBFX_PHS7: return (r2 * r2 * r2) * sqrt(r2);
    case RBFX_IMQ: return 1.0 / sqrt((b.eps * b.eps) * r2 + 1.0);
    default: return exp(-(b.eps * b.eps) * r2);
    }
}

// phi_r2 with the kind a compile-time constant (same arithmetic), for loops instantiated
// once per kind
template <int KIND>
__device__ __forceinline__ double phi_r2_k(double eps, double r2)
{
    if constexpr (KIND == RBFX_PHS1) return sqrt(r2);
    else if constexpr (KIND == RBFX_PHS3) return r2 * sqrt(r2);
    else if constexpr (KIND == RBFX_PHS5) return (r2 * r2) * sqrt(r2);
    else if constexpr (KIND == RBFX_PHS7) return (r2 * r2 * r2) * sqrt(r2);
    else if constexpr (KIND == RBFX_IMQ) return 1.0 / sqrt((eps * eps) * r2 + 1.0);
    else return exp(-(eps * eps) * r2);
}

// One operator term applied to phi at (x, xi), D = x - xi, r2 = |D|^2.
template <int DIM>
__device__ __forceinline__ double phi_term(const BasisDev& b, int kind, int d1, int d2,
                                           const double (&D)[DIM], double r2)
{
    if (kind == RBFX_OP_VALUE) return phi_r2(b, r2);
    const double r = sqrt(r2);
    double dd1 = 0.0, dd2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (d == d1) dd1 = D[d];
        if (d == d2) dd2 = D[d];
    }
    if (kind == RBFX_OP_MIXED && d1 == d2) kind = RBFX_OP_D2;
    switch (b.kind) {
    case RBFX_PHS1:
        if (kind == RBFX_OP_D1) return dd1 / (r + kAvoidInf);
        if (kind == RBFX_OP_D2) return (-(dd1 * dd1) + r2) / (r * r * r + kAvoidInf);
        if (kind == RBFX_OP_MIXED) return -dd1 * dd2 / (r * r * r + kAvoidInf);
        return 2.0 / (r + kAvoidInf);
    case RBFX_PHS3:
        if (kind == RBFX_OP_D1) return 3.0 * dd1 * r;
        if (kind == RBFX_OP_D2) return 3.0 * (r + (dd1 * dd1) / (r + kAvoidInf));
        if (kind == RBFX_OP_MIXED) return 3.0 * (dd1 * dd2 / (r + kAvoidInf));
        return 12.0 * r;
    case RBFX_PHS5:
        if (kind == RBFX_OP_D1) return 5.0 * dd1 * (r * r * r);
        if (kind == RBFX_OP_D2) return 5.0 * r * (3.0 * (dd1 * dd1) + r2);
        if (kind == RBFX_OP_MIXED) return 5.0 * (3.0 * dd1 * dd2 * r);
        return 30.0 * (r * r * r);
    case RBFX_PHS7:
        if (kind == RBFX_OP_D1) return 7.0 * dd1 * (r * r * r * r * r);
        if (kind == RBFX_OP_D2) return 7.0 * (r * r * r) * (5.0 * (dd1 * dd1) + r2);
        if (kind == RBFX_OP_MIXED) return 7.0 * (5.0 * dd1 * dd2 * (r * r * r));
        return 56.0 * (r * r * r * r * r);
    case RBFX_IMQ: {
        const double e2 = b.eps * b.eps, e4 = e2 * e2, s = e2 * r2 + 1.0;
        const double s3 = sqrt(s * s * s);
        if (kind == RBFX_OP_D1) return (-dd1) * (e2 / s3);
        const double s5 = sqrt(s * s * s * s * s);
        if (kind == RBFX_OP_D2) return (3.0 * e4 * (dd1 * dd1)) / s5 - e2 / s3;
        if (kind == RBFX_OP_MIXED) return 3.0 * e4 * dd1 * dd2 / s5;
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc += (3.0 * e4 * (D[d] * D[d])) / s5 - e2 / s3;
        return acc;
    }
    default: {  // Gaussian
        const double e2 = b.eps * b.eps, ph = exp(-e2 * r2);
        if (kind == RBFX_OP_D1) return -2.0 * e2 * dd1 * ph;
        if (kind == RBFX_OP_D2) return (4.0 * (e2 * e2) * (dd1 * dd1) - 2.0 * e2) * ph;
        if (kind == RBFX_OP_MIXED) return 4.0 * (e2 * e2) * dd1 * dd2 * ph;
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc += (4.0 * (e2 * e2) * (D[d] * D[d]) - 2.0 * e2) * ph;
        return acc;
    }
    }
}

// Hermite normal form n . grad_{xi} L phi (PHS with D1, D2, LAP only).
template <int DIM>
__device__ __forceinline__ double phi_term_normal(const BasisDev& b, int kind, int d1,
                                                  const double (&D)[DIM], double r2,
                                                  const double (&n)[DIM])
{
    const double r = sqrt(r2), r3 = r * r * r, r5 = r3 * r * r;
    double dn = 0.0, nd = 0.0, Dd = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        dn += n[d] * D[d];
        if (d == d1) { nd = n[d]; Dd = D[d]; }
    }
    if (kind == RBFX_OP_LAP) { nd = 0.0; Dd = 0.0; }
    switch (b.kind) {
    case RBFX_PHS1:
        if (kind == RBFX_OP_D1) return -nd / (r + kAvoidInf) + dn * Dd / (r3 + kAvoidInf);
        if (kind == RBFX_OP_D2)
            return (2.0 * nd * Dd + dn) / (r3 + kAvoidInf) - 3.0 * (Dd * Dd) * dn / (r5 + kAvoidInf);
        return 2.0 * dn / (r3 + kAvoidInf);
    case RBFX_PHS3:
        if (kind == RBFX_OP_D1) return -3.0 * (nd * r + dn * Dd / (r + kAvoidInf));
        if (kind == RBFX_OP_D2)
            return -3.0 * (2.0 * Dd * nd + dn - dn * (Dd * Dd) / (r2 + kAvoidInf)) / (r + kAvoidInf);
        return -12.0 * dn / (r + kAvoidInf);
    case RBFX_PHS5:
        if (kind == RBFX_OP_D1) return -5.0 * (nd * r3 + 3.0 * dn * Dd * r);
        if (kind == RBFX_OP_D2) return -5.0 * (6.0 * nd * Dd * r + 3.0 * dn * (r + (Dd * Dd) / (r + kAvoidInf)));
        return -90.0 * dn * r;
    default:  // PHS7
        if (kind == RBFX_OP_D1) return -7.0 * (nd * r5 + 5.0 * r3 * Dd * dn);
        if (kind == RBFX_OP_D2) return -7.0 * (10.0 * nd * Dd * r3 + 5.0 * dn * (3.0 * r * (Dd * Dd) + r3));
        return -280.0 * dn * r3;
    }
}

// n_i . grad_x phi and n_i' H n_j at D = x_i - x_j (Hermite collocation entries,
// src/solve/assembly.jl:240-332).
template <int DIM>
__device__ __forceinline__ void phi_grad_hess_dirs(const BasisDev& b, const double (&D)[DIM], double r2,
                                                   const double (&ni)[DIM], const double (&nj)[DIM],
                                                   double& Di, double& Dj, double& D2)
{
    const double r = sqrt(r2);
    double g = 0.0, hd = 0.0, ho = 0.0;  // grad = g*D ; H = hd*I + ho*D D'
    switch (b.kind) {
    case RBFX_PHS1: g = 1.0 / (r + kAvoidInf); hd = 1.0 / (r + kAvoidInf); ho = -1.0 / (r * r * r + kAvoidInf); break;
    case RBFX_PHS3: g = 3.0 * r; hd = 3.0 * r; ho = 3.0 / (r + kAvoidInf); break;
    case RBFX_PHS5: g = 5.0 * (r * r * r); hd = 5.0 * (r * r * r); ho = 15.0 * r; break;
    case RBFX_PHS7: g = 7.0 * (r * r * r * r * r); hd = 7.0 * (r * r * r * r * r); ho = 35.0 * (r * r * r); break;
    case RBFX_IMQ: {
        const double e2 = b.eps * b.eps, s = e2 * r2 + 1.0;
        const double s3 = sqrt(s * s * s), s5 = sqrt(s * s * s * s * s);
        g = -e2 / s3; hd = -e2 / s3; ho = 3.0 * e2 * e2 / s5;
        break;
    }
    default: {
        const double e2 = b.eps * b.eps, ph = exp(-e2 * r2);
        g = -2.0 * e2 * ph; hd = -2.0 * e2 * ph; ho = 4.0 * e2 * e2 * ph;
        break;
    }
    }
    double niD = 0.0, njD = 0.0, ninj = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        niD += ni[d] * D[d];
        njD += nj[d] * D[d];
        ninj += ni[d] * nj[d];
    }
    Di = g * niD;
    Dj = g * njD;
    D2 = hd * ninj + ho * niD * njD;
}

// ---------------------------------------------------------------------------
// Monomials of total degree <= DEG in DIM variables. Any ordering spans the same
// polynomial space, and the stencil weights do not depend on the ordering (only
// the discarded Lagrange multipliers do), so the device uses graded order.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int n_mono(int dim, int deg)
{
    if (deg < 0) return 0;
    int num = 1, den = 1;
    for (int i = 1; i <= dim; ++i) { num *= (deg + i); den *= i; }
    return num / den;
}

template <int DIM, int DEG>
struct MonoTable {
    static constexpr int M = n_mono(DIM, DEG);
    int e[M > 0 ? M : 1][DIM];
    constexpr MonoTable() : e{}
    {
        int c = 0;
        for (int tot = 0; tot <= DEG; ++tot) {
            if (DIM == 1) {
                e[c][0] = tot; ++c;
            } else if (DIM == 2) {
                for (int a = tot; a >= 0; --a) { e[c][0] = a; e[c][DIM - 1] = tot - a; ++c; }
            } else {
                for (int a = tot; a >= 0; --a)
                    for (int b = tot - a; b >= 0; --b) {
                        e[c][0] = a; e[c][1 % DIM] = b; e[c][DIM - 1] = tot - a - b; ++c;
                    }
            }
        }
    }
};

// powers pw[d][p] = x_d^p, p = 0..DEG
template <int DIM, int DEG>
__device__ __forceinline__ void mono_powers(const double (&x)[DIM], double (&pw)[DIM][(DEG > 0 ? DEG : 0) + 1])
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        pw[d][0] = 1.0;
#pragma unroll
        for (int p = 1; p <= DEG; ++p) pw[d][p] = pw[d][p - 1] * x[d];
    }
}

// Operator term `kind` (dims d1,d2) applied to monomial a, evaluated from powers.
template <int DIM, int DEG>
__device__ __forceinline__ double mono_term(int a_static_e0, const int (&ea)[DIM], int kind, int d1, int d2,
                                            const double (&pw)[DIM][(DEG > 0 ? DEG : 0) + 1])
{
    (void)a_static_e0;
    double prod = 1.0;
    if (kind == RBFX_OP_VALUE) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) prod *= pw[d][ea[d]];
        return prod;
    }
    if (kind == RBFX_OP_LAP) {
        double acc = 0.0;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
            if (ea[dd] >= 2) {
                double t = (double)(ea[dd] * (ea[dd] - 1));
#pragma unroll
                for (int d = 0; d < DIM; ++d) t *= (d == dd) ? pw[d][ea[d] - 2 < 0 ? 0 : ea[d] - 2] : pw[d][ea[d]];
                acc += t;
            }
        }
        return acc;
    }
    if (kind == RBFX_OP_MIXED && d1 == d2) kind = RBFX_OP_D2;
    int o[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) o[d] = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (kind == RBFX_OP_D1 && d == d1) o[d] = 1;
        if (kind == RBFX_OP_D2 && d == d1) o[d] = 2;
        if (kind == RBFX_OP_MIXED && (d == d1 || d == d2)) o[d] = 1;
    }
    double coef = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int ed = ea[d];
        if (o[d] > ed) return 0.0;
        if (o[d] == 1) coef *= (double)ed;
        if (o[d] == 2) coef *= (double)(ed * (ed - 1));
        int p = ed - o[d];
        prod *= pw[d][p < 0 ? 0 : p];
    }
    return coef * prod;
}

}  // namespace rbfx
