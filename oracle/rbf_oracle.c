/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * CPU restatement of the RadialBasisFunctions.jl v0.7.0 RBF-FD weight build:
 * basis functions (src/basis/), monomial bases and the operators applied to
 * them (src/basis/monomial.jl, src/operators/monomial/), collocation/RHS
 * assembly and Hermite dispatch (src/solve/assembly.jl), orchestration and CSR
 * layout (src/solve/execution.jl, src/solve/types.jl) and find_neighbors
 * (src/utils.jl:12-18, restated as brute force).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker. Built with -ffp-contract=off so that no
 * multiply-add is fused, as Julia does not fuse without @fastmath/muladd.
 */
#include "rbf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* avoid_inf(Float64) = max(1e-16, floatmin(Float64))  src/RadialBasisFunctions.jl:111-114 */
#define AI 1.0e-16

static double sqeucl(int dim, const double* x, const double* y)
{
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = x[d] - y[d];
        s += t * t;
    }
    return s;
}

static double dotn(int dim, const double* n, const double* x, const double* xi)
{
    /* LinearAlgebra.dot(normal, x .- xi) */
    double s = 0.0;
    for (int d = 0; d < dim; ++d) s += n[d] * (x[d] - xi[d]);
    return s;
}

static int is_phs(int kind)
{
    return kind == RBFX_PHS1 || kind == RBFX_PHS3 || kind == RBFX_PHS5 || kind == RBFX_PHS7;
}

/* ---------------------------------------------------------------------------
 * basis values: polyharmonic_spline.jl:142,227,313,399; inverse_multiquadric.jl:33;
 * gaussian.jl:122
 * ------------------------------------------------------------------------- */
double orc_phi(const rbfx_basis* b, int dim, const double* x, const double* xi)
{
    double r2 = sqeucl(dim, x, xi);
    switch (b->kind) {
    case RBFX_PHS1: return sqrt(r2);
    case RBFX_PHS3: return r2 * sqrt(r2);
    case RBFX_PHS5: return (r2 * r2) * sqrt(r2);
    case RBFX_PHS7: return (r2 * r2 * r2) * sqrt(r2);
    case RBFX_IMQ: return 1.0 / sqrt((b->eps * b->eps) * r2 + 1.0);
    case RBFX_GAUSSIAN: return exp(-(b->eps * b->eps) * r2);
    }
    return NAN;
}

/* d/dx_d  (polyharmonic_spline.jl:146-149,231-234,317-320,403-406;
 *          inverse_multiquadric.jl:37-40; gaussian.jl:126-129) */
static double phi_d1(const rbfx_basis* b, int dim, int d, const double* x, const double* xi)
{
    double r2 = sqeucl(dim, x, xi), r = sqrt(r2);
    double dd = x[d] - xi[d];
    switch (b->kind) {
    case RBFX_PHS1: return dd / (r + AI);
    case RBFX_PHS3: return 3.0 * dd * r;
    case RBFX_PHS5: return 5.0 * dd * (r * r * r);
    case RBFX_PHS7: return 7.0 * dd * (r * r * r * r * r);
    case RBFX_IMQ: {
        double e2 = b->eps * b->eps, s = e2 * r2 + 1.0;
        return (xi[d] - x[d]) * (e2 / sqrt(s * s * s));
    }
    case RBFX_GAUSSIAN: {
        double e2 = b->eps * b->eps;
        return -2.0 * e2 * dd * exp(-e2 * r2);
    }
    }
    return NAN;
}

/* d^2/dx_d^2  (polyharmonic_spline.jl:186-190,272-275,358-362,444-448;
 *              inverse_multiquadric.jl:68-74; gaussian.jl:154-157) */
static double phi_d2(const rbfx_basis* b, int dim, int d, const double* x, const double* xi)
{
    double r2 = sqeucl(dim, x, xi), r = sqrt(r2);
    double dd = x[d] - xi[d];
    switch (b->kind) {
    case RBFX_PHS1: return (-(dd * dd) + r2) / (r * r * r + AI);
    case RBFX_PHS3: return 3.0 * (r + (dd * dd) / (r + AI));
    case RBFX_PHS5: return 5.0 * r * (3.0 * (dd * dd) + r2);
    case RBFX_PHS7: return 7.0 * (r * r * r) * (5.0 * (dd * dd) + r2);
    case RBFX_IMQ: {
        double e2 = b->eps * b->eps, e4 = e2 * e2;
        double num1 = 3.0 * e4 * (dd * dd);
        double s = e2 * r2 + 1.0;
        return num1 / sqrt(s * s * s * s * s) - e2 / sqrt(s * s * s);
    }
    case RBFX_GAUSSIAN: {
        double e2 = b->eps * b->eps;
        return (4.0 * (e2 * e2) * (dd * dd) - 2.0 * e2) * exp(-e2 * r2);
    }
    }
    return NAN;
}

/* Laplacian (polyharmonic_spline.jl:202-205,288-291,373-376,459-462 -- the
 * reference's constants 2, 12, 30, 56 are the 3-D values and are used in every
 * dimension; inverse_multiquadric.jl:77-83; gaussian.jl:160-163) */
static double phi_lap(const rbfx_basis* b, int dim, const double* x, const double* xi)
{
    double r2 = sqeucl(dim, x, xi), r = sqrt(r2);
    switch (b->kind) {
    case RBFX_PHS1: return 2.0 / (r + AI);
    case RBFX_PHS3: return 12.0 * r;
    case RBFX_PHS5: return 30.0 * (r * r * r);
    case RBFX_PHS7: return 56.0 * (r * r * r * r * r);
    case RBFX_IMQ: {
        double e2 = b->eps * b->eps, e4 = e2 * e2;
        double s = e2 * r2 + 1.0;
        double s5 = sqrt(s * s * s * s * s), s3 = sqrt(s * s * s);
        double acc = 0.0;
        for (int d = 0; d < dim; ++d) {
            double dd = x[d] - xi[d];
            acc += (3.0 * e4 * (dd * dd)) / s5 - e2 / s3;
        }
        return acc;
    }
    case RBFX_GAUSSIAN: {
        double e2 = b->eps * b->eps;
        double ph = exp(-e2 * r2);
        double acc = 0.0;
        for (int d = 0; d < dim; ++d) {
            double dd = x[d] - xi[d];
            acc += (4.0 * (e2 * e2) * (dd * dd) - 2.0 * e2) * ph;
        }
        return acc;
    }
    }
    return NAN;
}

/* gradient w.r.t. x (polyharmonic_spline.jl:159-162,245-248,331-334,417-420;
 * inverse_multiquadric.jl:43-46; gaussian.jl:132-135) */
void orc_grad(const rbfx_basis* b, int dim, const double* x, const double* xi, double* g)
{
    for (int d = 0; d < dim; ++d) g[d] = phi_d1(b, dim, d, x, xi);
}

/* Hessian w.r.t. x (polyharmonic_spline.jl:171-183,257-269,343-355,429-441;
 * inverse_multiquadric.jl:49-65; gaussian.jl:138-151) */
void orc_hess(const rbfx_basis* b, int dim, const double* x, const double* xi, double* h)
{
    double r2 = sqeucl(dim, x, xi), r = sqrt(r2);
    double D[3];
    for (int d = 0; d < dim; ++d) D[d] = x[d] - xi[d];
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            double dl = (i == j) ? 1.0 : 0.0, v = 0.0;
            switch (b->kind) {
            case RBFX_PHS1: v = dl / (r + AI) - D[i] * D[j] / (r * r * r + AI); break;
            case RBFX_PHS3: v = 3.0 * (dl * r + D[i] * D[j] / (r + AI)); break;
            case RBFX_PHS5: v = 5.0 * (dl * (r * r * r) + 3.0 * D[i] * D[j] * r); break;
            case RBFX_PHS7: v = 7.0 * (dl * (r * r * r * r * r) + 5.0 * D[i] * D[j] * (r * r * r)); break;
            case RBFX_IMQ: {
                double e2 = b->eps * b->eps, e4 = e2 * e2, s = e2 * r2 + 1.0;
                double s3 = sqrt(s * s * s), s5 = sqrt(s * s * s * s * s);
                v = 3.0 * e4 * D[i] * D[j] / s5 - e2 * dl / s3;
                break;
            }
            case RBFX_GAUSSIAN: {
                double e2 = b->eps * b->eps, ph = exp(-e2 * r2);
                v = (4.0 * (e2 * e2) * D[i] * D[j] - 2.0 * e2 * dl) * ph;
                break;
            }
            }
            h[i * dim + j] = v;
        }
}

double orc_lrbf(const rbfx_basis* b, int dim, int kind, int d1, int d2,
                const double* x, const double* xi)
{
    switch (kind) {
    case RBFX_OP_VALUE: return orc_phi(b, dim, x, xi);
    case RBFX_OP_D1: return phi_d1(b, dim, d1, x, xi);
    case RBFX_OP_D2: return phi_d2(b, dim, d1, x, xi);
    case RBFX_OP_MIXED: {
        /* basis.jl:87-92: equal dims -> d2, else Hessian entry */
        if (d1 == d2) return phi_d2(b, dim, d1, x, xi);
        double h[9];
        orc_hess(b, dim, x, xi, h);
        return h[d1 * dim + d2];
    }
    case RBFX_OP_LAP: return phi_lap(b, dim, x, xi);
    }
    return NAN;
}

/* Hermite normal forms, PHS only (polyharmonic_spline.jl:151-156,192-199,207-210,
 * 236-242,277-285,293-296,322-328,364-370,378-381,408-414,450-456,464-467). */
double orc_lrbf_normal(const rbfx_basis* b, int dim, int kind, int d1,
                       const double* x, const double* xi, const double* n, int* ok)
{
    *ok = 1;
    if (!is_phs(b->kind) || !(kind == RBFX_OP_D1 || kind == RBFX_OP_D2 || kind == RBFX_OP_LAP)) {
        *ok = 0;
        return 0.0;
    }
    double r2 = sqeucl(dim, x, xi), r = sqrt(r2);
    double dn = dotn(dim, n, x, xi);
    double nd = (kind == RBFX_OP_LAP) ? 0.0 : n[d1];
    double Dd = (kind == RBFX_OP_LAP) ? 0.0 : x[d1] - xi[d1];
    double r3 = r * r * r, r5 = r3 * r * r;
    switch (b->kind) {
    case RBFX_PHS1:
        if (kind == RBFX_OP_D1) return -nd / (r + AI) + dn * Dd / (r3 + AI);
        if (kind == RBFX_OP_D2) return (2.0 * nd * Dd + dn) / (r3 + AI) - 3.0 * (Dd * Dd) * dn / (r5 + AI);
        return 2.0 * dn / (r3 + AI);
    case RBFX_PHS3:
        if (kind == RBFX_OP_D1) return -3.0 * (nd * r + dn * Dd / (r + AI));
        if (kind == RBFX_OP_D2)
            return -3.0 * (2.0 * Dd * nd + dn - dn * (Dd * Dd) / (r2 + AI)) / (r + AI);
        return -12.0 * dn / (r + AI);
    case RBFX_PHS5:
        if (kind == RBFX_OP_D1) return -5.0 * (nd * r3 + 3.0 * dn * Dd * r);
        if (kind == RBFX_OP_D2) return -5.0 * (6.0 * nd * Dd * r + 3.0 * dn * (r + (Dd * Dd) / (r + AI)));
        return -90.0 * dn * r;
    case RBFX_PHS7:
        if (kind == RBFX_OP_D1) return -7.0 * (nd * r5 + 5.0 * r3 * Dd * dn);
        if (kind == RBFX_OP_D2) return -7.0 * (10.0 * nd * Dd * r3 + 5.0 * dn * (3.0 * r * (Dd * Dd) + r3));
        return -280.0 * dn * r3;
    }
    *ok = 0;
    return 0.0;
}

/* ---------------------------------------------------------------------------
 * monomials (src/basis/monomial.jl:6-123, src/operators/monomial/)
 * ------------------------------------------------------------------------- */
int orc_n_monomials(int dim, int deg)
{
    if (deg < 0) return 0;
    long num = 1, den = 1;
    for (int i = 1; i <= dim; ++i) {
        num *= (deg + i);
        den *= i;
    }
    return (int)(num / den);
}

static int hand_coded(int dim, int deg)
{
    /* monomial.jl:223-268 hand-coded evaluators */
    return (dim == 2 && (deg == 1 || deg == 2)) || (dim == 3 && (deg == 1 || deg == 2));
}

/* Combinatorics.multiexponents(dim+1, deg), first `dim` components; the
 * enumeration order is the one fixed by test/basis/monomial.jl:123-139:
 * descending in the first exponent, then the second, ... */
static int multiexponents(int dim, int deg, int (*e)[3])
{
    int cnt = 0, v[4];
    /* recursive enumeration over dim+1 slots with descending order */
    int m = dim + 1;
    int stack[4];
    for (int i = 0; i < 4; ++i) stack[i] = 0;
    /* simple nested loops for m <= 4 */
    (void)stack;
    if (m == 2) {
        for (int a = deg; a >= 0; --a) { e[cnt][0] = a; ++cnt; }
    } else if (m == 3) {
        for (int a = deg; a >= 0; --a)
            for (int b2 = deg - a; b2 >= 0; --b2) { e[cnt][0] = a; e[cnt][1] = b2; ++cnt; }
    } else if (m == 4) {
        for (int a = deg; a >= 0; --a)
            for (int b2 = deg - a; b2 >= 0; --b2)
                for (int c = deg - a - b2; c >= 0; --c) {
                    e[cnt][0] = a; e[cnt][1] = b2; e[cnt][2] = c; ++cnt;
                }
    }
    (void)v;
    return cnt;
}

/* generic evaluator: db .= 1; products over dims/levels; db .*= c
 * (monomial.jl:270-295, operators/monomial/monomial.jl:110-142) */
static void generic_apply(int dim, int deg, int order1, int dim1, int order2, int dim2,
                          const double* x, double* out)
{
    int e[64][3];
    memset(e, 0, sizeof(e));
    if (orc_n_monomials(dim, deg) > 64) { for (int j = 0; j < 64; ++j) out[j] = NAN; return; }
    int nt = multiexponents(dim, deg, e);
    double c[64];
    for (int j = 0; j < nt; ++j) c[j] = 1.0;
    /* ∂exponents! (monomial.jl:132-142), applied `order` times per dim */
    int passes[2][2] = {{order1, dim1}, {order2, dim2}};
    for (int p = 0; p < 2; ++p)
        for (int o = 0; o < passes[p][0]; ++o) {
            int d = passes[p][1];
            for (int j = 0; j < nt; ++j) {
                if (e[j][d] == 0) c[j] = 0.0;
                if (e[j][d] > 0) {
                    c[j] *= e[j][d];
                    e[j][d] -= 1;
                }
            }
        }
    for (int j = 0; j < nt; ++j) out[j] = 1.0;
    for (int i = 0; i < dim; ++i)
        for (int l = 1; l <= deg; ++l)
            for (int j = 0; j < nt; ++j)
                if (e[j][i] >= l) out[j] *= x[i];
    if (order1 + order2 > 0)
        for (int j = 0; j < nt; ++j) out[j] *= c[j];
}

static double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

void orc_mon(int dim, int deg, const double* x, double* out)
{
    if (deg < 0) return;
    if (deg == 0) { out[0] = 1.0; return; }
    if (dim == 1) {   /* monomial.jl:209-221 */
        out[0] = 1.0;
        for (int n = 1; n <= deg; ++n) out[n] = ipow(x[0], n);
        return;
    }
    if (dim == 2 && deg == 1) { out[0] = 1.0; out[1] = x[0]; out[2] = x[1]; return; }
    if (dim == 2 && deg == 2) {
        out[0] = 1.0; out[1] = x[0]; out[2] = x[1];
        out[3] = x[0] * x[1]; out[4] = x[0] * x[0]; out[5] = x[1] * x[1];
        return;
    }
    if (dim == 3 && deg == 1) { out[0] = 1.0; out[1] = x[0]; out[2] = x[1]; out[3] = x[2]; return; }
    if (dim == 3 && deg == 2) {
        out[0] = 1.0; out[1] = x[0]; out[2] = x[1]; out[3] = x[2];
        out[4] = x[0] * x[1]; out[5] = x[0] * x[2]; out[6] = x[1] * x[2];
        out[7] = x[0] * x[0]; out[8] = x[1] * x[1]; out[9] = x[2] * x[2];
        return;
    }
    generic_apply(dim, deg, 0, 0, 0, 0, x, out);
}

/* first derivative, hand-coded tables partial.jl:1-167 else generic */
static void mon_d1(int dim, int deg, int d, const double* x, double* out)
{
    int m = orc_n_monomials(dim, deg);
    if (deg == 0) { out[0] = 0.0; return; }
    if (dim == 1) {   /* partial.jl:14-28 */
        out[0] = 0.0; out[1] = 1.0;
        for (int q = 1; q <= deg - 1; ++q) out[q + 1] = (q + 1) * ipow(x[0], q);
        return;
    }
    if (hand_coded(dim, deg)) {
        for (int i = 0; i < m; ++i) out[i] = 0.0;
        out[1 + d] = 1.0;
        if (deg == 2 && dim == 2) {
            out[3] = (d == 0) ? x[1] : x[0];
            out[4 + d] = 2.0 * x[d];
        } else if (deg == 2 && dim == 3) {
            /* [1,x,y,z,xy,xz,yz,x2,y2,z2] */
            if (d == 0) { out[4] = x[1]; out[5] = x[2]; }
            if (d == 1) { out[4] = x[0]; out[6] = x[2]; }
            if (d == 2) { out[5] = x[0]; out[6] = x[1]; }
            out[7 + d] = 2.0 * x[d];
        }
        return;
    }
    generic_apply(dim, deg, 1, d, 0, 0, x, out);
}

static void mon_d2(int dim, int deg, int d, const double* x, double* out)
{
    int m = orc_n_monomials(dim, deg);
    if (deg == 0) { out[0] = 0.0; return; }
    if (dim == 1) {   /* partial.jl:175-200 */
        if (deg == 1) { out[0] = 0.0; out[1] = 0.0; return; }
        out[0] = 0.0; out[1] = 0.0; out[2] = 2.0;
        for (int q = 1; q <= deg - 2; ++q) out[q + 2] = (q + 1) * (q + 2) * ipow(x[0], q);
        return;
    }
    if (hand_coded(dim, deg)) {
        for (int i = 0; i < m; ++i) out[i] = 0.0;
        if (deg == 2) out[(dim == 2 ? 4 : 7) + d] = 2.0;
        return;
    }
    generic_apply(dim, deg, 2, d, 0, 0, x, out);
}

static void mon_mixed(int dim, int deg, int d1, int d2, const double* x, double* out)
{
    int m = orc_n_monomials(dim, deg);
    if (d1 == d2) { mon_d2(dim, deg, d1, x, out); return; }   /* monomial.jl:40-43 */
    if (dim == 2 && deg == 2) {   /* partial.jl:347-356 */
        for (int i = 0; i < m; ++i) out[i] = 0.0;
        out[3] = 1.0;
        return;
    }
    if (dim == 3 && deg == 2) {   /* partial.jl:358-369 */
        for (int i = 0; i < m; ++i) out[i] = 0.0;
        out[(d1 + 1) + (d2 + 1) + 2 - 1] = 1.0;
        return;
    }
    if (deg == 0) { out[0] = 0.0; return; }
    generic_apply(dim, deg, 1, d1, 1, d2, x, out);
}

void orc_lmon(int dim, int deg, int kind, int d1, int d2, const double* x, double* out)
{
    int m = orc_n_monomials(dim, deg);
    if (m == 0) return;
    switch (kind) {
    case RBFX_OP_VALUE: orc_mon(dim, deg, x, out); return;
    case RBFX_OP_D1: mon_d1(dim, deg, d1, x, out); return;
    case RBFX_OP_D2: mon_d2(dim, deg, d1, x, out); return;
    case RBFX_OP_MIXED: mon_mixed(dim, deg, d1, d2, x, out); return;
    case RBFX_OP_LAP: {   /* monomial.jl:91-103: b = 0; b += d2_d for d = 1..dim */
        double tmp[128];
        for (int i = 0; i < m; ++i) out[i] = 0.0;
        for (int d = 0; d < dim; ++d) {
            mon_d2(dim, deg, d, x, tmp);
            for (int i = 0; i < m; ++i) out[i] += tmp[i];
        }
        return;
    }
    }
}

/* ∂_normal! (partial.jl:377-402) */
void orc_mon_normal(int dim, int deg, const double* n, const double* x, double* out)
{
    int m = orc_n_monomials(dim, deg);
    double tmp[128];
    for (int i = 0; i < m; ++i) out[i] = 0.0;
    for (int d = 0; d < dim; ++d) {
        if (n[d] != 0.0) {
            mon_d1(dim, deg, d, x, tmp);
            for (int i = 0; i < m; ++i) out[i] += n[d] * tmp[i];
        }
    }
}

/* ---------------------------------------------------------------------------
 * Symmetric indefinite solve with rook pivoting.
 *
 * The reference solves each stencil with bunchkaufman!(Symmetric(A,:U), true)
 * (assembly.jl:220-222), i.e. LAPACK ?sytrf_rook / ?sytrs_rook with uplo='U',
 * which eliminates from the last column toward the first. This is a restatement
 * of that published algorithm (Ashcraft, Grimes & Lewis, SIAM J. Matrix Anal.
 * Appl. 20(2), 1998; pivot tests as in LAPACK dsytf2_rook), written on full
 * symmetric storage with the right-hand sides eliminated alongside.
 * ------------------------------------------------------------------------- */
#define DEFINE_SYM_SOLVE(NAME, REAL, FABS, SQRT)                                          \
    static int NAME(int n, int r, REAL* A, REAL* B, REAL* X)                              \
    {                                                                                     \
        const REAL alpha = ((REAL)1 + SQRT((REAL)17)) / (REAL)8;                          \
        int perm[128], blk[128];                                                          \
        for (int i = 0; i < n; ++i) { perm[i] = i; blk[i] = 0; }                          \
        int k = n - 1;                                                                    \
        while (k >= 0) {                                                                  \
            int kstep = 1, p = k, kp;                                                     \
            REAL absakk = FABS(A[k * n + k]);                                             \
            int imax = -1;                                                                \
            REAL colmax = 0;                                                              \
            for (int i = 0; i < k; ++i)                                                   \
                if (FABS(A[i * n + k]) > colmax) { colmax = FABS(A[i * n + k]); imax = i; } \
            if ((absakk > colmax ? absakk : colmax) == 0) return 1;                       \
            if (absakk >= alpha * colmax) {                                               \
                kp = k;                                                                   \
            } else {                                                                      \
                int guard = 0;                                                            \
                for (;;) {                                                                \
                    REAL rowmax = 0;                                                      \
                    int jmax = -1;                                                        \
                    for (int j = imax + 1; j <= k; ++j)                                   \
                        if (jmax < 0 || FABS(A[imax * n + j]) > rowmax) {                 \
                            rowmax = FABS(A[imax * n + j]); jmax = j;                     \
                        }                                                                 \
                    for (int i = 0; i < imax; ++i) {                                      \
                        REAL t = FABS(A[i * n + imax]);                                   \
                        if (t > rowmax) { rowmax = t; jmax = i; }                         \
                    }                                                                     \
                    if (!(FABS(A[imax * n + imax]) < alpha * rowmax)) { kp = imax; break; } \
                    if (p == jmax || rowmax <= colmax || ++guard > 4 * n) {               \
                        kp = imax; kstep = 2; break;                                      \
                    }                                                                     \
                    p = imax; colmax = rowmax; imax = jmax;                               \
                }                                                                         \
            }                                                                             \
            int kk = k - kstep + 1;                                                       \
            for (int s = 0; s < 2; ++s) {                                                 \
                int u, v;                                                                 \
                if (s == 0) { if (!(kstep == 2 && p != k)) continue; u = p; v = k; }      \
                else { if (kp == kk) continue; u = kp; v = kk; }                          \
                for (int j = 0; j < n; ++j) { REAL t = A[u * n + j]; A[u * n + j] = A[v * n + j]; A[v * n + j] = t; } \
                for (int i = 0; i < n; ++i) { REAL t = A[i * n + u]; A[i * n + u] = A[i * n + v]; A[i * n + v] = t; } \
                for (int c = 0; c < r; ++c) { REAL t = B[u * r + c]; B[u * r + c] = B[v * r + c]; B[v * r + c] = t; } \
                int t2 = perm[u]; perm[u] = perm[v]; perm[v] = t2;                        \
            }                                                                             \
            if (kstep == 1) {                                                             \
                REAL d = A[k * n + k];                                                    \
                if (d == 0) return 1;                                                     \
                for (int i = 0; i < k; ++i) {                                             \
                    REAL l = A[i * n + k] / d;                                            \
                    for (int j = 0; j < k; ++j) A[i * n + j] -= l * A[k * n + j];         \
                    for (int c = 0; c < r; ++c) B[i * r + c] -= l * B[k * r + c];         \
                }                                                                         \
                blk[k] = 1;                                                               \
            } else {                                                                      \
                REAL a = A[(k - 1) * n + (k - 1)], bb = A[(k - 1) * n + k], c2 = A[k * n + k]; \
                REAL det = a * c2 - bb * bb;                                              \
                if (det == 0) return 1;                                                   \
                for (int i = 0; i < k - 1; ++i) {                                         \
                    REAL u0 = A[i * n + (k - 1)], u1 = A[i * n + k];                      \
                    REAL l0 = (u0 * c2 - u1 * bb) / det;                                  \
                    REAL l1 = (u1 * a - u0 * bb) / det;                                   \
                    for (int j = 0; j < k - 1; ++j)                                       \
                        A[i * n + j] -= l0 * A[(k - 1) * n + j] + l1 * A[k * n + j];      \
                    for (int c = 0; c < r; ++c)                                           \
                        B[i * r + c] -= l0 * B[(k - 1) * r + c] + l1 * B[k * r + c];      \
                }                                                                         \
                blk[k - 1] = 2; blk[k] = 0;                                               \
            }                                                                             \
            k -= kstep;                                                                   \
        }                                                                                 \
        /* substitution in elimination-reverse order: positions 0..n-1 */                 \
        REAL x[128 * RBFX_MAX_COLS];                                                      \
        int pos = 0;                                                                      \
        while (pos < n) {                                                                 \
            if (blk[pos] == 1) {                                                          \
                for (int c = 0; c < r; ++c) {                                             \
                    REAL s = B[pos * r + c];                                              \
                    for (int j = 0; j < pos; ++j) s -= A[pos * n + j] * x[j * r + c];     \
                    x[pos * r + c] = s / A[pos * n + pos];                                \
                }                                                                         \
                pos += 1;                                                                 \
            } else {                                                                      \
                REAL a = A[pos * n + pos], bb = A[pos * n + pos + 1];                     \
                REAL c2 = A[(pos + 1) * n + pos + 1], det = a * c2 - bb * bb;             \
                for (int c = 0; c < r; ++c) {                                             \
                    REAL s0 = B[pos * r + c], s1 = B[(pos + 1) * r + c];                  \
                    for (int j = 0; j < pos; ++j) {                                       \
                        s0 -= A[pos * n + j] * x[j * r + c];                              \
                        s1 -= A[(pos + 1) * n + j] * x[j * r + c];                        \
                    }                                                                     \
                    x[pos * r + c] = (s0 * c2 - s1 * bb) / det;                           \
                    x[(pos + 1) * r + c] = (s1 * a - s0 * bb) / det;                      \
                }                                                                         \
                pos += 2;                                                                 \
            }                                                                             \
        }                                                                                 \
        for (int i = 0; i < n; ++i)                                                       \
            for (int c = 0; c < r; ++c) X[perm[i] * r + c] = x[i * r + c];                \
        return 0;                                                                         \
    }

DEFINE_SYM_SOLVE(sym_solve_d, double, fabs, sqrt)
DEFINE_SYM_SOLVE(sym_solve_ld, long double, fabsl, sqrtl)

int orc_sym_solve(int n, int r, double* A, double* B, double* X)
{
    if (n > 128 || r > RBFX_MAX_COLS) return 2;
    return sym_solve_d(n, r, A, B, X);
}

int orc_sym_solve_ld(int n, int r, const double* A, const double* B, double* X)
{
    if (n > 128 || r > RBFX_MAX_COLS) return 2;
    long double* AL = (long double*)malloc(sizeof(long double) * n * n);
    long double* BL = (long double*)malloc(sizeof(long double) * n * r);
    long double* XL = (long double*)malloc(sizeof(long double) * n * r);
    for (int i = 0; i < n * n; ++i) AL[i] = A[i];
    for (int i = 0; i < n * r; ++i) BL[i] = B[i];
    int rc = sym_solve_ld(n, r, AL, BL, XL);
    for (int i = 0; i < n * r; ++i) X[i] = (double)XL[i];
    free(AL); free(BL); free(XL);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Stencil assembly and orchestration
 * ------------------------------------------------------------------------- */

/* point types, types.jl:246-258 */
enum { PT_INTERIOR = 0, PT_DIRICHLET = 1, PT_NR = 2 };

typedef struct {
    int type;
    double alpha, beta;
    double n[3];
} pinfo;

static int is_dirichlet_bc(double a, double b) { return a == 1.0 && b == 0.0; }   /* types.jl:33 */

static pinfo point_info(const rbfx_build_args* A, int64_t g)
{
    pinfo p;
    p.type = PT_INTERIOR; p.alpha = 0; p.beta = 0; p.n[0] = p.n[1] = p.n[2] = 0;
    if (A->is_boundary && A->is_boundary[g]) {
        int64_t bi = A->bnd_index[g];
        p.alpha = A->bc_alpha[bi];
        p.beta = A->bc_beta[bi];
        for (int d = 0; d < A->dim; ++d) p.n[d] = A->normals[bi * A->dim + d];
        p.type = is_dirichlet_bc(p.alpha, p.beta) ? PT_DIRICHLET : PT_NR;
    }
    return p;
}

/* entry (i,j) of the collocation matrix (assembly.jl:95-115, 240-332) */
static double coll_entry(const rbfx_basis* b, int dim, const double* xi, const double* xj,
                         const pinfo* pi, const pinfo* pj, int hermite)
{
    double phi = orc_phi(b, dim, xi, xj);
    if (!hermite) return phi;
    int ti = pi->type, tj = pj->type;
    if (ti != PT_NR && tj != PT_NR) return phi;
    double g[3];
    orc_grad(b, dim, xi, xj, g);
    double Di = 0.0, Dj = 0.0;
    for (int d = 0; d < dim; ++d) { Di += pi->n[d] * g[d]; Dj += pj->n[d] * g[d]; }
    if (ti != PT_NR && tj == PT_NR) return pj->alpha * phi - pj->beta * Dj;   /* :261-268, :281-290 */
    if (ti == PT_NR && tj != PT_NR) return pi->alpha * phi + pi->beta * Di;   /* :293-300, :303-312 */
    /* NR-NR (:315-332): + beta_i beta_j n_i' H n_j as the reference writes it */
    double h[9], D2 = 0.0;
    orc_hess(b, dim, xi, xj, h);
    for (int a = 0; a < dim; ++a)
        for (int c = 0; c < dim; ++c) D2 += pi->n[a] * h[a * dim + c] * pj->n[c];
    return pi->alpha * pj->alpha * phi - pi->alpha * pj->beta * Dj + pi->beta * pj->alpha * Di +
           pi->beta * pj->beta * D2;
}

/* scale factor of an operator term of derivative order o in the radius-scaled frame */
static int term_order(int kind)
{
    switch (kind) {
    case RBFX_OP_VALUE: return 0;
    case RBFX_OP_D1: return 1;
    default: return 2;
    }
}

typedef struct {
    double* A;   /* n*n */
    double* B;   /* n*r */
    double* X;   /* n*r */
    double* xs;  /* k*dim local coords */
    pinfo* pinf; /* k */
    double* dumpA; /* optional: receives the assembled n*n matrix before the solve */
} workspace;

/* one stencil: returns status code */
static int build_stencil(const rbfx_build_args* a, int64_t e, int scaled, int use_ld,
                         workspace* ws, double* wout /* k*r */)
{
    const int dim = a->dim;
    const rbfx_basis* basis0 = &a->basis;
    int64_t s0 = a->adj_ptr ? a->adj_ptr[e] : e * (int64_t)a->k;
    int k = a->adj_ptr ? (int)(a->adj_ptr[e + 1] - s0) : a->k;
    const int32_t* nb = a->adj_idx + s0;
    const int m = orc_n_monomials(dim, basis0->poly_deg);
    const int n = k + m;
    const int r = a->op.n_cols;
    const double* xe = a->eval + e * dim;

    /* classify_stencil, types.jl:208-223. A Dirichlet eval node is classified before
     * the neighbour scan: count_nonzeros (execution.jl:39-69) gives it a one-entry row
     * whatever its neighbours are, and the reference's neighbour-first order would, for
     * a stencil holding no boundary point, fill k entries into that row (a write past
     * it under @inbounds). Every other stencil is classified as the reference does. */
    int anyb = 0;
    if (a->is_boundary)
        for (int i = 0; i < k; ++i) anyb |= a->is_boundary[nb[i]] ? 1 : 0;
    if (a->is_boundary && e < a->n_data && a->is_boundary[e]) {
        pinfo pe = point_info(a, e);
        if (pe.type == PT_DIRICHLET) return RBFX_ST_DIRICHLET;
    }
    int hermite = anyb;

    for (int i = 0; i < k; ++i) {
        if (hermite) {
            ws->pinf[i] = point_info(a, nb[i]);
        } else {
            memset(&ws->pinf[i], 0, sizeof(pinfo));
            ws->pinf[i].type = PT_INTERIOR;
        }
    }

    /* local frame */
    double h = 1.0;
    rbfx_basis b = *basis0;
    double xc[3] = {0, 0, 0};
    if (scaled) {
        double r2max = 0.0;
        for (int i = 0; i < k; ++i) {
            double t = sqeucl(dim, a->data + (int64_t)nb[i] * dim, xe);
            if (t > r2max) r2max = t;
        }
        h = sqrt(r2max);
        if (h == 0.0) h = 1.0;
        for (int d = 0; d < dim; ++d) xc[d] = xe[d];
        if (b.kind == RBFX_IMQ || b.kind == RBFX_GAUSSIAN) b.eps = b.eps * h;
    }
    for (int i = 0; i < k; ++i)
        for (int d = 0; d < dim; ++d) {
            double v = a->data[(int64_t)nb[i] * dim + d];
            ws->xs[i * dim + d] = scaled ? (v - xc[d]) / h : v;
        }
    double xel[3];
    for (int d = 0; d < dim; ++d) xel[d] = scaled ? (xe[d] - xc[d]) / h : xe[d];
    /* Neumann/Robin beta acts through a first derivative: beta/h in the scaled frame */
    if (scaled && hermite)
        for (int i = 0; i < k; ++i) ws->pinf[i].beta /= h;

    /* eval point local index for the F7 mono RHS (types.jl:151-154, assembly.jl:34-46) */
    int eval_local = -1;
    if (hermite) {
        for (int i = 0; i < k; ++i) {
            int eq = 1;
            for (int d = 0; d < dim; ++d) eq &= (a->data[(int64_t)nb[i] * dim + d] == xe[d]);
            if (eq) { eval_local = i; break; }
        }
        if (eval_local < 0) return RBFX_ST_EVAL_NOT_FOUND;
    }

    /* collocation matrix, full symmetric storage (the reference fills the upper
     * triangle and wraps it in Symmetric(A,:U), assembly.jl:102-112) */
    double* A = ws->A;
    memset(A, 0, sizeof(double) * n * n);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i <= j; ++i) {
            double v = coll_entry(&b, dim, ws->xs + i * dim, ws->xs + j * dim, &ws->pinf[i], &ws->pinf[j], hermite);
            A[i * n + j] = v;
            A[j * n + i] = v;
        }
    if (m > 0) {
        double pv[128], pn[128];
        for (int i = 0; i < k; ++i) {
            orc_mon(dim, b.poly_deg, ws->xs + i * dim, pv);
            if (hermite && ws->pinf[i].type == PT_NR) {   /* assembly.jl:362-386 */
                orc_mon_normal(dim, b.poly_deg, ws->pinf[i].n, ws->xs + i * dim, pn);
                for (int q = 0; q < m; ++q) pv[q] = ws->pinf[i].alpha * pv[q] + ws->pinf[i].beta * pn[q];
            }
            for (int q = 0; q < m; ++q) {
                A[i * n + k + q] = pv[q];
                A[(k + q) * n + i] = pv[q];
            }
        }
    }

    /* RHS (assembly.jl:127-185, 399-467) */
    double* B = ws->B;
    memset(B, 0, sizeof(double) * n * r);
    for (int t = 0; t < a->op.n_terms; ++t) {
        const rbfx_term* T = &a->op.terms[t];
        double sc = T->coeff;
        if (scaled) sc *= pow(h, -(double)term_order(T->kind));
        for (int i = 0; i < k; ++i) {
            const pinfo* pi = &ws->pinf[i];
            double v = orc_lrbf(&b, dim, T->kind, T->dim1, T->dim2, xel, ws->xs + i * dim);
            if (hermite && pi->type == PT_NR) {
                int ok = 1;
                double vn = orc_lrbf_normal(&b, dim, T->kind, T->dim1, xel, ws->xs + i * dim, pi->n, &ok);
                if (!ok) return -RBFX_EUNSUPPORTED;
                v = pi->alpha * v + pi->beta * vn;
            }
            B[i * r + T->col] += sc * v;
        }
    }
    if (m > 0) {
        double lm[128];
        if (hermite && ws->pinf[eval_local].type == PT_NR) {
            /* F7: the BC operator applied to P at a Neumann/Robin eval node,
             * the same for every operator column (assembly.jl:452-467) */
            double pv[128], pn[128];
            const pinfo* pe = &ws->pinf[eval_local];
            orc_mon(dim, b.poly_deg, xel, pv);
            orc_mon_normal(dim, b.poly_deg, pe->n, xel, pn);
            for (int c = 0; c < r; ++c) {
                /* sum-of-terms columns still get one copy per column */
                for (int q = 0; q < m; ++q) B[(k + q) * r + c] = pe->alpha * pv[q] + pe->beta * pn[q];
            }
        } else {
            for (int t = 0; t < a->op.n_terms; ++t) {
                const rbfx_term* T = &a->op.terms[t];
                double sc = T->coeff;
                if (scaled) sc *= pow(h, -(double)term_order(T->kind));
                orc_lmon(dim, b.poly_deg, T->kind, T->dim1, T->dim2, xel, lm);
                for (int q = 0; q < m; ++q) B[(k + q) * r + T->col] += sc * lm[q];
            }
        }
    }

    if (ws->dumpA) memcpy(ws->dumpA, A, sizeof(double) * n * n);
    int rc;
    if (use_ld) rc = orc_sym_solve_ld(n, r, A, B, ws->X);
    else rc = orc_sym_solve(n, r, A, B, ws->X);
    if (rc) return RBFX_ST_SINGULAR;
    for (int i = 0; i < k; ++i)
        for (int c = 0; c < r; ++c) wout[i * r + c] = ws->X[i * r + c];
    return RBFX_ST_OK;
}

int64_t orc_count_nonzeros(int64_t n_eval, const int64_t* adj_ptr, int k,
                           const uint8_t* is_boundary, const int32_t* bnd_index,
                           const double* bc_alpha, const double* bc_beta, int64_t* out_ptr)
{
    int64_t tot = 0;
    out_ptr[0] = 0;
    for (int64_t e = 0; e < n_eval; ++e) {
        int64_t len = adj_ptr ? adj_ptr[e + 1] - adj_ptr[e] : k;
        if (is_boundary && is_boundary[e]) {
            int64_t bi = bnd_index[e];
            if (is_dirichlet_bc(bc_alpha[bi], bc_beta[bi])) len = 1;
        }
        tot += len;
        out_ptr[e + 1] = tot;
    }
    return tot;
}

int orc_build(const rbfx_build_args* a, int scaled, int long_double, int n_threads)
{
    const int dim = a->dim;
    const int m = orc_n_monomials(dim, a->basis.poly_deg);
    const int r = a->op.n_cols;
    if (dim < 1 || dim > 3 || r < 1 || r > RBFX_MAX_COLS) return RBFX_EINVAL;
    int kmax = a->k;
    if (a->adj_ptr) {
        kmax = 0;
        for (int64_t e = 0; e < a->n_eval; ++e) {
            int kk = (int)(a->adj_ptr[e + 1] - a->adj_ptr[e]);
            if (kk > kmax) kmax = kk;
        }
    }
    const int nmax = kmax + m;
    if (nmax > 128) return RBFX_EUNSUPPORTED;
    int err = 0;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        workspace ws;
        ws.A = (double*)malloc(sizeof(double) * nmax * nmax);
        ws.B = (double*)malloc(sizeof(double) * nmax * r);
        ws.X = (double*)malloc(sizeof(double) * nmax * r);
        ws.xs = (double*)malloc(sizeof(double) * kmax * dim + 8);
        ws.pinf = (pinfo*)malloc(sizeof(pinfo) * (kmax + 1));
        ws.dumpA = NULL;
        double* w = (double*)malloc(sizeof(double) * kmax * r + 8);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t e = 0; e < a->n_eval; ++e) {
            int st = build_stencil(a, e, scaled, long_double, &ws, w);
            int64_t o0 = a->out_ptr ? a->out_ptr[e] : e * (int64_t)a->k;
            int64_t s0 = a->adj_ptr ? a->adj_ptr[e] : e * (int64_t)a->k;
            int k = a->adj_ptr ? (int)(a->adj_ptr[e + 1] - s0) : a->k;
            if (st < 0) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                err = -st;
                st = RBFX_ST_SINGULAR;
            }
            if (a->status) a->status[e] = st;
            if (st == RBFX_ST_DIRICHLET) {   /* fill_dirichlet_entry!, execution.jl:297-303 */
                if (a->out_col) a->out_col[o0] = (int32_t)e;
                for (int c = 0; c < r; ++c) a->out_val[c * a->out_nnz + o0] = 1.0;
            } else {
                for (int i = 0; i < k; ++i) {   /* fill_entries!, execution.jl:279-294 */
                    if (a->out_col) a->out_col[o0 + i] = a->adj_idx[s0 + i];
                    for (int c = 0; c < r; ++c)
                        a->out_val[c * a->out_nnz + o0 + i] = (st == RBFX_ST_OK) ? w[i * r + c] : NAN;
                }
            }
        }
        free(ws.A); free(ws.B); free(ws.X); free(ws.xs); free(ws.pinf); free(w);
    }
    return err;
}

/* The assembled system of stencil e (the matrix build_stencil factorises), for
 * conditioning diagnostics: writes n*n doubles to A_out, returns n, or 0 for an
 * identity (Dirichlet) row, or -1 on error.                                        */
int orc_assemble(const rbfx_build_args* a, int64_t e, int scaled, double* A_out)
{
    const int m = orc_n_monomials(a->dim, a->basis.poly_deg);
    const int r = a->op.n_cols;
    int k = a->adj_ptr ? (int)(a->adj_ptr[e + 1] - a->adj_ptr[e]) : a->k;
    const int n = k + m;
    if (n > 128) return -1;
    workspace ws;
    ws.A = (double*)malloc(sizeof(double) * n * n);
    ws.B = (double*)malloc(sizeof(double) * n * r);
    ws.X = (double*)malloc(sizeof(double) * n * r);
    ws.xs = (double*)malloc(sizeof(double) * k * a->dim + 8);
    ws.pinf = (pinfo*)malloc(sizeof(pinfo) * (k + 1));
    ws.dumpA = A_out;
    double* w = (double*)malloc(sizeof(double) * k * r + 8);
    int st = build_stencil(a, e, scaled, 0, &ws, w);
    free(ws.A); free(ws.B); free(ws.X); free(ws.xs); free(ws.pinf); free(w);
    if (st == RBFX_ST_DIRICHLET) return 0;
    return n;
}

/* ---------------------------------------------------------------------------
 * brute-force k-NN (restates find_neighbors' result, utils.jl:12-18)
 * ------------------------------------------------------------------------- */
int orc_knn(const double* data, int64_t n_data, const double* query, int64_t n_query,
            int dim, int k, int32_t* out_idx, int n_threads)
{
    if (k > n_data || k < 1) return RBFX_EINVAL;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        double* bd = (double*)malloc(sizeof(double) * k);
        int32_t* bi = (int32_t*)malloc(sizeof(int32_t) * k);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (int64_t q = 0; q < n_query; ++q) {
            int cnt = 0;
            const double* xq = query + q * dim;
            for (int64_t j = 0; j < n_data; ++j) {
                double d2 = sqeucl(dim, data + j * dim, xq);
                if (cnt == k && !(d2 < bd[k - 1])) continue;   /* (d2, j) ordering: j increasing */
                int pos = (cnt < k) ? cnt++ : k - 1;
                while (pos > 0 && d2 < bd[pos - 1]) {
                    bd[pos] = bd[pos - 1];
                    bi[pos] = bi[pos - 1];
                    --pos;
                }
                bd[pos] = d2;
                bi[pos] = (int32_t)j;
            }
            for (int i = 0; i < k; ++i) out_idx[q * k + i] = bi[i];
        }
        free(bd); free(bi);
    }
    return 0;
}

void orc_spmv(int64_t n_rows, const int64_t* row_ptr, const int32_t* col, const double* val,
              int F, const double* x, double* y, int n_threads)
{
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for num_threads(n_threads) schedule(static)
#endif
    for (int64_t i = 0; i < n_rows; ++i) {
        for (int f = 0; f < F; ++f) {
            double s = 0.0;
            for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) s += val[e] * x[(int64_t)col[e] * F + f];
            y[i * F + f] = s;
        }
    }
}

/* struct sizes, so tests can check the ctypes mirror of rbfx.h */
int64_t orc_sizeof_build_args(void) { return (int64_t)sizeof(rbfx_build_args); }
int64_t orc_sizeof_spmv_args(void) { return (int64_t)sizeof(rbfx_spmv_args); }
int64_t orc_sizeof_operator(void) { return (int64_t)sizeof(rbfx_operator); }
int64_t orc_sizeof_mlp_model(void) { return (int64_t)sizeof(rbfx_mlp_model); }
int64_t orc_sizeof_mlp_run_args(void) { return (int64_t)sizeof(rbfx_mlp_run_args); }
