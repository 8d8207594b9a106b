/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A CPU restatement of the RadialBasisFunctions.jl (v0.7.0) weight build, used only
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only
 * as the checker (never as the thing measured or shipped). Each function cites the
 * reference file:line it follows.
 */
#ifndef RBF_ORACLE_H
#define RBF_ORACLE_H

#include <stdint.h>
#include "../include/rbfx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* number of monomials binomial(dim+deg, dim); 0 when deg < 0 */
int orc_n_monomials(int dim, int deg);

/* basis value phi(x, xi) */
double orc_phi(const rbfx_basis* b, int dim, const double* x, const double* xi);
/* one operator term applied to the basis, (x, xi) form */
double orc_lrbf(const rbfx_basis* b, int dim, int kind, int d1, int d2,
                const double* x, const double* xi);
/* Hermite normal form (x, xi, n): n . grad_{xi} L phi. Returns 0 and sets *ok=0
 * where the reference has no such method (IMQ, Gaussian, identity, mixed). */
double orc_lrbf_normal(const rbfx_basis* b, int dim, int kind, int d1,
                       const double* x, const double* xi, const double* n, int* ok);
/* gradient (dim values) and Hessian (dim*dim, row-major) of phi w.r.t. x */
void orc_grad(const rbfx_basis* b, int dim, const double* x, const double* xi, double* g);
void orc_hess(const rbfx_basis* b, int dim, const double* x, const double* xi, double* h);

/* monomial basis and operators applied to it, reference orderings */
void orc_mon(int dim, int deg, const double* x, double* out);
void orc_lmon(int dim, int deg, int kind, int d1, int d2, const double* x, double* out);
void orc_mon_normal(int dim, int deg, const double* n, const double* x, double* out);

/* Symmetric indefinite solve A X = B (A n*n row-major full storage, B n*r
 * row-major), rook pivoting from the last column. Returns 0, or 1 when singular.
 * A and B are overwritten. */
int orc_sym_solve(int n, int r, double* A, double* B, double* X);
int orc_sym_solve_ld(int n, int r, const double* A, const double* B, double* X);

/* Full build as execution.jl:128-303. Host pointers inside args.
 * scaled = 0: the reference's arithmetic (uncentred, unscaled system);
 * scaled = 1: the centred, radius-scaled formulation the GPU kernels use.
 * long_double = 1 solves each stencil in 80-bit extended precision.
 * Returns 0, or an RBFX_E* code. n_threads <= 0 uses all OpenMP threads. */
int orc_build(const rbfx_build_args* a, int scaled, int long_double, int n_threads);

/* Brute-force k nearest neighbours (squared distance in dimension order, no FMA),
 * ties broken by lower index. out_idx n_query*k. */
int orc_knn(const double* data, int64_t n_data, const double* query, int64_t n_query,
            int dim, int k, int32_t* out_idx, int n_threads);

/* exact nnz count and row offsets (execution.jl:39-69), 0-based, ptr has n_eval+1 */
int64_t orc_count_nonzeros(int64_t n_eval, const int64_t* adj_ptr, int k,
                           const uint8_t* is_boundary, const int32_t* bnd_index,
                           const double* bc_alpha, const double* bc_beta, int64_t* out_ptr);

/* y = W x for one CSR value array with F fields, X point-major (n_cols x F) */
void orc_spmv(int64_t n_rows, const int64_t* row_ptr, const int32_t* col, const double* val,
              int F, const double* x, double* y, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
