/*
 * rbfx.h -- C ABI of the MI355X (gfx950) RBF-FD weight engine.
 *
 * The library replaces three call sites of RadialBasisFunctions.jl (v0.7.0):
 *
 *   rbfx_knn            find_neighbors(data, eval_points, k)         src/utils.jl:12-18
 *                       (KDTree + knn(..., sortres=true); adjl is 1-based in Julia,
 *                       0-based here)
 *   rbfx_build_weights  build_weights_kernel(data, eval_points, adjl, basis,
 *                       Lrbf, Lmon, mon, boundary_data; batch_size, device)
 *                                                                    src/solve/execution.jl:128-174
 *                       (the device gate at execution.jl:140-148 is where a Julia
 *                       caller routes a non-CPU `device` to this entry point)
 *   rbfx_spmv           _eval_op(op, x) = op.weights * x  and  mul!(y, op, x, a, b)
 *                                                                    src/operators/operators.jl:217-218, 529-542
 *
 * Conventions
 *   - Every pointer inside an args struct is a DEVICE pointer (HBM) unless noted.
 *   - Coordinates are FP64, array-of-structs (point-major), `dim` doubles per point:
 *     the memory layout of Julia's Vector{SVector{dim,Float64}}.
 *   - Indices are 0-based int32 (the Julia side subtracts 1).
 *   - Calls are blocking: the library synchronises its stream before returning,
 *     as KernelAbstractions.synchronize does at execution.jl:264.
 *   - Return value 0 = success; otherwise an RBFX_E* code, with a message from
 *     rbfx_last_error() (thread-local).
 *   - No exceptions cross this boundary. No allocator types appear in it.
 */
#ifndef RBFX_H
#define RBFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
enum {
    RBFX_OK = 0,
    RBFX_EINVAL = 1,     /* bad argument: maps to Julia ArgumentError           */
    RBFX_EDIM = 2,       /* size mismatch: maps to Julia DimensionMismatch      */
    RBFX_EDEVICE = 3,    /* HIP runtime failure                                 */
    RBFX_ESINGULAR = 4,  /* at least one stencil system was singular: maps to
                            LinearAlgebra.SingularException from bunchkaufman!  */
    RBFX_EUNSUPPORTED = 5
};

/* ---- basis kinds (src/basis/) -------------------------------------- */
enum {
    RBFX_PHS1 = 1,       /* phi = r    polyharmonic_spline.jl:142-143 */
    RBFX_PHS3 = 3,       /* phi = r^3  polyharmonic_spline.jl:227-228 */
    RBFX_PHS5 = 5,       /* phi = r^5  polyharmonic_spline.jl:313-314 */
    RBFX_PHS7 = 7,       /* phi = r^7  polyharmonic_spline.jl:399-400 */
    RBFX_IMQ = 10,       /* phi = 1/sqrt((eps r)^2+1)  inverse_multiquadric.jl:33-34 */
    RBFX_GAUSSIAN = 11   /* phi = exp(-(eps r)^2)      gaussian.jl:122-123 */
};

/* ---- operator term kinds ------------------------------------------------
 * The reference represents an operator L as a Julia callable applied to the
 * basis (api.jl:30-35). The closed set that its named operators and algebra
 * produce is lowered to a table of terms; one RHS column is the sum of its terms
 * (SumKernel/ScaledKernel, operator_algebra.jl:97-114).                        */
enum {
    RBFX_OP_VALUE = 0,   /* identity: Regrid()(b)=b regridding.jl:7, Identity operator_algebra.jl:83-84 */
    RBFX_OP_D1 = 1,      /* d/dx_dim1          partial.jl:53 -> basis d functor          */
    RBFX_OP_D2 = 2,      /* d^2/dx_dim1^2      partial.jl:53 -> basis d2 functor         */
    RBFX_OP_MIXED = 3,   /* d^2/dx_dim1 dx_dim2  mixed_partial.jl:393-400, basis.jl:87-92 */
    RBFX_OP_LAP = 4      /* laplacian          laplacian.jl:7 (PHS constants as in the
                            reference: 2/r, 12r, 30r^3, 56r^5 in every dimension)          */
};

/* ---- precision / method ------------------------------------------------- */
enum {
    RBFX_PREC_FP64 = 0,          /* plain method: FP64 assembly + pivoted LDL^T      */
    RBFX_PREC_FP32_REFINED = 1,  /* FP32 factorisation, FP64 residual + refinement,
                                    FP64 re-solve of stencils that do not converge  */
    RBFX_PREC_LEARNED = 2        /* MLP surrogate, FP64 residual acceptance,
                                    FP32_REFINED fallback (needs a loaded model)    */
};

/* ---- per-stencil status codes (written to args.status) ------------------ */
enum {
    RBFX_ST_OK = 0,              /* solved by the requested method                  */
    RBFX_ST_DIRICHLET = 1,       /* Dirichlet eval node: identity row               */
    RBFX_ST_REFINED = 2,         /* FP32 factorisation + FP64 refinement accepted   */
    RBFX_ST_FALLBACK64 = 3,      /* refinement did not converge: FP64 re-solve      */
    RBFX_ST_LEARNED = 4,         /* surrogate prediction accepted                   */
    RBFX_ST_FP64 = 5,            /* FP32_REFINED / LEARNED fallback: solved directly in
                                    FP64, the faster kernel for its size on this device */
    RBFX_ST_SINGULAR = 8,        /* zero pivot: system singular                     */
    RBFX_ST_EVAL_NOT_FOUND = 9,  /* Hermite stencil without its eval point
                                    (assembly.jl:35-46 raises ArgumentError)        */
    RBFX_ST_TOO_LARGE = 10,      /* n = k + m beyond the supported maximum          */
    RBFX_ST_INVALID = 11         /* per-stencil input out of range (degree map)     */
};

#define RBFX_MAX_TERMS 16
#define RBFX_MAX_COLS 8
#define RBFX_MAX_N 80            /* largest k + m: up to 64 rows factorised in one
                                    wavefront's registers, 65..80 with the system in
                                    LDS (within 64 KB at up to 8 operator columns)   */

typedef struct rbfx_basis {
    int32_t kind;        /* RBFX_PHS* / RBFX_IMQ / RBFX_GAUSSIAN             */
    int32_t poly_deg;    /* -1 disables the polynomial block (utils.jl:57-66) */
    double eps;          /* shape parameter (IMQ, Gaussian)                   */
} rbfx_basis;

typedef struct rbfx_term {
    int32_t col;         /* RHS column this term contributes to */
    int32_t kind;        /* RBFX_OP_*                            */
    int32_t dim1;        /* 0-based                              */
    int32_t dim2;        /* 0-based (MIXED only)                 */
    double coeff;        /* ScaledOperator alpha                 */
} rbfx_term;

typedef struct rbfx_operator {
    int32_t n_cols;      /* r: 1 for scalar operators, dim for Jacobian/gradient */
    int32_t n_terms;
    rbfx_term terms[RBFX_MAX_TERMS];
} rbfx_operator;

/* Arguments of rbfx_build_weights (execution.jl:128-174).
 * Hermite data mirrors BoundaryData (types.jl:178-182) plus the
 * global_to_boundary map (execution.jl:78-93), 0-based, -1 for interior.     */
struct rbfx_mlp_model;

typedef struct rbfx_build_args {
    int32_t dim;                 /* 1..3                                              */
    int32_t k;                   /* stencil size; the largest one when adj_ptr != NULL */
    int64_t n_data;
    int64_t n_eval;
    const double* data;          /* n_data*dim                                        */
    const double* eval;          /* n_eval*dim                                        */
    const int64_t* adj_ptr;      /* n_eval+1 offsets into adj_idx, or NULL (=i*k)     */
    const int32_t* adj_idx;      /* stencil members, nearest first                    */
    rbfx_basis basis;
    rbfx_operator op;
    /* Hermite boundary data (all NULL for interior-only builds) */
    const uint8_t* is_boundary;  /* n_data                                            */
    const int32_t* bnd_index;    /* n_data: index into bc arrays or -1                */
    const double* bc_alpha;      /* n_bnd                                             */
    const double* bc_beta;       /* n_bnd                                             */
    const double* normals;       /* n_bnd*dim                                         */
    int64_t n_bnd;               /* number of boundary points                         */
    /* output CSR (row pointers exact as count_nonzeros, execution.jl:39-69) */
    const int64_t* out_ptr;      /* n_eval+1, or NULL (=i*k)                          */
    int32_t* out_col;            /* nnz, or NULL to skip (fixed k: equals adj_idx)    */
    double* out_val;             /* n_cols * out_nnz, column c at out_val + c*out_nnz */
    int64_t out_nnz;
    int32_t* status;             /* n_eval per-stencil status, or NULL                */
    int32_t precision;           /* RBFX_PREC_*                                       */
    int32_t refine_steps;        /* FP32_REFINED: refinement updates (0 -> 3)         */
    double accept_tol;           /* relative residual tolerance (0 -> default)        */
    void* stream;                /* hipStream_t, NULL = default stream                */
    /* ABI 2 */
    double* out_lambda;          /* optional n_cols*n_eval*m: polynomial multipliers
                                    of the radius-scaled system (the reference keeps
                                    only the first k rows, assembly.jl:225-226)      */
    const struct rbfx_mlp_model* model;  /* RBFX_PREC_LEARNED: host struct          */
    const int8_t* poly_deg_map;  /* optional n_eval: polynomial degree of each stencil
                                    (-1..basis.poly_deg; no reference counterpart: one
                                    operator mixing degrees, e.g. 3 near a boundary) */
    const int32_t* order;        /* optional n_eval: the order in which stencils are
                                    solved (a permutation, e.g. rbfx_spatial_order of
                                    the eval points), so neighbouring wavefronts gather
                                    neighbouring coordinates; results do not depend on
                                    it. Entries outside 0..n_eval-1 are skipped.      */
} rbfx_build_args;

/* Surrogate for RBFX_PREC_LEARNED (no reference counterpart). A BF16 MLP
 * (4k -> 256 -> 256 -> 256 -> k+m, GELU, FP32 accumulation) maps the canonical,
 * radius-scaled stencil offsets to (w~, lambda) r1^2 of the scaled system (r1: the
 * distance of the canonical reference neighbour). The host
 * struct holds device pointers; the signature fields must match the build.     */
typedef struct rbfx_mlp_model {
    int32_t dim;                 /* 2                                                  */
    int32_t k;                   /* stencil size the network was trained for (20)      */
    int32_t basis_kind;          /* RBFX_PHS3                                          */
    int32_t poly_deg;            /* 2                                                  */
    int32_t op_kind;             /* RBFX_OP_LAP (rotation-invariant operators only)    */
    int32_t n_out;               /* k + m                                              */
    const float* params;         /* rbfx_mlp_param_count() FP32 master parameters      */
    const uint16_t* wbf16;       /* rbfx_mlp_bf16_count() BF16 weights                 */
    const double* out_mean;      /* n_out: output de-standardisation y = mean + std*z  */
    const double* out_std;
    double accept_tol;           /* accept when max|collocation residual| <=
                                    accept_tol * max|L phi| (FP64, per stencil)        */
    double* out_rho;             /* optional n_eval: the relative residual per stencil */
} rbfx_mlp_model;

/* One forward (and, with target != NULL, backward) pass of the surrogate on a batch. */
typedef struct rbfx_mlp_run_args {
    int32_t batch;               /* M stencils                                         */
    int32_t n_out;               /* outputs used by the loss (k + m <= 32)             */
    const uint16_t* x;           /* M x 128 BF16 inputs (rbfx_stencil_features)        */
    const float* target;         /* M x n_out standardised targets, or NULL            */
    const float* params;
    const uint16_t* wbf16;
    float* grads;                /* rbfx_mlp_param_count() FP32 (with target)           */
    float* loss;                 /* 1 FP32: batch mean squared error (with target)      */
    float* y_out;                /* optional M x 32 FP32 network outputs                */
    void* workspace;             /* rbfx_mlp_workspace_bytes(batch) bytes               */
    void* stream;
    const float* row_weight;     /* optional M per-stencil loss weights (1 if NULL)    */
} rbfx_mlp_run_args;

/* Arguments of rbfx_spmv: Y_v = alpha * W_v X + beta * Y_v for every value array v.
 * W_v share row pointers and column indices (one stencil set, several operators).  */
typedef struct rbfx_spmv_args {
    int64_t n_rows;
    int64_t n_cols;              /* rows of X (data points)                            */
    const int64_t* row_ptr;      /* n_rows+1, or NULL (=i*k)                           */
    int32_t k;                   /* fixed row length when row_ptr == NULL              */
    const int32_t* col_idx;
    const double* vals;          /* n_valarrays * nnz                                  */
    int64_t nnz;
    int32_t n_valarrays;
    int32_t n_fields;            /* F right-hand fields                                */
    const double* x;             /* X[p, f] at x[p*x_ld_point + f*x_ld_field]          */
    int64_t x_ld_point;
    int64_t x_ld_field;
    double* y;                   /* Y_v[i, f] at y[v*y_ld_val + i*y_ld_point + f*y_ld_field] */
    int64_t y_ld_point;
    int64_t y_ld_field;
    int64_t y_ld_val;
    double alpha;
    double beta;
    void* stream;
    const int32_t* row_order;    /* ABI 2, optional n_rows: the order in which rows are
                                    processed (e.g. from rbfx_spatial_order) so the X
                                    gathers of consecutive rows share cache lines; the
                                    result does not depend on it                        */
    int32_t rows_permuted;       /* ABI 2: nonzero when row_ptr / col_idx / vals are
                                    stored in row_order (rbfx_permute_rows)             */
    int64_t max_row_len;         /* ABI 2, optional: longest row (0 = unknown; k is used
                                    without row_ptr). Lets a workgroup stage its rows'
                                    entries through LDS; an underestimate is safe (those
                                    workgroups read HBM directly)                       */
    int32_t mapping;             /* ABI 2: 0 automatic, 1 lanes per row reading HBM,
                                    2 staged through LDS (needs rows stored in processing
                                    order), 3 vector loads per lane (needs rows of one
                                    length, no row_ptr, fewer than 8 fields), 4 source
                                    tiles (ABI 3: needs the plan below, 8+ fields, one
                                    value array);
                                    RBFX_EUNSUPPORTED when the mapping does not serve   */
    /* ABI 3, optional: a source-tile plan from rbfx_source_tiles for these rows as
       stored (all NULL / 0 without one). Tile t holds rows [t*tile_rows,
       (t+1)*tile_rows); its distinct columns are tile_src[tile_src_off[t] ..
       + tile_src_n[t]) and entry q's column is tile_src[tile_src_off[t] + tile_col[q]].
       With mapping 4, an apply of 8 or more fields and one value array then copies
       each tile's X rows into LDS once (not the automatic choice: slower than the
       staged mapping on the applies measured).                                        */
    int32_t tile_rows;
    const int32_t* tile_src;
    const int64_t* tile_src_off;
    const int32_t* tile_src_n;
    const uint16_t* tile_col;
} rbfx_spmv_args;

/* ---- entry points -------------------------------------------------------- */

/* Select the HIP device for this thread and check it is a gfx950 part. */
int rbfx_init(int32_t device);

/* Message describing the last failure on this thread ("" if none). */
const char* rbfx_last_error(void);

/* ABI version: bumped whenever a struct above changes layout. */
int32_t rbfx_abi_version(void);

/* find_neighbors (utils.jl:12-18): for each query, the k nearest data points by
 * Euclidean distance, ascending, ties broken by lower index. out_idx is
 * n_query*k int32. With q_ptr != NULL, query i receives q_ptr[i+1]-q_ptr[i] <= k
 * neighbours at out_idx + q_ptr[i] (variable-size stencils).                    */
int rbfx_knn(const double* data, int64_t n_data, const double* query, int64_t n_query,
             int32_t dim, int32_t k, const int64_t* q_ptr, int32_t* out_idx,
             void* stream);

/* build_weights_kernel (execution.jl:128-174) and the per-stencil solve it drives
 * (_build_stencil!, assembly.jl:201-222; Hermite dispatch assembly.jl:240-467).  */
int rbfx_build_weights(const rbfx_build_args* args);

/* op.weights * X (operators.jl:217-218) and mul!(y, op, x, alpha, beta)
 * (operators.jl:536-542) for one or several value arrays sharing a pattern.      */
int rbfx_spmv(const rbfx_spmv_args* args);

/* Capturable variants (no reference counterpart): the same work, enqueued on
 * args->stream without synchronising, reading back or allocating, so a caller can
 * record build + apply in a HIP graph. The asynchronous build reports no status
 * counts and skips the host-side Hermite normal-form check; it needs one earlier
 * blocking rbfx_build_weights with the same sizes (which sizes the library's lists).
 * Its fallback launches cover every stencil and read the device counters.        */
int rbfx_build_weights_async(const rbfx_build_args* args);
int rbfx_spmv_async(const rbfx_spmv_args* args);

/* Build that allocates nothing (no reference counterpart), for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): the asynchronous build on args->stream with no earlier blocking
 * build needed. The null-space kernels' hand-back list lives in a static device area
 * of the library (one per GPU, 16 MiB, held until the process ends). Limits: plain
 * FP64 (RBFX_PREC_FP64), at most 4,194,304 stencils (RBFX_EINVAL beyond), and
 * stencils of varying size (adj_ptr or poly_deg_map) only when k + m <= 32
 * (RBFX_EUNSUPPORTED). No autotuning: the default or an already tuned kernel runs;
 * status words are written only to args->status when given. Calls that overlap on
 * one GPU share the area and race; serialise them (one stream, or events).       */
int rbfx_build_weights_noalloc(const rbfx_build_args* args);

/* Morton (Z-curve) order of n points (FP64 AoS, dim doubles each): out_order[r]
 * is the index of the r-th point along the curve. The device-side counterpart of
 * reorder_points! (src/utils.jl:45-55), used to order the apply's rows without
 * moving user data.                                                               */
int rbfx_spatial_order(const double* points, int64_t n, int32_t dim, int32_t* out_order, void* stream);

/* Apply plan for fixed-length rows (no reference counterpart): out row r = in row
 * row_order[r], for the column indices and every value array ([n_valarrays][nnz]),
 * so an apply with rows_permuted streams its weights while walking row_order.     */
int rbfx_permute_rows(int64_t n_rows, int32_t k, const int32_t* row_order, const int32_t* col_idx,
                      const double* vals, int64_t nnz, int32_t n_valarrays, int32_t* out_col, double* out_vals,
                      void* stream);

/* Source-tile plan (no reference counterpart) for fixed-length rows stored back to back
 * (tile_rows * k <= 4096): per tile of tile_rows consecutive rows its distinct column
 * indices, sorted (tile_src, at tile_src_off[t], tile_src_n[t] of them), and each
 * entry's position in its tile's list (tile_col, nnz = n_rows * k entries). Blocking.
 * Writes tile_src_n, tile_src_off (n_tiles = ceil(n_rows / tile_rows) each) and
 * *n_src (their total); fills tile_src and tile_col only when tile_src != NULL and
 * src_capacity >= *n_src (a call with tile_src == NULL sizes tile_src).            */
int rbfx_source_tiles(int64_t n_rows, int32_t k, const int32_t* col_idx, int32_t tile_rows, int32_t* tile_src_n,
                      int64_t* tile_src_off, uint16_t* tile_col, int32_t* tile_src, int64_t src_capacity,
                      int64_t* n_src, void* stream);

/* Number of stencils per status code from the last rbfx_build_weights on this
 * thread (counts[0..15]); lets the host raise the reference's exception types. */
int rbfx_last_build_counts(int64_t counts[16]);

/* Device time (ms, HIP events on the build's stream) of the stages of the last
 * rbfx_build_weights on this thread: [0] the main kernel (plain solve, FP32 solve
 * or surrogate), [1] the FP32 solve of the surrogate's rejected stencils, [2] the
 * FP64 re-solve of non-converged stencils, [3] the whole build. No reference
 * counterpart (instrumentation).                                                  */
int rbfx_last_build_timing(double ms[4]);

/* Size buckets of the last build on this thread (no reference counterpart): stencils
 * with n = k + m <= 32, 33..64 and 65..RBFX_MAX_N, and the kernel each ran on
 * (0: registers, two stencils per wavefront; 1: registers, one; 2: system in LDS,
 * FP64; 3, 4: kernels 0, 1 in FP64 within an FP32 + refinement build; 5, 6: null-space
 * Cholesky; 7: packed lower triangle in LDS, FP64; -1: none). Uneven builds (ragged k or a degree map) classify stencils into these
 * buckets; the kernel of a bucket is timed on a sample the first time a
 * configuration is seen, and the fastest is kept (counts of -1: not read back).  */
int rbfx_last_build_buckets(int64_t n[3], int32_t kernel[3]);

/* Pin the kernel of a size bucket for later builds in this process (kernel -1:
 * automatic choice again). For tests and tuning; no reference counterpart.
 * Kernel ids: 0 rook LDL^T in registers, two stencils per wavefront (n <= 32);
 * 1 the same, one stencil per wavefront (n <= 64); 2 rook LDL^T with the system in
 * LDS; 3, 4 kernels 0, 1 in FP64 within
 * an FP32 + refinement build; 5 null-space Cholesky with one FP64 refinement step
 * (interior stencils of bases definite on the polynomial null space: PHS1 with degree
 * >= 0, PHS3 >= 1, PHS5 >= 2, PHS7 >= 3, IMQ, Gaussian; the stencils whose pivots or
 * refinement fail are re-solved by kernel 0 or 7); 6 the same with 8 lanes of three
 * rows per stencil (k <= 20, at most 6 monomials); 7 rook LDL^T with the packed lower
 * triangle of the system in LDS (symmetric pivot moves; half the LDS of kernel 2, so
 * more wavefronts per CU). A pinned kernel that cannot serve a build is ignored by that
 * build.                                                                        */
int rbfx_tune_bucket(int32_t bucket, int32_t kernel);

/* Counters of the last rbfx_build_weights on this thread (no reference counterpart):
 * [0] stencils the null-space kernels handed to the rook kernels (breakdown, refinement
 * not settling, Hermite or out-of-shape stencils), [1] autotuner sample launches, [2]
 * the launch plan of an uneven build (1 one launch per size bucket, 2 one launch for
 * all stencils; 0 for a build of one size), the rest 0.                          */
int rbfx_last_build_stats(int64_t stats[8]);

/* Pin the launch plan of uneven builds (ragged k or a degree map) for later builds in
 * this process: 0 automatic, 1 one launch per size bucket after a classify pass (each
 * bucket on its own kernel), 2 one launch of kernel 7 over every stencil in processing
 * order, with no classify pass. The automatic choice is made the first time a
 * configuration is timed (every nonempty bucket holding 4096+ stencils) by timing
 * kernel 7, sized for the largest stencil, on each bucket's sample against that bucket's
 * own choice; it is kept in the tuning map (bucket field 3, kernel field the plan).
 * No reference counterpart (one launch for stencils of uneven size, which the reference
 * builds one stencil at a time on the CPU). */
int rbfx_tune_plan(int32_t plan);

/* Save / load the autotuner's choices (no reference counterpart; the checkpoint of
 * the tuning map): one text line per configuration (dimension, operator columns,
 * precision, size bucket, largest k, monomials, basis kind, boundary data, kernel).
 * A loaded choice is used instead of timing samples, on every device, for the
 * configurations it names.                                                       */
int rbfx_tune_save(const char* path);
int rbfx_tune_load(const char* path);
/* Forget every choice made or loaded so far: later builds time their samples again. */
int rbfx_tune_reset(void);

/* ---- surrogate training and inputs (no reference counterpart) ------------ */

/* Sizes of the flat FP32 parameter array, its BF16 weight copy, and the run
 * workspace for a batch of M stencils.                                          */
int64_t rbfx_mlp_param_count(void);
int64_t rbfx_mlp_bf16_count(void);
int64_t rbfx_mlp_workspace_bytes(int32_t batch);

/* Forward pass; with a target also the mean-squared-error loss and the gradient
 * of every parameter (BF16 operands, FP32 accumulation on the matrix cores).    */
int rbfx_mlp_run(const rbfx_mlp_run_args* args);

/* Adam step on the FP32 parameters, then refresh of the BF16 weight copy.
 * step is the 1-based update count (bias correction).                           */
int rbfx_mlp_adam(float* params, const float* grads, float* m, float* v, uint16_t* wbf16, float lr, float beta1,
                  float beta2, float eps, int32_t step, void* stream);

/* BF16 weight copy of a parameter array (after initialising or loading it).    */
int rbfx_mlp_sync_bf16(const float* params, uint16_t* wbf16, void* stream);

/* Canonical radius-scaled offsets of fixed-size 2-D stencils: out_xi (n_eval*k*2
 * FP64), the network inputs out_x (n_eval*128 BF16: the 2k offsets' BF16 high parts,
 * their BF16 remainders, zeros), the stencil radius out_h and the distance of the
 * canonical reference neighbour out_r1 (n_eval each). Any output may be NULL.    */
int rbfx_stencil_features(const double* data, const double* eval, int64_t n_eval, int32_t dim, int32_t k,
                          const int32_t* adj_idx, double* out_xi, uint16_t* out_x, double* out_h, double* out_r1,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RBFX_H */
