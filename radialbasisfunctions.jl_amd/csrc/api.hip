// C ABI of the RBF-FD weight engine (declared in include/rbfx.h).
//
// Argument validation follows the reference's error behaviour:
//   ArgumentError for unsupported normal forms (src/solve/api.jl:382-397),
//   ArgumentError for a Hermite stencil without its eval point (assembly.jl:35-46),
//   SingularException from bunchkaufman! for singular stencils (assembly.jl:220-222),
// reported here as RBFX_EINVAL / RBFX_ESINGULAR with a message.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "build.h"
#include "knn.h"
#include "learned.h"
#include "mlp.h"
#include "order.h"
#include "rbfx.h"
#include "spmv.h"

namespace {

thread_local std::string g_err;
thread_local int64_t g_counts[16];
thread_local double g_stage_ms[4];  // last build: main kernel, FP32 list solve, FP64 fallback, total

int set_err(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return set_err(RBFX_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// per-device library scratch (grow-only)
struct DeviceState {
    rbfx::KnnWorkspace knn;
    int32_t* fail_list = nullptr;
    size_t cap_fail = 0;
    int32_t* fail_count = nullptr;
    int32_t* status = nullptr;
    size_t cap_status = 0;
    unsigned long long* hist = nullptr;
    int32_t* reject_list = nullptr;
    size_t cap_reject = 0;
    int32_t* reject_count = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // stage boundaries of a build
};

std::mutex g_mu;
DeviceState g_dev[64];

DeviceState& dev_state()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return g_dev[d & 63];
}

hipError_t grow(void** p, size_t& cap, size_t need)
{
    if (need <= cap) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(p, need);
    if (e == hipSuccess) cap = need;
    return e;
}

__global__ void k_hist16(const int32_t* __restrict__ st, int64_t n, unsigned long long* __restrict__ h)
{
    __shared__ unsigned int loc[16];
    if (threadIdx.x < 16) loc[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&loc[st[i] & 15], 1u);
    __syncthreads();
    if (threadIdx.x < 16 && loc[threadIdx.x]) atomicAdd(&h[threadIdx.x], (unsigned long long)loc[threadIdx.x]);
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ void k_zero_i32(int32_t* p) { *p = 0; }

int n_mono_host(int dim, int deg)
{
    if (deg < 0) return 0;
    int num = 1, den = 1;
    for (int i = 1; i <= dim; ++i) { num *= (deg + i); den *= i; }
    return num / den;
}

bool is_phs(int k) { return k == RBFX_PHS1 || k == RBFX_PHS3 || k == RBFX_PHS5 || k == RBFX_PHS7; }

// a device counter read back to size the next launch (blocking)
hipError_t read_count(const int32_t* d, int32_t* h, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(h, d, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace

static int spmv_impl(const rbfx_spmv_args* a, bool async)
{
    if (!a) return set_err(RBFX_EINVAL, "rbfx_spmv: null args");
    if (!a->col_idx || !a->vals || !a->x || !a->y) return set_err(RBFX_EINVAL, "rbfx_spmv: null pointer argument");
    if (a->n_fields < 1 || a->n_fields > 64) return set_err(RBFX_EUNSUPPORTED, "rbfx_spmv: n_fields=%d out of range 1..64", a->n_fields);
    if (a->n_valarrays < 1) return set_err(RBFX_EINVAL, "rbfx_spmv: n_valarrays must be >= 1");
    if (!a->row_ptr && a->k < 1) return set_err(RBFX_EINVAL, "rbfx_spmv: k must be >= 1 without row_ptr");
    if (a->n_rows == 0) return RBFX_OK;
    hipStream_t st = (hipStream_t)a->stream;
    rbfx::SpmvDev P;
    P.n_rows = a->n_rows;
    P.row_ptr = a->row_ptr;
    P.k = a->k;
    P.col_idx = a->col_idx;
    P.nnz = a->nnz > 0 ? a->nnz : a->n_rows * (int64_t)a->k;
    P.n_fields = a->n_fields;
    P.x = a->x;
    P.x_ld_point = a->x_ld_point;
    P.x_ld_field = a->x_ld_field;
    P.y_ld_point = a->y_ld_point;
    P.y_ld_field = a->y_ld_field;
    P.y_ld_val = a->y_ld_val;
    P.alpha = a->alpha;
    P.beta = a->beta;
    P.row_order = a->row_order;
    P.rows_permuted = a->rows_permuted;
    if (a->rows_permuted && !a->row_order) return set_err(RBFX_EINVAL, "rbfx_spmv: rows_permuted needs row_order");
    if (a->mapping < 0 || a->mapping > 2) return set_err(RBFX_EINVAL, "rbfx_spmv: mapping must be 0, 1 or 2");
    if (a->max_row_len < 0) return set_err(RBFX_EINVAL, "rbfx_spmv: max_row_len must be >= 0");
    P.max_row = a->row_ptr ? a->max_row_len : a->k;
    P.mapping = a->mapping;
    if (a->mapping == 2) {
        const int64_t L = P.max_row;
        const int g = L <= 20 ? 4 : (L <= 40 ? 8 : 16);
        const int64_t lds = (256 / g) * L * (8 * (int64_t)std::min(4, a->n_valarrays) + 4);
        if ((a->row_order && !a->rows_permuted) || a->n_fields >= 8 || L <= 0 || lds > 48 * 1024)
            return set_err(RBFX_EUNSUPPORTED,
                           "rbfx_spmv: staged mapping needs rows stored in processing order, n_fields < 8 and "
                           "a longest row that fits LDS (got %lld)", (long long)L);
    }
    for (int v0 = 0; v0 < a->n_valarrays; v0 += 4) {
        const int nv = std::min(4, a->n_valarrays - v0);
        P.n_valarrays = nv;
        P.vals = a->vals + (int64_t)v0 * P.nnz;
        P.y = a->y + (int64_t)v0 * a->y_ld_val;
        HIPCHK(rbfx::launch_spmv(P, st));
    }
    if (!async) HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}


static int build_impl(const rbfx_build_args* a, bool async)
{
    if (!a) return set_err(RBFX_EINVAL, "rbfx_build_weights: null args");
    for (int i = 0; i < 16; ++i) g_counts[i] = 0;
    for (int i = 0; i < 4; ++i) g_stage_ms[i] = 0.0;
    if (a->dim < 1 || a->dim > 3) return set_err(RBFX_EINVAL, "dim must be 1, 2 or 3 (got %d)", a->dim);
    if (!a->data || !a->eval || !a->adj_idx || !a->out_val)
        return set_err(RBFX_EINVAL, "rbfx_build_weights: data, eval, adj_idx and out_val are required");
    const int deg = a->basis.poly_deg;
    if (deg < -1) return set_err(RBFX_EINVAL, "poly_deg must be >= -1 (got %d)", deg);
    if (deg > rbfx::kMaxPolyDeg)
        return set_err(RBFX_EUNSUPPORTED, "poly_deg %d above %d is not built for the device", deg, rbfx::kMaxPolyDeg);
    const int kind = a->basis.kind;
    if (!(is_phs(kind) || kind == RBFX_IMQ || kind == RBFX_GAUSSIAN))
        return set_err(RBFX_EINVAL, "unknown basis kind %d", kind);
    if ((kind == RBFX_IMQ || kind == RBFX_GAUSSIAN) && !(a->basis.eps > 0))
        return set_err(RBFX_EINVAL, "Shape parameter eps must be > 0 (got eps=%g)", a->basis.eps);
    const rbfx_operator& op = a->op;
    if (op.n_cols < 1 || op.n_cols > RBFX_MAX_COLS) return set_err(RBFX_EINVAL, "n_cols=%d out of range", op.n_cols);
    if (op.n_terms < 1 || op.n_terms > RBFX_MAX_TERMS) return set_err(RBFX_EINVAL, "n_terms=%d out of range", op.n_terms);
    for (int t = 0; t < op.n_terms; ++t) {
        const rbfx_term& tm = op.terms[t];
        if (tm.col < 0 || tm.col >= op.n_cols) return set_err(RBFX_EINVAL, "term %d: column %d out of range", t, tm.col);
        if (tm.kind < RBFX_OP_VALUE || tm.kind > RBFX_OP_LAP) return set_err(RBFX_EINVAL, "term %d: unknown kind", t);
        if ((tm.kind == RBFX_OP_D1 || tm.kind == RBFX_OP_D2 || tm.kind == RBFX_OP_MIXED) &&
            (tm.dim1 < 0 || tm.dim1 >= a->dim))
            return set_err(RBFX_EINVAL, "term %d: dimension %d out of range", t, tm.dim1);
        if (tm.kind == RBFX_OP_MIXED && (tm.dim2 < 0 || tm.dim2 >= a->dim))
            return set_err(RBFX_EINVAL, "term %d: dimension %d out of range", t, tm.dim2);
    }
    const int M = n_mono_host(a->dim, deg);
    if (a->k < 1) return set_err(RBFX_EINVAL, "stencil size k must be >= 1");
    if (a->k + M > RBFX_MAX_N)
        return set_err(RBFX_EUNSUPPORTED, "stencil system k+m=%d exceeds the device maximum %d", a->k + M, RBFX_MAX_N);
    if (a->n_eval == 0) return RBFX_OK;
    if (a->precision < RBFX_PREC_FP64 || a->precision > RBFX_PREC_LEARNED)
        return set_err(RBFX_EUNSUPPORTED, "precision mode %d is not available through rbfx_build_weights", a->precision);
    const rbfx_mlp_model* mdl = a->model;
    if (a->precision == RBFX_PREC_LEARNED) {
        // the surrogate covers the configuration it was trained on
        if (!mdl) return set_err(RBFX_EINVAL, "learned build needs a model");
        if (!mdl->params || !mdl->wbf16 || !mdl->out_mean || !mdl->out_std)
            return set_err(RBFX_EINVAL, "learned build: model pointers missing");
        if (a->dim != 2 || mdl->dim != 2) return set_err(RBFX_EUNSUPPORTED, "learned build supports dim 2 only");
        if (a->adj_ptr || a->out_ptr || a->k != mdl->k || mdl->k != 20)
            return set_err(RBFX_EUNSUPPORTED, "learned build needs fixed stencils of the trained size k=%d", mdl->k);
        if (a->basis.kind != mdl->basis_kind || a->basis.poly_deg != mdl->poly_deg || mdl->poly_deg != 2)
            return set_err(RBFX_EUNSUPPORTED, "learned build: basis differs from the model's");
        if (a->is_boundary) return set_err(RBFX_EUNSUPPORTED, "learned build: Hermite stencils are not covered");
        if (a->op.n_cols != 1) return set_err(RBFX_EUNSUPPORTED, "learned build: one operator column only");
        for (int t = 0; t < a->op.n_terms; ++t)
            if (a->op.terms[t].kind != mdl->op_kind || mdl->op_kind != RBFX_OP_LAP)
                return set_err(RBFX_EUNSUPPORTED, "learned build: only the Laplacian the model was trained on");
        if (mdl->n_out != a->k + n_mono_host(2, 2)) return set_err(RBFX_EINVAL, "learned build: model n_out mismatch");
    }
    hipStream_t st = (hipStream_t)a->stream;

    // Hermite: the normal-form guard of _check_normal_form_support (api.jl:382-397)
    if (a->is_boundary) {
        if (!a->bnd_index || (a->n_bnd > 0 && (!a->bc_alpha || !a->bc_beta || !a->normals)))
            return set_err(RBFX_EINVAL, "Hermite build needs bnd_index, bc_alpha, bc_beta and normals");
        if (a->n_bnd > 0 && !async) {
            std::vector<double> al(a->n_bnd), be(a->n_bnd);
            HIPCHK(hipMemcpy(al.data(), a->bc_alpha, sizeof(double) * a->n_bnd, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(be.data(), a->bc_beta, sizeof(double) * a->n_bnd, hipMemcpyDeviceToHost));
            bool needs_normal = false;
            for (int64_t i = 0; i < a->n_bnd; ++i)
                if (!(al[i] == 1.0 && be[i] == 0.0)) needs_normal = true;
            bool supported = is_phs(kind);
            for (int t = 0; t < op.n_terms; ++t) {
                const int kd = op.terms[t].kind;
                if (!(kd == RBFX_OP_D1 || kd == RBFX_OP_D2 || kd == RBFX_OP_LAP)) supported = false;
            }
            if (needs_normal && !supported)
                return set_err(RBFX_EINVAL,
                               "Neumann/Robin boundary conditions require the operator's basis action to support the "
                               "normal-derivative form Lrbf(x, xi, normal); normal-form methods exist only for PHS "
                               "bases with the d, grad, d2 and laplacian operators");
        }
    }

    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState& ds = dev_state();
    // an asynchronous (capturable) build allocates nothing: the library's lists must
    // already have room for n_eval stencils from an earlier blocking build
    const size_t need = sizeof(int32_t) * (size_t)a->n_eval;
    if (async) {
        if (!a->status && ds.cap_status < need)
            return set_err(RBFX_EINVAL, "asynchronous build needs args.status or an earlier blocking build");
        if (a->precision != RBFX_PREC_FP64 && (ds.cap_fail < need || !ds.fail_count))
            return set_err(RBFX_EINVAL, "asynchronous build: run one blocking build with these sizes first");
        if (a->precision == RBFX_PREC_LEARNED && (ds.cap_reject < need || !ds.reject_count))
            return set_err(RBFX_EINVAL, "asynchronous build: run one blocking build with these sizes first");
    }
    int32_t* status = a->status;
    if (!status) {
        if (!async) HIPCHK(grow((void**)&ds.status, ds.cap_status, need));
        status = ds.status;
    }
    if (!ds.hist && !async) HIPCHK(hipMalloc((void**)&ds.hist, 16 * sizeof(unsigned long long)));

    rbfx::BuildDev P;
    std::memset(&P, 0, sizeof(P));
    P.n_data = a->n_data;
    P.n_eval = a->n_eval;
    P.data = a->data;
    P.eval = a->eval;
    P.adj_ptr = a->adj_ptr;
    P.adj_idx = a->adj_idx;
    P.order = a->order;
    P.k = a->k;
    P.basis_kind = kind;
    P.eps = a->basis.eps;
    P.n_cols = op.n_cols;
    P.n_terms = op.n_terms;
    for (int t = 0; t < op.n_terms; ++t) {
        P.terms[t].col = op.terms[t].col;
        P.terms[t].kind = op.terms[t].kind;
        P.terms[t].d1 = op.terms[t].dim1;
        P.terms[t].d2 = op.terms[t].dim2;
        P.terms[t].coeff = op.terms[t].coeff;
    }
    P.poly_deg = deg;
    P.n_mono = deg >= 0 ? rbfx::fill_mono_table(a->dim, deg, P.mono_e) : 0;
    P.is_boundary = a->is_boundary;
    P.bnd_index = a->bnd_index;
    P.bc_alpha = a->bc_alpha;
    P.bc_beta = a->bc_beta;
    P.normals = a->normals;
    P.out_ptr = a->out_ptr;
    P.out_col = a->out_col;
    P.out_val = a->out_val;
    P.out_nnz = a->out_nnz > 0 ? a->out_nnz : a->n_eval * (int64_t)a->k;
    P.status = status;
    P.out_lambda = a->out_lambda;
    P.refine_steps = a->refine_steps > 0 ? a->refine_steps : 4;
    P.accept_tol = a->accept_tol > 0 ? a->accept_tol : 1e-12;
    P.ok_status = RBFX_ST_OK;

    const int nmax = a->k + M;
    // stage boundaries: ev0 | main kernel | ev1 | FP32 list solve | ev2 | FP64 fallback | ev3
    // (blocking builds only; an asynchronous build records nothing but kernels)
    auto mark = [&](int i) -> hipError_t {
        if (async) return hipSuccess;
        if (!ds.ev[i]) {
            hipError_t e = hipEventCreate(&ds.ev[i]);
            if (e != hipSuccess) return e;
        }
        return hipEventRecord(ds.ev[i], st);
    };
    HIPCHK(mark(0));
    if (a->precision == RBFX_PREC_FP64) {
        HIPCHK(rbfx::launch_build(P, a->dim, nmax, false, a->n_eval, st));
        HIPCHK(mark(1));
        HIPCHK(mark(2));
    } else {
        if (!async) {
            HIPCHK(grow((void**)&ds.fail_list, ds.cap_fail, need));
            if (!ds.fail_count) HIPCHK(hipMalloc((void**)&ds.fail_count, sizeof(int32_t)));
        }
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.fail_count);
        P.fail_list = ds.fail_list;
        P.fail_count = ds.fail_count;
        int64_t n_solve = a->n_eval;
        if (a->precision == RBFX_PREC_LEARNED) {
            // surrogate + FP64 acceptance; the rejected stencils go to the solver
            if (!async) {
                HIPCHK(grow((void**)&ds.reject_list, ds.cap_reject, need));
                if (!ds.reject_count) HIPCHK(hipMalloc((void**)&ds.reject_count, sizeof(int32_t)));
            }
            hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.reject_count);
            rbfx::LearnedDev Q;
            std::memset(&Q, 0, sizeof(Q));
            Q.n_eval = a->n_eval;
            Q.data = a->data;
            Q.eval = a->eval;
            Q.adj_idx = a->adj_idx;
            Q.order = a->order;
            Q.k = a->k;
            Q.basis_kind = kind;
            Q.eps = a->basis.eps;
            Q.n_mono = P.n_mono;
            std::memcpy(Q.mono_e, P.mono_e, sizeof(Q.mono_e));
            std::memcpy(Q.terms, P.terms, sizeof(Q.terms));
            Q.n_terms = P.n_terms;
            Q.op_scale_order = 2.0;
            Q.wbf16 = reinterpret_cast<const __bf16*>(mdl->wbf16);
            Q.params = mdl->params;
            Q.out_mean = mdl->out_mean;
            Q.out_std = mdl->out_std;
            Q.accept_tol = mdl->accept_tol;
            Q.out_col = a->out_col;
            Q.out_val = a->out_val;
            Q.status = status;
            Q.reject_list = ds.reject_list;
            Q.reject_count = ds.reject_count;
            Q.out_rho = mdl->out_rho;
            HIPCHK(rbfx::launch_learned_build(Q, st));
            HIPCHK(mark(1));
            P.list = ds.reject_list;
            P.list_len = ds.reject_count;
            if (!async) {
                int32_t nrej = 0;
                HIPCHK(read_count(ds.reject_count, &nrej, st));
                n_solve = nrej;
            }  // asynchronous: the grid covers every stencil, the kernel reads the count
        }
        if (n_solve > 0) HIPCHK(rbfx::launch_build(P, a->dim, nmax, true, n_solve, st));
        // FP32_REFINED: the FP32 solve is the main kernel (stage 1 stays empty)
        if (a->precision != RBFX_PREC_LEARNED) HIPCHK(mark(1));
        HIPCHK(mark(2));
        // FP64 re-solve of the stencils whose refinement did not converge
        int32_t nfail = (int32_t)a->n_eval;
        if (!async) HIPCHK(read_count(ds.fail_count, &nfail, st));
        rbfx::BuildDev F = P;
        F.fail_list = nullptr;
        F.fail_count = nullptr;
        F.list = ds.fail_list;
        F.list_len = ds.fail_count;
        F.ok_status = RBFX_ST_FALLBACK64;
        if (nfail > 0) HIPCHK(rbfx::launch_build(F, a->dim, nmax, false, nfail, st));
    }
    HIPCHK(mark(3));
    if (async) {
        HIPCHK(hipGetLastError());
        g_err.clear();
        return RBFX_OK;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, ds.hist, 16);
    const int64_t hb = std::min<int64_t>((a->n_eval + 255) / 256, 4096);
    hipLaunchKernelGGL(k_hist16, dim3((unsigned)hb), dim3(256), 0, st, status, a->n_eval, ds.hist);
    unsigned long long h[16];
    HIPCHK(hipMemcpyAsync(h, ds.hist, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < 3; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ds.ev[i], ds.ev[i + 1]));
        g_stage_ms[i] = ms;
    }
    {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ds.ev[0], ds.ev[3]));
        g_stage_ms[3] = ms;
    }
    for (int i = 0; i < 16; ++i) g_counts[i] = (int64_t)h[i];
    if (h[RBFX_ST_TOO_LARGE])
        return set_err(RBFX_EUNSUPPORTED, "%llu stencils exceed the device maximum k+m=%d", h[RBFX_ST_TOO_LARGE], RBFX_MAX_N);
    if (h[RBFX_ST_EVAL_NOT_FOUND])
        return set_err(RBFX_EINVAL, "evaluation point not found in stencil data (%llu Hermite stencils)",
                       h[RBFX_ST_EVAL_NOT_FOUND]);
    if (h[RBFX_ST_SINGULAR])
        return set_err(RBFX_ESINGULAR, "%llu stencil systems are singular", h[RBFX_ST_SINGULAR]);
    g_err.clear();
    return RBFX_OK;
}


extern "C" {

int32_t rbfx_abi_version(void) { return 2; }

const char* rbfx_last_error(void) { return g_err.c_str(); }

int rbfx_init(int32_t device)
{
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(RBFX_EINVAL, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(RBFX_EDEVICE, "rbfx kernels are built for gfx950; device %d is %s", device, prop.gcnArchName);
    g_err.clear();
    return RBFX_OK;
}

int rbfx_last_build_counts(int64_t counts[16])
{
    for (int i = 0; i < 16; ++i) counts[i] = g_counts[i];
    return RBFX_OK;
}

int rbfx_last_build_timing(double ms[4])
{
    for (int i = 0; i < 4; ++i) ms[i] = g_stage_ms[i];
    return RBFX_OK;
}

int rbfx_knn(const double* data, int64_t n_data, const double* query, int64_t n_query, int32_t dim, int32_t k,
             const int64_t* q_ptr, int32_t* out_idx, void* stream)
{
    if (!data || !query || !out_idx) return set_err(RBFX_EINVAL, "rbfx_knn: null pointer argument");
    if (dim < 1 || dim > 3) return set_err(RBFX_EINVAL, "rbfx_knn: dim must be 1, 2 or 3 (got %d)", dim);
    if (k < 1 || k > n_data) return set_err(RBFX_EINVAL, "rbfx_knn: need 1 <= k <= n_data (k=%d, n_data=%lld)", k, (long long)n_data);
    if (k > 64) return set_err(RBFX_EUNSUPPORTED, "rbfx_knn: k=%d exceeds the supported maximum of 64", k);
    if (n_query == 0) return RBFX_OK;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState& ds = dev_state();
    HIPCHK(rbfx::run_knn(ds.knn, data, n_data, query, n_query, dim, k, q_ptr, out_idx, st));
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_build_weights(const rbfx_build_args* a) { return build_impl(a, false); }

int rbfx_build_weights_async(const rbfx_build_args* a) { return build_impl(a, true); }

int64_t rbfx_mlp_param_count(void) { return rbfx::mlp_n_params(); }

int64_t rbfx_mlp_bf16_count(void) { return rbfx::mlp_n_wbf16(); }

}  // extern "C"

namespace {

// run workspace: Z_l (FP32), H_l (BF16), dY and two dZ buffers (BF16), 256-B aligned
struct MlpWs {
    float* Z[rbfx::kMlpLayers];
    __bf16* H[rbfx::kMlpLayers - 1];
    __bf16* dY;
    __bf16* dZ[2];
    int64_t bytes;
};

MlpWs mlp_ws(void* base, int64_t M)
{
    MlpWs w;
    char* p = static_cast<char*>(base);
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char* q = p ? p + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return q;
    };
    for (int l = 0; l < rbfx::kMlpLayers; ++l) w.Z[l] = reinterpret_cast<float*>(take(M * rbfx::mlp_n(l) * 4));
    for (int l = 0; l < rbfx::kMlpLayers - 1; ++l) w.H[l] = reinterpret_cast<__bf16*>(take(M * rbfx::kHid * 2));
    w.dY = reinterpret_cast<__bf16*>(take(M * rbfx::kOutPad * 2));
    w.dZ[0] = reinterpret_cast<__bf16*>(take(M * rbfx::kHid * 2));
    w.dZ[1] = reinterpret_cast<__bf16*>(take(M * rbfx::kHid * 2));
    w.bytes = off;
    return w;
}

}  // namespace

extern "C" {

int64_t rbfx_mlp_workspace_bytes(int32_t batch) { return batch > 0 ? mlp_ws(nullptr, batch).bytes : 0; }

int rbfx_mlp_run(const rbfx_mlp_run_args* a)
{
    using namespace rbfx;
    if (!a) return set_err(RBFX_EINVAL, "rbfx_mlp_run: null args");
    if (a->batch < 1) return set_err(RBFX_EINVAL, "rbfx_mlp_run: batch must be >= 1");
    if (!a->x || !a->params || !a->wbf16 || !a->workspace)
        return set_err(RBFX_EINVAL, "rbfx_mlp_run: x, params, wbf16 and workspace are required");
    if (a->target && (!a->grads || !a->loss || a->n_out < 1 || a->n_out > kOutPad))
        return set_err(RBFX_EINVAL, "rbfx_mlp_run: training needs grads, loss and 1 <= n_out <= %d", kOutPad);
    hipStream_t st = (hipStream_t)a->stream;
    const int M = a->batch;
    MlpWs w = mlp_ws(a->workspace, M);
    const __bf16* X = reinterpret_cast<const __bf16*>(a->x);
    const __bf16* WB = reinterpret_cast<const __bf16*>(a->wbf16);
    float* Zout = a->y_out ? a->y_out : w.Z[kMlpLayers - 1];
    for (int l = 0; l < kMlpLayers; ++l) {
        GemmArgs g;
        std::memset(&g, 0, sizeof(g));
        g.M = M;
        g.N = mlp_n(l);
        g.K = mlp_k(l);
        g.A = l == 0 ? X : w.H[l - 1];
        g.sAm = mlp_k(l);
        g.sAk = 1;
        g.B = WB + mlp_wb_off(l);
        g.sBn = mlp_k(l);
        g.sBk = 1;
        g.C = l == kMlpLayers - 1 ? Zout : w.Z[l];
        g.ldc = mlp_n(l);
        g.bias = a->params + mlp_b_off(l);
        if (l < kMlpLayers - 1) {
            g.epi = EPI_BIAS_GELU;
            g.H = w.H[l];
            g.ldh = kHid;
        } else {
            g.epi = EPI_BIAS;
        }
        HIPCHK(launch_gemm(g, st));
    }
    if (a->target) {
        HIPCHK(launch_fill_f32(a->loss, 1, 0.0f, st));
        HIPCHK(launch_fill_f32(a->grads, mlp_n_params(), 0.0f, st));
        HIPCHK(launch_mse_grad(M, a->n_out, Zout, kOutPad, a->target, a->n_out, a->row_weight, w.dY, kOutPad,
                               a->loss, st));
        const __bf16* dZ = w.dY;
        int64_t ld_dz = kOutPad;
        for (int l = kMlpLayers - 1; l >= 0; --l) {
            const __bf16* Hin = l == 0 ? X : w.H[l - 1];
            GemmArgs gw;
            std::memset(&gw, 0, sizeof(gw));
            gw.M = mlp_n(l);
            gw.N = mlp_k(l);
            gw.K = M;
            gw.A = dZ;
            gw.sAm = 1;
            gw.sAk = ld_dz;
            gw.B = Hin;
            gw.sBn = 1;
            gw.sBk = mlp_k(l);
            gw.C = a->grads + mlp_w_off(l);
            gw.ldc = mlp_k(l);
            gw.epi = EPI_F32;
            gw.accumulate = 1;
            gw.k_split = 512;
            HIPCHK(launch_gemm(gw, st));
            HIPCHK(launch_colsum_bf16(M, mlp_n(l), dZ, ld_dz, a->grads + mlp_b_off(l), st));
            if (l > 0) {
                __bf16* dZn = w.dZ[l & 1];
                GemmArgs gd;
                std::memset(&gd, 0, sizeof(gd));
                gd.M = M;
                gd.N = mlp_k(l);
                gd.K = mlp_n(l);
                gd.A = dZ;
                gd.sAm = ld_dz;
                gd.sAk = 1;
                gd.B = WB + mlp_wb_off(l);
                gd.sBn = 1;
                gd.sBk = mlp_k(l);
                gd.epi = EPI_GELU_BWD;
                gd.Z = w.Z[l - 1];
                gd.ldz = kHid;
                gd.H = dZn;
                gd.ldh = kHid;
                HIPCHK(launch_gemm(gd, st));
                dZ = dZn;
                ld_dz = kHid;
            }
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_mlp_adam(float* params, const float* grads, float* m, float* v, uint16_t* wbf16, float lr, float beta1,
                  float beta2, float eps, int32_t step, void* stream)
{
    if (!params || !grads || !m || !v || !wbf16) return set_err(RBFX_EINVAL, "rbfx_mlp_adam: null pointer argument");
    if (step < 1) return set_err(RBFX_EINVAL, "rbfx_mlp_adam: step must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(rbfx::launch_adam(rbfx::mlp_n_params(), params, grads, m, v, lr, beta1, beta2, eps, step, st));
    HIPCHK(rbfx::launch_params_to_bf16(params, reinterpret_cast<__bf16*>(wbf16), st));
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_mlp_sync_bf16(const float* params, uint16_t* wbf16, void* stream)
{
    if (!params || !wbf16) return set_err(RBFX_EINVAL, "rbfx_mlp_sync_bf16: null pointer argument");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(rbfx::launch_params_to_bf16(params, reinterpret_cast<__bf16*>(wbf16), st));
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_stencil_features(const double* data, const double* eval, int64_t n_eval, int32_t dim, int32_t k,
                          const int32_t* adj_idx, double* out_xi, uint16_t* out_x, double* out_h, double* out_r1,
                          void* stream)
{
    if (!data || !eval || !adj_idx) return set_err(RBFX_EINVAL, "rbfx_stencil_features: null pointer argument");
    if (dim != 2) return set_err(RBFX_EUNSUPPORTED, "rbfx_stencil_features: dim 2 only");
    if (k < 1 || 4 * k > rbfx::kInPad) return set_err(RBFX_EUNSUPPORTED, "rbfx_stencil_features: need 1 <= k <= 32");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(rbfx::launch_stencil_features(data, eval, n_eval, k, adj_idx, out_xi, reinterpret_cast<__bf16*>(out_x),
                                         out_h, out_r1, st));
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_permute_rows(int64_t n_rows, int32_t k, const int32_t* row_order, const int32_t* col_idx,
                      const double* vals, int64_t nnz, int32_t n_valarrays, int32_t* out_col, double* out_vals,
                      void* stream)
{
    if (!row_order || !col_idx || !vals || !out_col || !out_vals)
        return set_err(RBFX_EINVAL, "rbfx_permute_rows: null pointer argument");
    if (k < 1 || n_valarrays < 1 || nnz < n_rows * (int64_t)k)
        return set_err(RBFX_EINVAL, "rbfx_permute_rows: need k >= 1, n_valarrays >= 1, nnz >= n_rows*k");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(rbfx::launch_permute_rows(n_rows, k, row_order, col_idx, vals, nnz, n_valarrays, out_col, out_vals, st));
    HIPCHK(hipStreamSynchronize(st));
    return RBFX_OK;
}

int rbfx_spatial_order(const double* points, int64_t n, int32_t dim, int32_t* out_order, void* stream)
{
    if (!points || !out_order) return set_err(RBFX_EINVAL, "rbfx_spatial_order: null pointer argument");
    if (dim < 1 || dim > 3) return set_err(RBFX_EINVAL, "rbfx_spatial_order: dim must be 1, 2 or 3");
    if (n > INT32_MAX) return set_err(RBFX_EUNSUPPORTED, "rbfx_spatial_order: n above 2^31-1");
    HIPCHK(rbfx::spatial_order(points, n, dim, out_order, (hipStream_t)stream));
    return RBFX_OK;
}

int rbfx_spmv(const rbfx_spmv_args* a) { return spmv_impl(a, false); }

int rbfx_spmv_async(const rbfx_spmv_args* a) { return spmv_impl(a, true); }

}  // extern "C"
