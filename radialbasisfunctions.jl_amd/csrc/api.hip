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
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "build.h"
#include "build_nsc.h"
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

// Autotuner configuration key: dimension, operator columns, precision mode, size
// bucket, largest stencil (k points, m monomials), basis kind, whether the build has
// boundary data and whether every stencil has the same shape (no ragged k, no degree
// map: the null-space kernels serve only such builds).
typedef std::tuple<int, int, int, int, int, int, int, int, int> TuneKey;

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
    // size buckets of an uneven build: three lists of stencil ids (n_eval each) + counts
    int32_t* bucket_list = nullptr;
    size_t cap_bucket = 0;
    int32_t* bucket_count = nullptr;  // [3] list lengths, [3..5] tuning sample sizes
    hipEvent_t tev[2] = {nullptr, nullptr};
    std::map<TuneKey, int> tuned;  // configuration -> kernel the autotuner picked
    int32_t* nsc_list = nullptr;   // stencils the null-space kernels hand back
    size_t cap_nsc = 0;
    int32_t* nsc_count = nullptr;
    double* packed = nullptr;  // per-position stencil offsets of a register-kernel launch
    size_t cap_packed = 0;
};

std::mutex g_mu;
DeviceState g_dev[64];
std::map<TuneKey, int> g_tune_saved;  // choices loaded with rbfx_tune_load (any device)

// The static device area of rbfx_build_weights_noalloc: the null-space kernels'
// hand-back list (up to kNoAllocStencils stencils) and its count. Every GPU the library
// runs on holds its own copy (16 MiB) until the process ends; calls on one GPU share it.
constexpr int64_t kNoAllocStencils = int64_t(1) << 22;
static __device__ __align__(256) int32_t scratch[kNoAllocStencils + 64];
int32_t* g_scratch_ptr[64];  // the area's address on each device (hipGetSymbolAddress)

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
__global__ void k_set_i32(int32_t* p, int32_t v) { *p = v; }

// Size buckets of an uneven build, in processing order: n = k + m <= 32, <= 64, above.
// Lanes of a wavefront append with one atomic per bucket, so each list keeps the
// (spatial) order of the stencils at wavefront granularity.
__global__ __launch_bounds__(256) void k_classify(const rbfx::BuildDev P, int dim, int32_t* lists, int32_t* counts)
{
    const int64_t wi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int b = -1;
    int32_t e = 0;
    if (wi < P.n_eval) {
        e = P.order ? P.order[wi] : (int32_t)wi;
        if (e >= 0 && e < P.n_eval) {
            const int kk = P.adj_ptr ? (int)(P.adj_ptr[e + 1] - P.adj_ptr[e]) : P.k;
            const int m = rbfx::stencil_n_mono(P, e, dim);
            const int nn = kk + (m < 0 ? 0 : m);
            b = nn <= 32 ? 0 : (nn <= 64 ? 1 : 2);
        }
    }
    const int lane = threadIdx.x & 63;
    for (int bb = 0; bb < 3; ++bb) {
        const unsigned long long mask = __ballot(b == bb);
        if (!mask) continue;
        const int leader = __builtin_ctzll(mask);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counts[bb], __popcll(mask));
        base = __shfl(base, leader, 64);
        if (b == bb) {
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            lists[(int64_t)bb * P.n_eval + base + rank] = e;
        }
    }
}

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
    if (a->mapping < 0 || a->mapping > 4) return set_err(RBFX_EINVAL, "rbfx_spmv: mapping must be 0..4");
    P.tile_rows = a->tile_rows;
    P.tile_src = a->tile_src;
    P.tile_src_off = a->tile_src_off;
    P.tile_src_n = a->tile_src_n;
    P.tile_col = a->tile_col;
    if (a->tile_col && (a->tile_rows < 1 || !a->tile_src || !a->tile_src_off || !a->tile_src_n))
        return set_err(RBFX_EINVAL, "rbfx_spmv: a source-tile plan needs tile_rows, tile_src, tile_src_off and tile_src_n");
    if (a->tile_col && a->n_fields < 8) {
        P.tile_col = nullptr;  // (few fields: the plan does not apply; the other mappings run)
    }
    if (a->mapping == 4 && (!P.tile_col || a->n_valarrays != 1))
        return set_err(RBFX_EUNSUPPORTED,
                       "rbfx_spmv: the source-tile mapping needs a plan, 8 or more fields and one value array");
    if (a->max_row_len < 0) return set_err(RBFX_EINVAL, "rbfx_spmv: max_row_len must be >= 0");
    P.max_row = a->row_ptr ? a->max_row_len : a->k;
    P.mapping = a->mapping;
    P.y = a->y;
    if (a->mapping == 3) {
        for (int v0 = 0; v0 < a->n_valarrays; v0 += 4) {
            rbfx::SpmvDev Q = P;
            Q.vals = a->vals + (int64_t)v0 * P.nnz;
            if (!rbfx::vec_entries(Q, std::min(4, a->n_valarrays - v0)))
                return set_err(RBFX_EUNSUPPORTED,
                               "rbfx_spmv: vector mapping needs rows of one length (no row_ptr), fewer than 8 "
                               "fields and at most 64 lanes per row");
        }
    }
    if (a->mapping == 2 && !rbfx::spmv_staged_fits(P, std::min(4, a->n_valarrays)))
        return set_err(RBFX_EUNSUPPORTED,
                       "rbfx_spmv: staged mapping needs rows stored in processing order and a longest row "
                       "that fits LDS (got %lld)", (long long)P.max_row);
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


// ---------------------------------------------------------------- size buckets
// Kernels a bucket can run on:
//   K_REG32   rook LDL^T in registers, two stencils per wavefront (n <= 32)
//   K_REG64   the same, one stencil per wavefront (n <= 64; one wave per SIMD, the
//             register overflow held in AGPRs, no scratch memory)
//   K_WIDE    rook LDL^T with the system in LDS (n <= RBFX_MAX_N)
//   K_REG32_F64, K_REG64_F64  the register kernels in FP64 within an FP32 + refinement
//             build
//   K_NSC, K_NSC2  null-space Cholesky (no pivot search) with one FP64 refinement
//             step, for interior stencils of bases definite on the polynomial null
//             space; the stencils it hands back are re-solved by K_REG32 / K_SYM.
//             K_NSC2 maps 8 lanes of three rows to a stencil (eight stencils per
//             wavefront, k <= 20, m <= 6).
//   K_SYM     rook LDL^T with the packed lower triangle in LDS (n <= RBFX_MAX_N; half the
//             LDS of K_WIDE, so more wavefronts per CU)
enum {
    K_REG32 = 0,
    K_REG64 = 1,
    K_WIDE = 2,
    K_REG32_F64 = 3,
    K_REG64_F64 = 4,
    K_NSC = 5,
    K_NSC2 = 6,
    K_SYM = 7,
    kNumKernels = 8
};
int g_force_kernel[3] = {-1, -1, -1};  // rbfx_tune_bucket overrides
// Launch plans of an uneven build: one launch per size bucket, each on its own kernel,
// after a classify pass (kPlanBucketed), or one launch of the packed LDS kernel over
// every stencil in processing order, no classify pass (kPlanSingle). The autotuner keeps
// its choice in the tuning map under the pseudo-bucket kPlanKey.
constexpr int kPlanBucketed = 1, kPlanSingle = 2, kPlanKey = 3;
int g_force_plan = 0;  // rbfx_tune_plan override (0: automatic)
thread_local int64_t g_bucket_n[3];
thread_local int32_t g_bucket_kernel[3];
thread_local int64_t g_stats[8];       // rbfx_last_build_stats
constexpr int64_t kTuneMin = 4096;     // stencils in a bucket before its mapping is timed
constexpr int32_t kTuneSample = 65536;

// How a build runs: blocking (sizes its buffers, reads back statuses), asynchronous on
// the caller's stream (capturable; needs buffers an earlier blocking build sized), or
// asynchronous with no allocation at all (buffers in the library's static area).
enum class Mode { kBlocking, kAsync, kNoAlloc };

// What decides which kernels can serve a launch: the largest stencil of the bucket
// (kmax points, m monomials), the operator columns, the basis and whether the build has
// boundary data (Hermite stencils are never solved by the null-space kernels; they hand
// them back).
struct Shape {
    int dim, kmax, m, n_cols, basis_kind, poly_deg;
    bool hermite;
    bool uniform;  // every stencil has k points and m monomials (no ragged k, no degree map)
};

static bool is_nsc(int kid) { return kid == K_NSC || kid == K_NSC2; }

// In FP32 + refinement mode the FP64 kernels are candidates too (their stencils get
// status FP64): with packed FP32 arithmetic unused, FP32 issues at the FP64 rate on
// gfx950 and the FP32 kernels hold extra registers for the refinement, so an FP64
// solve can be the faster one for a size class; the tuner measures which.
static bool kernel_serves(int kid, int bucket, bool fp32, const Shape& sh, bool pinned)
{
    switch (kid) {
    case K_REG32: return bucket == 0;
    // (in FP32 mode the 3-D kernel with 8 columns still spills at one wave: not served)
    case K_REG64: return bucket <= 1 && !(fp32 && sh.dim == 3 && sh.n_cols > 4);
    case K_WIDE: return bucket == 2 || (bucket == 1 && rbfx::wide_fits(64, sh.n_cols));
    case K_SYM: return rbfx::sym_fits(bucket == 0 ? 32 : (bucket == 1 ? 64 : RBFX_MAX_N), sh.n_cols);
    case K_REG32_F64: return fp32 && bucket == 0;
    case K_REG64_F64: return fp32 && bucket <= 1;
    case K_NSC:
    case K_NSC2: {
        if (!sh.uniform || rbfx::nsc_sign(sh.basis_kind, sh.poly_deg) == 0) return false;
        const int cap = bucket == 0 ? 32 : (bucket == 1 ? 64 : RBFX_MAX_N);
        const int kb = std::min(sh.kmax, cap);
        return rbfx::nsc_fits(sh.dim, kb, sh.m, sh.n_cols, kid == K_NSC2 ? 1 : 0);
    }
    default: return false;
    }
}

// The rook kernel that re-solves what a null-space launch handed back (FP64).
static int rook_for(int bucket, int n_cols)
{
    if (bucket == 0) return K_REG32;
    return rbfx::sym_fits(bucket == 1 ? 64 : RBFX_MAX_N, n_cols) ? K_SYM : K_REG64;
}

// Where a null-space launch queues the stencils it hands back (n_eval entries) and
// their count: the library's grow-only list, or the no-allocation build's static area.
struct HandBack {
    int32_t* list;
    int32_t* count;
};

static hipError_t launch_kernel_id(int kid, const rbfx::BuildDev& P, const Shape& sh, const HandBack& hb, int bucket,
                                   bool fp32, int rows_cap, int64_t n_work, bool async, hipStream_t st)
{
    rbfx::BuildDev Q = P;
    if (fp32 && (kid == K_WIDE || kid == K_SYM || kid == K_REG32_F64 || kid == K_REG64_F64 || is_nsc(kid)))
        Q.ok_status = RBFX_ST_FP64;
    if (is_nsc(kid)) {
        // the null-space launch, then the rook kernel on the stencils it handed back
        const int cap = bucket == 0 ? 32 : (bucket == 1 ? 64 : RBFX_MAX_N);
        Q.fail_list = hb.list;
        Q.fail_count = hb.count;
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, hb.count);
        hipError_t e = rbfx::launch_build_nsc(Q, sh.dim, std::min(sh.kmax, cap), sh.m, kid == K_NSC2 ? 1 : 0, n_work, st);
        if (e != hipSuccess) return e;
        int32_t nf = (int32_t)n_work;
        if (!async) {
            e = read_count(hb.count, &nf, st);
            if (e != hipSuccess) return e;
            g_stats[0] += nf;
        }
        if (nf <= 0) return hipSuccess;
        rbfx::BuildDev F = Q;
        F.fail_list = P.fail_list;  // (FP32 mode: the caller's list, unused by FP64 kernels)
        F.fail_count = P.fail_count;
        F.list = hb.list;
        F.list_len = hb.count;
        F.packed = nullptr;         // packed offsets follow the launch's own work order
        return launch_kernel_id(rook_for(bucket, Q.n_cols), F, sh, hb, bucket, false, rows_cap, nf, async, st);
    }
    if (kid == K_WIDE) return rbfx::launch_build_wide(Q, sh.dim, rows_cap, n_work, st);
    if (kid == K_SYM) return rbfx::launch_build_sym(Q, sh.dim, rows_cap, n_work, st);
    const bool f32 = fp32 && (kid == K_REG32 || kid == K_REG64);
    return rbfx::launch_build(Q, sh.dim, (kid == K_REG32 || kid == K_REG32_F64) ? 32 : 64, f32, n_work, st);
}

// Kernel for a size bucket: forced (rbfx_tune_bucket), cached for this configuration,
// or timed on a sample of the bucket's stencils (the autotuner: each kernel that serves
// the bucket runs the sample twice, the second run timed; the fastest is kept), or the
// default (a null-space kernel where one serves, else the narrowest rook kernel that
// fits). Q describes the sample: a work list whose length is sample_len (uneven
// builds), or the first `sample` stencils in processing order.
static int pick_kernel(const rbfx_build_args* a, const rbfx::BuildDev& Q, const Shape& sh, DeviceState& ds,
                       const HandBack& hb, int bucket, bool fp32, int rows_cap, int64_t count, int32_t* sample_len,
                       Mode mode, hipStream_t st, int* kid, float* timed_per_stencil = nullptr)
{
    if (timed_per_stencil) *timed_per_stencil = -1.f;  // (set only when this call times a sample)
    const bool async = mode != Mode::kBlocking;
    int dflt = bucket == 0 ? K_REG32 : K_SYM;
    if (bucket == 1 && !rbfx::sym_fits(64, sh.n_cols)) dflt = K_REG64;
    if (kernel_serves(K_NSC, bucket, fp32, sh, false)) dflt = K_NSC;
    if (kernel_serves(K_NSC2, bucket, fp32, sh, false)) dflt = K_NSC2;  // (the faster mapping where it fits)
    *kid = dflt;
    const int forced = g_force_kernel[bucket];
    if (forced >= 0 && kernel_serves(forced, bucket, fp32, sh, true)) {
        *kid = forced;
        return RBFX_OK;
    }
    const TuneKey key{a->dim,         sh.n_cols, a->precision, bucket, sh.kmax, sh.m, sh.basis_kind, sh.hermite ? 1 : 0,
                      sh.uniform ? 1 : 0};
    // a cached or loaded choice is used only where that kernel serves this build
    // (looked up, never inserted, by a build that allocates nothing)
    auto it = ds.tuned.find(key);
    if (it != ds.tuned.end() && kernel_serves(it->second, bucket, fp32, sh, false)) {
        *kid = it->second;
        return RBFX_OK;
    }
    auto jt = g_tune_saved.find(key);  // a map loaded with rbfx_tune_load
    if (jt != g_tune_saved.end() && kernel_serves(jt->second, bucket, fp32, sh, false)) {
        if (mode != Mode::kNoAlloc) ds.tuned[key] = jt->second;
        *kid = jt->second;
        return RBFX_OK;
    }
    if (async || count < kTuneMin) return RBFX_OK;
    // FP32-factorising candidates last: their time is charged with the FP64 re-solve of
    // the sample stencils they fail on, priced at the fastest FP64 candidate's rate
    const auto f32_factor = [&](int c) { return fp32 && (c == K_REG32 || c == K_REG64); };
    int cands[kNumKernels], nc = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < kNumKernels; ++c)
            if (kernel_serves(c, bucket, fp32, sh, false) && f32_factor(c) == (pass == 1)) cands[nc++] = c;
    if (nc < 2) {
        ds.tuned[key] = dflt;
        return RBFX_OK;
    }
    const int32_t sample = (int32_t)std::min<int64_t>(count, kTuneSample);
    rbfx::BuildDev S = Q;  // the sample: the first `sample` entries of the list, or stencils
    if (sample_len) {
        hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, st, sample_len, sample);
        S.list_len = sample_len;
    }
    for (int i = 0; i < 2; ++i)
        if (!ds.tev[i]) HIPCHK(hipEventCreate(&ds.tev[i]));
    float best = 0.f, f64_per_stencil = -1.f;
    for (int c = 0; c < nc; ++c) {
        const bool f32c = f32_factor(cands[c]) && Q.fail_count;
        HIPCHK(launch_kernel_id(cands[c], S, sh, hb, bucket, fp32, rows_cap, sample, false, st));  // warm-up
        if (f32c) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, Q.fail_count);
        HIPCHK(hipEventRecord(ds.tev[0], st));
        HIPCHK(launch_kernel_id(cands[c], S, sh, hb, bucket, fp32, rows_cap, sample, false, st));
        HIPCHK(hipEventRecord(ds.tev[1], st));
        HIPCHK(hipEventSynchronize(ds.tev[1]));
        g_stats[1] += 2;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ds.tev[0], ds.tev[1]));
        if (f32c) {
            int32_t fails = 0;
            HIPCHK(hipMemcpy(&fails, Q.fail_count, sizeof(fails), hipMemcpyDeviceToHost));
            const float rate = f64_per_stencil >= 0.f ? f64_per_stencil : ms / (float)sample;
            ms += (float)std::min(fails, sample) * rate;
        } else {
            const float r = ms / (float)sample;
            if (f64_per_stencil < 0.f || r < f64_per_stencil) f64_per_stencil = r;
        }
        if (c == 0 || ms < best) {
            best = ms;
            *kid = cands[c];
        }
    }
    g_stats[0] = 0;  // the sample runs' hand-backs are not the build's
    ds.tuned[key] = *kid;
    if (timed_per_stencil) *timed_per_stencil = best / (float)sample;
    return RBFX_OK;
}

// Device time of one launch of kernel kid on a sample (a warm-up run, then a timed one).
static int time_sample(int kid, const rbfx::BuildDev& S, const Shape& sh, DeviceState& ds, const HandBack& hb,
                       int bucket, bool fp32, int rows_cap, int32_t sample, hipStream_t st, float* ms)
{
    for (int i = 0; i < 2; ++i)
        if (!ds.tev[i]) HIPCHK(hipEventCreate(&ds.tev[i]));
    HIPCHK(launch_kernel_id(kid, S, sh, hb, bucket, fp32, rows_cap, sample, false, st));
    HIPCHK(hipEventRecord(ds.tev[0], st));
    HIPCHK(launch_kernel_id(kid, S, sh, hb, bucket, fp32, rows_cap, sample, false, st));
    HIPCHK(hipEventRecord(ds.tev[1], st));
    HIPCHK(hipEventSynchronize(ds.tev[1]));
    g_stats[1] += 2;
    HIPCHK(hipEventElapsedTime(ms, ds.tev[0], ds.tev[1]));
    return RBFX_OK;
}

// Offsets of the stencils of a launch of up to 32 rows with fixed k and no Hermite rows,
// packed by work position (launch_pack_offsets), or nullptr when the launch gathers
// them itself (other shapes, or an asynchronous build whose buffer is too small: a
// blocking build of the same size sizes it).
static const double* pack_offsets(const rbfx_build_args* a, const rbfx::BuildDev& P, DeviceState& ds, int nmax,
                                  int64_t n_work, bool async, hipStream_t st, int* rc)
{
    *rc = RBFX_OK;
    if (nmax > 32 || a->adj_ptr || P.out_ptr || P.is_boundary || P.k < 1 || n_work <= 0) return nullptr;
    const size_t need = sizeof(double) * (size_t)n_work * (size_t)P.k * (size_t)a->dim;
    if (!async) {
        if (grow((void**)&ds.packed, ds.cap_packed, need) != hipSuccess) {
            *rc = set_err(RBFX_EDEVICE, "rbfx_build_weights: device allocation of %zu bytes failed", need);
            return nullptr;
        }
    } else if (ds.cap_packed < need) {
        return nullptr;
    }
    if (rbfx::launch_pack_offsets(P, a->dim, n_work, ds.packed, st) != hipSuccess) {
        *rc = set_err(RBFX_EDEVICE, "rbfx_build_weights: pack launch failed");
        return nullptr;
    }
    return ds.packed;
}

// The hand-back list of the null-space kernels (n_eval entries), grown by blocking builds.
static int reserve_nsc_list(DeviceState& ds, int64_t n, bool async)
{
    const size_t need = sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1);
    if (async) {
        if (ds.cap_nsc < need || !ds.nsc_count)
            return set_err(RBFX_EINVAL, "asynchronous build: run one blocking build with these sizes first");
        return RBFX_OK;
    }
    HIPCHK(grow((void**)&ds.nsc_list, ds.cap_nsc, need));
    if (!ds.nsc_count) HIPCHK(hipMalloc((void**)&ds.nsc_count, sizeof(int32_t)));
    return RBFX_OK;
}

// Main solve of all stencils (plain FP64, or FP32 + refinement). Stencils of one size
// run in one launch. Uneven stencils (ragged k, per-stencil degree) are classified into
// size buckets in processing order, and each bucket runs on the kernel chosen for it.
// (A tuning sample's queued FP32 failures are discarded with the old count.)
static int solve_all(const rbfx_build_args* a, const rbfx::BuildDev& P, const Shape& sh, DeviceState& ds, int nmax,
                     bool fp32, Mode mode, const HandBack* static_hb, hipStream_t st)
{
    const int dim = a->dim;
    const bool async = mode != Mode::kBlocking;
    for (int b = 0; b < 3; ++b) {
        g_bucket_n[b] = 0;
        g_bucket_kernel[b] = -1;
    }
    HandBack hb{ds.nsc_list, ds.nsc_count};
    if (static_hb) {
        hb = *static_hb;
    } else {
        const int rc = reserve_nsc_list(ds, a->n_eval, async);
        if (rc != RBFX_OK) return rc;
        hb = HandBack{ds.nsc_list, ds.nsc_count};
    }
    const bool uneven = (a->adj_ptr || a->poly_deg_map) && nmax > 32;
    if (!uneven) {
        const int b = nmax <= 32 ? 0 : (nmax <= 64 ? 1 : 2);
        int kid = K_REG32;
        rbfx::BuildDev Q = P;
        Q.packed = nullptr;
        const int rc = pick_kernel(a, Q, sh, ds, hb, b, fp32, nmax, a->n_eval, nullptr, mode, st, &kid);
        if (rc != RBFX_OK) return rc;
        // Packed offsets (the gather done ahead by a pack kernel) serve the register
        // kernels. The null-space kernels gather their coordinates themselves: on the 4M
        // Laplacian that measured 12.02 ms per build against 12.42 ms with the pack (whose
        // 1.28 GB of offsets are written once and read back). No pack without allocation.
        if (mode != Mode::kNoAlloc && !is_nsc(kid)) {
            int prc = RBFX_OK;
            Q.packed = pack_offsets(a, P, ds, nmax, a->n_eval, async, st, &prc);
            if (prc != RBFX_OK) return prc;
        }
        g_bucket_n[b] = a->n_eval;
        g_bucket_kernel[b] = kid;
        if (fp32) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.fail_count);
        HIPCHK(launch_kernel_id(kid, Q, sh, hb, b, fp32, nmax, a->n_eval, async, st));
        return RBFX_OK;
    }
    if (mode == Mode::kNoAlloc)  // (the entry point refuses these builds before any work)
        return set_err(RBFX_EUNSUPPORTED, "a build without allocation needs one size class (k + m <= 32 or fixed k)");
    // the launch plan: pinned, or the tuned choice for this configuration, or decided below
    const int cap_single = std::min(nmax, RBFX_MAX_N);
    const bool single_ok = rbfx::sym_fits(cap_single, sh.n_cols);
    const TuneKey pkey{a->dim,        sh.n_cols, a->precision, kPlanKey, sh.kmax, sh.m, sh.basis_kind,
                       sh.hermite ? 1 : 0, sh.uniform ? 1 : 0};
    int plan = g_force_plan;
    if (plan == 0) {
        auto it = ds.tuned.find(pkey);
        auto jt = g_tune_saved.find(pkey);
        if (it != ds.tuned.end()) plan = it->second;
        else if (jt != g_tune_saved.end()) plan = jt->second;
    }
    if (plan == kPlanSingle && !single_ok) plan = kPlanBucketed;
    auto launch_single = [&]() -> int {
        // one launch over every stencil in processing order, each wavefront sized by
        // its own stencil (the packed kernel's lane slots above the stencil stay idle)
        for (int b = 0; b < 3; ++b) {
            g_bucket_n[b] = -1;
            g_bucket_kernel[b] = K_SYM;
        }
        g_stats[2] = kPlanSingle;
        if (fp32) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.fail_count);
        HIPCHK(launch_kernel_id(K_SYM, P, sh, hb, 2, fp32, cap_single, a->n_eval, async, st));
        return RBFX_OK;
    };
    if (plan == kPlanSingle) return launch_single();
    const size_t need = sizeof(int32_t) * 3 * (size_t)a->n_eval;
    if (!async) {
        HIPCHK(grow((void**)&ds.bucket_list, ds.cap_bucket, need));
        if (!ds.bucket_count) HIPCHK(hipMalloc((void**)&ds.bucket_count, 6 * sizeof(int32_t)));
    } else if (ds.cap_bucket < need || !ds.bucket_count) {
        return set_err(RBFX_EINVAL, "asynchronous build: run one blocking build with these sizes first");
    }
    for (int b = 0; b < 3; ++b) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.bucket_count + b);
    hipLaunchKernelGGL(k_classify, dim3((unsigned)((a->n_eval + 255) / 256)), dim3(256), 0, st, P, dim,
                       ds.bucket_list, ds.bucket_count);
    HIPCHK(hipGetLastError());
    int32_t cnt[3] = {(int32_t)a->n_eval, (int32_t)a->n_eval, (int32_t)a->n_eval};
    if (!async) {
        HIPCHK(hipMemcpyAsync(cnt, ds.bucket_count, sizeof(cnt), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    // (the LDS kernels size their tables by rows_cap: the bucket's bound or the largest
    // stencil of the build, whichever is smaller)
    const int rows_cap[3] = {std::min(32, nmax), std::min(64, nmax), std::min(nmax, RBFX_MAX_N)};
    int kid[3];
    float per_stencil[3];
    for (int b = 0; b < 3; ++b) {
        rbfx::BuildDev Q = P;
        Q.list = ds.bucket_list + (int64_t)b * a->n_eval;
        Q.list_len = ds.bucket_count + 3 + b;
        const int rc = pick_kernel(a, Q, sh, ds, hb, b, fp32, rows_cap[b], cnt[b], ds.bucket_count + 3 + b, mode, st,
                                   &kid[b], &per_stencil[b]);
        if (rc != RBFX_OK) return rc;
    }
    // Plan choice, once per configuration, when every nonempty bucket was just timed:
    // the single launch's kernel on each bucket's sample against the bucket's own pick,
    // weighted by the bucket counts.
    if (plan == 0 && !async && single_ok) {
        bool all_timed = true;
        for (int b = 0; b < 3; ++b)
            if (cnt[b] > 0 && per_stencil[b] < 0.f) all_timed = false;
        if (all_timed) {
            double est_bucketed = 0.0, est_single = 0.0;
            for (int b = 0; b < 3; ++b) {
                if (cnt[b] <= 0) continue;
                const int32_t sample = (int32_t)std::min<int64_t>(cnt[b], kTuneSample);
                rbfx::BuildDev S = P;
                S.list = ds.bucket_list + (int64_t)b * a->n_eval;
                S.list_len = ds.bucket_count + 3 + b;
                hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, st, ds.bucket_count + 3 + b, sample);
                float ms = 0.f;
                const int rc = time_sample(K_SYM, S, sh, ds, hb, 2, fp32, cap_single, sample, st, &ms);
                if (rc != RBFX_OK) return rc;
                est_bucketed += (double)cnt[b] * per_stencil[b];
                est_single += (double)cnt[b] * ms / sample;
            }
            g_stats[0] = 0;  // (sample runs)
            plan = est_single < est_bucketed ? kPlanSingle : kPlanBucketed;
            ds.tuned[pkey] = plan;
            if (plan == kPlanSingle) return launch_single();
        }
    }
    if (fp32) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.fail_count);
    for (int b = 0; b < 3; ++b) {
        g_bucket_n[b] = async ? -1 : cnt[b];
        g_bucket_kernel[b] = kid[b];
        if (cnt[b] <= 0) continue;
        rbfx::BuildDev Q = P;
        Q.list = ds.bucket_list + (int64_t)b * a->n_eval;
        Q.list_len = ds.bucket_count + b;
        HIPCHK(launch_kernel_id(kid[b], Q, sh, hb, b, fp32, rows_cap[b], cnt[b], async, st));
    }
    g_stats[2] = kPlanBucketed;
    return RBFX_OK;
}

// Address of the static area on the current device, looked up once per device (host
// code never passes the array's own name, which is not a device address).
static int scratch_area(int32_t** out)
{
    int d = 0;
    HIPCHK(hipGetDevice(&d));
    if (d < 0 || d >= 64) return set_err(RBFX_EDEVICE, "device index %d out of range", d);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[d]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[d] = static_cast<int32_t*>(ptr);
    }
    *out = g_scratch_ptr[d];
    return RBFX_OK;
}

static int build_impl(const rbfx_build_args* a, Mode mode)
{
    const bool async = mode != Mode::kBlocking;
    if (!a) return set_err(RBFX_EINVAL, "rbfx_build_weights: null args");
    for (int i = 0; i < 16; ++i) g_counts[i] = 0;
    for (int i = 0; i < 4; ++i) g_stage_ms[i] = 0.0;
    for (int i = 0; i < 8; ++i) g_stats[i] = 0;
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
    HandBack static_hb{nullptr, nullptr};
    if (mode == Mode::kNoAlloc) {
        if (a->precision != RBFX_PREC_FP64)
            return set_err(RBFX_EUNSUPPORTED, "rbfx_build_weights_noalloc: plain FP64 builds only");
        if (a->n_eval > kNoAllocStencils)
            return set_err(RBFX_EINVAL, "rbfx_build_weights_noalloc: %lld stencils exceed its maximum of %lld",
                           (long long)a->n_eval, (long long)kNoAllocStencils);
        if ((a->adj_ptr || a->poly_deg_map) && a->k + M > 32)
            return set_err(RBFX_EUNSUPPORTED,
                           "rbfx_build_weights_noalloc: stencils of varying size need k + m <= 32 (got up to %d)",
                           a->k + M);
        int32_t* area = nullptr;
        const int rc = scratch_area(&area);
        if (rc != RBFX_OK) return rc;
        static_hb = HandBack{area, area + kNoAllocStencils};
    }
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
        if (a->poly_deg_map) return set_err(RBFX_EUNSUPPORTED, "learned build: one polynomial degree only");
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
    if (!status && mode != Mode::kNoAlloc) {  // (without allocation: no status words unless passed)
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
    P.deg_map = a->poly_deg_map;
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
    Shape sh;
    sh.dim = a->dim;
    sh.kmax = a->k;
    sh.m = M;
    sh.n_cols = op.n_cols;
    sh.basis_kind = kind;
    sh.poly_deg = deg;
    sh.hermite = a->is_boundary != nullptr;
    sh.uniform = !a->adj_ptr && !a->poly_deg_map;
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
        const int rc = solve_all(a, P, sh, ds, nmax, false, mode, mode == Mode::kNoAlloc ? &static_hb : nullptr, st);
        if (rc != RBFX_OK) return rc;
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
        if (a->precision == RBFX_PREC_LEARNED) {
            // the rejected stencils (one size class) on the kernel picked for it
            if (!async && !ds.bucket_count) HIPCHK(hipMalloc((void**)&ds.bucket_count, 6 * sizeof(int32_t)));
            const int b = nmax <= 32 ? 0 : 1;
            int kid = b == 0 ? K_REG32 : K_REG64;
            rbfx::BuildDev Q = P;  // (P itself stays unpacked for the FP64 re-solve below)
            int prc = RBFX_OK;
            Q.packed = pack_offsets(a, P, ds, nmax, async ? a->n_eval : n_solve, async, st, &prc);
            if (prc != RBFX_OK) return prc;
            {
                const int rr = reserve_nsc_list(ds, a->n_eval, async);
                if (rr != RBFX_OK) return rr;
            }
            const HandBack hb{ds.nsc_list, ds.nsc_count};
            const int rc = pick_kernel(a, Q, sh, ds, hb, b, true, nmax, n_solve,
                                       ds.bucket_count ? ds.bucket_count + 3 : nullptr, mode, st, &kid);
            if (rc != RBFX_OK) return rc;
            hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, ds.fail_count);  // drop tuning failures
            for (int i = 0; i < 3; ++i) {
                g_bucket_n[i] = 0;
                g_bucket_kernel[i] = -1;
            }
            g_bucket_n[b] = async ? -1 : n_solve;
            g_bucket_kernel[b] = kid;
            if (n_solve > 0) HIPCHK(launch_kernel_id(kid, Q, sh, hb, b, true, nmax, n_solve, async, st));
        } else {
            const int rc = solve_all(a, P, sh, ds, nmax, true, mode, nullptr, st);
            if (rc != RBFX_OK) return rc;
        }
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
        // (stencils of more than 32 rows re-solve on the LDS kernel: the FP64 register
        // kernel for up to 64 rows spills)
        if (nfail > 0) {
            if (nmax > 32 && rbfx::wide_fits(std::min(nmax, RBFX_MAX_N), F.n_cols))
                HIPCHK(rbfx::launch_build_wide(F, a->dim, std::min(nmax, RBFX_MAX_N), nfail, st));
            else
                HIPCHK(rbfx::launch_build(F, a->dim, std::min(nmax, 64), false, nfail, st));
        }
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
    if (h[RBFX_ST_INVALID])
        return set_err(RBFX_EINVAL, "%llu stencils have a polynomial degree outside -1..%d", h[RBFX_ST_INVALID],
                       a->basis.poly_deg);
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

int32_t rbfx_abi_version(void) { return 3; }

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
    int32_t* area = nullptr;  // the no-allocation build's static area on this device
    const int rc = scratch_area(&area);
    if (rc != RBFX_OK) return rc;
    g_err.clear();
    return RBFX_OK;
}

int rbfx_last_build_counts(int64_t counts[16])
{
    for (int i = 0; i < 16; ++i) counts[i] = g_counts[i];
    return RBFX_OK;
}

int rbfx_tune_bucket(int32_t bucket, int32_t kernel)
{
    if (bucket < 0 || bucket > 2) return set_err(RBFX_EINVAL, "bucket must be 0, 1 or 2");
    if (kernel < -1 || kernel >= kNumKernels)
        return set_err(RBFX_EINVAL, "kernel %d is not a kernel id (0..%d)", kernel, kNumKernels - 1);
    if (kernel >= 0) {
        // a kernel that can serve the bucket for some configuration (a pinned kernel that
        // cannot serve a particular build is ignored by that build)
        Shape any{2, 20, 6, 1, RBFX_PHS3, 2, false, true};
        bool ok = false;
        for (int kmax : {20, 30, 40, 60})
            for (int fp32 = 0; fp32 < 2; ++fp32) {
                any.kmax = kmax;
                ok = ok || kernel_serves(kernel, bucket, fp32 != 0, any, true);
            }
        if (!ok) return set_err(RBFX_EINVAL, "kernel %d cannot serve bucket %d", kernel, bucket);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_force_kernel[bucket] = kernel;
    g_err.clear();
    return RBFX_OK;
}

int rbfx_last_build_buckets(int64_t n[3], int32_t kernel[3])
{
    for (int b = 0; b < 3; ++b) {
        n[b] = g_bucket_n[b];
        kernel[b] = g_bucket_kernel[b];
    }
    return RBFX_OK;
}

int rbfx_last_build_stats(int64_t stats[8])
{
    for (int i = 0; i < 8; ++i) stats[i] = g_stats[i];
    return RBFX_OK;
}

int rbfx_tune_plan(int32_t plan)
{
    if (plan < 0 || plan > kPlanSingle) return set_err(RBFX_EINVAL, "plan must be 0, 1 or 2 (got %d)", plan);
    std::lock_guard<std::mutex> lk(g_mu);
    g_force_plan = plan;
    g_err.clear();
    return RBFX_OK;
}

// Text format, one configuration per line:
//   dim n_cols precision bucket kmax m basis_kind hermite uniform kernel
int rbfx_tune_save(const char* path)
{
    if (!path) return set_err(RBFX_EINVAL, "rbfx_tune_save: null path");
    std::map<TuneKey, int> all;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        all = g_tune_saved;
        for (const DeviceState& d : g_dev)
            for (const auto& kv : d.tuned) all[kv.first] = kv.second;
    }
    FILE* f = std::fopen(path, "w");
    if (!f) return set_err(RBFX_EINVAL, "rbfx_tune_save: cannot open %s for writing", path);
    std::fprintf(f, "# rbfx autotuner choices: dim n_cols precision bucket kmax m basis hermite uniform kernel\n");
    for (const auto& kv : all) {
        const TuneKey& k = kv.first;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
                     std::get<3>(k), std::get<4>(k), std::get<5>(k), std::get<6>(k), std::get<7>(k), std::get<8>(k),
                     kv.second);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) return set_err(RBFX_EDEVICE, "rbfx_tune_save: write to %s failed", path);
    g_err.clear();
    return RBFX_OK;
}

int rbfx_tune_reset(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_tune_saved.clear();
    for (DeviceState& d : g_dev) d.tuned.clear();
    g_err.clear();
    return RBFX_OK;
}

int rbfx_tune_load(const char* path)
{
    if (!path) return set_err(RBFX_EINVAL, "rbfx_tune_load: null path");
    FILE* f = std::fopen(path, "r");
    if (!f) return set_err(RBFX_EINVAL, "rbfx_tune_load: cannot open %s", path);
    std::map<TuneKey, int> got;
    char line[256];
    int lineno = 0;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineno;
        if (line[0] == '#' || line[0] == '\n') continue;
        int v[10];
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                        &v[8], &v[9]) != 10 ||
            v[3] < 0 || v[3] > kPlanKey || v[9] < 0 || v[9] >= kNumKernels ||
            (v[3] == kPlanKey && v[9] != kPlanBucketed && v[9] != kPlanSingle)) {
            std::fclose(f);
            return set_err(RBFX_EINVAL, "rbfx_tune_load: %s line %d is not a tuning entry", path, lineno);
        }
        got[TuneKey{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]}] = v[9];
    }
    std::fclose(f);
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& kv : got) g_tune_saved[kv.first] = kv.second;
    g_err.clear();
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

int rbfx_build_weights(const rbfx_build_args* a) { return build_impl(a, Mode::kBlocking); }

int rbfx_build_weights_async(const rbfx_build_args* a) { return build_impl(a, Mode::kAsync); }

int rbfx_build_weights_noalloc(const rbfx_build_args* a) { return build_impl(a, Mode::kNoAlloc); }

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

int rbfx_source_tiles(int64_t n_rows, int32_t k, const int32_t* col_idx, int32_t tile_rows, int32_t* tile_src_n,
                      int64_t* tile_src_off, uint16_t* tile_col, int32_t* tile_src, int64_t src_capacity,
                      int64_t* n_src, void* stream)
{
    if (!col_idx || !tile_src_n || !tile_src_off || !tile_col || !n_src)
        return set_err(RBFX_EINVAL, "rbfx_source_tiles: null pointer argument");
    if (n_rows < 0 || k < 1 || tile_rows < 1 || (int64_t)tile_rows * k > rbfx::kTileEntries)
        return set_err(RBFX_EINVAL, "rbfx_source_tiles: need n_rows >= 0, k >= 1 and 1 <= tile_rows * k <= %d",
                       rbfx::kTileEntries);
    *n_src = 0;
    if (n_rows == 0) return RBFX_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n_tiles = (n_rows + tile_rows - 1) / tile_rows;
    HIPCHK(rbfx::launch_source_tiles(n_rows, k, nullptr, col_idx, tile_rows, 0, tile_src_n, nullptr, nullptr, nullptr,
                                     st));
    std::vector<int32_t> cnt((size_t)n_tiles);
    std::vector<int64_t> off((size_t)n_tiles);
    HIPCHK(hipMemcpyAsync(cnt.data(), tile_src_n, sizeof(int32_t) * n_tiles, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int64_t t = 0; t < n_tiles; ++t) {
        off[(size_t)t] = total;
        total += cnt[(size_t)t];
    }
    *n_src = total;
    HIPCHK(hipMemcpyAsync(tile_src_off, off.data(), sizeof(int64_t) * n_tiles, hipMemcpyHostToDevice, st));
    if (tile_src) {
        if (src_capacity < total)
            return set_err(RBFX_EINVAL, "rbfx_source_tiles: tile_src holds %lld entries, the plan needs %lld",
                           (long long)src_capacity, (long long)total);
        HIPCHK(rbfx::launch_source_tiles(n_rows, k, nullptr, col_idx, tile_rows, 1, tile_src_n, tile_src_off, tile_src,
                                         tile_col, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    g_err.clear();
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
