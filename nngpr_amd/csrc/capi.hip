// extern "C" entry points of libnngp_hip.so (declared in include/nngp.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "knn.h"
#include "conditioning.h"
#include "vecchia.h"
#include "vecchia_reg.h"

namespace nngp {

static thread_local char g_last_error[256] = "";

void set_last_hip_error(hipError_t e, const char* what, int line) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s (%d) at capi line %d: %s",
                  hipGetErrorString(e), (int)e, line, what);
}

namespace {

struct TimingState {
    std::mutex mu;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[TK_COUNT];
    hipEvent_t pending[TK_COUNT] = {};
    unsigned long long* work = nullptr;  // device: one work counter per timed kernel
};
TimingState& timing() {
    static TimingState t;
    return t;
}

}  // namespace

void timing_begin(int id, hipStream_t st) {
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, st);
    t.pending[id] = e;
}

unsigned long long* timing_work_counter(int id) {
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    return (t.on && t.work) ? t.work + id : nullptr;
}

void timing_end(int id, hipStream_t st) {
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.on || !t.pending[id]) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, st);
    t.ev[id].push_back({t.pending[id], e});
    t.pending[id] = nullptr;
}

namespace {

int to_kern_params(const nngp_kernel* k, int d, KernParams* out) {
    if (!k) return NNGP_ERR_INVALID_ARGUMENT;
    if (k->n_length_scale != 0 && k->n_length_scale != 1 && k->n_length_scale != d)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (k->n_theta < 0 || k->n_theta > NNGP_MAX_DIM + 3) return NNGP_ERR_INVALID_ARGUMENT;
    auto in_range = [&](int t, int span) { return t < 0 || (t + span <= k->n_theta); };
    if (!in_range(k->theta_noise_level, 1) || !in_range(k->theta_constant_value, 1) ||
        !in_range(k->theta_rbf_level, 1) ||
        !in_range(k->theta_length_scale, k->n_length_scale > 0 ? k->n_length_scale : 1))
        return NNGP_ERR_INVALID_ARGUMENT;
    KernParams kp{};
    kp.w = k->noise_level;
    kp.C = k->constant_value;
    kp.s = k->n_length_scale > 0 ? k->rbf_level : 0.0;
    kp.n_ls = k->n_length_scale;
    kp.t_w = k->theta_noise_level;
    kp.t_C = k->theta_constant_value;
    kp.t_s = k->n_length_scale > 0 ? k->theta_rbf_level : -1;
    kp.t_l = k->n_length_scale > 0 ? k->theta_length_scale : -1;
    kp.n_theta = k->n_theta;
    if (k->covariance < NNGP_COV_RBF || k->covariance > NNGP_COV_MATERN_52)
        return NNGP_ERR_INVALID_ARGUMENT;
    kp.fam = k->covariance;
    for (int f = 0; f < kp.n_ls; ++f) {
        const double l = k->length_scale[f];
        if (!(l > 0.0)) return NNGP_ERR_INVALID_ARGUMENT;
        kp.h[f] = 1.0 / (2.0 * l * l);
    }
    *out = kp;
    return NNGP_OK;
}

struct VecLayout {
    WsHeader* hdr;
    double* pt;
    double* partial;
    double* wsL;
    double* wsv;
    int64_t P;
    int64_t* flag_list;
    unsigned long long* flag_count;
};

// Largest dynamic LDS a block may use on the current device.
size_t max_block_lds() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 64 * 1024;
    return (size_t)v;
}

// Which fused kernel serves (m, d): the register kernel when its LDS fits.
bool use_register_kernel(int m, int d, int precision) {
    if (reg_mr(m) == 0) return false;
    const bool fp32 = precision == NNGP_PRECISION_FP32_REFINED;
    // The FP32 form re-runs flagged blocks with the FP64 kernel: both must fit.
    size_t need = reg_lds_bytes(m, d, false);
    if (fp32) need = need > reg_lds_bytes(m, d, true) ? need : reg_lds_bytes(m, d, true);
    return need <= max_block_lds();
}

VecLayout carve_vec(Carver& c, int64_t npts, int m, int d, int n_theta, bool need_pt,
                    int precision) {
    VecLayout L{};
    L.hdr = c.take<WsHeader>(1);
    const int cols = 1 + n_theta;
    if (need_pt) {
        L.pt = c.take<double>((size_t)npts * cols + 1);
        L.partial = c.take<double>(reduce_partial_doubles(cols));
    }
    L.flag_count = c.take<unsigned long long>(4);
    if (use_register_kernel(m, d, precision)) {
        L.P = 0;
        if (precision == NNGP_PRECISION_FP32_REFINED) L.flag_list = c.take<int64_t>((size_t)npts + 1);
    } else {
        L.P = generic_chunk(npts, m);
        L.wsL = c.take<double>((size_t)L.P * ((size_t)m * (m + 1) / 2));
        L.wsv = c.take<double>((size_t)L.P * 5 * (size_t)m);
    }
    return L;
}

__global__ void zero_flag_kernel(unsigned long long* c) {
    if (threadIdx.x == 0) *c = 0ULL;
}

__global__ void add_refactored_kernel(const unsigned long long* c, WsHeader* h) {
    if (threadIdx.x == 0) h->points_refactored_fp64 += (int64_t)*c;
}

// Runs the fused pass over a.npts points with the kernel that fits (m, d).
int run_fused(const VecArgs& a, const VecLayout& L, int precision, hipStream_t st) {
    if (a.npts == 0) return NNGP_OK;
    if (!use_register_kernel(a.m, a.d, precision)) {
        launch_generic(a, L.wsL, L.wsv, L.P, st);
        return NNGP_OK;
    }
    RegExtra ex{};
    const int tk = a.ys ? TK_LOGLIK : TK_PREDICT;
    if (precision == NNGP_PRECISION_FP32_REFINED) {
        hipLaunchKernelGGL(zero_flag_kernel, dim3(1), dim3(64), 0, st, L.flag_count);
        ex.flag_list = L.flag_list;
        ex.flag_count = L.flag_count;
        const bool seg = seg_supported(a.m, a.d);
        timing_begin(tk, st);
        const int rc = seg ? launch_seg(a, ex, true, a.npts, st) : launch_reg(a, ex, true, a.npts, st);
        timing_end(tk, st);
        if (rc) return NNGP_ERR_UNSUPPORTED;
        RegExtra redo{};
        redo.plist = L.flag_list;
        redo.pcount = L.flag_count;
        // the flagged points (a device-side count): a grid sized for the bound
        if ((seg ? launch_seg(a, redo, false, a.npts, st) : launch_reg(a, redo, false, a.npts < 4096 ? a.npts : 4096, st)))
            return NNGP_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(add_refactored_kernel, dim3(1), dim3(64), 0, st, L.flag_count, L.hdr);
    } else {
        timing_begin(tk, st);
        const int rc = seg_supported(a.m, a.d) ? launch_seg(a, ex, false, a.npts, st)
                                               : launch_reg(a, ex, false, a.npts, st);
        timing_end(tk, st);
        if (rc) return NNGP_ERR_UNSUPPORTED;
    }
    return NNGP_OK;
}

__global__ void extract_col_kernel(const double* pt, int stride, int64_t n, double* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pt[i * stride];
}

}  // namespace
}  // namespace nngp

// Resident prediction model (opaque to callers): copies of the training set, the
// kNN ref-side prepared state and the kernel parameters, in one device allocation.
struct nngp_model {
    int device;
    int64_t n;
    int d, ny, m, precision;
    double alpha;
    nngp::KernParams kp;
    double* x;
    double* y;
    nngp::KnnRefs refs;
    void* mem;
    size_t mem_bytes;
};

using namespace nngp;

extern "C" {

const char* nngp_status_string(int status) {
    switch (status) {
        case NNGP_OK: return "success";
        case NNGP_ERR_INVALID_ARGUMENT: return "invalid argument";
        case NNGP_ERR_UNSUPPORTED: return "unsupported configuration";
        case NNGP_ERR_NOT_POSITIVE_DEFINITE: return "matrix is not positive definite";
        case NNGP_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
        case NNGP_ERR_HIP: return "HIP runtime error";
        case NNGP_ERR_KNN_OVERFLOW: return "kNN candidate overflow";
        default: return "unknown status";
    }
}

const char* nngp_last_error_detail(void) { return g_last_error; }

int nngp_abi_version(void) { return NNGP_ABI_VERSION; }

int nngp_timing_enable(int on) {
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    if (on && !t.work) {
        // measurement support only: the counters are allocated once, zeroed from
        // the host
        void* p = nullptr;
        NNGP_HIP_CHECK(hipMalloc(&p, sizeof(unsigned long long) * TK_COUNT));
        const unsigned long long zero[TK_COUNT] = {};
        NNGP_HIP_CHECK(hipMemcpy(p, zero, sizeof(zero), hipMemcpyHostToDevice));
        t.work = static_cast<unsigned long long*>(p);
    }
    t.on = on != 0;
    return NNGP_OK;
}

int nngp_timing_work(int kernel, int reset, int64_t* units) {
    if (kernel < 0 || kernel >= TK_COUNT || !units) return NNGP_ERR_INVALID_ARGUMENT;
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    *units = 0;
    if (!t.work) return NNGP_OK;
    NNGP_HIP_CHECK(hipDeviceSynchronize());
    unsigned long long v = 0;
    NNGP_HIP_CHECK(hipMemcpy(&v, t.work + kernel, sizeof(v), hipMemcpyDeviceToHost));
    *units = (int64_t)v;
    if (reset) {
        const unsigned long long zero = 0;
        NNGP_HIP_CHECK(hipMemcpy(t.work + kernel, &zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    return NNGP_OK;
}

int nngp_timing_read(int kernel, int reset, double* total_ms, int64_t* launches) {
    if (kernel < 0 || kernel >= TK_COUNT || !total_ms || !launches) return NNGP_ERR_INVALID_ARGUMENT;
    TimingState& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    double tot = 0.0;
    for (auto& pr : t.ev[kernel]) {
        NNGP_HIP_CHECK(hipEventSynchronize(pr.second));
        float ms = 0.0f;
        NNGP_HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)t.ev[kernel].size();
    if (reset) {
        for (auto& pr : t.ev[kernel]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        t.ev[kernel].clear();
    }
    return NNGP_OK;
}

size_t nngp_knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision) {
    // Separate query buffers are always budgeted so one size serves both forms.
    return knn_workspace_bytes(n_ref, n_query, d, k, precision, false);
}

int nngp_knn(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
             const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut,
             int precision, int32_t* out_idx, void* workspace, size_t workspace_bytes,
             void* stream) {
    if ((n_ref > 0 && !ref) || (n_query > 0 && (!query || !out_idx)) || (n_scale > 0 && !scale))
        return NNGP_ERR_INVALID_ARGUMENT;
    return knn_run(ref, n_ref, query, n_query, d, scale, n_scale, k, has_cut, ref_cut, precision,
                   out_idx, workspace, workspace_bytes, (hipStream_t)stream, true);
}

size_t nngp_knn_rescale_state_bytes(int64_t n_ref, int64_t n_query, int d, int k, int K) {
    if (n_ref < 0 || n_query < 0 || d < 1 || d > NNGP_MAX_DIM || k < 1) return 0;
    return rescale_state_bytes(n_ref, n_query, d, k, K);
}

int nngp_knn_rescale_init(const double* ref, int64_t n_ref, const double* query, int64_t n_query,
                          int d, const double* scale, int n_scale, const int32_t* cand, int K,
                          int k, void* state, size_t state_bytes, void* stream) {
    if ((n_ref > 0 && !ref) || (n_query > 0 && (!query || !cand))) return NNGP_ERR_INVALID_ARGUMENT;
    return rescale_init(ref, n_ref, query, n_query, d, scale, n_scale, cand, K, k, state,
                        state_bytes, (hipStream_t)stream);
}

int nngp_knn_rescale(const double* ref, int64_t n_ref, const double* query, int64_t n_query,
                     int d, const double* scale_old, const double* scale_new, int n_scale,
                     int has_cut, int64_t ref_cut, const int32_t* cand, int K, int k,
                     int32_t* out_idx, void* state, size_t state_bytes, void* stream) {
    if ((n_ref > 0 && !ref) || (n_query > 0 && (!query || !cand || !out_idx)))
        return NNGP_ERR_INVALID_ARGUMENT;
    return rescale_run(ref, n_ref, query, n_query, d, scale_old, scale_new, n_scale, has_cut,
                       ref_cut, cand, K, k, out_idx, state, state_bytes, (hipStream_t)stream);
}

int nngp_knn_set_method(int method) { return knn_set_method(method); }

int nngp_knn_get_method(void) { return knn_method(); }

int nngp_fill_nn_kernel(const double* x, int64_t n_x, int d, const int32_t* nn,
                        int64_t n_points, int width, const nngp_kernel* kernel, double* out_k,
                        double* out_grad, void* stream) {
    if (d < 1 || d > NNGP_MAX_DIM || width < 1 || n_points < 0 || n_x < 0 || !out_k || !nn || !x)
        return NNGP_ERR_INVALID_ARGUMENT;
    KernParams kp;
    int rc = to_kern_params(kernel, d, &kp);
    if (rc) return rc;
    launch_fill(x, n_x, d, nn, n_points, width, kp, out_k, out_grad, (hipStream_t)stream);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

size_t nngp_vecchia_workspace_bytes(int64_t n_rows, int m, int d, int ny, int n_theta,
                                    int precision, int eval_gradient) {
    (void)ny;
    (void)eval_gradient;
    Carver c(nullptr, 0);
    carve_vec(c, n_rows, m, d, n_theta, true, precision);
    return c.used + 256;
}

}  // extern "C"

namespace {

// Argument checks of the log-likelihood entry points.
int check_loglik_args(int64_t n, int d, int ny, int m, int64_t row0, int64_t row1,
                      const nngp_kernel* kernel, int precision, int eval_gradient,
                      const double* out_loglik, const double* out_grad, KernParams* kp) {
    if (d < 1 || d > NNGP_MAX_DIM || ny < 1 || m < 1 || m > NNGP_MAX_NN || row0 < 0 ||
        row1 < row0 || row1 > n || !out_loglik)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (precision != NNGP_PRECISION_FP64 && precision != NNGP_PRECISION_FP32_REFINED)
        return NNGP_ERR_INVALID_ARGUMENT;
    const int rc = to_kern_params(kernel, d, kp);
    if (rc) return rc;
    if (eval_gradient && kp->n_theta > 0 && !out_grad) return NNGP_ERR_INVALID_ARGUMENT;
    return NNGP_OK;
}

// out_loglik and out_grad[0, n_theta) become NaN when the call met a block that
// could not be factored or a neighbour index outside [-1, n).
__global__ void mark_failed_kernel(const WsHeader* h, double* out_ll, double* out_grad,
                                   int n_theta) {
    const bool failed = h->not_pd > 0 || h->bad_index > 0;
    if (!failed) return;
    const int i = threadIdx.x;
    if (i == 0) *out_ll = __builtin_nan("");
    if (out_grad)
        for (int q = i; q < n_theta; q += blockDim.x) out_grad[q] = __builtin_nan("");
}

// The fused log-likelihood pass of rows [row0, row1) over a checked argument set,
// with its scratch carved from (workspace, workspace_bytes).
int run_loglik(const double* x, const double* y, int64_t n, int d, int ny, const int32_t* nn,
               int m, int64_t row0, int64_t row1, const KernParams& kp, double alpha,
               int precision, int eval_gradient, double* out_loglik, double* out_grad,
               double* out_point_loglik, void* workspace, size_t workspace_bytes,
               bool mark_failed, hipStream_t st) {
    const int64_t npts = row1 - row0;
    Carver c(workspace, workspace_bytes);
    VecLayout L = carve_vec(c, npts, m, d, kp.n_theta, true, precision);
    if (!c.ok() || !workspace) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, L.hdr);
    VecArgs a{};
    a.x = x;
    a.y = y;
    a.xs = x + row0 * d;
    a.ys = y + row0 * ny;
    a.nn = nn + row0 * m;
    a.n_src = n;
    a.npts = npts;
    a.d = d;
    a.ny = ny;
    a.m = m;
    a.alpha = alpha;
    a.alpha_self = 1;
    a.kp = kp;
    a.n_theta = kp.n_theta;
    a.grad = eval_gradient ? 1 : 0;
    a.pt = L.pt;
    a.pt_stride = 1 + kp.n_theta;
    a.hdr = L.hdr;
    const int rc = run_fused(a, L, precision, st);
    if (rc) return rc;
    launch_reduce(L.pt, npts, 1 + kp.n_theta, L.partial, out_loglik,
                  eval_gradient ? out_grad : nullptr, st);
    if (out_point_loglik && npts > 0)
        hipLaunchKernelGGL(extract_col_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0,
                           st, L.pt, 1 + kp.n_theta, npts, out_point_loglik);
    if (mark_failed)
        hipLaunchKernelGGL(mark_failed_kernel, dim3(1), dim3(64), 0, st, L.hdr, out_loglik,
                           eval_gradient ? out_grad : nullptr, kp.n_theta);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

// Scratch of nngp_vecchia_loglik_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the host
// it names a shadow variable): it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[NNGP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the current device's scratch, looked up once per device.
int scratch_address(void** out) {
    int dev = 0;
    NNGP_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return NNGP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        NNGP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return NNGP_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NNGP_OK;
}

}  // namespace

extern "C" {

int nngp_vecchia_loglik(const double* x, const double* y, int64_t n, int d, int ny,
                        const int32_t* nn, int m, int64_t row0, int64_t row1,
                        const nngp_kernel* kernel, double alpha, int precision,
                        int eval_gradient, double* out_loglik, double* out_grad,
                        double* out_point_loglik, void* workspace, size_t workspace_bytes,
                        void* stream) {
    KernParams kp;
    const int rc = check_loglik_args(n, d, ny, m, row0, row1, kernel, precision, eval_gradient,
                                     out_loglik, out_grad, &kp);
    if (rc) return rc;
    return run_loglik(x, y, n, d, ny, nn, m, row0, row1, kp, alpha, precision, eval_gradient,
                      out_loglik, out_grad, out_point_loglik, workspace, workspace_bytes, false,
                      (hipStream_t)stream);
}

int nngp_vecchia_loglik_noalloc(const double* x, const double* y, int64_t n, int d, int ny,
                                const int32_t* nn, int m, int64_t row0, int64_t row1,
                                const nngp_kernel* kernel, double alpha, int precision,
                                int eval_gradient, double* out_loglik, double* out_grad,
                                double* out_point_loglik, void* stream) {
    KernParams kp;
    int rc = check_loglik_args(n, d, ny, m, row0, row1, kernel, precision, eval_gradient,
                               out_loglik, out_grad, &kp);
    if (rc) return rc;
    Carver c(nullptr, 0);
    carve_vec(c, row1 - row0, m, d, kp.n_theta, true, precision);
    const size_t need = c.used + 256;
    if (need > NNGP_NOALLOC_SCRATCH_BYTES) {
        std::snprintf(g_last_error, sizeof(g_last_error),
                      "nngp_vecchia_loglik_noalloc: %zu rows x %d neighbours need %zu bytes of "
                      "scratch; the library holds %zu per GPU",
                      (size_t)(row1 - row0), m, need, (size_t)NNGP_NOALLOC_SCRATCH_BYTES);
        return NNGP_ERR_WORKSPACE_TOO_SMALL;
    }
    void* ptr = nullptr;
    rc = scratch_address(&ptr);
    if (rc) return rc;
    return run_loglik(x, y, n, d, ny, nn, m, row0, row1, kp, alpha, precision, eval_gradient,
                      out_loglik, out_grad, out_point_loglik, ptr, NNGP_NOALLOC_SCRATCH_BYTES, true,
                      (hipStream_t)stream);
}

size_t nngp_predict_workspace_bytes(int64_t n_query, int m, int d, int ny, int precision) {
    (void)ny;
    Carver c(nullptr, 0);
    carve_vec(c, n_query, m, d, 0, false, precision);
    return c.used + 256;
}

int nngp_predict(const double* x_train, const double* y_train, int64_t n_train, int d, int ny,
                 const double* x_query, int64_t n_query, const int32_t* nn, int m,
                 const nngp_kernel* kernel, double alpha, int precision, double* out_mean,
                 double* out_std, void* workspace, size_t workspace_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (d < 1 || d > NNGP_MAX_DIM || ny < 1 || m < 1 || m > NNGP_MAX_NN || n_query < 0 ||
        n_train < 0 || !out_mean)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (precision != NNGP_PRECISION_FP64 && precision != NNGP_PRECISION_FP32_REFINED)
        return NNGP_ERR_INVALID_ARGUMENT;
    KernParams kp;
    int rc = to_kern_params(kernel, d, &kp);
    if (rc) return rc;
    Carver c(workspace, workspace_bytes);
    VecLayout L = carve_vec(c, n_query, m, d, 0, false, precision);
    if (!c.ok() || !workspace) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, L.hdr);
    VecArgs a{};
    a.x = x_train;
    a.y = y_train;
    a.xs = x_query;
    a.ys = nullptr;
    a.nn = nn;
    a.n_src = n_train;
    a.npts = n_query;
    a.d = d;
    a.ny = ny;
    a.m = m;
    a.alpha = alpha;
    a.alpha_self = 0;
    a.kp = kp;
    a.n_theta = 0;
    a.grad = 0;
    a.mean = out_mean;
    a.stdv = out_std;
    a.hdr = L.hdr;
    rc = run_fused(a, L, precision, st);
    if (rc) return rc;
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

size_t nngp_predict_sequential_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d,
                                               int ny, int output) {
    if (n_train < 0 || n_query < 0 || m < 1 || d < 1 || ny < 1) return 0;
    return cond_workspace_bytes(n_train, n_query, m, d, ny, output);
}

namespace {

int check_sequential(const double* x_train, const double* y_train, int64_t n_train,
                     const double* x_query, int64_t n_query, int d, int ny, const int32_t* nn,
                     int m, const nngp_kernel* kernel, KernParams* kp) {
    if (d < 1 || d > NNGP_MAX_DIM || ny < 1 || m < 1 || m > NNGP_MAX_NN || n_train < 0 ||
        n_query < 0 || n_train + n_query > INT32_MAX)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (n_train > 0 && (!x_train || !y_train)) return NNGP_ERR_INVALID_ARGUMENT;
    if (n_query > 0 && (!x_query || !nn)) return NNGP_ERR_INVALID_ARGUMENT;
    return to_kern_params(kernel, d, kp);
}

}  // namespace

int nngp_predict_sequential(const double* x_train, const double* y_train, int64_t n_train,
                            const double* x_query, int64_t n_query, int d, int ny,
                            const int32_t* nn, int m, const nngp_kernel* kernel, double alpha,
                            int output, double* out_mean, double* out_std, double* out_cov,
                            void* workspace, size_t workspace_bytes, void* stream) {
    KernParams kp;
    int rc = check_sequential(x_train, y_train, n_train, x_query, n_query, d, ny, nn, m, kernel,
                              &kp);
    if (rc) return rc;
    if (output != NNGP_COND_MEAN && output != NNGP_COND_STD && output != NNGP_COND_COV)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (n_query > 0 && (!out_mean || (output == NNGP_COND_STD && !out_std) ||
                        (output == NNGP_COND_COV && !out_cov)))
        return NNGP_ERR_INVALID_ARGUMENT;
    CondProblem P{};
    P.x_train = x_train;
    P.y_train = y_train;
    P.n_train = n_train;
    P.x_query = x_query;
    P.n_query = n_query;
    P.d = d;
    P.ny = ny;
    P.m = m;
    P.nn = nn;
    P.kp = kp;
    P.alpha = alpha;
    P.output = output == NNGP_COND_MEAN ? COND_MEAN : (output == NNGP_COND_STD ? COND_STD : COND_COV);
    P.out_mean = out_mean;
    P.out_std = out_std;
    P.out_cov = out_cov;
    return cond_run(P, workspace, workspace_bytes, (hipStream_t)stream);
}

size_t nngp_sample_sequential_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d,
                                              int ny) {
    if (n_train < 0 || n_query < 0 || m < 1 || d < 1 || ny < 1) return 0;
    return cond_workspace_bytes(n_train, n_query, m, d, ny, COND_SAMPLE);
}

int nngp_sample_sequential(const double* x_train, const double* y_train, int64_t n_train,
                           const double* x_query, int64_t n_query, int d, int ny,
                           const int32_t* nn, int m, const nngp_kernel* kernel, double alpha,
                           int n_samples, const double* normals, double* out_samples,
                           void* workspace, size_t workspace_bytes, void* stream) {
    KernParams kp;
    int rc = check_sequential(x_train, y_train, n_train, x_query, n_query, d, ny, nn, m, kernel,
                              &kp);
    if (rc) return rc;
    if (n_samples < 1 || (n_query > 0 && (!normals || !out_samples)))
        return NNGP_ERR_INVALID_ARGUMENT;
    CondProblem P{};
    P.x_train = x_train;
    P.y_train = y_train;
    P.n_train = n_train;
    P.x_query = x_query;
    P.n_query = n_query;
    P.d = d;
    P.ny = ny;
    P.m = m;
    P.nn = nn;
    P.kp = kp;
    P.alpha = alpha;
    P.output = COND_SAMPLE;
    P.n_samples = n_samples;
    P.normals = normals;
    P.out_samples = out_samples;
    return cond_run(P, workspace, workspace_bytes, (hipStream_t)stream);
}

int nngp_model_create(const double* x_train, const double* y_train, int64_t n_train, int d, int ny,
                      const nngp_kernel* kernel, double alpha, int num_nn, const double* scale,
                      int n_scale, int precision, nngp_model** out_model, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (!out_model || !x_train || !y_train || n_train < 1 || d < 1 || d > NNGP_MAX_DIM || ny < 1 ||
        num_nn < 1)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (precision != NNGP_PRECISION_FP64 && precision != NNGP_PRECISION_FP32_REFINED)
        return NNGP_ERR_INVALID_ARGUMENT;
    KernParams kp;
    int rc = to_kern_params(kernel, d, &kp);
    if (rc) return rc;
    const int m = (int)(num_nn < n_train ? num_nn : n_train);
    if (m > NNGP_MAX_K) return NNGP_ERR_UNSUPPORTED;
    Carver c(nullptr, 0);
    c.take<double>((size_t)n_train * d);
    c.take<double>((size_t)n_train * ny);
    const size_t rb = knn_refs_bytes(n_train, d, precision);
    c.take<char>(rb);
    nngp_model* M = new nngp_model();
    M->mem_bytes = c.used + 256;
    if (hipMalloc(&M->mem, M->mem_bytes) != hipSuccess) {
        delete M;
        return NNGP_ERR_HIP;
    }
    Carver cm(M->mem, M->mem_bytes);
    M->x = cm.take<double>((size_t)n_train * d);
    M->y = cm.take<double>((size_t)n_train * ny);
    char* rmem = cm.take<char>(rb);
    (void)hipGetDevice(&M->device);
    M->n = n_train;
    M->d = d;
    M->ny = ny;
    M->m = m;
    M->precision = precision;
    M->alpha = alpha;
    M->kp = kp;
    hipError_t e1 = hipMemcpyAsync(M->x, x_train, sizeof(double) * n_train * d,
                                   hipMemcpyDeviceToDevice, st);
    hipError_t e2 = hipMemcpyAsync(M->y, y_train, sizeof(double) * n_train * ny,
                                   hipMemcpyDeviceToDevice, st);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        (void)hipFree(M->mem);
        delete M;
        return NNGP_ERR_HIP;
    }
    rc = knn_prepare_refs(M->x, n_train, d, scale, n_scale, precision, rmem, rb, &M->refs, st);
    if (rc == NNGP_OK && hipStreamSynchronize(st) != hipSuccess) rc = NNGP_ERR_HIP;
    if (rc) {
        (void)hipFree(M->mem);
        delete M;
        return rc;
    }
    *out_model = M;
    return NNGP_OK;
}

int nngp_model_destroy(nngp_model* model) {
    if (!model) return NNGP_ERR_INVALID_ARGUMENT;
    hipError_t e = hipFree(model->mem);
    delete model;
    return e == hipSuccess ? NNGP_OK : NNGP_ERR_HIP;
}

static void carve_model_ws(Carver& c, const nngp_model* M, int64_t nq, WsHeader** hdr, char** kmem,
                           size_t* kb, int32_t** nn, VecLayout* L) {
    *hdr = c.take<WsHeader>(1);
    *kb = knn_search_bytes(M->n, nq, M->d, M->m, M->precision, false, false, M->refs.layout);
    *kmem = c.take<char>(*kb);
    *nn = c.take<int32_t>((size_t)nq * M->m + 1);
    *L = carve_vec(c, nq, M->m, M->d, 0, false, M->precision);
}

size_t nngp_model_predict_workspace_bytes(const nngp_model* model, int64_t max_query) {
    if (!model || max_query < 0) return 0;
    Carver c(nullptr, 0);
    WsHeader* hdr;
    char* kmem;
    size_t kb;
    int32_t* nn;
    VecLayout L;
    carve_model_ws(c, model, max_query, &hdr, &kmem, &kb, &nn, &L);
    return c.used + 256;
}

int nngp_model_predict(const nngp_model* model, const double* x_query, int64_t n_query,
                       double* out_mean, double* out_std, void* workspace, size_t workspace_bytes,
                       void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (!model || n_query < 0 || (n_query > 0 && (!x_query || !out_mean)) || !workspace)
        return NNGP_ERR_INVALID_ARGUMENT;
    const nngp_model* M = model;
    Carver c(workspace, workspace_bytes);
    WsHeader* hdr;
    char* kmem;
    size_t kb;
    int32_t* nn;
    VecLayout L;
    carve_model_ws(c, M, n_query, &hdr, &kmem, &kb, &nn, &L);
    if (!c.ok()) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    L.hdr = hdr;
    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, hdr);
    if (n_query == 0) return NNGP_OK;
    int rc = knn_search(M->refs, x_query, n_query, M->m, 0, 0, nn, kmem, kb, hdr, st);
    if (rc) return rc;
    VecArgs a{};
    a.x = M->x;
    a.y = M->y;
    a.xs = x_query;
    a.ys = nullptr;
    a.nn = nn;
    a.n_src = M->n;
    a.npts = n_query;
    a.d = M->d;
    a.ny = M->ny;
    a.m = M->m;
    a.alpha = M->alpha;
    a.alpha_self = 0;
    a.kp = M->kp;
    a.n_theta = 0;
    a.grad = 0;
    a.mean = out_mean;
    a.stdv = out_std;
    a.hdr = hdr;
    rc = run_fused(a, L, M->precision, st);
    if (rc) return rc;
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

int nngp_read_stats(const void* workspace, nngp_stats* out, void* stream) {
    if (!workspace || !out) return NNGP_ERR_INVALID_ARGUMENT;
    WsHeader h;
    NNGP_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    NNGP_HIP_CHECK(hipMemcpy(&h, workspace, sizeof(h), hipMemcpyDeviceToHost));
    out->knn_rows_recomputed = h.knn_rows_recomputed;
    out->points_refactored_fp64 = h.points_refactored_fp64;
    if (h.bad_index > 0) return NNGP_ERR_INVALID_ARGUMENT;
    return h.not_pd > 0 ? NNGP_ERR_NOT_POSITIVE_DEFINITE : NNGP_OK;
}

}  // extern "C"
