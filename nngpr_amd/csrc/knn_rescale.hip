// Exact neighbour lists at a new length scale from exact lists at an old one.
//
// The reference recomputes the kNN at every optimizer evaluation
// (base_nngpr.py:530-548) because the kernel-space key depends on theta.  Given a
// search's exact top-K lists at scale s0 (K > k, the lists of nngp_knn), the
// top-k at a scale s1 follows from the candidates alone whenever a bound shows
// that no other ref can enter:
//
//   rounding of a key  the computed key of a pair at distance D (in length-scale
//      units) is D^2 (1 + e)^2 with |e| <= eta(D) = 2 u sqrt(d) M / D + (d + 2) u,
//      u = 2^-53, M the largest |coordinate| in the same units (the scaled
//      coordinates are rounded once, the differences and the FMA sum of squares
//      once per term).  M / D does not depend on an isotropic scale.
//   isotropic change (one length scale, or input space): every true distance
//      scales by s0 / s1, so a row whose first k + 1 keys at s0 are separated by
//      more than their rounding bounds (and none is 0) keeps the same list, in
//      the same order, at every isotropic scale: the base pass marks such rows
//      "stable" once, and the rescale copies them.
//   any change: a ref outside the K candidates had key >= T0 (the K-th candidate
//      key) at s0, so its distance at s1 is at least
//      rho sqrt(T0) / (1 + eta0) with rho = min_f s0_f / s1_f, and its computed
//      key at s1 at least that times (1 - eta1).  When the k-th of the
//      candidates re-ranked at s1 is strictly below this bound, the row is exact;
//      otherwise the row is recomputed exactly (one wave per row over its refs).
// Rounding bounds carry a safety factor of 4.  Rows whose candidate list holds
// every admissible ref (an entry -1) are exact at every scale.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "knn.h"
#include "knn_device.h"

namespace nngp {

namespace {

constexpr int kRescaleMaxK = 72;
constexpr double kUnit = 1.1102230246251565e-16;  // 2^-53

struct ScaleArgs {
    int n_scale;                // 0 (raw), 1 (isotropic) or d
    double s[NNGP_MAX_DIM];
};

__device__ __forceinline__ double scaled_coord(const double* x, int64_t i, int d, int f,
                                               const ScaleArgs& s) {
    double v = x[i * d + f];
    if (s.n_scale == 1) v = v / s.s[0];
    else if (s.n_scale > 1) v = v / s.s[f];
    return v;
}

// Largest |coordinate| in length-scale units over the finite coordinates.
__device__ __forceinline__ double max_scaled(const unsigned long long* maxabs, int d,
                                             const ScaleArgs& s) {
    double m = 0.0;
    for (int f = 0; f < d; ++f) {
        const double a = __longlong_as_double((long long)maxabs[f]);
        const double sf = s.n_scale == 0 ? 1.0 : (s.n_scale == 1 ? s.s[0] : s.s[f]);
        m = fmax(m, a / sf);
    }
    return m;
}

// Rounding bound of a key whose distance is about `dist` (safety factor 4).  The
// relative bounds hold away from subnormal keys and from overflow of the squares:
// outside dist > 1e-140 and M < 1e150 no bound is given.
__device__ __forceinline__ double key_eta(double dist, double M, int d) {
    if (!(dist > 1e-140) || !(M < 1e150)) return __builtin_inf();
    return 4.0 * (2.0 * kUnit * sqrt((double)d) * M / (dist * (1.0 - 1e-3)) + (d + 2) * kUnit);
}

struct RescaleState {
    WsHeader* hdr;
    unsigned long long* maxabs;  // d: bits of max |x_f| over finite ref / query coordinates
    double* t0;                  // n_query: K-th candidate key at s0, +inf for complete lists
    int32_t* stable;             // n_query: list unchanged under isotropic rescaling
    double* ur;                  // n_ref x d: refs at s1
    double* uq;                  // n_query x d: queries at s1
    int64_t* flag_rows;
    double* flag_key;
    int32_t* flag_idx;
    unsigned long long* flag_count;
    int64_t* over_rows;
    unsigned long long* over_count;
    double* lkey;
    int32_t* lidx;
};

RescaleState carve_state(Carver& c, int64_t n_ref, int64_t n_query, int d, int k) {
    RescaleState S{};
    S.hdr = c.take<WsHeader>(1);
    S.maxabs = c.take<unsigned long long>(NNGP_MAX_DIM);
    S.t0 = c.take<double>((size_t)n_query + 1);
    S.stable = c.take<int32_t>((size_t)n_query + 1);
    S.ur = c.take<double>((size_t)n_ref * d + 1);
    S.uq = c.take<double>((size_t)n_query * d + 1);
    S.flag_rows = c.take<int64_t>((size_t)n_query + 1);
    S.flag_key = c.take<double>((size_t)n_query + 1);
    S.flag_idx = c.take<int32_t>((size_t)n_query + 1);
    S.flag_count = c.take<unsigned long long>(4);
    S.over_rows = c.take<int64_t>((size_t)n_query + 1);
    S.over_count = c.take<unsigned long long>(4);
    S.lkey = c.take<double>(exact_rows_scratch_count(k));
    S.lidx = c.take<int32_t>(exact_rows_scratch_count(k));
    return S;
}

__global__ void zero_words_kernel(unsigned long long* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ULL;
}

// max |x_f| over the finite coordinates of n points (block maxima, one atomic per
// block and feature; non-negative doubles order like their bit patterns).
__global__ __launch_bounds__(256) void maxabs_kernel(const double* x, int64_t n, int d,
                                                     unsigned long long* maxabs) {
    __shared__ double sm[256];
    for (int f = 0; f < d; ++f) {
        double m = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            const double v = fabs(x[i * d + f]);
            if (__builtin_isfinite(v)) m = fmax(m, v);
        }
        sm[threadIdx.x] = m;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0 && sm[0] > 0.0)
            atomicMax(&maxabs[f], (unsigned long long)__double_as_longlong(sm[0]));
        __syncthreads();
    }
}

struct BaseArgs {
    const double* ref;
    const double* qry;
    int d;
    int64_t n_ref;           // candidate entries outside [0, n_ref) read as empty
    int64_t n_query;
    ScaleArgs s0;
    const int32_t* cand;
    int K;
    int k;
    const unsigned long long* maxabs;
    double* t0;
    int32_t* stable;
};

// Per row: T0 and the isotropic stability of the first k + 1 candidates.
__global__ __launch_bounds__(256) void rescale_base_kernel(BaseArgs a) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n_query) return;
    const int d = a.d;
    double uq[NNGP_MAX_DIM];
    for (int f = 0; f < d; ++f) uq[f] = scaled_coord(a.qry, q, d, f, a.s0);
    const double M = max_scaled(a.maxabs, d, a.s0);
    const int32_t* cq = a.cand + q * a.K;
    auto key_of = [&](int id) {
        double t = scaled_coord(a.ref, id, d, 0, a.s0) - uq[0];
        double key = t * t;
        for (int f = 1; f < d; ++f) {
            t = scaled_coord(a.ref, id, d, f, a.s0) - uq[f];
            key = fma(t, t, key);
        }
        return key;
    };
    auto cid = [&](int j) {
        const int id = cq[j];
        return (id >= 0 && (int64_t)id < a.n_ref) ? id : -1;
    };
    const int last = cid(a.K - 1);
    a.t0[q] = last >= 0 ? key_of(last) : __builtin_inf();
    // pairs (j, j + 1) of the first min(k + 1, count) candidates; with K = k and a
    // full list the (k + 1)-th key is unknown, so the boundary is not certified
    bool stable = !(a.K <= a.k && last >= 0);
    double prev = 0.0;
    for (int j = 0; stable && j <= a.k && j < a.K; ++j) {
        const int id = cid(j);
        if (id < 0) break;  // every admissible ref is listed: nothing beyond
        const double c = key_of(id);
        const double dist = sqrt(c);
        const double e = key_eta(dist, M, d);
        const double lo = dist * (1.0 - e) / (1.0 + e);
        const double hi = dist * (1.0 + e) / (1.0 - e);
        if (!(e < 0.5) || (j > 0 && !(prev < lo))) {
            stable = false;
            break;
        }
        prev = hi;
    }
    a.stable[q] = stable ? 1 : 0;
}

__global__ __launch_bounds__(256) void scale_rows_kernel(const double* x, int64_t n, int d,
                                                         ScaleArgs s, double* u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d) return;
    const int f = (int)(i % d);
    u[i] = scaled_coord(x, i / d, d, f, s);
}

struct RunArgs {
    const double* ur;        // refs at s1 (AoS)
    const double* uq;        // queries at s1 (AoS)
    int d;
    int64_t n_ref;           // candidate entries outside [0, n_ref) read as empty
    int64_t n_query;
    const int32_t* cand;
    int K;
    int k;
    int iso;                 // the change is isotropic: stable rows are copied
    int all_stable;          // identical scales: every row is copied
    double rho;              // min_f s0_f / s1_f
    ScaleArgs s0, s1;
    const unsigned long long* maxabs;
    const double* t0;
    const int32_t* stable;
    int32_t* out_idx;
    int64_t* flag_rows;
    double* flag_key;
    int32_t* flag_idx;
    unsigned long long* flag_count;
};

template <int KP>
__global__ __launch_bounds__(256) void rescale_kernel(RunArgs a) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n_query) return;
    const int32_t* cq = a.cand + q * a.K;
    int32_t* oq = a.out_idx + q * a.k;
    // a stable row keeps its list at any isotropic scale whose coordinates stay
    // inside the range of the rounding bound
    if (a.all_stable ||
        (a.iso && a.stable[q] && max_scaled(a.maxabs, a.d, a.s1) < 1e150)) {
        for (int j = 0; j < a.k; ++j) oq[j] = cq[j];
        return;
    }
    double key[KP];
    int ids[KP];
    int count = 0, nfin = 0;
    const double* uq = a.uq + q * a.d;
#pragma unroll
    for (int t = 0; t < KP; ++t) {
        int id = t < a.K ? cq[t] : -1;
        if ((int64_t)id >= a.n_ref) id = -1;
        key[t] = __builtin_inf();
        if (id >= 0) {
            key[t] = exact_key(a.ur + (int64_t)id * a.d, uq, a.d);
            ++count;
            if (!(key[t] < __builtin_inf())) id = -1;
        }
        ids[t] = id;
        nfin += id >= 0 ? 1 : 0;
    }
    double dm = __builtin_inf();
    int dmi = -1;
#pragma unroll
    for (int t = 0; t < KP; ++t) {
        if (ids[t] < 0) continue;
        int rank = 0;
#pragma unroll
        for (int s = 0; s < KP; ++s)
            rank += (ids[s] >= 0 && pair_less(key[s], ids[s], key[t], ids[t])) ? 1 : 0;
        if (rank < a.k) oq[rank] = ids[t];
        if (rank == a.k - 1) {
            dm = key[t];
            dmi = ids[t];
        }
    }
    for (int j = nfin; j < a.k; ++j) oq[j] = -1;
    if (count < a.K) return;  // every admissible ref was a candidate
    const double T0 = a.t0[q];
    if (!(T0 < __builtin_inf())) return;
    // lower bound of the computed key at s1 of any ref outside the candidates
    const double d0 = sqrt(T0);
    const double e0 = key_eta(d0, max_scaled(a.maxabs, a.d, a.s0), a.d);
    const double d1 = a.rho * d0 / (1.0 + e0);
    const double e1 = key_eta(d1, max_scaled(a.maxabs, a.d, a.s1), a.d);
    const double lb = d1 * (1.0 - e1);
    const bool ok = e0 < 0.5 && e1 < 0.5 && sqrt(dm) < lb;
    if (!ok) {
        const unsigned long long slot = atomicAdd(a.flag_count, 1ULL);
        a.flag_rows[slot] = q;
        a.flag_key[slot] = dm;
        a.flag_idx[slot] = dmi;
    }
}

__global__ void add_flagged_kernel(const unsigned long long* c, WsHeader* h) {
    if (threadIdx.x == 0) h->knn_rows_recomputed += (int64_t)*c;
}

ScaleArgs scale_args(const double* scale, int n_scale) {
    ScaleArgs s{};
    s.n_scale = n_scale;
    for (int f = 0; f < n_scale; ++f) s.s[f] = scale[f];
    return s;
}

int check_args(int64_t n_ref, int64_t n_query, int d, const double* scale, int n_scale, int K,
               int k) {
    if (d < 1 || d > NNGP_MAX_DIM || n_ref < 0 || n_query < 0 || n_ref >= (int64_t)INT32_MAX)
        return NNGP_ERR_INVALID_ARGUMENT;
    if (n_scale != 0 && n_scale != 1 && n_scale != d) return NNGP_ERR_INVALID_ARGUMENT;
    if (n_scale > 0 && !scale) return NNGP_ERR_INVALID_ARGUMENT;
    for (int f = 0; f < n_scale; ++f)
        if (!(scale[f] > 0.0) || !std::isfinite(scale[f])) return NNGP_ERR_INVALID_ARGUMENT;
    if (k < 1 || K < k || K > kRescaleMaxK) return NNGP_ERR_UNSUPPORTED;
    return NNGP_OK;
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)std::max<int64_t>((n + per - 1) / per, 1); }

}  // namespace

int rescale_max_candidates() { return kRescaleMaxK; }

size_t rescale_state_bytes(int64_t n_ref, int64_t n_query, int d, int k, int K) {
    (void)K;
    Carver c(nullptr, 0);
    carve_state(c, n_ref, n_query, d, k);
    return c.used + 256;
}

int rescale_init(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
                 const double* scale, int n_scale, const int32_t* cand, int K, int k, void* state,
                 size_t state_bytes, hipStream_t st) {
    int rc = check_args(n_ref, n_query, d, scale, n_scale, K, k);
    if (rc) return rc;
    Carver c(state, state_bytes);
    RescaleState S = carve_state(c, n_ref, n_query, d, k);
    if (!c.ok() || !state) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, S.hdr);
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st, S.maxabs, NNGP_MAX_DIM);
    if (n_ref > 0)
        hipLaunchKernelGGL(maxabs_kernel, dim3(std::min(blocks_for(n_ref, 256), 256u)), dim3(256),
                           0, st, ref, n_ref, d, S.maxabs);
    if (n_query > 0 && query != ref)
        hipLaunchKernelGGL(maxabs_kernel, dim3(std::min(blocks_for(n_query, 256), 256u)),
                           dim3(256), 0, st, query, n_query, d, S.maxabs);
    if (n_query > 0) {
        BaseArgs b{};
        b.ref = ref;
        b.qry = query;
        b.d = d;
        b.n_ref = n_ref;
        b.n_query = n_query;
        b.s0 = scale_args(scale, n_scale);
        b.cand = cand;
        b.K = K;
        b.k = k;
        b.maxabs = S.maxabs;
        b.t0 = S.t0;
        b.stable = S.stable;
        hipLaunchKernelGGL(rescale_base_kernel, dim3(blocks_for(n_query, 256)), dim3(256), 0, st, b);
    }
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

int rescale_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
                const double* scale_old, const double* scale_new, int n_scale, int has_cut,
                int64_t ref_cut, const int32_t* cand, int K, int k, int32_t* out_idx, void* state,
                size_t state_bytes, hipStream_t st) {
    int rc = check_args(n_ref, n_query, d, scale_old, n_scale, K, k);
    if (rc) return rc;
    rc = check_args(n_ref, n_query, d, scale_new, n_scale, K, k);
    if (rc) return rc;
    Carver c(state, state_bytes);
    RescaleState S = carve_state(c, n_ref, n_query, d, k);
    if (!c.ok() || !state) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, S.hdr);
    if (n_query == 0) {
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    RunArgs a{};
    a.s0 = scale_args(scale_old, n_scale);
    a.s1 = scale_args(scale_new, n_scale);
    bool same_scale = true;
    double rho = n_scale == 0 ? 1.0 : HUGE_VAL;
    for (int f = 0; f < n_scale; ++f) {
        same_scale = same_scale && scale_old[f] == scale_new[f];
        rho = std::min(rho, scale_old[f] / scale_new[f]);
    }
    // the division is rounded: shrink rho by a relative 1e-15 so it stays a bound
    a.rho = rho * (1.0 - 1e-15);
    a.iso = n_scale <= 1 ? 1 : 0;
    a.all_stable = same_scale ? 1 : 0;
    a.d = d;
    a.n_ref = n_ref;
    a.n_query = n_query;
    a.cand = cand;
    a.K = K;
    a.k = k;
    a.maxabs = S.maxabs;
    a.t0 = S.t0;
    a.stable = S.stable;
    a.out_idx = out_idx;
    a.flag_rows = S.flag_rows;
    a.flag_key = S.flag_key;
    a.flag_idx = S.flag_idx;
    a.flag_count = S.flag_count;
    const bool same = query == ref && n_query == n_ref;
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st, S.flag_count, 1);
    if (n_ref > 0)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks_for(n_ref * d, 256)), dim3(256), 0, st,
                           ref, n_ref, d, a.s1, S.ur);
    const double* uq = S.ur;
    if (!same) {
        hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks_for(n_query * d, 256)), dim3(256), 0, st,
                           query, n_query, d, a.s1, S.uq);
        uq = S.uq;
    }
    a.ur = S.ur;
    a.uq = uq;
    const unsigned g = blocks_for(n_query, 256);
    switch ((K + 7) / 8 * 8) {
        case 8: hipLaunchKernelGGL(rescale_kernel<8>, dim3(g), dim3(256), 0, st, a); break;
        case 16: hipLaunchKernelGGL(rescale_kernel<16>, dim3(g), dim3(256), 0, st, a); break;
        case 24: hipLaunchKernelGGL(rescale_kernel<24>, dim3(g), dim3(256), 0, st, a); break;
        case 32: hipLaunchKernelGGL(rescale_kernel<32>, dim3(g), dim3(256), 0, st, a); break;
        case 40: hipLaunchKernelGGL(rescale_kernel<40>, dim3(g), dim3(256), 0, st, a); break;
        case 48: hipLaunchKernelGGL(rescale_kernel<48>, dim3(g), dim3(256), 0, st, a); break;
        case 56: hipLaunchKernelGGL(rescale_kernel<56>, dim3(g), dim3(256), 0, st, a); break;
        case 64: hipLaunchKernelGGL(rescale_kernel<64>, dim3(g), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL(rescale_kernel<72>, dim3(g), dim3(256), 0, st, a); break;
    }
    ExactRows er{};
    er.rows = S.flag_rows;
    er.count = S.flag_count;
    er.bound_key = S.flag_key;
    er.bound_idx = S.flag_idx;
    er.over_rows = S.over_rows;
    er.over_count = S.over_count;
    er.lkey = S.lkey;
    er.lidx = S.lidx;
    launch_exact_rows(S.ur, uq, d, n_ref, n_query, has_cut, ref_cut, k, er, out_idx, st);
    hipLaunchKernelGGL(add_flagged_kernel, dim3(1), dim3(64), 0, st, S.flag_count, S.hdr);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

}  // namespace nngp
