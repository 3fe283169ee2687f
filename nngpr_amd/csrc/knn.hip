// Brute-force k-nearest-neighbour search for the NNGP engine (gfx950).
//
// Key of a (query, ref) pair: FP64 squared distance of the scaled coordinates,
// key = fma(t_{d-1}, t_{d-1}, ... fma(t_1, t_1, t_0 * t_0)), t_f = u_ref,f - u_query,f;
// order = ascending (key, ref index).  This is the reference's calc_nn ordering
// (cuda_kernel_accelerators.py:711-742) evaluated in FP64.
//
// Two search forms produce identical indices:
//   * exact:  one lane per query streams the refs (wave-uniform, read through the
//             scalar cache), keeps a sorted top-K in VGPRs and an append buffer in
//             LDS; FP64 keys throughout.
//   * screen: the same streaming loop in FP32 on centred coordinates with the
//             expansion |r|^2 - 2 q.r (3 FMAs + 1 compare per pair for d = 3),
//             keeping K = m + margin candidates; a re-rank pass computes the exact
//             FP64 keys of the candidates, takes the top m and certifies the row
//             with a rigorous bound on the FP32 error.  Rows that fail the
//             certificate are recomputed exactly by the per-row fallback kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "common.h"
#include "knn.h"
#include "knn_device.h"

namespace nngp {

namespace {

constexpr int kFallbackThreads = 16384;

struct PrepArgs {
    const double* x;   // n x d row-major
    int64_t n;
    int d;
    int n_scale;
    double scale[NNGP_MAX_DIM];
};


__device__ __forceinline__ double scaled(const PrepArgs& a, int64_t i, int f) {
    double v = a.x[i * a.d + f];
    if (a.n_scale == 1) v = v / a.scale[0];
    else if (a.n_scale > 1) v = v / a.scale[f];
    return v;
}

constexpr int kCentreBlocks = 256;

__global__ void init_ref_kernel(double* centre, int d, unsigned int* smax) {
    int f = threadIdx.x;
    if (f < d) centre[f] = 0.0;
    if (f == 0) *smax = 0u;
}

// Centre of the FP32 screen: per feature, the mean of the finite scaled
// coordinates.  The mean (unlike the box midpoint) moves by only D / n for a point
// at distance D, so a few far points leave the centred coordinates of the others
// small.  Each block writes (sum of v / n, finite count) per feature; the final
// kernel adds the block partials in block order, so the centre is deterministic.
__global__ __launch_bounds__(256) void centre_partial_kernel(PrepArgs a, double* part) {
    __shared__ double ssum[256], scnt[256];
    const double inv_n = 1.0 / (double)a.n;
    for (int f = 0; f < a.d; ++f) {
        double sum = 0.0, cnt = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
             i += (int64_t)gridDim.x * blockDim.x) {
            const double v = scaled(a, i, f);
            if (__builtin_isfinite(v)) {  // NaN and infinite coordinates do not enter
                sum += v * inv_n;
                cnt += 1.0;
            }
        }
        ssum[threadIdx.x] = sum;
        scnt[threadIdx.x] = cnt;
        __syncthreads();
        for (int st = blockDim.x / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) {
                ssum[threadIdx.x] += ssum[threadIdx.x + st];
                scnt[threadIdx.x] += scnt[threadIdx.x + st];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[(blockIdx.x * a.d + f) * 2] = ssum[0];
            part[(blockIdx.x * a.d + f) * 2 + 1] = scnt[0];
        }
        __syncthreads();
    }
}

__global__ void centre_final_kernel(const double* part, int nb, int d, int64_t n,
                                    double* centre) {
    const int f = threadIdx.x;
    if (f >= d) return;
    double sum = 0.0, cnt = 0.0;
    for (int b = 0; b < nb; ++b) {
        sum += part[(b * d + f) * 2];
        cnt += part[(b * d + f) * 2 + 1];
    }
    // sum = (sum of the finite v) / n: no overflow for finite v
    const double c = cnt > 0.0 ? sum * ((double)n / cnt) : 0.0;
    centre[f] = __builtin_isfinite(c) ? c : 0.0;
}

// Index of feature f of point i in the blocked layout used by the scans: points
// are grouped in blocks of 8, each block stores its nv feature rows contiguously
// (f * 8 + lane), so one wave-uniform block is a few wide scalar loads.
__host__ __device__ __forceinline__ int64_t blk8(int64_t i, int f, int nv) {
    return (i >> 3) * (int64_t)(nv * 8) + f * 8 + (i & 7);
}

// Writes, for points [0, n) (entries up to npad are padding):
//   u64  (n x d AoS): scaled FP64 coordinates (u = x / scale)
//   s64  (blocked, nv = dp): same values for the exact scan; padding points +inf
//   s32  (blocked, nv = dp + 1): fl32(u - centre), and -|fl32(u - centre)|^2 as
//        feature dp; padding points have -inf there so they never pass a test
//   p2   (n, optional): |u - centre|^2 in FP64
//   smax (optional): max_i |u_i - centre| over these points (atomic, float bits)
// Features d .. dp-1 are zero: they leave every key unchanged.
__global__ void prep_kernel(PrepArgs a, const double* centre, int64_t npad, int dp, double* u64,
                            double* s64, float* s32, double* p2, unsigned int* smax) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= a.n) {
        for (int f = 0; f < dp; ++f) {
            if (s64) s64[blk8(i, f, dp)] = f == 0 ? __builtin_inf() : 0.0;
            if (s32) s32[blk8(i, f, dp + 1)] = 0.0f;
        }
        if (s32) s32[blk8(i, dp, dp + 1)] = -__builtin_inff();
        return;
    }
    for (int f = a.d; f < dp; ++f) {
        if (s64) s64[blk8(i, f, dp)] = 0.0;
        if (s32) s32[blk8(i, f, dp + 1)] = 0.0f;
    }
    double nrm = 0.0, pp = 0.0;
    for (int f = 0; f < a.d; ++f) {
        double v = scaled(a, i, f);
        if (u64) u64[i * a.d + f] = v;
        if (s64) s64[blk8(i, f, dp)] = v;
        if (s32) {
            double p = v - centre[f];
            float pf = (float)p;
            s32[blk8(i, f, dp + 1)] = pf;
            nrm += (double)pf * (double)pf;
            pp += p * p;
        }
    }
    if (s32) {
        s32[blk8(i, dp, dp + 1)] = -(float)nrm;
        if (p2) p2[i] = pp;
        if (smax) {
            // a point with a non-finite coordinate is never admitted by the
            // screen, so it does not enter the radius bound
            float r = (float)sqrt(fmax(nrm, pp)) * 1.0001f;
            if (__builtin_isfinite(r)) atomicMax(smax, __float_as_uint(r));
        }
    }
}

// |u_i - centre|^2 in FP64 for points whose scaled coordinates are already in u64.
__global__ void p2_kernel(const double* u64, int64_t n, int d, const double* centre, double* p2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pp = 0.0;
    for (int f = 0; f < d; ++f) {
        double p = u64[i * d + f] - centre[f];
        pp += p * p;
    }
    p2[i] = pp;
}

// ---------------------------------------------------------------------------
// Streaming scan (exact FP64 or FP32 screen)
// ---------------------------------------------------------------------------
template <typename T>
struct ScanArgs {
    const T* ref;        // blocked layout (blk8), nv = D (exact) or D + 1 (screen)
    int32_t n_ref;
    const T* qry;        // blocked layout, same nv
    int64_t n_query;
    int has_cut;
    int64_t cut;
    int reverse;         // process the longest rows first
    int k_out;           // exact: neighbours written per row (<= K)
    int32_t* out_idx;    // exact: n_query x k_out ; screen: n_query x K candidates
    float* out_thr;      // screen: K-th candidate score per row
    // screen with gridDim.y > 1 ref stripes: stripe s scans refs
    // [s * stripe_w, (s + 1) * stripe_w) (the last stripe to the end) and writes
    // its sorted packed list to out_pk[(s * n_query + q) * K ...] instead of
    // out_idx / out_thr; stripe_merge_kernel merges the stripes
    int32_t stripe_w;    // multiple of 8 (0: one stripe)
    uint64_t* out_pk;
    // exact scan with ref stripes: stripe s writes its sorted (key, index) list
    // to out_skey / out_sidx[(s * n_query + q) * K ...]; stripe_merge_exact_kernel
    // writes the k_out smallest pairs
    T* out_skey;
    int32_t* out_sidx;
};


// Score of ref u of an 8-ref block for one lane's query.
//   screen: -thr + sum_f (-2 q_f) r_f, compared against -|r|^2 (feature D);
//   exact:  sum_f (r_f - q_f)^2 with the FMA order of the reference key.
template <typename T, int D, bool SCREEN>
__device__ __forceinline__ T block_score(const T* blk, int u, const T (&qr)[D], T negthr) {
    if constexpr (SCREEN) {
        T acc = negthr;
#pragma unroll
        for (int f = 0; f < D; ++f) acc = fma(qr[f], blk[f * 8 + u], acc);
        return acc;
    } else {
        T t0 = blk[u] - qr[0];
        T acc = t0 * t0;
#pragma unroll
        for (int f = 1; f < D; ++f) {
            T t = blk[f * 8 + u] - qr[f];
            acc = fma(t, t, acc);
        }
        return acc;
    }
}

template <typename T, int D, int K, int B, bool SCREEN>
__global__ __launch_bounds__(64) void scan_kernel(ScanArgs<T> a) {
    constexpr int NV = SCREEN ? D + 1 : D;
    __shared__ T sbk[B][64];
    __shared__ int sbi[B][64];
    const int lane = threadIdx.x;
    const int64_t blk = a.reverse ? (int64_t)(gridDim.x - 1 - blockIdx.x) : (int64_t)blockIdx.x;
    const int64_t q = blk * 64 + lane;
    const bool qv = q < a.n_query;
    // ref stripe [rs, re) of this block (the whole range with one stripe)
    const bool striped = gridDim.y > 1;
    const int32_t rs = striped ? (int32_t)blockIdx.y * a.stripe_w : 0;
    const int64_t re = (striped && blockIdx.y + 1 < gridDim.y) ? (int64_t)rs + a.stripe_w
                                                                : (int64_t)a.n_ref;
    int64_t lim64 = a.n_ref;
    if (a.has_cut) {
        lim64 = q + a.cut;
        lim64 = lim64 < 0 ? 0 : (lim64 > a.n_ref ? a.n_ref : lim64);
    }
    lim64 = lim64 < re ? lim64 : re;
    const int32_t lim = qv ? (int32_t)lim64 : 0;
    const int32_t lim_lo = wave_min32(qv ? lim : (int32_t)re);
    const int32_t lim_hi = wave_max32(lim);
    const int32_t full_end = lim_lo & ~7;  // bulk region [rs, full_end), whole 8-ref blocks

    T qr[D];
    const int64_t qc = qv ? q : 0;
#pragma unroll
    for (int f = 0; f < D; ++f) {
        T v = a.qry[blk8(qc, f, NV)];
        qr[f] = SCREEN ? (T)(-2) * v : v;
    }
    // Sorted top-K list: (key, index) pairs (exact), or packed (score, index)
    // uint64 values (screen).
    constexpr int KE = SCREEN ? 1 : K;
    constexpr int KP = SCREEN ? K : 1;
    T keys[KE];
    int ids[KE];
    uint64_t pk[KP];
#pragma unroll
    for (int t = 0; t < KE; ++t) {
        keys[t] = dev_inf<T>();
        ids[t] = -1;
    }
    // Empty slot: (+inf, index -1); its unpacked score is +inf, never NaN.
#pragma unroll
    for (int t = 0; t < KP; ++t) pk[t] = kPackedEmpty;
    // Admission threshold: screen compares score - thr < -|r|^2, exact key < thr.
    T thr = qv ? dev_inf<T>() : -dev_inf<T>();
    T negthr = -thr;
    int cnt = 0;

    auto merge = [&]() {
#pragma unroll 1
        for (int b = 0; b < B; ++b) {
            if (b < cnt) {
                T xk = sbk[b][lane];
                int xi = sbi[b][lane];
                if constexpr (SCREEN) {
                    const uint64_t x = pack_key((float)xk, xi);
                    if (x < pk[K - 1]) topk_insert_packed<K>(pk, x);
                } else {
                    if (xk < keys[K - 1]) topk_insert<T, K>(keys, ids, xk, xi);
                }
            }
        }
        cnt = 0;
        if constexpr (SCREEN) {
            thr = qv ? (T)unord_f32((uint32_t)(pk[K - 1] >> 32)) : -dev_inf<T>();
        } else {
            thr = qv ? keys[KE - 1] : -dev_inf<T>();
        }
        negthr = -thr;
    };

    // Appends the lanes' hits of one block (rare path); the buffer is merged
    // before an append could overflow it.
    auto take_block = [&](const T* bp, int32_t r0, int nref, bool masked) {
#pragma unroll 1
        for (int u = 0; u < nref; ++u) {
            if (__any(cnt >= B)) merge();
            const int32_t r = r0 + u;
            T sc = block_score<T, D, SCREEN>(bp, u, qr, negthr);
            bool hit = SCREEN ? (sc < bp[D * 8 + u]) : (sc < thr);
            if (masked) hit = hit && (r < lim);
            if (hit) {
                T st;
                if constexpr (SCREEN) {
                    st = -bp[D * 8 + u];
#pragma unroll
                    for (int f = 0; f < D; ++f) st = fma(qr[f], bp[f * 8 + u], st);
                } else {
                    st = sc;
                }
                sbk[cnt][lane] = st;
                sbi[cnt][lane] = r;
                ++cnt;
            }
        }
    };

    // Bulk: every lane may take every ref.  Refs are read as sub-blocks of SUB
    // refs through wave-uniform scalar loads into two ping-pong register sets,
    // the next sub-block prefetched while the current one is scored (the padded
    // layout keeps a prefetch past full_end in bounds).  The per-ref test value
    // A sub-block with no hit in any lane costs its scores and one branch.
    // refs per sub-block: the two ping-pong sets stay within the scalar registers
    constexpr int SUB = (NV * sizeof(T) <= 16) ? 8
                        : (NV * sizeof(T) <= 40) ? 4
                        : (NV * sizeof(T) <= 72) ? 2 : 1;
    auto sub_ptr = [&](int32_t r) { return a.ref + (int64_t)(r >> 3) * (NV * 8) + (r & 7); };
    auto load_sub = [&](T (&rv)[NV][SUB], const T* bp) {
#pragma unroll
        for (int f = 0; f < NV; ++f)
#pragma unroll
            for (int u = 0; u < SUB; ++u) rv[f][u] = bp[f * 8 + u];
    };
    auto score_sub = [&](const T (&rv)[NV][SUB], int32_t r0) {
        bool hm[SUB];
        T kv[SUB];
        bool any = false;
#pragma unroll
        for (int u = 0; u < SUB; ++u) {
            if constexpr (SCREEN) {
                T acc = negthr;
#pragma unroll
                for (int f = 0; f < D; ++f) acc = fma(qr[f], rv[f][u], acc);
                hm[u] = acc < rv[D][u];
            } else {
                T t0 = rv[0][u] - qr[0];
                T acc = t0 * t0;
#pragma unroll
                for (int f = 1; f < D; ++f) {
                    T t = rv[f][u] - qr[f];
                    acc = fma(t, t, acc);
                }
                kv[u] = acc;
                hm[u] = acc < thr;
            }
            any = any | hm[u];
        }
        if (__builtin_amdgcn_ballot_w64(any)) {
            // Room for this sub-block's hits: at most SUB per lane.
            if (__any(cnt > B - SUB)) merge();
#pragma unroll
            for (int u = 0; u < SUB; ++u) {
                if (hm[u]) {
                    T st;
                    if constexpr (SCREEN) {
                        st = -rv[D][u];
#pragma unroll
                        for (int f = 0; f < D; ++f) st = fma(qr[f], rv[f][u], st);
                    } else {
                        st = kv[u];
                    }
                    sbk[cnt][lane] = st;
                    sbi[cnt][lane] = r0 + u;
                    ++cnt;
                }
            }
        }
    };
    {
        T ra[NV][SUB], rb[NV][SUB];
        load_sub(ra, sub_ptr(rs));
        int32_t r0 = rs;
        while (r0 < full_end) {
            load_sub(rb, sub_ptr(r0 + SUB));
            score_sub(ra, r0);
            r0 += SUB;
            if (r0 >= full_end) break;
            load_sub(ra, sub_ptr(r0 + SUB));
            score_sub(rb, r0);
            r0 += SUB;
        }
    }
    // Tail: refs past some lanes' Vecchia limit.
    for (int32_t r0 = full_end > rs ? full_end : rs; r0 < lim_hi; r0 += 8) {
        const T* bp = a.ref + (int64_t)(r0 >> 3) * (NV * 8);
        take_block(bp, r0, 8, true);
    }
    if (__any(cnt > 0)) merge();
    if (!qv) return;
    if constexpr (SCREEN) {
        if (striped) {
            uint64_t* o = a.out_pk + ((int64_t)blockIdx.y * a.n_query + q) * K;
#pragma unroll
            for (int t = 0; t < K; ++t) o[t] = pk[t];
        } else {
#pragma unroll
            for (int t = 0; t < K; ++t) a.out_idx[q * K + t] = (int)(uint32_t)pk[t];
            a.out_thr[q] = unord_f32((uint32_t)(pk[K - 1] >> 32));
        }
    } else {
        if (striped) {
            const int64_t o = ((int64_t)blockIdx.y * a.n_query + q) * K;
#pragma unroll
            for (int t = 0; t < KE; ++t) {
                a.out_skey[o + t] = keys[t];
                a.out_sidx[o + t] = ids[t];
            }
        } else {
#pragma unroll
            for (int t = 0; t < KE; ++t)
                if (t < a.k_out) a.out_idx[q * a.k_out + t] = ids[t];
        }
    }
}

// Merge of the ref stripes' screen lists: per row, the K smallest packed
// (score, index) values of the stripes' sorted lists, as candidates and the K-th
// score.  Each stripe list holds its stripe's K smallest, so the merge holds the
// K smallest over all refs, and every ref left out scored at least the merged
// K-th value: the certificate of the unstriped screen applies unchanged.
constexpr int kMaxStripes = 8;
template <int K>
__global__ __launch_bounds__(256) void stripe_merge_kernel(const uint64_t* pk, int n_stripes,
                                                           int64_t n_query, int32_t* cand,
                                                           float* thr) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_query) return;
    int h[kMaxStripes];
#pragma unroll
    for (int s = 0; s < kMaxStripes; ++s) h[s] = 0;
    uint64_t last = 0;
    for (int t = 0; t < K; ++t) {
        uint64_t best = ~0ull;
        int bs = 0;
#pragma unroll
        for (int s = 0; s < kMaxStripes; ++s) {
            if (s < n_stripes && h[s] < K) {
                const uint64_t v = pk[((int64_t)s * n_query + q) * K + h[s]];
                if (v < best) {
                    best = v;
                    bs = s;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kMaxStripes; ++s) h[s] += s == bs ? 1 : 0;
        cand[q * K + t] = (int)(uint32_t)best;
        last = best;
    }
    thr[q] = unord_f32((uint32_t)(last >> 32));
}

// Merge of the exact scan's stripe lists: per row, the k_out smallest (key,
// index) pairs of the stripes' sorted lists (empty slots (+inf, -1) last).
template <int K>
__global__ __launch_bounds__(256) void stripe_merge_exact_kernel(const double* sk,
                                                                 const int32_t* si,
                                                                 int n_stripes, int64_t n_query,
                                                                 int k_out, int32_t* out_idx) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_query) return;
    int h[kMaxStripes];
#pragma unroll
    for (int s = 0; s < kMaxStripes; ++s) h[s] = 0;
    for (int t = 0; t < k_out; ++t) {
        double bk = __builtin_inf();
        uint32_t bi = 0xFFFFFFFFu;
        int bs = -1;
#pragma unroll
        for (int s = 0; s < kMaxStripes; ++s) {
            if (s < n_stripes && h[s] < K) {
                const int64_t o = ((int64_t)s * n_query + q) * K + h[s];
                const double kk = sk[o];
                const uint32_t ii = (uint32_t)si[o];
                if (bs < 0 || kk < bk || (kk == bk && ii < bi)) {
                    bk = kk;
                    bi = ii;
                    bs = s;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kMaxStripes; ++s) h[s] += s == bs ? 1 : 0;
        out_idx[q * k_out + t] = (int32_t)bi;
    }
}

// ---------------------------------------------------------------------------
// Re-rank of screened candidates in FP64 + certificate
// ---------------------------------------------------------------------------
struct RerankArgs {
    const double* ru64;   // ref AoS scaled
    const double* qu64;   // query AoS scaled
    const double* qp2;    // |u_q - c|^2
    const unsigned int* smax_bits;
    int d;
    int64_t n_query;
    int k;                // neighbours to output
    const int32_t* cand;  // n_query x K
    const float* thr;
    int32_t* out_idx;     // n_query x k
    int64_t* flag_rows;   // flagged rows list
    unsigned long long* flag_count;
    double* flag_key;     // per flagged slot: (key, index) of the m-th candidate,
    int32_t* flag_idx;    //   an upper bound of the row's true m-th pair
    double eps_factor;    // (4 d + 16) * 2^-24
};


// Certificate of a screened row whose K candidates were all kept.  A
// non-candidate r, of centred norm rho (the larger of its exact and FP32-rounded
// norms, rho <= S, the largest over the refs), had a screen score >= thr, so its
// exact key is (the rounding of the screen's score and of its threshold
// comparison, E = eps_factor)
//   >= f(rho) = thr + |p|^2 - (E + 1e-12) (|thr| + (|p| + rho)^2),
// and by the reverse triangle inequality (the exact norm is >= rho / 1.0002)
//   >= g(rho) = (rho / 1.0002 - |p|)^2 - 1e-12 (|p| + rho)^2  for rho >= 1.0002 |p|.
// f decreases, and g increases on [1.0003 |p|, inf), so for any rc in
// [1.0003 |p|, S) every non-candidate key is >= min(f(rc), g(rc)) (rc is taken
// at the crossing of f and g); if the crossing or 1.0003 |p| is >= S it is >= f(S).
// Far refs then raise S without loosening the bound of the rows near the centre.
// The row's top m is exact when its m-th candidate key is strictly below this.
__device__ __forceinline__ double rerank_bound(const RerankArgs& a, int64_t q) {
    const double s = (double)__uint_as_float(*a.smax_bits);
    const double p2 = a.qp2[q];
    const double p = sqrt(p2);
    const double cc = a.eps_factor + 1e-12;
    const double A = (double)a.thr[q] + p2 - cc * fabs((double)a.thr[q]);
    const double rlo = 1.0003 * p;
    if (!(s > rlo)) return A - cc * (p + s) * (p + s);
    // f(r) = (r - p)^2  <=>  (1 + cc) r^2 - 2 p (1 - cc) r + (1 + cc) p^2 - A = 0
    const double disc = (1.0 + cc) * A - 4.0 * cc * p2;
    double rc = disc > 0.0 ? (p * (1.0 - cc) + sqrt(disc)) / (1.0 + cc) : rlo;
    rc = fmax(rc, rlo);
    if (!(rc < s)) return A - cc * (p + s) * (p + s);  // every ref has rho <= rc: f alone
    const double t = rc / 1.0002 - p;
    const double m2 = (p + rc) * (p + rc);
    return fmin(A - cc * m2, t * t - 1e-12 * m2);
}

// One wave per row: lane j re-keys candidates j and j + 64 in FP64 (the exact
// key, gathers of all candidates in flight together), ranks its candidates among
// all K by (key, index) through shuffles, writes the top k and certifies the row;
// uncertified rows are listed with their m-th pair for the exact recompute.
template <int K>
__global__ __launch_bounds__(256) void rerank_kernel(RerankArgs a) {
    constexpr int J = (K + 63) / 64;   // candidate slots per lane
    const int lane = threadIdx.x & 63;
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (q >= a.n_query) return;        // uniform over the wave
    const double* uq = a.qu64 + q * a.d;
    double key[J];
    int ids[J];
    int count = 0;  // candidates the screen kept
    int nfin = 0;   // of which with a finite exact key (the others are never neighbours)
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int t = lane + 64 * j;
        int id = t < K ? a.cand[q * K + t] : -1;
        count += __popcll(__ballot(id >= 0));
        key[j] = __builtin_inf();
        if (id >= 0) {
            key[j] = exact_key(a.ru64 + (int64_t)id * a.d, uq, a.d);
            if (!(key[j] < __builtin_inf())) id = -1;
        }
        ids[j] = id;
        nfin += __popcll(__ballot(id >= 0));
    }
    int rank[J];
#pragma unroll
    for (int j = 0; j < J; ++j) rank[j] = 0;
#pragma unroll
    for (int js = 0; js < J; ++js) {
        const int smax = js + 1 < J ? 64 : K - 64 * js;
#pragma unroll 4
        for (int sl = 0; sl < smax; ++sl) {
            const double ks = __shfl(key[js], sl, 64);
            const int is = __shfl(ids[js], sl, 64);
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const bool before = is >= 0 && (ks < key[j] || (ks == key[j] && is < ids[j]));
                rank[j] += before ? 1 : 0;
            }
        }
    }
    // m-th (key, index) pair; +inf when fewer than m finite keys (never certified)
    double dm = __builtin_inf();
    int dmi = -1;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (ids[j] >= 0 && rank[j] < a.k) a.out_idx[q * a.k + rank[j]] = ids[j];
        const uint64_t m = __ballot(ids[j] >= 0 && rank[j] == a.k - 1);
        if (m) {
            const int src = __ffsll((unsigned long long)m) - 1;
            dm = __shfl(key[j], src, 64);
            dmi = __shfl(ids[j], src, 64);
        }
    }
    for (int jj = nfin + lane; jj < a.k; jj += 64) a.out_idx[q * a.k + jj] = -1;
    if (count < K) return;  // every ref of the row was a candidate: exact
    const double bound = rerank_bound(a, q);
    if (lane == 0 && !(dm < bound)) {
        unsigned long long slot = atomicAdd(a.flag_count, 1ULL);
        a.flag_rows[slot] = q;
        a.flag_key[slot] = dm;
        a.flag_idx[slot] = dmi;
    }
}

// ---------------------------------------------------------------------------
// Exact recomputation of flagged rows.  One block per row: its threads stream the
// row's refs (each keying kFbRefs refs per step, their loads in flight together),
// keep every (key, index) <= the row's bound pair in LDS, then rank the kept
// entries.  The true top-k pairs are all <= the bound, so they are all kept.  Rows
// that keep more than kFbCap entries (massive exact ties) go to the per-row serial
// kernel instead.
// ---------------------------------------------------------------------------
struct WaveFbArgs {
    const double* ru64;
    const double* qu64;
    int d;
    int64_t n_ref;
    int has_cut;
    int64_t cut;
    int k;
    const int64_t* rows;
    const unsigned long long* row_count;
    const double* bound_key;
    const int32_t* bound_idx;
    int32_t* out_idx;
    int64_t* over_rows;                 // rows handed to the serial kernel
    unsigned long long* over_count;
};

constexpr int kFbCap = 1024;
constexpr int kFbThreads = 256;
constexpr int kFbRefs = 4;   // refs each thread keys per step (their loads in flight together)

__global__ __launch_bounds__(kFbThreads) void row_fallback_kernel(WaveFbArgs a) {
    __shared__ double skey[kFbCap];
    __shared__ int sidx[kFbCap];
    __shared__ int scnt;
    const int tid = threadIdx.x;
    const int64_t nrows = (int64_t)*a.row_count;
    for (int64_t slot = blockIdx.x; slot < nrows; slot += gridDim.x) {
        const int64_t q = a.rows[slot];
        const double bk = a.bound_key[slot];
        const int bi = a.bound_idx[slot];
        int64_t lim = a.n_ref;
        if (a.has_cut) {
            lim = q + a.cut;
            lim = lim < 0 ? 0 : (lim > a.n_ref ? a.n_ref : lim);
        }
        if (tid == 0) scnt = 0;
        __syncthreads();
        const double* uq = a.qu64 + q * a.d;
        for (int64_t r0 = 0; r0 < lim; r0 += (int64_t)kFbThreads * kFbRefs) {
            double key[kFbRefs];
#pragma unroll
            for (int u = 0; u < kFbRefs; ++u) {
                const int64_t r = r0 + (int64_t)u * kFbThreads + tid;
                key[u] = r < lim ? exact_key(a.ru64 + r * a.d, uq, a.d) : __builtin_inf();
            }
#pragma unroll
            for (int u = 0; u < kFbRefs; ++u) {
                const int64_t r = r0 + (int64_t)u * kFbThreads + tid;
                if (r < lim && (key[u] < bk || (key[u] == bk && (int)r <= bi))) {
                    const int pos = atomicAdd(&scnt, 1);
                    if (pos < kFbCap) {
                        skey[pos] = key[u];
                        sidx[pos] = (int)r;
                    }
                }
            }
        }
        __syncthreads();
        const int cnt = scnt;
        if (cnt > kFbCap) {
            if (tid == 0) {
                unsigned long long o = atomicAdd(a.over_count, 1ULL);
                a.over_rows[o] = q;
            }
        } else {
            for (int e = tid; e < cnt; e += kFbThreads) {
                const double ke = skey[e];
                const int ie = sidx[e];
                int rank = 0;
                for (int f = 0; f < cnt; ++f) {
                    const double kf = skey[f];
                    const int ifx = sidx[f];
                    rank += (kf < ke || (kf == ke && ifx < ie)) ? 1 : 0;
                }
                if (rank < a.k) a.out_idx[q * a.k + rank] = ie;
            }
            for (int j = cnt + tid; j < a.k; j += kFbThreads) a.out_idx[q * a.k + j] = -1;
        }
        __syncthreads();  // the list is reused by the block's next row
    }
}

// ---------------------------------------------------------------------------
// Per-row exact fallback (one thread per row, sorted list in global memory).
// Used for rows that failed the certificate, and for k > 64 or d > 32.
// ---------------------------------------------------------------------------
struct SimpleArgs {
    const double* ru64;
    const double* qu64;
    int d;
    int64_t n_ref;
    int64_t n_query;
    int has_cut;
    int64_t cut;
    int k;
    const int64_t* rows;                   // NULL: row = slot
    const unsigned long long* row_count;   // NULL: n_query rows
    double* lkey;                          // k x nthreads
    int32_t* lidx;                         // k x nthreads
    int32_t* out_idx;                      // n_query x k
};

__global__ __launch_bounds__(256) void simple_knn_kernel(SimpleArgs a) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    const int64_t nrows = a.row_count ? (int64_t)*a.row_count : a.n_query;
    for (int64_t slot = tid; slot < nrows; slot += nth) {
        const int64_t q = a.rows ? a.rows[slot] : slot;
        int64_t lim = a.n_ref;
        if (a.has_cut) {
            lim = q + a.cut;
            lim = lim < 0 ? 0 : (lim > a.n_ref ? a.n_ref : lim);
        }
        for (int j = 0; j < a.k; ++j) {
            a.lkey[j * nth + tid] = __builtin_inf();
            a.lidx[j * nth + tid] = -1;
        }
        double kth = __builtin_inf();
        const double* uq = a.qu64 + q * a.d;
        for (int64_t r = 0; r < lim; ++r) {
            double key = exact_key(a.ru64 + r * a.d, uq, a.d);
            if (key < kth) {
                int ii = a.k - 1;
                while (ii > 0) {
                    double prev = a.lkey[(ii - 1) * nth + tid];
                    if (prev <= key) break;
                    a.lkey[ii * nth + tid] = prev;
                    a.lidx[ii * nth + tid] = a.lidx[(ii - 1) * nth + tid];
                    --ii;
                }
                a.lkey[ii * nth + tid] = key;
                a.lidx[ii * nth + tid] = (int32_t)r;
                kth = a.lkey[(a.k - 1) * nth + tid];
            }
        }
        for (int j = 0; j < a.k; ++j) a.out_idx[q * a.k + j] = a.lidx[j * nth + tid];
    }
}

__global__ void add_count_kernel(const unsigned long long* src, int64_t* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst += (int64_t)*src;
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exact top-k (k <= 64) of query rows [qlo, qlo + gridDim.x) whose ref limit
// (row + cut, at most a few hundred refs) is below the first grid band: one wave
// per row, the refs in chunks of 64 lanes (exact_key), each chunk sorted by (key,
// index) with a bitonic network and merged into the row's 64 smallest so far.
// Refs or queries with a non-finite coordinate give no key (never neighbours);
// missing neighbours are -1.
__global__ __launch_bounds__(64) void prefix_knn_kernel(const double* ru64, const double* qu64,
                                                        int d, int64_t qlo, int64_t cut,
                                                        int64_t n_ref, int k, int32_t* out_idx) {
    const int64_t i = qlo + blockIdx.x;
    const int lane = threadIdx.x;
    const double* uq = qu64 + i * d;
    int64_t lim = i + cut;
    lim = lim < 0 ? 0 : (lim > n_ref ? n_ref : lim);
    double bk = __builtin_inf();
    int bi = -1;
    for (int64_t base = 0; base < lim; base += 64) {
        const int64_t r = base + lane;
        double key = __builtin_inf();
        int idx = -1;
        if (r < lim) {
            const double kr = exact_key(ru64 + r * d, uq, d);
            if (kr < __builtin_inf()) {  // NaN and +inf are never admitted
                key = kr;
                idx = (int)r;
            }
        }
        for (int size = 2; size <= 64; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1)
                wave_pair_step(key, idx, lane, size, stride);
        // the 64 smallest of the running list and the chunk: elementwise minimum
        // against the reversed chunk (a bitonic sequence), then a bitonic merge
        const double rk = __shfl(key, 63 - lane, 64);
        const int ri = __shfl(idx, 63 - lane, 64);
        if (pair_less(rk, ri, bk, bi)) {
            bk = rk;
            bi = ri;
        }
        for (int stride = 32; stride > 0; stride >>= 1) wave_pair_step(bk, bi, lane, 64, stride);
    }
    if (lane < k) out_idx[i * k + lane] = bi;
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------

// Screen candidate count for k neighbours: k plus a margin of 8, rounded to 8.
// Rows whose margin turns out too thin fail the certificate and are redone.
inline int screen_K(int k) { return (int)round_up(k + 8, 8); }

// Ref stripes of the scans (screen and exact): with few query blocks (one lane per
// query) a scan cannot fill the chip, so each block's refs are split over up to 8 stripes
// (blocks) of at least 2,048 refs, for about 2,048 blocks in all.
inline int screen_stripes(int64_t n_ref, int64_t n_query) {
    const int64_t nblk = std::max<int64_t>((n_query + 63) / 64, 1);
    int64_t s = std::min<int64_t>(2048 / nblk, kMaxStripes);
    s = std::min<int64_t>(s, n_ref / 2048);
    return (int)std::max<int64_t>(s, 1);
}
inline int32_t screen_stripe_w(int64_t n_ref, int n_stripes) {
    return (int32_t)round_up((n_ref + n_stripes - 1) / n_stripes, 8);
}

bool use_fast_scan(int d, int k) { return d >= 1 && d <= 32 && k >= 1 && k <= 64; }

// Feature count the scan kernels are compiled for; extra features are zero in
// both refs and queries, which leaves every key unchanged (t = 0 adds fma(0,0,acc)).
inline int scan_dim(int d) { return d <= 3 ? 3 : (d == 4 ? 4 : (d <= 8 ? 8 : (d <= 16 ? 16 : 32))); }

// Exact-scan list length for k neighbours.
inline int exact_K(int k) { return k <= 16 ? 16 : (k <= 32 ? 32 : 64); }

template <typename T, int D, int K, bool SCREEN>
void launch_scan_u(const ScanArgs<T>& a, hipStream_t st) {
    constexpr int B = SCREEN ? 32 : 16;
    const int64_t nblk = (a.n_query + 63) / 64;
    const unsigned ns = a.stripe_w > 0 ? (unsigned)((a.n_ref + a.stripe_w - 1) / a.stripe_w) : 1u;
    hipLaunchKernelGGL((scan_kernel<T, D, K, B, SCREEN>), dim3((unsigned)nblk, ns), dim3(64), 0, st,
                       a);
    if (ns > 1) {
        const dim3 mg((unsigned)((a.n_query + 255) / 256));
        if constexpr (SCREEN)
            hipLaunchKernelGGL((stripe_merge_kernel<K>), mg, dim3(256), 0, st, a.out_pk, (int)ns,
                               a.n_query, a.out_idx, a.out_thr);
        else
            hipLaunchKernelGGL((stripe_merge_exact_kernel<K>), mg, dim3(256), 0, st, a.out_skey,
                               a.out_sidx, (int)ns, a.n_query, a.k_out, a.out_idx);
    }
}

template <typename T, int D, bool SCREEN>
int launch_scan_d(const ScanArgs<T>& a, int K, hipStream_t st) {
    if constexpr (SCREEN) {
        switch (K) {
            case 16: launch_scan_u<T, D, 16, SCREEN>(a, st); return 0;
            case 24: launch_scan_u<T, D, 24, SCREEN>(a, st); return 0;
            case 32: launch_scan_u<T, D, 32, SCREEN>(a, st); return 0;
            case 40: launch_scan_u<T, D, 40, SCREEN>(a, st); return 0;
            case 48: launch_scan_u<T, D, 48, SCREEN>(a, st); return 0;
            case 56: launch_scan_u<T, D, 56, SCREEN>(a, st); return 0;
            case 64: launch_scan_u<T, D, 64, SCREEN>(a, st); return 0;
            case 72: launch_scan_u<T, D, 72, SCREEN>(a, st); return 0;
            default: return -1;
        }
    } else {
        switch (K) {
            case 16: launch_scan_u<T, D, 16, SCREEN>(a, st); return 0;
            case 32: launch_scan_u<T, D, 32, SCREEN>(a, st); return 0;
            case 64: launch_scan_u<T, D, 64, SCREEN>(a, st); return 0;
            default: return -1;
        }
    }
}

template <typename T, bool SCREEN>
int launch_scan(const ScanArgs<T>& a, int d, int K, hipStream_t st) {
    switch (scan_dim(d)) {
        case 3: return launch_scan_d<T, 3, SCREEN>(a, K, st);
        case 4: return launch_scan_d<T, 4, SCREEN>(a, K, st);
        case 8: return launch_scan_d<T, 8, SCREEN>(a, K, st);
        case 16: return launch_scan_d<T, 16, SCREEN>(a, K, st);
        case 32: return launch_scan_d<T, 32, SCREEN>(a, K, st);
        default: return -1;
    }
}

template <int K>
void launch_rerank_k(const RerankArgs& a, hipStream_t st) {
    const unsigned grid = (unsigned)((a.n_query + 3) / 4);   // one wave per row
    hipLaunchKernelGGL(rerank_kernel<K>, dim3(grid), dim3(256), 0, st, a);
}

int launch_rerank(const RerankArgs& a, int K, hipStream_t st) {
    switch (K) {
        case 16: launch_rerank_k<16>(a, st); return 0;
        case 24: launch_rerank_k<24>(a, st); return 0;
        case 32: launch_rerank_k<32>(a, st); return 0;
        case 40: launch_rerank_k<40>(a, st); return 0;
        case 48: launch_rerank_k<48>(a, st); return 0;
        case 56: launch_rerank_k<56>(a, st); return 0;
        case 64: launch_rerank_k<64>(a, st); return 0;
        case 72: launch_rerank_k<72>(a, st); return 0;
        default: return -1;
    }
}

// Merge of n_lists sorted screen lists per row (the matrix-core screen's ref
// stripe halves) into candidates and K-th scores.
int launch_stripe_merge(int K, const uint64_t* pk, int n_lists, int64_t n_query, int32_t* cand,
                        float* thr, hipStream_t st) {
    if (n_lists > kMaxStripes) return -1;
    const dim3 g((unsigned)((n_query + 255) / 256));
    switch (K) {
        case 24: hipLaunchKernelGGL(stripe_merge_kernel<24>, g, dim3(256), 0, st, pk, n_lists, n_query, cand, thr); return 0;
        case 40: hipLaunchKernelGGL(stripe_merge_kernel<40>, g, dim3(256), 0, st, pk, n_lists, n_query, cand, thr); return 0;
        case 56: hipLaunchKernelGGL(stripe_merge_kernel<56>, g, dim3(256), 0, st, pk, n_lists, n_query, cand, thr); return 0;
        case 72: hipLaunchKernelGGL(stripe_merge_kernel<72>, g, dim3(256), 0, st, pk, n_lists, n_query, cand, thr); return 0;
        default: return -1;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Ref-side preparation (kept resident by serving models) and query-side search
// ---------------------------------------------------------------------------
namespace {

struct RefMem {
    double* centre;
    double* cpart;
    unsigned int* smax;
    double* u64;
    double* s64;
    float* s32;
    uint16_t* frag;
    GridRefMem grid;
};

// Every layout a method could choose for these refs is budgeted (the scan
// layout, the grid for d <= 3, the matrix-core fragments for d >= 4 in
// FP32_REFINED), so that one size serves every search method.
RefMem carve_refs(Carver& c, int64_t n_ref, int d, int precision) {
    RefMem M{};
    const int64_t rpad = round_up(std::max<int64_t>(n_ref, 1), 64) + 64;
    const int dp = scan_dim(d);
    M.centre = c.take<double>(NNGP_MAX_DIM);
    M.cpart = c.take<double>((size_t)kCentreBlocks * 2 * NNGP_MAX_DIM);
    M.smax = c.take<unsigned int>(4);
    M.u64 = c.take<double>((size_t)n_ref * d + 1);
    if (d <= 32) {
        if (precision == NNGP_PRECISION_FP32_REFINED) M.s32 = c.take<float>((size_t)rpad * (dp + 1));
        else M.s64 = c.take<double>((size_t)rpad * dp);
    }
    if (precision == NNGP_PRECISION_FP32_REFINED && mfma_screen_supported(d))
        M.frag = c.take<uint16_t>(mfma_frag_elems(n_ref, d));
    if (d <= 3) carve_grid_refs(c, n_ref, &M.grid);
    return M;
}

struct QueryMem {
    double* u64;
    double* s64;
    float* s32;
    uint16_t* qfrag;  // matrix-core screen: query operand fragments
    double* p2;
    int32_t* cand;
    float* thr;
    uint64_t* spk;    // per-stripe screen lists (more than one stripe)
    double* sxk;      // per-stripe exact-scan lists: keys
    int32_t* sxi;     //   and indices
    int64_t* flag_rows;
    unsigned long long* flag_count;
    double* flag_key;
    int32_t* flag_idx;
    int64_t* over_rows;
    unsigned long long* over_count;
    double* lkey;
    int32_t* lidx;
    GridSearchMem grid;
};

// Which search a knn_search over refs of this layout runs for k neighbours.
struct SearchKind {
    bool grid;    // uniform-grid search
    bool mfma;    // matrix-core screen + re-rank
    bool fast;    // a streaming scan (screen or exact)
    bool screen;  // a screen (VALU or matrix-core) + re-rank
};
SearchKind search_kind(int layout, int d, int k, int precision) {
    SearchKind s{};
    s.grid = layout == kRefGrid && grid_supported(d, k);
    s.mfma = !s.grid && layout == kRefMfma && k >= 1 && k <= 64;
    s.fast = s.mfma || (!s.grid && layout != kRefMfma && use_fast_scan(d, k));
    s.screen = s.mfma || (s.fast && precision == NNGP_PRECISION_FP32_REFINED);
    return s;
}

// Query-side buffers of a search over refs of the given layout: only the
// buffers of the search it runs for k (search_kind) are carved.
QueryMem carve_query(Carver& c, int64_t n_ref, int64_t n_query, int d, int k, int precision,
                     bool same, bool with_cut, int layout) {
    QueryMem Q{};
    const SearchKind sk = search_kind(layout, d, k, precision);
    if (sk.grid) {
        if (!same) Q.u64 = c.take<double>((size_t)n_query * d + 1);
        carve_grid_search(c, n_ref, n_query, k, with_cut, &Q.grid);
        return Q;
    }
    const int64_t qpad = round_up(std::max<int64_t>(n_query, 1), 64) + 64;
    const int dp = scan_dim(d);
    const bool fast = sk.fast, screen = sk.screen;
    if (!same) {
        Q.u64 = c.take<double>((size_t)n_query * d + 1);
        if (fast && !screen) Q.s64 = c.take<double>((size_t)qpad * dp);
        if (screen && !sk.mfma) Q.s32 = c.take<float>((size_t)qpad * (dp + 1));
    }
    if (sk.mfma) Q.qfrag = c.take<uint16_t>(mfma_frag_elems(n_query, d));
    if (fast && !screen && n_ref >= 2 * 2048) {  // exact-scan stripe lists (bound as below)
        const size_t ne = (size_t)std::min<int64_t>(8 * n_query, 2048 * 64) * exact_K(k) + 1;
        Q.sxk = c.take<double>(ne);
        Q.sxi = c.take<int32_t>(ne);
    }
    if (screen) {
        Q.p2 = c.take<double>((size_t)n_query + 1);
        const int K = sk.mfma ? mfma_screen_K(k) : screen_K(k);
        Q.cand = c.take<int32_t>((size_t)n_query * K + 1);
        Q.thr = c.take<float>((size_t)n_query + 1);
        // stripe lists: stripes x rows <= min(8 n_query, 2048 x 64) for every
        // n_query' <= n_query, so a workspace sized for a batch serves smaller ones
        if (n_ref >= 2 * 2048)
            Q.spk = c.take<uint64_t>((size_t)std::min<int64_t>(8 * n_query, 2048 * 64) * K + 1);
        Q.flag_rows = c.take<int64_t>((size_t)n_query + 1);
        Q.flag_key = c.take<double>((size_t)n_query + 1);
        Q.flag_idx = c.take<int32_t>((size_t)n_query + 1);
        Q.over_rows = c.take<int64_t>((size_t)n_query + 1);
        Q.flag_count = c.take<unsigned long long>(4);
        Q.over_count = c.take<unsigned long long>(4);
    }
    Q.lkey = c.take<double>((size_t)k * kFallbackThreads);
    Q.lidx = c.take<int32_t>((size_t)k * kFallbackThreads);
    return Q;
}

std::atomic<int> g_knn_method{NNGP_KNN_AUTO};

int check_common(int64_t n_ref, int d, int n_scale, int precision) {
    if (d < 1 || d > NNGP_MAX_DIM || n_ref < 0) return NNGP_ERR_INVALID_ARGUMENT;
    if (n_scale != 0 && n_scale != 1 && n_scale != d) return NNGP_ERR_INVALID_ARGUMENT;
    if (n_ref >= (int64_t)INT32_MAX) return NNGP_ERR_UNSUPPORTED;
    if (precision != NNGP_PRECISION_FP64 && precision != NNGP_PRECISION_FP32_REFINED)
        return NNGP_ERR_INVALID_ARGUMENT;
    return NNGP_OK;
}

}  // namespace

size_t exact_rows_scratch_count(int k) { return (size_t)k * kFallbackThreads; }

void launch_exact_rows(const double* ru64, const double* qu64, int d, int64_t n_ref,
                       int64_t n_query, int has_cut, int64_t cut, int k, const ExactRows& er,
                       int32_t* out_idx, hipStream_t st) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, er.over_count, 1);
    WaveFbArgs wf{};
    wf.ru64 = ru64;
    wf.qu64 = qu64;
    wf.d = d;
    wf.n_ref = n_ref;
    wf.has_cut = has_cut;
    wf.cut = cut;
    wf.k = k;
    wf.rows = er.rows;
    wf.row_count = er.count;
    wf.bound_key = er.bound_key;
    wf.bound_idx = er.bound_idx;
    wf.out_idx = out_idx;
    wf.over_rows = er.over_rows;
    wf.over_count = er.over_count;
    hipLaunchKernelGGL(row_fallback_kernel, dim3(2048), dim3(kFbThreads), 0, st, wf);
    SimpleArgs sa{};
    sa.ru64 = ru64;
    sa.qu64 = qu64;
    sa.d = d;
    sa.n_ref = n_ref;
    sa.n_query = n_query;
    sa.has_cut = has_cut;
    sa.cut = cut;
    sa.k = k;
    sa.rows = er.over_rows;
    sa.row_count = er.over_count;
    sa.lkey = er.lkey;
    sa.lidx = er.lidx;
    sa.out_idx = out_idx;
    hipLaunchKernelGGL(simple_knn_kernel, dim3(kFallbackThreads / 256), dim3(256), 0, st, sa);
}

int knn_method() { return g_knn_method.load(); }

int knn_set_method(int method) {
    if (method != NNGP_KNN_AUTO && method != NNGP_KNN_BRUTE_FORCE &&
        method != NNGP_KNN_BRUTE_FORCE_VALU && method != NNGP_KNN_GRID_LANES &&
        method != NNGP_KNN_GRID_WAVES)
        return NNGP_ERR_INVALID_ARGUMENT;
    g_knn_method.store(method);
    return NNGP_OK;
}

// The layout refs of d features prepared now get (the method is read when refs
// are prepared).
int refs_layout(int d, int precision) {
    const int m = knn_method();
    if (d <= 3 && (m == NNGP_KNN_AUTO || m == NNGP_KNN_GRID_LANES || m == NNGP_KNN_GRID_WAVES))
        return kRefGrid;
    if (precision == NNGP_PRECISION_FP32_REFINED && mfma_screen_supported(d) &&
        m != NNGP_KNN_BRUTE_FORCE_VALU)
        return kRefMfma;
    return kRefScan;
}

size_t knn_refs_bytes(int64_t n_ref, int d, int precision) {
    Carver c(nullptr, 0);
    carve_refs(c, n_ref, d, precision);
    return c.used + 256;
}

int knn_prepare_refs(const double* ref, int64_t n_ref, int d, const double* scale, int n_scale,
                     int precision, void* mem, size_t bytes, KnnRefs* out, hipStream_t st) {
    int rc = check_common(n_ref, d, n_scale, precision);
    if (rc) return rc;
    Carver c(mem, bytes);
    RefMem M = carve_refs(c, n_ref, d, precision);
    if (!c.ok() || !mem) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    const int layout = refs_layout(d, precision);
    const bool grid = layout == kRefGrid;
    if (layout != kRefScan) {  // the grid search and the matrix-core screen read neither
        M.s64 = nullptr;
        M.s32 = nullptr;
    }
    if (layout != kRefMfma) M.frag = nullptr;
    KnnRefs R{};
    R.u64 = M.u64;
    R.s64 = M.s64;
    R.s32 = M.s32;
    R.frag = M.frag;
    R.grid = M.grid;
    R.layout = layout;
    R.use_grid = grid ? 1 : 0;
    R.centre = M.centre;
    R.smax = M.smax;
    R.n = n_ref;
    R.d = d;
    R.precision = precision;
    R.n_scale = n_scale;
    for (int f = 0; f < n_scale; ++f) R.scale[f] = scale[f];
    *out = R;
    hipLaunchKernelGGL(init_ref_kernel, dim3(1), dim3(64), 0, st, M.centre, d, M.smax);
    if (n_ref == 0) {
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    PrepArgs pr{};
    pr.x = ref;
    pr.n = n_ref;
    pr.d = d;
    pr.n_scale = n_scale;
    for (int f = 0; f < n_scale; ++f) pr.scale[f] = scale[f];
    if (M.s32 || M.frag) {
        hipLaunchKernelGGL(centre_partial_kernel, dim3(kCentreBlocks), dim3(256), 0, st, pr,
                           M.cpart);
        hipLaunchKernelGGL(centre_final_kernel, dim3(1), dim3(64), 0, st, M.cpart, kCentreBlocks,
                           d, n_ref, M.centre);
    }
    const int64_t rpad = round_up(n_ref, 64) + 64;
    hipLaunchKernelGGL(prep_kernel, dim3((unsigned)((rpad + 255) / 256)), dim3(256), 0, st, pr,
                       M.centre, rpad, scan_dim(d), M.u64, M.s64, M.s32, (double*)nullptr, M.smax);
    if (M.frag) mfma_prep(M.u64, n_ref, d, M.centre, 0, M.frag, nullptr, M.smax, st);
    NNGP_HIP_CHECK(hipGetLastError());
    if (grid) return grid_build(M.u64, n_ref, d, M.grid, st);
    return NNGP_OK;
}

size_t knn_search_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision, bool same,
                        bool with_cut, int layout) {
    Carver c(nullptr, 0);
    carve_query(c, n_ref, n_query, d, k, precision, same, with_cut, layout);
    return c.used + 256;
}

int knn_search(const KnnRefs& R, const double* query, int64_t n_query, int k, int has_cut,
               int64_t ref_cut, int32_t* out_idx, void* ws, size_t ws_bytes, WsHeader* hdr,
               hipStream_t st) {
    const int d = R.d;
    if (k < 1 || k > NNGP_MAX_K || n_query < 0) return NNGP_ERR_INVALID_ARGUMENT;
    const bool same = query == nullptr;
    const SearchKind sk = search_kind(R.layout, d, k, R.precision);
    const bool grid_path = sk.grid;
    Carver c(ws, ws_bytes);
    QueryMem Q = carve_query(c, R.n, n_query, d, k, R.precision, same, has_cut != 0, R.layout);
    if (!c.ok() || (!ws && c.used > 0)) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    if (n_query == 0) return NNGP_OK;
    if (R.n == 0) {
        const int64_t tot = n_query * (int64_t)k;
        hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                           out_idx, tot, -1);
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    // the grid path reads only the query's scaled coordinates (no scan layouts)
    const bool mfma = sk.mfma;
    const bool fast = sk.fast && (mfma || R.s32 || R.s64);
    const bool screen = fast && sk.screen;
    const int dp = scan_dim(d);
    const int64_t qpad = round_up(n_query, 64) + 64;
    const double* qu64 = R.u64;
    const double* qs64 = R.s64;
    const float* qs32 = R.s32;
    double* qp2 = Q.p2;
    PrepArgs pr{};
    pr.x = query;
    pr.n = n_query;
    pr.d = d;
    pr.n_scale = R.n_scale;
    for (int f = 0; f < R.n_scale; ++f) pr.scale[f] = R.scale[f];
    if (!same) {
        // only the layouts the chosen search reads are written
        double* qs64w = fast && !screen ? Q.s64 : nullptr;
        float* qs32w = screen && !mfma ? Q.s32 : nullptr;
        hipLaunchKernelGGL(prep_kernel, dim3((unsigned)((qpad + 255) / 256)), dim3(256), 0, st, pr,
                           R.centre, qpad, dp, Q.u64, qs64w, qs32w,
                           screen && !mfma ? Q.p2 : nullptr, (unsigned int*)nullptr);
        qu64 = Q.u64;
        qs64 = qs64w;
        qs32 = qs32w;
    } else if (screen && !mfma) {
        // |u_q - centre|^2 of the refs themselves (the query set is the ref set)
        pr.x = nullptr;
        hipLaunchKernelGGL(p2_kernel, dim3((unsigned)((n_query + 255) / 256)), dim3(256), 0, st,
                           R.u64, n_query, d, R.centre, Q.p2);
    }
    // query operand fragments (and |u_q - centre|^2) of the matrix-core screen
    if (mfma) mfma_prep(qu64, n_query, d, R.centre, 1, Q.qfrag, Q.p2, nullptr, st);
    NNGP_HIP_CHECK(hipGetLastError());

    if (grid_path) {
        int64_t row0 = 0;
        if (has_cut) {
            // Rows whose limit q + cut is below the first grid band: brute-force
            // exact scan over their few refs; rows with no ref at all get -1.
            const int64_t L0 = grid_prefix_limit(R.n, k);
            const int64_t qlo = std::min<int64_t>(std::max<int64_t>(-ref_cut, 0), n_query);
            row0 = std::min<int64_t>(std::max<int64_t>(L0 - ref_cut, 0), n_query);
            if (qlo > 0) {
                const int64_t tot = qlo * (int64_t)k;
                hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                                   0, st, out_idx, tot, -1);
            }
            if (row0 > qlo) {
                timing_begin(TK_KNN_SCAN, st);
                hipLaunchKernelGGL(prefix_knn_kernel, dim3((unsigned)(row0 - qlo)), dim3(64), 0,
                                   st, R.u64, qu64, d, qlo, ref_cut, R.n, k, out_idx);
                timing_end(TK_KNN_SCAN, st);
            }
            NNGP_HIP_CHECK(hipGetLastError());
        }
        return grid_search(R.grid, R.u64, R.n, d, qu64, n_query, k, has_cut, ref_cut, row0,
                           Q.grid, out_idx, hdr ? &hdr->knn_rows_recomputed : nullptr, st);
    }
    if (!fast) {
        SimpleArgs sa{};
        sa.ru64 = R.u64;
        sa.qu64 = qu64;
        sa.d = d;
        sa.n_ref = R.n;
        sa.n_query = n_query;
        sa.has_cut = has_cut;
        sa.cut = ref_cut;
        sa.k = k;
        sa.lkey = Q.lkey;
        sa.lidx = Q.lidx;
        sa.out_idx = out_idx;
        hipLaunchKernelGGL(simple_knn_kernel, dim3(kFallbackThreads / 256), dim3(256), 0, st, sa);
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    if (!screen) {
        ScanArgs<double> a{};
        a.ref = R.s64;
        a.n_ref = (int32_t)R.n;
        a.qry = qs64;
        a.n_query = n_query;
        a.has_cut = has_cut;
        a.cut = ref_cut;
        a.reverse = has_cut;
        a.k_out = k;
        const int n_stripes = screen_stripes(R.n, n_query);
        if (n_stripes > 1 && Q.sxk) {
            a.stripe_w = screen_stripe_w(R.n, n_stripes);
            a.out_skey = Q.sxk;
            a.out_sidx = Q.sxi;
        }
        a.out_idx = out_idx;
        timing_begin(TK_KNN_SCAN, st);
        const int rc = launch_scan<double, false>(a, d, exact_K(k), st);
        timing_end(TK_KNN_SCAN, st);
        if (rc != 0) return NNGP_ERR_UNSUPPORTED;
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    const int K = mfma ? mfma_screen_K(k) : screen_K(k);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, Q.flag_count, 1);
    if (mfma) {
        timing_begin(TK_KNN_SCAN, st);
        const int lists = launch_mfma_screen(R.frag, R.n, Q.qfrag, n_query, d, K, has_cut, ref_cut,
                                             Q.cand, Q.thr, Q.spk, st);
        if (lists > 0 && launch_stripe_merge(K, Q.spk, lists, n_query, Q.cand, Q.thr, st) != 0)
            return NNGP_ERR_UNSUPPORTED;
        timing_end(TK_KNN_SCAN, st);
        if (lists < 0) return NNGP_ERR_UNSUPPORTED;
    } else {
        ScanArgs<float> a{};
        a.ref = R.s32;
        a.n_ref = (int32_t)R.n;
        a.qry = qs32;
        a.n_query = n_query;
        a.has_cut = has_cut;
        a.cut = ref_cut;
        a.reverse = has_cut;
        a.out_idx = Q.cand;
        a.out_thr = Q.thr;
        const int n_stripes = screen_stripes(R.n, n_query);
        if (n_stripes > 1 && Q.spk) {
            a.stripe_w = screen_stripe_w(R.n, n_stripes);
            a.out_pk = Q.spk;
        }
        timing_begin(TK_KNN_SCAN, st);
        const int rcs = launch_scan<float, true>(a, d, K, st);
        timing_end(TK_KNN_SCAN, st);
        if (rcs != 0) return NNGP_ERR_UNSUPPORTED;
    }

    RerankArgs rr{};
    rr.ru64 = R.u64;
    rr.qu64 = qu64;
    rr.qp2 = qp2;
    rr.smax_bits = R.smax;
    rr.d = d;
    rr.n_query = n_query;
    rr.k = k;
    rr.cand = Q.cand;
    rr.thr = Q.thr;
    rr.out_idx = out_idx;
    rr.flag_rows = Q.flag_rows;
    rr.flag_count = Q.flag_count;
    rr.flag_key = Q.flag_key;
    rr.flag_idx = Q.flag_idx;
    rr.eps_factor = mfma ? mfma_screen_eps(d) : (4.0 * d + 16.0) * std::ldexp(1.0, -24);
    timing_begin(TK_KNN_RERANK, st);
    const int rcr = launch_rerank(rr, K, st);
    timing_end(TK_KNN_RERANK, st);
    if (rcr != 0) return NNGP_ERR_UNSUPPORTED;

    ExactRows er{};
    er.rows = Q.flag_rows;
    er.count = Q.flag_count;
    er.bound_key = Q.flag_key;
    er.bound_idx = Q.flag_idx;
    er.over_rows = Q.over_rows;
    er.over_count = Q.over_count;
    er.lkey = Q.lkey;
    er.lidx = Q.lidx;
    launch_exact_rows(R.u64, qu64, d, R.n, n_query, has_cut, ref_cut, k, er, out_idx, st);
    if (hdr) hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(1), 0, st, Q.flag_count,
                                &hdr->knn_rows_recomputed);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

size_t knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision,
                           bool same) {
    Carver c(nullptr, 0);
    c.take<WsHeader>(1);
    return c.used + knn_refs_bytes(n_ref, d, precision) +
           knn_search_bytes(n_ref, n_query, d, k, precision, same, true,
                            refs_layout(d, precision)) + 512;
}

int knn_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
            const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut, int precision,
            int32_t* out_idx, void* ws, size_t ws_bytes, hipStream_t st, bool reset_stats) {
    int rc = check_common(n_ref, d, n_scale, precision);
    if (rc) return rc;
    if (k < 1 || k > NNGP_MAX_K || n_query < 0) return NNGP_ERR_INVALID_ARGUMENT;
    if (!ws) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    const bool same = (ref == query && n_ref == n_query);
    Carver c(ws, ws_bytes);
    WsHeader* hdr = c.take<WsHeader>(1);
    const size_t rb = knn_refs_bytes(n_ref, d, precision);
    char* rmem = c.take<char>(rb);
    // sized for the method read now; knn_search carves for the refs' actual
    // method and reports a workspace too small if the method changed in between
    const size_t qb = knn_search_bytes(n_ref, n_query, d, k, precision, same, has_cut != 0,
                                       refs_layout(d, precision));
    char* qmem = c.take<char>(qb);
    if (!c.ok()) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    if (reset_stats) hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, hdr);
    KnnRefs R;
    rc = knn_prepare_refs(ref, n_ref, d, scale, n_scale, precision, rmem, rb, &R, st);
    if (rc) return rc;
    return knn_search(R, same ? nullptr : query, n_query, k, has_cut, ref_cut, out_idx, qmem, qb,
                      hdr, st);
}

}  // namespace nngp
