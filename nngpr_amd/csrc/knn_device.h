// Device helpers shared by the kNN kernels (brute-force scans and grid search).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nngp {

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Monotone map of a double onto int64 so that atomicMin/atomicMax order like the
// doubles (NaN excluded).
__device__ __forceinline__ long long dbl_key(double v) {
    long long b = __double_as_longlong(v);
    return b >= 0 ? b : (b ^ 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double key_dbl(long long k) {
    long long b = k >= 0 ? k : (k ^ 0x7fffffffffffffffLL);
    return __longlong_as_double(b);
}

__device__ __forceinline__ int32_t wave_min32(int32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int32_t wave_max32(int32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// Order-preserving map of an FP32 score onto uint32, packed with the ref index:
// comparing the packed uint64 values orders by (score, index) and puts empty
// slots (+inf, index -1) last.
__device__ __forceinline__ uint32_t ord_f32(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ uint64_t pack_key(float s, int idx) {
    return ((uint64_t)ord_f32(s) << 32) | (uint32_t)idx;
}
// Empty slot of a packed screen list: score +inf, index -1.
constexpr uint64_t kPackedEmpty = ((uint64_t)0xFF800000u << 32) | 0xFFFFFFFFu;

// Insertion into a packed screen list sorted by score.  Screening needs the
// candidate set, not an order among equal scores: only the score half is
// compared (one 32-bit compare and four selects per step).
template <int K>
__device__ __forceinline__ void topk_insert_packed(uint64_t (&e)[K], uint64_t x) {
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const bool lt = (uint32_t)(x >> 32) < (uint32_t)(e[t] >> 32);
        const uint64_t lo = lt ? x : e[t];
        x = lt ? e[t] : x;
        e[t] = lo;
    }
}

// Compare-exchange ladder over a sorted register list, ordered by (key, index):
// the carried element moves ahead of an equal key with a higher index.  Indices
// compare unsigned, so an empty slot (index -1) sorts after every real entry of
// the same key.
template <typename T, int K>
__device__ __forceinline__ void topk_insert(T (&keys)[K], int (&ids)[K], T xk, int xi) {
#pragma unroll
    for (int t = 0; t < K; ++t) {
        // bitwise, not short-circuit: keeps the ladder straight-line code
        const bool lt = (xk < keys[t]) | ((xk == keys[t]) & ((unsigned)xi < (unsigned)ids[t]));
        T nk = lt ? xk : keys[t];
        T ok = lt ? keys[t] : xk;
        int ni = lt ? xi : ids[t];
        int oi = lt ? ids[t] : xi;
        keys[t] = nk;
        ids[t] = ni;
        xk = ok;
        xi = oi;
    }
}

// Sorting networks over (key, index) pairs in registers, for lists whose keys are
// never NaN (-inf / +inf allowed): keys move with fmin / fmax (equal keys are
// interchangeable), indices follow the pair order (indices unsigned, as above).
__device__ __forceinline__ bool pair_less(double a, int ai, double b, int bi) {
    return (a < b) | ((a == b) & ((unsigned)ai < (unsigned)bi));
}

// Ascending compare-exchange: afterwards (ka, ia) <= (kb, ib).
__device__ __forceinline__ void pair_cex(double& ka, int& ia, double& kb, int& ib) {
    const bool sw = pair_less(kb, ib, ka, ia);
    const double lo = fmin(ka, kb), hi = fmax(ka, kb);
    const int il = sw ? ib : ia, ih = sw ? ia : ib;
    ka = lo;
    kb = hi;
    ia = il;
    ib = ih;
}

// Bitonic sort of N (a power of two) pairs, ascending.
template <int N>
__device__ __forceinline__ void pair_sort(double (&k)[N], int (&id)[N]) {
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    if ((i & size) == 0) pair_cex(k[i], id[i], k[j], id[j]);
                    else pair_cex(k[j], id[j], k[i], id[i]);
                }
            }
        }
    }
}

// Merges B pairs (any order) into a sorted list of K pairs, keeping the K
// smallest: sort the B pairs, take C_t = min(A_t, B_{K-1-t}) over the last B
// slots (with the sorted B padded by +inf this is min(A_t, B'_{K-1-t}) over all
// t, a bitonic sequence holding the K smallest of both), then half-cleaners.
template <int K, int B>
__device__ __forceinline__ void pair_merge(double (&keys)[K], int (&ids)[K], double (&bk)[B],
                                           int (&bi)[B]) {
    pair_sort<B>(bk, bi);
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int t = K - B + i, j = B - 1 - i;
        const bool lt = pair_less(bk[j], bi[j], keys[t], ids[t]);
        keys[t] = fmin(bk[j], keys[t]);
        ids[t] = lt ? bi[j] : ids[t];
    }
#pragma unroll
    for (int stride = K >> 1; stride > 0; stride >>= 1) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int j = i ^ stride;
            if (j > i) pair_cex(keys[i], ids[i], keys[j], ids[j]);
        }
    }
}

// Key-only forms of the networks above: the index follows the key comparison
// alone (5 instructions per compare-exchange instead of 7 plus 2 scalar mask
// operations), so a list comes out in key order with equal keys in whatever order
// the network left them.  key_merge also returns the smallest key it dropped:
// when that equals the kept list's last key, the K kept pairs are right as keys
// but which of the equal-key pairs were kept is arbitrary.
__device__ __forceinline__ void key_cex(double& ka, int& ia, double& kb, int& ib) {
    const bool sw = kb < ka;
    const double lo = fmin(ka, kb), hi = fmax(ka, kb);
    const int il = sw ? ib : ia, ih = sw ? ia : ib;
    ka = lo;
    kb = hi;
    ia = il;
    ib = ih;
}

template <int N>
__device__ __forceinline__ void key_sort(double (&k)[N], int (&id)[N]) {
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    if ((i & size) == 0) key_cex(k[i], id[i], k[j], id[j]);
                    else key_cex(k[j], id[j], k[i], id[i]);
                }
            }
        }
    }
}

template <int K, int B>
__device__ __forceinline__ void key_merge(double (&keys)[K], int (&ids)[K], double (&bk)[B],
                                          int (&bi)[B], double& dropped_min) {
    key_sort<B>(bk, bi);
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int t = K - B + i, j = B - 1 - i;
        const bool lt = bk[j] < keys[t];
        dropped_min = fmin(dropped_min, fmax(bk[j], keys[t]));
        keys[t] = fmin(bk[j], keys[t]);
        ids[t] = lt ? bi[j] : ids[t];
    }
#pragma unroll
    for (int stride = K >> 1; stride > 0; stride >>= 1) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int j = i ^ stride;
            if (j > i) key_cex(keys[i], ids[i], keys[j], ids[j]);
        }
    }
}

// Wave-wide bitonic stage over (key, index) pairs, one pair per lane: lanes l and
// l ^ stride exchange; the lower lane of an ascending block keeps the smaller pair.
__device__ __forceinline__ void wave_pair_step(double& key, int& idx, int lane, int size,
                                               int stride) {
    const double ok = __shfl_xor(key, stride, 64);
    const int oi = __shfl_xor(idx, stride, 64);
    const bool up = (lane & size) == 0;
    const bool lower = (lane & stride) == 0;
    const bool other_less = pair_less(ok, oi, key, idx);
    const bool other_more = pair_less(key, idx, ok, oi);
    if (lower == up ? other_less : other_more) {
        key = ok;
        idx = oi;
    }
}

// Value of lane l ^ stride (stride a power of two below 64, a constant once
// unrolled): DPP moves for strides 1, 2 and 8 (quad permutes, a rotation by 8 within
// each row of 16), ds_swizzle xor patterns for 4 and 16, a permute for 32.
__device__ __forceinline__ int lane_xor(int v, int stride) {
    switch (stride) {
        case 1: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, true);   // quad_perm [1,0,3,2]
        case 2: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, true);   // quad_perm [2,3,0,1]
        case 4: return __builtin_amdgcn_ds_swizzle(v, 0x101F);              // xor 4 (32-lane groups)
        case 8: return __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, true);  // row_ror:8
        case 16: return __builtin_amdgcn_ds_swizzle(v, 0x401F);             // xor 16 (32-lane groups)
        default: return __shfl_xor(v, 32, 64);
    }
}
__device__ __forceinline__ double lane_xor(double v, int stride) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_xor((int)(b & 0xffffffffLL), stride);
    const int hi = lane_xor((int)(b >> 32), stride);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// wave_pair_step with the exchange of lane_xor; size and stride must be constants
// once unrolled.
__device__ __forceinline__ void wave_pair_step_x(double& key, int& idx, int lane, int size,
                                                 int stride) {
    const double ok = lane_xor(key, stride);
    const int oi = lane_xor(idx, stride);
    const bool up = (lane & size) == 0;
    const bool lower = (lane & stride) == 0;
    const bool other_less = pair_less(ok, oi, key, idx);
    const bool other_more = pair_less(key, idx, ok, oi);
    if (lower == up ? other_less : other_more) {
        key = ok;
        idx = oi;
    }
}

// FP64 key of a (ref, query) pair: fma(t_{d-1}, t_{d-1}, ... fma(t_1, t_1, t_0 * t_0)),
// t_f = u_ref,f - u_query,f.
__device__ __forceinline__ double exact_key(const double* ur, const double* uq, int d) {
    double t0 = ur[0] - uq[0];
    double key = t0 * t0;
    for (int f = 1; f < d; ++f) {
        double t = ur[f] - uq[f];
        key = fma(t, t, key);
    }
    return key;
}

}  // namespace nngp
