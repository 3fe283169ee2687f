// Matrix-core screen of the brute-force k-nearest-neighbour search (d >= 4), gfx950.
//
// Same contract as the FP32 VALU screen in knn.hip: per query row, the K refs of
// smallest screen score s = |r|^2 - 2 p.r (p, r = fl32 of the centred scaled
// coordinates; the exact key is s + |p|^2) and the K-th score, which the FP64
// re-rank turns into exact lists with a certificate, redoing uncertified rows
// exactly.  The screen only has to be accurate to a known bound.
//
// Scores come from v_mfma_f32_32x32x16_bf16 over K-vectors of 3 d + 2 entries:
//   ref   a = [hi(r), lo(r), hi(r), hi(rr), lo(rr), 0 ...],  rr = fl32(|r|^2)
//   query b = [hi(q), hi(q), lo(q), 1,      1,      0 ...],  q  = -2 p
// where hi(x) = bf16(x) and lo(x) = bf16(x - hi(x)), both rounded to nearest even,
// so |x - hi(x) - lo(x)| <= 2^-16 |x|.  Then a.b = hi(q).hi(r) + hi(q).lo(r) +
// lo(q).hi(r) + rr differs from s by the dropped lo(q).lo(r) and the split
// residuals (<= 3.02 * 2^-16 * 2 |p| |r| over the features) and the rr split
// (<= 2^-16 rr).  A bf16 product is exact in FP32, and the sum of the 3 d + 2 terms
// in the FP32 accumulator adds at most one FP32 ulp per term.  The error bound the
// certificate uses (mfma_screen_eps) is twice the sum of these terms plus the FP32
// screen's own bound for the rounding of p, r and rr.
//
// Tiling: one wave per 32 query rows (the B operand, kept in registers), ref tiles
// of 32 rows (the A operand) in the operand layout, NK = ceil((3 d + 2) / 16) matrix
// instructions per tile.  Output lane layout of the 32 x 32 tile: lane l holds
// column (query) l % 32, rows 8 (v / 4) + 4 (l / 32) + v % 4 for v = 0..15, so lanes
// l and l + 32 see disjoint halves of every tile and each keeps its own sorted list
// of K packed (score, index) candidates; a score below the lane's K-th is appended
// to the lane's LDS column and the columns are merged into the lists in lock-step
// (as in the VALU screen).  At the end the two halves of a row are combined: for
// sorted lists X, Y the elementwise min(X_t, Y_{K-1-t}) holds the K smallest of both.
//
// Each wave streams its ref tiles from L2 / HBM into a ring of register sets, up
// to three tiles ahead: the scan is bound by the latency of those loads, not by
// the matrix cores.  (Sharing each tile among the 4 waves of a block through LDS
// was measured 2x slower: a wave's list merge stalls the other three at the
// block's barriers.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "knn.h"
#include "knn_device.h"

namespace nngp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileRows = 32;   // refs per tile (A rows) and queries per wave (B columns)
constexpr int kChunkK = 16;     // K entries per matrix instruction
constexpr int kMaxListsMfma = 8;
constexpr int kAppendSlots = 32;  // LDS append slots per lane
constexpr int kPadTiles = 4;      // padding tiles past the last of a fragment array
constexpr int64_t kGenTiles = 2048;  // query tiles per launch of the scan

// bf16 bits of x rounded to nearest even (NaN stays NaN, overflow goes to inf).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr uint16_t kBf16One = 0x3f80;
constexpr uint16_t kBf16NaN = 0x7fc0;

// Writes the K-vectors of points [0, npad) (npad a multiple of 32) in the operand
// layout: point i, entry k lives in fragment ((i / 32) * nk + k / 16) at lane
// i % 32 + 32 * ((k % 16) / 8), element k % 8 (8 bf16 = 16 bytes per lane).
//   role 0 (refs):    a-vector; points with a non-finite coordinate and padding
//                     points get a NaN score (never admitted); raises *smax to the
//                     largest norm, as the VALU screen's preparation does.
//   role 1 (queries): b-vector; p2 = |u - centre|^2 in FP64.
__global__ void mfma_prep_kernel(const double* u64, int64_t n, int64_t npad, int d,
                                 const double* centre, int nk, int role, uint16_t* frag,
                                 double* p2, unsigned int* smax) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const int64_t t = i >> 5;
    const int lr = (int)(i & 31);
    auto put = [&](int k, uint16_t v) {
        const int c = k >> 4, kk = k & 15;
        frag[((t * nk + c) * 64 + lr + 32 * (kk >> 3)) * 8 + (kk & 7)] = v;
    };
    const int kv = nk * kChunkK;
    if (i >= n) {
        for (int k = 0; k < kv; ++k) put(k, (role == 0 && k == 3 * d) ? kBf16NaN : (uint16_t)0);
        return;
    }
    double nrm = 0.0, pp = 0.0;
    bool fin = true;
    for (int f = 0; f < d; ++f) {
        const double p = u64[i * d + f] - centre[f];
        const float pf = (float)p;
        fin = fin && __builtin_isfinite(pf);
        nrm += (double)pf * (double)pf;
        pp += p * p;
    }
    for (int f = 0; f < d; ++f) {
        const float pf = (float)(u64[i * d + f] - centre[f]);
        const float x = role == 0 ? pf : -2.0f * pf;
        const uint16_t hi = bf16_bits(x);
        const uint16_t lo = bf16_bits(x - bf16_value(hi));
        put(f, hi);
        put(d + f, role == 0 ? lo : hi);
        put(2 * d + f, role == 0 ? hi : lo);
    }
    if (role == 0) {
        const float rr = (float)nrm;
        const uint16_t rh = bf16_bits(rr);
        const uint16_t rl = bf16_bits(rr - bf16_value(rh));
        put(3 * d, fin ? rh : kBf16NaN);
        put(3 * d + 1, fin ? rl : (uint16_t)0);
        if (smax && fin) {
            const float r = (float)sqrt(fmax(nrm, pp)) * 1.0001f;
            if (__builtin_isfinite(r)) atomicMax(smax, __float_as_uint(r));
        }
    } else {
        put(3 * d, fin ? kBf16One : kBf16NaN);
        put(3 * d + 1, kBf16One);
        if (p2) p2[i] = pp;
    }
    for (int k = 3 * d + 2; k < kv; ++k) put(k, 0);
}

struct MfmaScanArgs {
    const bf16x8* rfrag;   // ref tiles (ceil(n_ref / 32) + kPadTiles, the last padding)
    int32_t n_ref;
    const bf16x8* qfrag;   // query tiles (ceil(n_query / 32) + kPadTiles)
    int64_t n_query;
    int has_cut;
    int64_t cut;
    int reverse;           // longest rows first
    int64_t qt_first;      // this launch scores query tiles [qt_first, qt_first + gridDim.x)
    int32_t stripe_w;      // refs per stripe, a multiple of 32 (gridDim.y stripes)
    int32_t* cand;         // one stripe: n_query x K candidates
    float* thr;            //   and the K-th score per row
    uint64_t* out_pk;      // stripes: sorted list (2 s + h) of row q at ((2 s + h) n_query + q) K
};

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// The refs a wave's rows scan: stripe [rs, re) of the refs, each lane's limit
// (the Vecchia cut, clipped to the stripe), and the wave's tiles: [t_begin, t_full)
// below every lane's limit, [t_full, t_end) masked per lane.  A wave without a
// valid row scans nothing.
struct RowWindow {
    int64_t q;
    bool qv;
    int32_t lim;
    int32_t t_begin, t_full, t_end;
};
__device__ __forceinline__ RowWindow row_window(const MfmaScanArgs& a, int64_t qt, int lane) {
    RowWindow w;
    w.q = qt * kTileRows + (lane & 31);
    w.qv = w.q < a.n_query;
    const bool striped = gridDim.y > 1;
    const int32_t rs = striped ? (int32_t)blockIdx.y * a.stripe_w : 0;
    const int64_t re = (striped && blockIdx.y + 1 < gridDim.y) ? (int64_t)rs + a.stripe_w
                                                                : (int64_t)a.n_ref;
    int64_t lim64 = a.n_ref;
    if (a.has_cut) {
        lim64 = w.q + a.cut;
        lim64 = lim64 < 0 ? 0 : (lim64 > a.n_ref ? a.n_ref : lim64);
    }
    lim64 = lim64 < re ? lim64 : re;
    w.lim = w.qv ? (int32_t)lim64 : 0;
    const int32_t lim_lo = wave_min32(w.qv ? w.lim : (int32_t)re);
    const int32_t lim_hi = wave_max32(w.lim);
    w.t_begin = rs / kTileRows;
    w.t_full = std::max(lim_lo / kTileRows, w.t_begin);
    w.t_end = (lim_hi + kTileRows - 1) / kTileRows;
    if (!__builtin_amdgcn_ballot_w64(w.qv)) w.t_full = w.t_end = w.t_begin;
    w.t_full = std::min(w.t_full, std::max(w.t_end, w.t_begin));
    return w;
}

// Per-lane candidate list of the screen: K packed (score, index) sorted by score,
// the admission threshold (the K-th score as of the last merge) and the lane's LDS
// append column.
template <int K, int B>
struct LaneList {
    uint64_t pk[K];
    float thr;
    int cnt;
    bool qv;
    int l, h;
    float (*sk)[64];
    int (*si)[64];

    __device__ __forceinline__ void init(bool valid, int lane, float (*k_)[64], int (*i_)[64]) {
#pragma unroll
        for (int t = 0; t < K; ++t) pk[t] = kPackedEmpty;
        qv = valid;
        thr = valid ? __builtin_inff() : -__builtin_inff();
        cnt = 0;
        l = lane;
        h = lane >> 5;
        sk = k_;
        si = i_;
    }
    __device__ __forceinline__ void merge() {
        const int nb = wave_max32(cnt);
#pragma unroll 1
        for (int b = 0; b < nb; ++b) {
            if (b < cnt) {
                const uint64_t x = pack_key(sk[b][l], si[b][l]);
                if (x < pk[K - 1]) topk_insert_packed<K>(pk, x);
            }
        }
        cnt = 0;
        thr = qv ? unord_f32((uint32_t)(pk[K - 1] >> 32)) : -__builtin_inff();
    }
    // ref row of output value v of tile t
    __device__ __forceinline__ int32_t row_of(int32_t t, int v) const {
        return t * kTileRows + 8 * (v >> 2) + 4 * h + (v & 3);
    }
    // Scores of tile t; masked: values of rows at or past the lane's limit drop out.
    __device__ __forceinline__ void take(f32x16 acc, int32_t t, bool masked, int32_t lim) {
        if (masked) {
#pragma unroll
            for (int v = 0; v < 16; ++v)
                if (!(row_of(t, v) < lim)) acc[v] = __builtin_inff();
        }
        float m = acc[0];
#pragma unroll
        for (int v = 1; v < 16; ++v) m = fminf(m, acc[v]);   // NaN scores drop out
        if (__builtin_amdgcn_ballot_w64(m < thr)) {
            // room for this tile's hits: at most 16 per lane
            if (__any(cnt > B - 16)) merge();
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                if (acc[v] < thr) {
                    sk[cnt][l] = acc[v];
                    si[cnt][l] = row_of(t, v);
                    ++cnt;
                }
            }
        }
    }
    // Final merge, then the row's output: one stripe - the K smallest of the two
    // halves' lists and their largest score; stripes - each half's sorted list.
    __device__ __forceinline__ void finish(const MfmaScanArgs& a, int64_t q) {
        if (__any(cnt > 0)) merge();
        if (gridDim.y == 1) {
            float kth = -__builtin_inff();
#pragma unroll
            for (int t = 0; t < K; ++t) {
                const uint64_t other = shfl_xor_u64(pk[K - 1 - t], 32);
                const uint64_t z =
                    (uint32_t)(other >> 32) < (uint32_t)(pk[t] >> 32) ? other : pk[t];
                kth = fmaxf(kth, unord_f32((uint32_t)(z >> 32)));
                if (h == 0 && qv) a.cand[q * K + t] = (int)(uint32_t)z;
            }
            if (h == 0 && qv) a.thr[q] = kth;
        } else if (qv) {
            uint64_t* o = a.out_pk + ((int64_t)(2 * blockIdx.y + h) * a.n_query + q) * K;
#pragma unroll
            for (int t = 0; t < K; ++t) o[t] = pk[t];
        }
    }
};

template <int NK>
__device__ __forceinline__ f32x16 tile_scores(const bf16x8 (&af)[NK], const bf16x8 (&qf)[NK]) {
    f32x16 acc = {};
#pragma unroll
    for (int c = 0; c < NK; ++c)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[c], qf[c], acc, 0, 0, 0);
    return acc;
}

template <int NK>
__device__ __forceinline__ void load_query(const MfmaScanArgs& a, int64_t qt, int l,
                                           bf16x8 (&qf)[NK]) {
    const int64_t n_qt = (a.n_query + kTileRows - 1) / kTileRows;
#pragma unroll
    for (int c = 0; c < NK; ++c) {
        if (qt < n_qt) qf[c] = a.qfrag[(qt * NK + c) * 64 + l];
        else qf[c] = bf16x8{};
    }
}

// Fragments of ref tile t through buffer loads: a scalar descriptor at the tile,
// the lane's 16-byte slot as the vector offset and the instruction's 1 KB block as
// the scalar offset, so the loads keep no per-lane address registers (whose reuse
// by later results would make the products wait for loads still in flight).
template <int NK>
__device__ __forceinline__ void load_tile(const bf16x8* rfrag, int32_t t, uint32_t lane_off,
                                          bf16x8 (&af)[NK]) {
    __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(rfrag + (int64_t)t * NK * 64), (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
    for (int c = 0; c < NK; ++c)
        af[c] = __builtin_bit_cast(
            bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)lane_off, c * 1024, 0));
}

// One wave per query tile.  The tiles below every lane's limit stream through a
// ring of P register sets, P - 1 tiles in flight ahead of the one being scored
// (the fragment arrays hold kPadTiles padding tiles past the last); the masked
// tail tiles load one at a time.
template <int NK, int K, int B, int P>
__global__ __launch_bounds__(64) void mfma_screen_kernel(MfmaScanArgs a) {
    __shared__ float sk[B][64];
    __shared__ int si[B][64];
    const int l = threadIdx.x;
    const int64_t qt =
        a.qt_first + (a.reverse ? (int64_t)(gridDim.x - 1 - blockIdx.x) : (int64_t)blockIdx.x);
    const RowWindow win = row_window(a, qt, l);
    bf16x8 qf[NK];
    load_query<NK>(a, qt, l, qf);
    LaneList<K, B> L;
    L.init(win.qv, l, sk, si);
    const uint32_t lane_off = (uint32_t)l * 16u;
    if (win.t_begin < win.t_full) {
        // At the start of a group of P tiles from t, ring[s] holds tile t + s for
        // s < P - 1; step s loads tile t + s + P - 1 into the set step s - 1 freed.
        // Groups run whole (no exit inside, so the load counts stay exact); the
        // last < P tiles are already in the ring.
        bf16x8 ring[P][NK];
#pragma unroll
        for (int p = 0; p < P - 1; ++p) load_tile<NK>(a.rfrag, win.t_begin + p, lane_off, ring[p]);
        int32_t t = win.t_begin;
#pragma unroll 1
        for (; t + P <= win.t_full; t += P) {
#pragma unroll
            for (int s = 0; s < P; ++s) {
                load_tile<NK>(a.rfrag, t + s + P - 1, lane_off, ring[(s + P - 1) % P]);
                L.take(tile_scores<NK>(ring[s], qf), t + s, false, win.lim);
            }
        }
#pragma unroll
        for (int s = 0; s < P - 1; ++s)
            if (t + s < win.t_full) L.take(tile_scores<NK>(ring[s], qf), t + s, false, win.lim);
    }
#pragma unroll 1
    for (int32_t t = win.t_full; t < win.t_end; ++t) {
        bf16x8 fa[NK];
        load_tile<NK>(a.rfrag, t, lane_off, fa);
        L.take(tile_scores<NK>(fa, qf), t, true, win.lim);
    }
    L.finish(a, win.q);
}

// Matrix instructions per tile for d features, rounded up to an instantiated count.
int mfma_nk(int d) {
    const int need = (3 * d + 2 + kChunkK - 1) / kChunkK;
    return need <= 1 ? 1 : (need <= 2 ? 2 : (need <= 4 ? 4 : (need <= 7 ? 7 : 13)));
}

template <int NK, int K>
void launch_nk_k(const MfmaScanArgs& a, unsigned qtiles, unsigned stripes, hipStream_t st) {
    // tiles in registers (measured): four for NK <= 2 with the long lists of K >= 56,
    // two with K <= 40 (where more loads in flight thrash L2), three for NK = 4, two
    // for the wide K-vectors (NK = 7, 13)
    constexpr int P = NK <= 2 ? (K >= 56 ? 4 : 2) : (NK <= 4 ? 3 : 2);
    hipLaunchKernelGGL((mfma_screen_kernel<NK, K, kAppendSlots, P>), dim3(qtiles, stripes),
                       dim3(64), 0, st, a);
}

template <int NK>
int launch_nk(const MfmaScanArgs& a, int K, unsigned qtiles, unsigned stripes, hipStream_t st) {
    switch (K) {
        case 24: launch_nk_k<NK, 24>(a, qtiles, stripes, st); return 0;
        case 40: launch_nk_k<NK, 40>(a, qtiles, stripes, st); return 0;
        case 56: launch_nk_k<NK, 56>(a, qtiles, stripes, st); return 0;
        case 72: launch_nk_k<NK, 72>(a, qtiles, stripes, st); return 0;
        default: return -1;
    }
}

}  // namespace

bool mfma_screen_supported(int d) { return d >= 4 && d <= NNGP_MAX_DIM; }

int mfma_screen_K(int k) { return (k + 15) / 16 * 16 + 8; }

size_t mfma_frag_elems(int64_t n, int d) {
    const int64_t tiles = (std::max<int64_t>(n, 1) + kTileRows - 1) / kTileRows + kPadTiles;
    return (size_t)tiles * mfma_nk(d) * 64 * 8;
}

double mfma_screen_eps(int d) {
    const double fp32 = (4.0 * d + 16.0) * std::ldexp(1.0, -24);
    const double split = 4.02 * std::ldexp(1.0, -16) + (3.0 * d + 3.0) * 1.02 * std::ldexp(1.0, -23);
    return fp32 + 2.0 * split;
}

int mfma_screen_stripes(int64_t n_ref, int64_t n_query) {
    const int64_t waves = std::max<int64_t>((n_query + kTileRows - 1) / kTileRows, 1);
    int64_t s = std::min<int64_t>(2048 / waves, kMaxListsMfma / 2);
    s = std::min<int64_t>(s, n_ref / 4096);
    return (int)std::max<int64_t>(s, 1);
}

void mfma_prep(const double* u64, int64_t n, int d, const double* centre, int role,
               uint16_t* frag, double* p2, unsigned int* smax, hipStream_t st) {
    const int64_t npad =
        ((std::max<int64_t>(n, 1) + kTileRows - 1) / kTileRows + kPadTiles) * kTileRows;
    hipLaunchKernelGGL(mfma_prep_kernel, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, st,
                       u64, n, npad, d, centre, mfma_nk(d), role, frag, p2, smax);
}

int launch_mfma_screen(const uint16_t* rfrag, int64_t n_ref, const uint16_t* qfrag,
                       int64_t n_query, int d, int K, int has_cut, int64_t cut, int32_t* cand,
                       float* thr, uint64_t* spk, hipStream_t st) {
    MfmaScanArgs a{};
    a.rfrag = reinterpret_cast<const bf16x8*>(rfrag);
    a.n_ref = (int32_t)n_ref;
    a.qfrag = reinterpret_cast<const bf16x8*>(qfrag);
    a.n_query = n_query;
    a.has_cut = has_cut;
    a.cut = cut;
    a.reverse = has_cut;
    a.cand = cand;
    a.thr = thr;
    int stripes = spk ? mfma_screen_stripes(n_ref, n_query) : 1;
    if (stripes > 1) {
        a.stripe_w = (int32_t)round_up((n_ref + stripes - 1) / stripes, kTileRows);
        stripes = (int)((n_ref + a.stripe_w - 1) / a.stripe_w);
        a.out_pk = spk;
    }
    // Generations of kGenTiles query tiles (about two waves per SIMD), one launch
    // each, the longest rows first: the waves of a launch start the ref stream
    // together, so they read the same tiles at about the same time and share L2.
    const int64_t qtiles = (n_query + kTileRows - 1) / kTileRows;
    for (int64_t done = 0; done < qtiles; done += kGenTiles) {
        const int64_t cnt = std::min<int64_t>(kGenTiles, qtiles - done);
        a.qt_first = has_cut ? qtiles - done - cnt : done;
        int rc = -1;
        switch (mfma_nk(d)) {
            case 1: rc = launch_nk<1>(a, K, (unsigned)cnt, (unsigned)stripes, st); break;
            case 2: rc = launch_nk<2>(a, K, (unsigned)cnt, (unsigned)stripes, st); break;
            case 4: rc = launch_nk<4>(a, K, (unsigned)cnt, (unsigned)stripes, st); break;
            case 7: rc = launch_nk<7>(a, K, (unsigned)cnt, (unsigned)stripes, st); break;
            case 13: rc = launch_nk<13>(a, K, (unsigned)cnt, (unsigned)stripes, st); break;
            default: break;
        }
        if (rc != 0) return -1;
    }
    return stripes > 1 ? 2 * stripes : 0;
}

}  // namespace nngp
