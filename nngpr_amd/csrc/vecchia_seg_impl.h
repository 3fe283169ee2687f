// Fused Vecchia segment kernel for m <= 64, d <= 32 (gfx950, wave64): S = MR / 2
// lanes per neighbour block (MR = 16, 32 or 64), G = 64 / S blocks per wave.
// Templates only: vecchia_seg.hip instantiates the RBF kernels and
// vecchia_seg_matern.hip the Matern ones (two translation units, compiled in
// parallel).
//
// Lane t of a block owns the two rows t and MR - 1 - t of the MR x MR neighbour
// covariance (rows past m are padding and behave like the reference's -1 slots:
// identity row/column, zero target), so every lane holds MR + 1 lower-triangle
// entries and the factor's work is balanced over the lanes.
//
//   fill     each lane computes the FP64 covariance of its rows' lower triangle
//            (and of the point itself); the strictly lower triangle stays in
//            LDS for the adjoint, the rows go to registers (type T);
//   factor   right-looking Cholesky with look-ahead: at step k every lane reads
//            column k from one of two LDS buffers, publishes column k + 1 into
//            the other and updates its rows (one barrier per step);
//   solves   forward substitution broadcasts z_k from the row's owner lane;
//            backward substitution (gradient only) sums the lanes' partial
//            products per step;
//   reduce   sigma^2, mu, the log-likelihood and the adjoint gradient
//            (dl/dtheta = g_s dc + dk.(a - 2 g_s b) + sum_ij dK_ij W_ij, each
//            unordered pair (i, j) counted once with weight 2).
// Same semantics as the generic and register kernels: no K or K_gradient ever
// reaches HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "common.h"
#include "vecchia.h"
#include "vecchia_reg.h"

namespace nngp {

namespace {

constexpr double kHalfLog2PiS = 0.91893853320467274178;
constexpr int kSegMaxDim = 32;
// Pivot-ratio threshold of the FP32 form: blocks with max(diag) / min(pivot)
// above it are re-factored in FP64.
constexpr double kSegPivotRatioMax = 1.0e3;

// Lane exchanges inside a segment of S lanes (a 16-lane DPP row for S = 16, half
// a row for S = 8) through DPP moves, 32 bits at a time.
// (Every pattern used here has a source lane for every lane, so bound_ctrl never
// applies; it spares the moves that would initialise an `old` operand.)
template <int CTRL>
__device__ __forceinline__ float dpp_x(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_x(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int kDppQuadSwap1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;   // lane i <-> 7 - i within each 8 lanes
constexpr int kDppMirror = 0x140;       // lane i <-> 15 - i within each 16 lanes
constexpr int kDppRowBcast = 0x150;     // + n: lane n of each 16 lanes to all of them

// Sum over the segment, the same value in every lane (each level adds two equal
// partial sums in either order).  A 32-lane segment adds its two rows' sums
// through one lane permute.
template <int S, typename T>
__device__ __forceinline__ T seg_sum(T v) {
    static_assert(S == 8 || S == 16 || S == 32, "segments of 8, 16 or 32 lanes");
    v += dpp_x<kDppQuadSwap1>(v);
    v += dpp_x<kDppQuadSwap2>(v);
    v += dpp_x<kDppHalfMirror>(v);
    if constexpr (S >= 16) v += dpp_x<kDppMirror>(v);
    if constexpr (S == 32) v += __shfl_xor(v, 16, 32);
    return v;
}

// Value of lane n (0 <= n < S, a constant once unrolled) of the segment.
template <int S, typename T>
__device__ __forceinline__ T seg_bcast(T v, int n, int lane) {
    static_assert(S == 8 || S == 16 || S == 32, "segments of 8, 16 or 32 lanes");
    if constexpr (S == 32) {
        return __shfl(v, n, 32);
    } else if constexpr (S == 16) {
        switch (n) {
            case 0: return dpp_x<kDppRowBcast + 0>(v);
            case 1: return dpp_x<kDppRowBcast + 1>(v);
            case 2: return dpp_x<kDppRowBcast + 2>(v);
            case 3: return dpp_x<kDppRowBcast + 3>(v);
            case 4: return dpp_x<kDppRowBcast + 4>(v);
            case 5: return dpp_x<kDppRowBcast + 5>(v);
            case 6: return dpp_x<kDppRowBcast + 6>(v);
            case 7: return dpp_x<kDppRowBcast + 7>(v);
            case 8: return dpp_x<kDppRowBcast + 8>(v);
            case 9: return dpp_x<kDppRowBcast + 9>(v);
            case 10: return dpp_x<kDppRowBcast + 10>(v);
            case 11: return dpp_x<kDppRowBcast + 11>(v);
            case 12: return dpp_x<kDppRowBcast + 12>(v);
            case 13: return dpp_x<kDppRowBcast + 13>(v);
            case 14: return dpp_x<kDppRowBcast + 14>(v);
            default: return dpp_x<kDppRowBcast + 15>(v);
        }
    } else {
        // lane n % 4 of every quad, then the other quad of the 8 if n / 4 differs
        T q;
        switch (n & 3) {
            case 0: q = dpp_x<0x00>(v); break;   // quad_perm [0,0,0,0]
            case 1: q = dpp_x<0x55>(v); break;   // quad_perm [1,1,1,1]
            case 2: q = dpp_x<0xAA>(v); break;   // quad_perm [2,2,2,2]
            default: q = dpp_x<0xFF>(v); break;  // quad_perm [3,3,3,3]
        }
        const T m = dpp_x<kDppHalfMirror>(q);
        return ((lane >> 2) & 1) == (n >> 2) ? q : m;
    }
}

// 1 / sqrt(x) for x > 0: the hardware estimate and Newton steps
// y <- y + y (1/2 - (x/2) y^2) (two in FP64, one in FP32).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = __builtin_amdgcn_rsqf(x);
    const float h = 0.5f * x;
    y = fmaf(y, fmaf(-h * y, y, 0.5f), y);
    return y;
}

// Validity of row j of this block: rows < S are bit j of vlo, rows >= S are bit
// MR - 1 - j of vhi (the lanes' ballots of their first and second rows).
template <int MR>
__device__ __forceinline__ bool row_valid(int j, uint32_t vlo, uint32_t vhi) {
    constexpr int S = MR / 2;
    return j < S ? ((vlo >> j) & 1u) : ((vhi >> (MR - 1 - j)) & 1u);
}

// Entry (i, j), j < i, of a strictly lower triangle stored by rows.
__device__ __forceinline__ int stri(int i, int j) { return i * (i - 1) / 2 + j; }

// Returns v unchanged but opaque to the compiler: comparisons of the lane's row
// indices against an unrolled step index are then formed inside each step (and
// consumed there) instead of being hoisted and kept live as lane masks.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// L z = rhs for a lane's two rows, NR right-hand sides at once.  With the
// strictly lower part of the rows in registers (L0: row t, L1: row MR - 1 - t) a
// row's running value stops changing once its own step has passed, and z_i is
// that value times 1 / L_ii.
template <typename T, int MR, int NR>
__device__ __forceinline__ void seg_fwd(const T (&L0)[MR / 2], const T (&L1)[MR], T dv0, T dv1,
                                        T (&r0)[NR], T (&r1)[NR], int lane) {
    constexpr int S = MR / 2;
#pragma unroll
    for (int k = 0; k < MR; ++k) {
        const int own = k < S ? k : MR - 1 - k;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            // the owner of row k scales its running value by its 1 / L_kk
            const T zk = seg_bcast<S>(k < S ? r0[q] * dv0 : r1[q] * dv1, own, lane);
            if (k < S) r0[q] = fma(-L0[k < S ? k : 0], zk, r0[q]);
            r1[q] = fma(-L1[k], zk, r1[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        r0[q] *= dv0;
        r1[q] *= dv1;
    }
}

// L^T x = z for a lane's two rows, NR right-hand sides at once (x starts at 0:
// rows not yet solved add nothing through the cleared entries).  z in r0 / r1 is
// replaced by x.  Only the owner of row k needs x_k, and it holds z_k and 1 / L_kk.
template <typename T, int MR, int NR>
__device__ __forceinline__ void seg_bwd(const T (&L0)[MR / 2], const T (&L1)[MR], T dv0, T dv1,
                                        T (&r0)[NR], T (&r1)[NR], int t) {
    constexpr int S = MR / 2;
    T x0[NR], x1[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        x0[q] = (T)0;
        x1[q] = (T)0;
    }
#pragma unroll
    for (int k = MR - 1; k >= 0; --k) {
        const int own = k < S ? k : MR - 1 - k;
        T part[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q)
            part[q] = seg_sum<S>(fma(L1[k], x1[q], k < S ? L0[k < S ? k : 0] * x0[q] : (T)0));
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const T xk = k < S ? (r0[q] - part[q]) * dv0 : (r1[q] - part[q]) * dv1;
            if (k < S) {
                if (t == k) x0[q] = xk;
            } else {
                if (t == own) x1[q] = xk;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        r0[q] = x0[q];
        r1[q] = x1[q];
    }
}

// Entry e of a lane (0 <= e < MR + 3): (i0, 0..i0), (i1, 0..i1), then (i0, MR) and
// (i1, MR), MR standing for the point itself.  `first`: the entry is on row i0.
struct SegEntry {
    int i, j;
    bool first;
};
template <int MR>
__device__ __forceinline__ SegEntry seg_entry(int e, int i0, int i1) {
    const bool in0 = e <= i0;
    const bool in1 = !in0 & (e <= MR);
    const bool self0 = e == MR + 1;
    SegEntry s;
    s.first = in0 | self0;
    s.i = s.first ? i0 : i1;
    s.j = in0 ? e : (in1 ? e - i0 - 1 : MR);
    return s;
}

// One group of G points starting at slot `base`.  Coordinates are held with DM
// features (DM >= d, zero padded).  MAT: Matern stationary part (kp.fam), else RBF.
// CW (FP64 prediction only): conditioning weights of a sequential prediction -
// alpha only on neighbours with index < a.alpha_limit, and w = K^-1 k and
// s2 = K_00 - k.w written to a.wout / a.s2out next to the mean.
template <typename T, int MR, int DM, bool MAT, bool CW = false>
__device__ __forceinline__ void seg_body(const VecArgs& a, const RegExtra& ex, int64_t base,
                                         int64_t count) {
    constexpr int S = MR / 2;
    constexpr int G = 64 / S;
    constexpr int TRI = MR * (MR - 1) / 2;  // strictly lower triangle
    constexpr int NE = MR + 3;  // entries per lane
    constexpr int U = 5;        // entries per step of the fill loop
    constexpr bool REFINE = sizeof(T) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int seg = lane / S;
    const int t = lane % S;
    const int i0 = t;               // first row of this lane
    const int i1 = MR - 1 - t;      // second row
    const int d = a.d;
    const int ny = a.ny;
    const KernParams& kp = a.kp;

    // LDS per block (G blocks): FP64 covariance (strictly lower triangle; the
    // diagonal is the same for every valid row), two FP64 vectors, X, the point's
    // own target.  The factor's column broadcast (T) shares the first vector's
    // space: it is dead before that vector is first written.
    double* Kt = reinterpret_cast<double*>(smem) + seg * TRI;
    double* xv = reinterpret_cast<double*>(smem) + G * TRI + seg * MR;
    double* vbuf = reinterpret_cast<double*>(smem) + G * (TRI + MR) + seg * MR;
    double* X = reinterpret_cast<double*>(smem) + G * (TRI + 2 * MR) + seg * (MR + 1) * DM;
    double* yself = reinterpret_cast<double*>(smem) + G * (TRI + 2 * MR + (MR + 1) * DM) + seg;

    const int64_t slot = base + seg;
    const bool pv = slot < count;
    const int64_t p = pv ? (ex.plist ? ex.plist[slot] : slot) : 0;

    // ---------------- gather ---------------------------------------------
    // Coordinates of the lane's rows (registers and LDS), the point's own
    // coordinates, and the first target column (registers; yself: own).  Every
    // load reads a clamped, valid address and is selected afterwards, so the
    // loads of each round are in flight together.
    const int64_t nrow = p * a.m;
    const int n0 = a.nn[nrow + min(i0, a.m - 1)];
    const int n1 = a.nn[nrow + min(i1, a.m - 1)];
    double xs_r[DM];
#pragma unroll
    for (int f = 0; f < DM; ++f) xs_r[f] = a.xs[p * d + min(f, d - 1)];
    const double ys_r = a.ys ? a.ys[p * ny] : 0.0;
    const int j0 = (pv && i0 < a.m) ? nbr_slot(n0, a.n_src, a.hdr) : -1;
    const int j1 = (pv && i1 < a.m) ? nbr_slot(n1, a.n_src, a.hdr) : -1;
    const bool v0 = j0 >= 0, v1 = j1 >= 0;
    const int64_t c0 = v0 ? j0 : 0, c1 = v1 ? j1 : 0;
    double xr0[DM], xr1[DM];
#pragma unroll
    for (int f = 0; f < DM; ++f) {
        const double a0 = a.x[c0 * d + min(f, d - 1)];
        const double a1 = a.x[c1 * d + min(f, d - 1)];
        xr0[f] = (v0 && f < d) ? a0 : 0.0;
        xr1[f] = (v1 && f < d) ? a1 : 0.0;
    }
    const double yl0 = a.y[c0 * ny], yl1 = a.y[c1 * ny];
    const double yn0 = v0 ? yl0 : 0.0;
    const double yn1 = v1 ? yl1 : 0.0;
#pragma unroll
    for (int f = 0; f < DM; ++f) {
        X[i0 * DM + f] = xr0[f];
        X[i1 * DM + f] = xr1[f];
    }
    if (t == 0) {
#pragma unroll
        for (int f = 0; f < DM; ++f) X[MR * DM + f] = (pv && f < d) ? xs_r[f] : 0.0;
        *yself = pv ? ys_r : 0.0;
    }
    const uint64_t b0 = __ballot(v0), b1 = __ballot(v1);
    const uint32_t vlo = (uint32_t)((b0 >> (seg * S)) & ((1ull << S) - 1ull));
    const uint32_t vhi = (uint32_t)((b1 >> (seg * S)) & ((1ull << S) - 1ull));
    // validity of row j as bit j (rows >= S from the second rows' ballot, reversed)
    const uint64_t vmask =
        (uint64_t)vlo | ((uint64_t)(__builtin_bitreverse32(vhi) >> (32 - S)) << S);
    __syncthreads();

    // per-feature weights of the squared distance (ARD: h_f; isotropic: applied after)
    const bool ard = kp.n_ls > 1;
    double hw[DM];
#pragma unroll
    for (int f = 0; f < DM; ++f) hw[f] = (ard && f < d) ? kp.h[f] : 0.0;
    // scaled squared distance between row coordinates xi (registers or LDS) and X[j];
    // ard_c: std::true_type / std::false_type (narrow rows: no branch in the
    // entry loop) or a run-time bool (wide rows)
    auto sqdist = [&](auto ard_c, const auto& xi, int j) -> double {
        bool is_ard;
        if constexpr (std::is_same_v<decltype(ard_c), bool>) is_ard = ard_c;
        else is_ard = decltype(ard_c)::value;
        const double* xj = X + j * DM;
        double u = 0.0;
        if (!is_ard) {
#pragma unroll
            for (int f = 0; f < DM; ++f) {
                const double tt = xi[f] - xj[f];
                u = fma(tt, tt, u);
            }
            u *= kp.h[0];
        } else {
#pragma unroll
            for (int f = 0; f < DM; ++f) {
                const double tt = xi[f] - xj[f];
                u = fma(tt * tt, hw[f], u);
            }
        }
        return u;
    };

    // ---------------- fill: FP64 covariance of the lane's rows -------------
    const double diag_add = kp.w + a.alpha;
    const double inv_s = kp.s != 0.0 ? 1.0 / kp.s : 0.0;
    // stationary part E of a valid off-diagonal pair from its covariance entry
    auto e_of = [&](double K) -> double { return kp.n_ls > 0 ? (K - kp.C) * inv_s : 0.0; };
    double e0s = 0.0, e1s = 0.0;
    // Branch-free: every entry's stationary value is evaluated and then selected, and
    // the entries that are not stored (the two diagonal ones, the two against the
    // point, the steps past the last entry) are written to this lane's slot of the
    // first vector, which the factorisation overwrites before it reads it.  (With
    // the evaluation and the store behind per-lane branches the fill issued about as
    // many scalar as vector instructions, and was 37 % of the pass.)
    double* const spare = xv + t;
    auto fill_entries = [&](auto ard_c) {
#pragma unroll 1
        for (int e0 = 0; e0 < NE; e0 += U) {
#pragma unroll
            for (int uu = 0; uu < U; ++uu) {
                const int e = e0 + uu;
                const SegEntry s = seg_entry<MR>(e, i0, i1);
                const bool vi = s.first ? v0 : v1;
                const bool vj = (s.j >= MR) | (((vmask >> (s.j & 63)) & 1ull) != 0ull);
                double u;
                if constexpr (DM <= 8) {
                    double xi[DM];
#pragma unroll
                    for (int f = 0; f < DM; ++f) xi[f] = s.first ? xr0[f] : xr1[f];
                    u = sqdist(ard_c, xi, s.j);
                } else {  // wide rows: the lane's coordinates are read back from LDS
                    u = sqdist(ard_c, X + s.i * DM, s.j);
                }
                const double Ex = MAT ? cov_e(kp.fam, u) : exp_neg(u);
                const bool live = vi & vj & (s.i != s.j) & (kp.n_ls > 0);
                const double Ev = live ? Ex : (s.i == s.j ? 1.0 : 0.0);
                const bool store = (e <= MR) & (s.j < s.i);
                double* const dst = store ? Kt + stri(s.i, s.j) : spare;
                *dst = (vi & vj) ? fma(kp.s, Ev, kp.C) : 0.0;
                e0s = e == MR + 1 ? Ev : e0s;
                e1s = e == MR + 2 ? Ev : e1s;
            }
        }
    };
    if constexpr (DM <= 8) {
        if (ard) fill_entries(std::true_type{});
        else fill_entries(std::false_type{});
    } else {
        fill_entries(ard);
    }
    const double k0 = v0 ? fma(kp.s, e0s, kp.C) : 0.0;
    const double k1 = v1 ? fma(kp.s, e1s, kp.C) : 0.0;
    const double c = kp.s + kp.C + kp.w + (a.alpha_self ? a.alpha : 0.0);
    double dmax = fmax(v0 ? kp.s + kp.C + diag_add : 1.0, v1 ? kp.s + kp.C + diag_add : 1.0);
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, S));  // once per point
    __syncthreads();

    // ---------------- rows into registers (type T) --------------------------
    // Entries right of a row's diagonal read neighbouring triangle entries: they
    // are never used before being cleared below.  Diagonal: K_ii + w + alpha for
    // a valid row, 1 for a padding row.
    const double kd = fma(kp.s, 1.0, kp.C) + diag_add;
    double kd0 = v0 ? kd : 1.0, kd1 = v1 ? kd : 1.0;
    if constexpr (CW) {
        const double kw = fma(kp.s, 1.0, kp.C) + kp.w;
        kd0 = v0 ? kw + (j0 < a.alpha_limit ? a.alpha : 0.0) : 1.0;
        kd1 = v1 ? kw + (j1 < a.alpha_limit ? a.alpha : 0.0) : 1.0;
    }
    T L0[S], L1[MR];
    {
        const double* r0p = Kt + stri(i0, 0);
        const double* r1p = Kt + stri(i1, 0);
#pragma unroll
        for (int j = 0; j < S; ++j) L0[j] = (T)(j == i0 ? kd0 : r0p[j]);
#pragma unroll
        for (int j = 0; j < MR; ++j) L1[j] = (T)(j == i1 ? kd1 : r1p[j]);
    }

    const double* krow0 = Kt + stri(i0, 0);
    const double* krow1 = Kt + stri(i1, 0);

    // ---------------- Cholesky (type T), no per-lane branches ---------------
    // Step k: the pivot and column k of every row are read from the published
    // buffer; each lane updates its rows with the unscaled column (l_ik l_jk =
    // a_ik a_jk / a_kk), column k + 1 first so that it is published into the
    // other buffer while the rest of the update runs; then it scales its column-k
    // entry (the owner's becomes sqrt(a_kk)).  Rows above k publish and update
    // entries that are never read.  The two buffers share the LDS of the vectors
    // xv and vbuf, both written only after the factorisation.
    T* colA = reinterpret_cast<T*>(xv);
    T* colB = reinterpret_cast<T*>(vbuf);
    bool bad = false;
    T dv0 = (T)1, dv1 = (T)1;
    double pmin = 1e300;
    colA[i0] = L0[0];
    colA[i1] = L1[0];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MR; ++k) {
        const T* cur = (k & 1) ? colB : colA;
        T* nxt = (k & 1) ? colA : colB;
        const T dkk = cur[k];
        bad = bad || !(dkk > (T)0);
        pmin = fmin(pmin, (double)dkk);
        const T inv = rsqrt_nr(dkk);
        // 1 / L_kk kept by the owner of row k
        if (k < S) {
            if (t == k) dv0 = inv;
        } else {
            if (t == MR - 1 - k) dv1 = inv;
        }
        const T tr0 = k < S ? L0[k < S ? k : 0] * inv * inv : (T)0;
        const T tr1 = L1[k] * inv * inv;
        if (k + 1 < MR) {
            const T cn = cur[k + 1];
            if (k + 1 < S) {
                L0[k + 1 < S ? k + 1 : 0] = fma(-tr0, cn, L0[k + 1 < S ? k + 1 : 0]);
                nxt[i0] = L0[k + 1 < S ? k + 1 : 0];
            }
            L1[k + 1 < MR ? k + 1 : 0] = fma(-tr1, cn, L1[k + 1 < MR ? k + 1 : 0]);
            nxt[i1] = L1[k + 1 < MR ? k + 1 : 0];
        }
        if (k < S) {
#pragma unroll
            for (int j = k + 2; j < S; ++j) L0[j] = fma(-tr0, cur[j], L0[j]);
            L0[k < S ? k : 0] = L0[k < S ? k : 0] * inv;
        }
#pragma unroll
        for (int j = k + 2; j < MR; ++j) L1[j] = fma(-tr1, cur[j], L1[j]);
        L1[k] = L1[k] * inv;
        __syncthreads();
    }
    // keep the strictly lower part only (diagonal and above cleared)
#pragma unroll
    for (int j = 0; j < S; ++j) L0[j] = j < i0 ? L0[j] : (T)0;
#pragma unroll
    for (int j = 0; j < MR; ++j) L1[j] = j < i1 ? L1[j] : (T)0;

    // FP32 form: ill-conditioned or failed blocks are redone in FP64
    const bool ill = REFINE && (bad || !(dmax / pmin <= kSegPivotRatioMax));
    if (ill && pv && t == 0) {
        const unsigned long long q = atomicAdd(ex.flag_count, 1ULL);
        ex.flag_list[q] = p;
    }
    if (!REFINE && bad && pv && t == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.hdr->not_pd), 1ULL);
        if (a.pt)
            for (int q = 0; q < a.pt_stride; ++q) a.pt[p * a.pt_stride + q] = __builtin_nan("");
        if (a.mean)
            for (int o = 0; o < ny; ++o) a.mean[p * ny + o] = __builtin_nan("");
        if (a.stdv) a.stdv[p] = __builtin_nan("");
    }
    const bool live = pv && !bad && !ill;

    // x = K^-1 rhs for the lane's rows; FP32 form: one FP64-residual refinement
    auto solve = [&](double rhs0, double rhs1, double& x0, double& x1) {
        T u0[1] = {(T)rhs0}, u1[1] = {(T)rhs1};
        seg_fwd<T, MR, 1>(L0, L1, dv0, dv1, u0, u1, lane);
        seg_bwd<T, MR, 1>(L0, L1, dv0, dv1, u0, u1, t);
        x0 = (double)u0[0];
        x1 = (double)u1[0];
        if constexpr (REFINE) {
            xv[i0] = x0;
            xv[i1] = x1;
            __syncthreads();
            double q0 = rhs0, q1 = rhs1;
#pragma unroll 1
            for (int j = 0; j < MR; ++j) {
                const double xj = xv[j];
                const double K0 = j == i0 ? kd0 : (j < i0 ? Kt[stri(i0, j)] : Kt[stri(j, i0)]);
                const double K1 = j == i1 ? kd1 : (j < i1 ? Kt[stri(i1, j)] : Kt[stri(j, i1)]);
                q0 = fma(-K0, xj, q0);
                q1 = fma(-K1, xj, q1);
            }
            __syncthreads();
            T w0[1] = {(T)q0}, w1[1] = {(T)q1};
            seg_fwd<T, MR, 1>(L0, L1, dv0, dv1, w0, w1, lane);
            seg_bwd<T, MR, 1>(L0, L1, dv0, dv1, w0, w1, t);
            x0 += (double)w0[0];
            x1 += (double)w1[0];
        }
    };

    const bool with_grad = a.grad && a.ys;
    double s2, sig, ll = 0.0, Gs = 0.0, at0 = 0.0, at1 = 0.0;
    double bv0 = 0.0, bv1 = 0.0;
    // one target column: mean, and the log-likelihood terms of the point
    auto emit = [&](int o, double mu, double av0, double av1) {
        if (live && a.mean && t == 0) a.mean[p * ny + o] = mu;
        if (a.ys) {
            const double yo = o == 0 ? *yself : (pv ? a.ys[p * ny + o] : 0.0);
            const double res = yo - mu;
            ll += -0.5 * res * res / s2 - kHalfLog2PiS - log(sig);
            if (a.grad) {
                Gs += -0.5 / s2 + 0.5 * res * res / (s2 * s2);
                at0 = fma(res / s2, av0, at0);
                at1 = fma(res / s2, av1, at1);
            }
        }
    };
    if constexpr (!REFINE) {
        // z_k = L^-1 k and z_y = L^-1 y (first column) together: sigma^2 =
        // c - z_k.z_k, mu = z_k.z_y; the gradient needs b = L^-T z_k, a = L^-T z_y
        T r0[2] = {(T)k0, (T)yn0}, r1[2] = {(T)k1, (T)yn1};
        seg_fwd<T, MR, 2>(L0, L1, dv0, dv1, r0, r1, lane);
        const double zk0 = r0[0], zk1 = r1[0];
        s2 = c - seg_sum<S>(zk0 * zk0 + zk1 * zk1);
        sig = sqrt(s2);
        const double mu0 = seg_sum<S>(zk0 * (double)r0[1] + zk1 * (double)r1[1]);
        if constexpr (CW) {
            T u0[1] = {r0[0]}, u1[1] = {r1[0]};
            seg_bwd<T, MR, 1>(L0, L1, dv0, dv1, u0, u1, t);
            if (live && i0 < a.m) a.wout[p * a.m + i0] = (double)u0[0];
            if (live && i1 < a.m) a.wout[p * a.m + i1] = (double)u1[0];
            if (live && t == 0) a.s2out[p] = s2;
        }
        double av0 = 0.0, av1 = 0.0;
        if (with_grad) {
            seg_bwd<T, MR, 2>(L0, L1, dv0, dv1, r0, r1, t);
            bv0 = r0[0];
            bv1 = r1[0];
            av0 = r0[1];
            av1 = r1[1];
        }
        emit(0, mu0, av0, av1);
        for (int o = 1; o < ny; ++o) {
            T u0[1] = {(T)(v0 ? a.y[(int64_t)j0 * ny + o] : 0.0)};
            T u1[1] = {(T)(v1 ? a.y[(int64_t)j1 * ny + o] : 0.0)};
            seg_fwd<T, MR, 1>(L0, L1, dv0, dv1, u0, u1, lane);
            const double mu = seg_sum<S>(zk0 * (double)u0[0] + zk1 * (double)u1[0]);
            double a0 = 0.0, a1 = 0.0;
            if (with_grad) {
                seg_bwd<T, MR, 1>(L0, L1, dv0, dv1, u0, u1, t);
                a0 = u0[0];
                a1 = u1[0];
            }
            emit(o, mu, a0, a1);
        }
    } else {
        // FP32 form: refined solves against the FP64 entries; k and the first target
        // column go through the factor together (each column's arithmetic is the
        // one-column solve's), the other columns one at a time
        double av0, av1;
        {
            T u0[2] = {(T)k0, (T)yn0}, u1[2] = {(T)k1, (T)yn1};
            seg_fwd<T, MR, 2>(L0, L1, dv0, dv1, u0, u1, lane);
            seg_bwd<T, MR, 2>(L0, L1, dv0, dv1, u0, u1, t);
            bv0 = (double)u0[0];
            bv1 = (double)u1[0];
            av0 = (double)u0[1];
            av1 = (double)u1[1];
            xv[i0] = bv0;
            xv[i1] = bv1;
            vbuf[i0] = av0;
            vbuf[i1] = av1;
            __syncthreads();
            double qk0 = k0, qk1 = k1, qy0 = yn0, qy1 = yn1;
#pragma unroll 1
            for (int j = 0; j < MR; ++j) {
                const double xj = xv[j], yj = vbuf[j];
                const double K0 = j == i0 ? kd0 : (j < i0 ? Kt[stri(i0, j)] : Kt[stri(j, i0)]);
                const double K1 = j == i1 ? kd1 : (j < i1 ? Kt[stri(i1, j)] : Kt[stri(j, i1)]);
                qk0 = fma(-K0, xj, qk0);
                qk1 = fma(-K1, xj, qk1);
                qy0 = fma(-K0, yj, qy0);
                qy1 = fma(-K1, yj, qy1);
            }
            __syncthreads();
            T w0[2] = {(T)qk0, (T)qy0}, w1[2] = {(T)qk1, (T)qy1};
            seg_fwd<T, MR, 2>(L0, L1, dv0, dv1, w0, w1, lane);
            seg_bwd<T, MR, 2>(L0, L1, dv0, dv1, w0, w1, t);
            bv0 += (double)w0[0];
            bv1 += (double)w1[0];
            av0 += (double)w0[1];
            av1 += (double)w1[1];
        }
        s2 = c - seg_sum<S>(k0 * bv0 + k1 * bv1);
        sig = sqrt(s2);
        emit(0, seg_sum<S>(k0 * av0 + k1 * av1), av0, av1);
        for (int o = 1; o < ny; ++o) {
            const double y0 = v0 ? a.y[(int64_t)j0 * ny + o] : 0.0;
            const double y1 = v1 ? a.y[(int64_t)j1 * ny + o] : 0.0;
            solve(y0, y1, av0, av1);
            emit(o, seg_sum<S>(k0 * av0 + k1 * av1), av0, av1);
        }
    }
    if (live && a.stdv && t == 0) a.stdv[p] = sig;
    if (live && a.pt && t == 0) a.pt[p * a.pt_stride] = ll;

    // ---------------- adjoint contraction ----------------------------------
    // Over each of the lane's rows i: the self term (i, point) against dk_i and
    // the pairs (i, j <= i) with weight W_ij (doubled off the diagonal) against
    // dK_ij.  Column j runs over the same j in every lane (b_j, a_j and X[j] are
    // broadcast reads).
    if (a.grad && a.pt) {
        xv[i0] = bv0;
        xv[i1] = bv1;
        vbuf[i0] = at0;
        vbuf[i1] = at1;
        __syncthreads();
        double acc_w = 0.0, acc_C = 0.0, acc_s = 0.0, acc_l = 0.0;
        double accf[DM];
#pragma unroll
        for (int f = 0; f < DM; ++f) accf[f] = 0.0;
        // Length-scale term of the pair (xi, X[j]) with dk / dK weight W and
        // stationary value E: W G against the squared distance (accumulated
        // without the factor 2 h_f, applied at the end).  RBF: W G = WE = W E.
        // ard_c: std::true_type / std::false_type (narrow rows: the pair loops
        // carry no branch on it) or a run-time bool (wide rows: one copy of the
        // loops keeps the register budget).
        auto add_dist = [&](auto ard_c, const double (&xi)[DM], int j, double W, double E,
                            double WE) {
            bool is_ard;
            if constexpr (std::is_same_v<decltype(ard_c), bool>) is_ard = ard_c;
            else is_ard = decltype(ard_c)::value;
            const double* xj = X + j * DM;
            if (!is_ard) {
                double u = 0.0;
#pragma unroll
                for (int f = 0; f < DM; ++f) {
                    const double tt = xi[f] - xj[f];
                    u = fma(tt, tt, u);
                }
                const double WG = MAT ? W * cov_g(kp.fam, u * kp.h[0], E) : WE;
                acc_l = fma(WG, u, acc_l);
            } else {
                double WG = WE;
                if constexpr (MAT) {
                    double uh = 0.0;
#pragma unroll
                    for (int f = 0; f < DM; ++f) {
                        const double tt = xi[f] - xj[f];
                        uh = fma(tt * tt, hw[f], uh);
                    }
                    WG = W * cov_g(kp.fam, uh, E);
                }
#pragma unroll
                for (int f = 0; f < DM; ++f) {
                    const double tt = xi[f] - xj[f];
                    accf[f] = fma(WG * tt, tt, accf[f]);
                }
            }
        };
        // Row i: the self term, the diagonal (W_ii = G_s b_i^2 - b_i a_i, with
        // dK_ii free of the length scales) and the off-diagonal pairs j < i with
        // W_ij doubled: 2 G_s b_i b_j - b_i a_j - a_i b_j = c1 b_j - b_i a_j.
        auto row_adj = [&](auto ard_c, auto ncols, int i, const double* krow, bool vi, double bi,
                           double ati, double eis) {
            constexpr int N = decltype(ncols)::value;
            double xi[DM];
#pragma unroll
            for (int f = 0; f < DM; ++f) xi[f] = X[i * DM + f];
            {
                const double v = vi ? ati - 2.0 * Gs * bi : 0.0;
                acc_C += v;
                acc_s += v * eis;
                add_dist(ard_c, xi, MR, v, eis, v * eis);
                const double wd = vi ? (Gs * bi - ati) * bi : 0.0;
                acc_w += wd;
                acc_C += wd;
                acc_s += wd;
            }
            const double c1 = 2.0 * Gs * bi - ati;
            // wide rows unroll less: the loads of unrolled iterations are hoisted
            constexpr int UA = DM <= 3 ? 4 : (DM <= 8 ? 2 : 1);
#pragma unroll UA
            for (int j = 0; j < N; ++j) {
                const bool ok = vi & (j < i) & row_valid<MR>(j, vlo, vhi);
                // b_j, a_j and the entry are loaded unconditionally and pinned in
                // registers, so that the mask is a select after the loads rather
                // than a branch around them.  For j >= i the entry read is the
                // next row's (for the last row of the last segment: the first
                // vector after the triangles), inside the block's LDS, and the
                // select discards it.
                double bj = xv[j], aj = vbuf[j], kij = krow[j];
                asm volatile("" : "+v"(bj), "+v"(aj), "+v"(kij));
                const double W = ok ? fma(c1, bj, -bi * aj) : 0.0;
                const double E = ok ? e_of(kij) : 0.0;
                const double WE = W * E;
                acc_C += W;
                acc_s += WE;
                add_dist(ard_c, xi, j, W, E, WE);
            }
        };
        auto adjoint_rows = [&](auto ard_c) {
            row_adj(ard_c, std::integral_constant<int, S>{}, i0, krow0, v0, bv0, at0, e0s);
            row_adj(ard_c, std::integral_constant<int, MR>{}, i1, krow1, v1, bv1, at1, e1s);
        };
        if constexpr (DM <= 8) {
            if (ard) adjoint_rows(std::true_type{});
            else adjoint_rows(std::false_type{});
        } else {
            adjoint_rows(ard);
        }
        acc_l *= 2.0 * kp.h[0];
#pragma unroll
        for (int f = 0; f < DM; ++f) accf[f] *= 2.0 * hw[f];
        acc_w = seg_sum<S>(acc_w);
        acc_C = seg_sum<S>(acc_C);
        acc_s = seg_sum<S>(acc_s);
        acc_l = seg_sum<S>(acc_l);
        if (ard) {
#pragma unroll
            for (int f = 0; f < DM; ++f) accf[f] = seg_sum<S>(accf[f]);
        }
        // every gradient entry written once
        if (live && t == 0) {
            double* gp = a.pt + p * a.pt_stride + 1;
            for (int q = 0; q < a.n_theta; ++q) {
                double v = 0.0;
                if (q == kp.t_w) v = kp.w * (acc_w + Gs);
                else if (q == kp.t_C) v = kp.C * (acc_C + Gs);
                else if (q == kp.t_s) v = kp.s * (acc_s + Gs);
                else if (kp.t_l >= 0 && q >= kp.t_l && q < kp.t_l + (ard ? d : 1)) {
                    if (!ard) {
                        v = kp.s * acc_l;
                    } else {
#pragma unroll
                        for (int f = 0; f < DM; ++f)
                            if (q - kp.t_l == f) v = kp.s * accf[f];
                    }
                }
                gp[q] = v;
            }
        }
    }
}

// One group of G points per block (no persistent loop: nothing loop-invariant is
// kept live across groups).
// (MR = 64: the LDS of a block keeps occupancy under one wave per SIMD, so the
// register budget is a whole SIMD's.)
// (RBF, Matern and conditioning-weight kernels carry distinct names so that
// profiles tell them apart.)
template <typename T, int MR, int DM, bool MAT>
__device__ __forceinline__ void seg_group(const VecArgs& a, const RegExtra& ex) {
    constexpr int G = 64 / (MR / 2);
    const int64_t count = ex.pcount ? (int64_t)*ex.pcount : a.npts;
    const int64_t base = (int64_t)blockIdx.x * G;
    if (base < count) seg_body<T, MR, DM, MAT>(a, ex, base, count);
}
template <typename T, int MR, int DM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MR <= 32 ? 2 : 1)))
void vecchia_seg_kernel(VecArgs a, RegExtra ex) {
    seg_group<T, MR, DM, false>(a, ex);
}
template <typename T, int MR, int DM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MR <= 32 ? 2 : 1)))
void vecchia_seg_matern_kernel(VecArgs a, RegExtra ex) {
    seg_group<T, MR, DM, true>(a, ex);
}

// Grid-stride form for point lists whose length is known on the device only.
template <typename T, int MR, int DM, bool MAT>
__device__ __forceinline__ void seg_list(const VecArgs& a, const RegExtra& ex) {
    constexpr int G = 64 / (MR / 2);
    const int64_t count = (int64_t)*ex.pcount;
#pragma unroll 1
    for (int64_t base = (int64_t)blockIdx.x * G; base < count; base += (int64_t)gridDim.x * G) {
        seg_body<T, MR, DM, MAT>(a, ex, base, count);
        __syncthreads();
    }
}
template <typename T, int MR, int DM>
__global__ __launch_bounds__(64) void vecchia_seg_list_kernel(VecArgs a, RegExtra ex) {
    seg_list<T, MR, DM, false>(a, ex);
}
template <typename T, int MR, int DM>
__global__ __launch_bounds__(64) void vecchia_seg_matern_list_kernel(VecArgs a, RegExtra ex) {
    seg_list<T, MR, DM, true>(a, ex);
}
// Conditioning weights (FP64 prediction pass, RBF or Matern).
template <int MR, int DM, bool MAT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MR <= 32 ? 2 : 1)))
void vecchia_seg_weights_kernel(VecArgs a) {
    constexpr int G = 64 / (MR / 2);
    const int64_t base = (int64_t)blockIdx.x * G;
    if (base < a.npts) seg_body<double, MR, DM, MAT, true>(a, RegExtra{}, base, a.npts);
}

template <typename T, int MR, int DM>
size_t seg_lds() {
    constexpr int S = MR / 2;
    constexpr int G = 64 / S;
    constexpr int TRI = MR * (MR - 1) / 2;
    const size_t doubles = (size_t)G * ((size_t)(MR + 1) * DM + TRI + 2 * MR + 1);
    return doubles * sizeof(double);
}

template <typename T, int MR, int DM, bool MAT>
void launch_seg_mr(const VecArgs& a, const RegExtra& ex, int64_t max_points, hipStream_t st) {
    constexpr int G = 64 / (MR / 2);
    if (ex.pcount) {
        // a device-side list: at most 1024 blocks striding over it
        int64_t blocks = (max_points + G - 1) / G;
        blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
        if constexpr (MAT)
            hipLaunchKernelGGL((vecchia_seg_matern_list_kernel<T, MR, DM>), dim3((unsigned)blocks),
                               dim3(64), (seg_lds<T, MR, DM>()), st, a, ex);
        else
            hipLaunchKernelGGL((vecchia_seg_list_kernel<T, MR, DM>), dim3((unsigned)blocks),
                               dim3(64), (seg_lds<T, MR, DM>()), st, a, ex);
        return;
    }
    int64_t blocks = (max_points + G - 1) / G;
    if (blocks < 1) blocks = 1;
    if constexpr (MAT)
        hipLaunchKernelGGL((vecchia_seg_matern_kernel<T, MR, DM>), dim3((unsigned)blocks), dim3(64),
                           (seg_lds<T, MR, DM>()), st, a, ex);
    else
        hipLaunchKernelGGL((vecchia_seg_kernel<T, MR, DM>), dim3((unsigned)blocks), dim3(64),
                           (seg_lds<T, MR, DM>()), st, a, ex);
}

template <typename T, int MR, bool MAT>
void launch_seg_d(const VecArgs& a, const RegExtra& ex, int64_t max_points, hipStream_t st) {
    if (a.d <= 3) launch_seg_mr<T, MR, 3, MAT>(a, ex, max_points, st);
    else if (a.d <= 8) launch_seg_mr<T, MR, 8, MAT>(a, ex, max_points, st);
    else launch_seg_mr<T, MR, kSegMaxDim, MAT>(a, ex, max_points, st);
}

template <int MR, bool MAT>
void launch_seg_weights_mr(const VecArgs& a, hipStream_t st) {
    constexpr int G = 64 / (MR / 2);
    const unsigned blocks = (unsigned)((a.npts + G - 1) / G);
    if (a.d <= 3)
        hipLaunchKernelGGL((vecchia_seg_weights_kernel<MR, 3, MAT>), dim3(blocks), dim3(64),
                           (seg_lds<double, MR, 3>()), st, a);
    else if (a.d <= 8)
        hipLaunchKernelGGL((vecchia_seg_weights_kernel<MR, 8, MAT>), dim3(blocks), dim3(64),
                           (seg_lds<double, MR, 8>()), st, a);
    else
        hipLaunchKernelGGL((vecchia_seg_weights_kernel<MR, kSegMaxDim, MAT>), dim3(blocks),
                           dim3(64), (seg_lds<double, MR, kSegMaxDim>()), st, a);
}

template <bool MAT>
void launch_seg_weights_family(const VecArgs& a, hipStream_t st) {
    if (a.m <= 16) launch_seg_weights_mr<16, MAT>(a, st);
    else if (a.m <= 32) launch_seg_weights_mr<32, MAT>(a, st);
    else launch_seg_weights_mr<64, MAT>(a, st);
}

// fp32: FP32 factor + FP64 refinement form; else the FP64 form.
template <bool MAT>
void launch_seg_family(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points,
                       hipStream_t st) {
    if (fp32) {
        if (a.m <= 16) launch_seg_d<float, 16, MAT>(a, ex, max_points, st);
        else if (a.m <= 32) launch_seg_d<float, 32, MAT>(a, ex, max_points, st);
        else launch_seg_d<float, 64, MAT>(a, ex, max_points, st);
    } else {
        if (a.m <= 16) launch_seg_d<double, 16, MAT>(a, ex, max_points, st);
        else if (a.m <= 32) launch_seg_d<double, 32, MAT>(a, ex, max_points, st);
        else launch_seg_d<double, 64, MAT>(a, ex, max_points, st);
    }
}

}  // namespace
}  // namespace nngp
