// Register-resident fused Vecchia kernel for m <= 64 (gfx950, wave64).
//
// One wave holds G = 64 / MR neighbour blocks (MR = 16, 32 or 64 rows; rows past
// m are padding and behave exactly like the reference's -1 slots: identity
// row/column, zero target).  Lane r of group g owns row r of its block:
//
//   fill     the (MR+1)(MR+2)/2 distinct covariance entries (neighbours + the
//            point itself) are computed once in FP64, spread over the group's
//            lanes, and kept in LDS as the stationary part E_ij;
//   factor   each lane loads its row into registers (type T) and the block is
//            factored right-looking with one LDS column broadcast per step;
//   solves   forward substitution with register rows and lane shuffles; back
//            substitution through a row-major copy of L in LDS;
//   T=float  (FP32 + refinement form): b = K^-1 k and a = K^-1 y are refined once
//            with an FP64 residual against the FP64 entries in LDS; blocks whose
//            pivot ratio flags ill-conditioning (or fail to factor) are queued
//            and recomputed by the T=double kernel;
//   reduce   sigma^2 = c - k.b, mu = k.a, log-det and quadratic-form terms, and
//            the adjoint gradient are accumulated in FP64.
// No K or K_gradient is ever written to HBM.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "vecchia.h"
#include "vecchia_reg.h"

namespace nngp {

namespace {

constexpr double kHalfLog2Pi = 0.91893853320467274178;
// Pivot-ratio threshold of the FP32 form: blocks with max(diag) / min(pivot)
// above it are re-factored in FP64.
constexpr double kPivotRatioMax = 1.0e3;

template <int MR>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
    for (int off = MR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, MR);
    return v;
}
template <int MR>
__device__ __forceinline__ double gmax(double v) {
#pragma unroll
    for (int off = MR / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, MR));
    return v;
}

struct Lds {
    double* E;    // G x MR x (MR+1)
    double* E0;   // G x MR
    double* X;    // G x (MR+1) x d
    double* B;    // G x MR
    double* A;    // G x MR
    int* I;       // G x MR
    void* L;      // G x MR x (MR+1) of T
    void* col;    // 2 x G x MR of T
};

template <typename T, int MR>
__device__ __forceinline__ T fwd_solve(T z, const T (&a)[MR], T dinv, int r) {
#pragma unroll
    for (int k = 0; k < MR; ++k) {
        T zk = __shfl(z * dinv, k, MR);
        if (r == k) z = zk;
        else if (r > k) z = fma(-a[k], zk, z);
    }
    return z;
}

template <typename T, int MR>
__device__ __forceinline__ T bwd_solve(T x, const T* Lg, T dinv, int r) {
#pragma unroll
    for (int k = MR - 1; k >= 0; --k) {
        T xk = __shfl(x * dinv, k, MR);
        if (r == k) x = xk;
        else if (r < k) x = fma(-Lg[k * (MR + 1) + r], xk, x);
    }
    return x;
}

template <typename T, int MR>
__global__ __launch_bounds__(64) void vecchia_reg_kernel(VecArgs a, RegExtra ex) {
    constexpr int G = 64 / MR;
    constexpr bool REFINE = sizeof(T) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int g = lane / MR;
    const int r = lane % MR;
    const int d = a.d;
    const int ny = a.ny;
    const KernParams& kp = a.kp;

    Lds s;
    {
        unsigned char* p = smem;
        s.E = reinterpret_cast<double*>(p); p += sizeof(double) * G * MR * (MR + 1);
        s.E0 = reinterpret_cast<double*>(p); p += sizeof(double) * G * MR;
        s.B = reinterpret_cast<double*>(p); p += sizeof(double) * G * MR;
        s.A = reinterpret_cast<double*>(p); p += sizeof(double) * G * MR;
        s.X = reinterpret_cast<double*>(p); p += sizeof(double) * G * (MR + 1) * d;
        s.L = p; p += sizeof(T) * G * MR * (MR + 1);
        s.col = p; p += sizeof(T) * 2 * G * MR;
        s.I = reinterpret_cast<int*>(p);
    }
    double* Eg = s.E + g * MR * (MR + 1);
    double* E0g = s.E0 + g * MR;
    double* Xg = s.X + g * (MR + 1) * d;
    double* Bg = s.B + g * MR;
    double* Ag = s.A + g * MR;
    int* Ig = s.I + g * MR;
    T* Lg = reinterpret_cast<T*>(s.L) + g * MR * (MR + 1);
    T* col0 = reinterpret_cast<T*>(s.col) + g * MR;
    T* col1 = reinterpret_cast<T*>(s.col) + (G + g) * MR;

    const int64_t count = ex.pcount ? (int64_t)*ex.pcount : a.npts;
    for (int64_t base = (int64_t)blockIdx.x * G; base < count; base += (int64_t)gridDim.x * G) {
        const int64_t slot = base + g;
        const bool pv = slot < count;
        const int64_t p = pv ? (ex.plist ? ex.plist[slot] : slot) : 0;

        // ---------------- gather --------------------------------------------
        int jr = -1;
        if (pv && r < a.m) jr = nbr_slot(a.nn[p * a.m + r], a.n_src, a.hdr);
        Ig[r] = jr;
        for (int f = 0; f < d; ++f) Xg[r * d + f] = jr >= 0 ? a.x[(int64_t)jr * d + f] : 0.0;
        for (int f = r; f < d; f += MR) Xg[MR * d + f] = pv ? a.xs[p * d + f] : 0.0;
        __syncthreads();
        const unsigned long long vball = __ballot(jr >= 0);
        const unsigned long long vbits =
            MR == 64 ? vball : ((vball >> (g * MR)) & ((1ULL << (MR & 63)) - 1ULL));
        const bool vr = jr >= 0;

        // ---------------- fill (FP64 stationary part) ------------------------
        constexpr int TRI = MR * (MR + 1) / 2;
        for (int e = r; e < TRI + MR; e += MR) {
            int i, j;
            bool self = e >= TRI;
            if (self) {
                i = e - TRI;
                j = MR;
            } else {
                i = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
                while ((i + 1) * (i + 2) / 2 <= e) ++i;
                while (i * (i + 1) / 2 > e) --i;
                j = e - i * (i + 1) / 2;
            }
            const bool vi = (vbits >> i) & 1ULL;
            const bool vj = self ? pv : ((vbits >> j) & 1ULL);
            double E = 0.0;
            if (vi && vj && kp.n_ls > 0) {
                double u = 0.0;
                const double* xi = Xg + i * d;
                const double* xj = Xg + j * d;
                if (kp.n_ls == 1) {
                    for (int f = 0; f < d; ++f) {
                        double t = xi[f] - xj[f];
                        u = fma(t, t, u);
                    }
                    u *= kp.h[0];
                } else {
                    for (int f = 0; f < d; ++f) {
                        double t = xi[f] - xj[f];
                        u = fma(t * t, kp.h[f], u);
                    }
                }
                E = cov_e(kp.fam, u);
            }
            if (self) {
                E0g[i] = E;
            } else {
                Eg[i * (MR + 1) + j] = E;
                Eg[j * (MR + 1) + i] = E;
            }
        }
        __syncthreads();

        // ---------------- rows into registers --------------------------------
        const double diag_add = kp.w + a.alpha;
        T row[MR];
#pragma unroll
        for (int j = 0; j < MR; ++j) {
            const bool vj = (vbits >> j) & 1ULL;
            double K = (vr && vj) ? fma(kp.s, Eg[r * (MR + 1) + j], kp.C) : 0.0;
            if (j == r) K = vr ? K + diag_add : 1.0;
            row[j] = (T)K;
        }
        const double kv = vr ? fma(kp.s, E0g[r], kp.C) : 0.0;
        const double c = kp.s + kp.C + kp.w + (a.alpha_self ? a.alpha : 0.0);
        const double dmax = gmax<MR>(vr ? kp.s + kp.C + diag_add : 1.0);

        // ---------------- Cholesky (right-looking, row ownership) -------------
        bool bad = false;
        double pmin = 1e300;
        T mydiag = (T)1;
#pragma unroll
        for (int k = 0; k < MR; ++k) {
            T* col = (k & 1) ? col1 : col0;
            col[r] = row[k];
            __syncthreads();
            const T dkk = col[k];
            bad = bad || !(dkk > (T)0);
            pmin = fmin(pmin, (double)dkk);
            const T Lkk = sqrt(dkk);
            const T inv = (T)1 / Lkk;
            if (r > k) {
                const T tr = row[k] * inv * inv;
#pragma unroll
                for (int j = k + 1; j < MR; ++j) row[j] = fma(-tr, col[j], row[j]);
                row[k] = row[k] * inv;
            } else if (r == k) {
                row[k] = Lkk;
                mydiag = Lkk;
            }
        }
        const T dinv = (T)1 / mydiag;
        const bool ill = REFINE && (bad || !(dmax / pmin <= kPivotRatioMax));
        if (ill) {
            if (pv && r == 0) {
                unsigned long long k = atomicAdd(ex.flag_count, 1ULL);
                ex.flag_list[k] = p;
            }
        }
        if (!REFINE && bad && pv && r == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.hdr->not_pd), 1ULL);
            if (a.pt)
                for (int q = 0; q < a.pt_stride; ++q) a.pt[p * a.pt_stride + q] = __builtin_nan("");
            if (a.mean)
                for (int o = 0; o < ny; ++o) a.mean[p * ny + o] = __builtin_nan("");
            if (a.stdv) a.stdv[p] = __builtin_nan("");
        }
        const bool live = pv && !bad && !ill;

        const bool need_back = REFINE || a.grad;
        if (need_back) {
#pragma unroll
            for (int j = 0; j < MR; ++j) Lg[r * (MR + 1) + j] = j <= r ? row[j] : (T)0;
            __syncthreads();
        }

        // FP64 residual rhs - K v of the block (v broadcast through Bg).
        auto residual = [&](double rhs, double v) -> double {
            Bg[r] = v;
            __syncthreads();
            double acc = rhs;
            for (int j = 0; j < MR; ++j) {
                const bool vj = (vbits >> j) & 1ULL;
                double K = (vr && vj) ? fma(kp.s, Eg[r * (MR + 1) + j], kp.C) : 0.0;
                if (j == r) K = vr ? K + diag_add : 1.0;
                acc = fma(-K, Bg[j], acc);
            }
            __syncthreads();
            return acc;
        };
        // x = K^-1 rhs (T factor; REFINE: one FP64-residual refinement step)
        auto solve = [&](double rhs) -> double {
            T z = fwd_solve<T, MR>((T)rhs, row, dinv, r);
            T x0 = bwd_solve<T, MR>(z, Lg, dinv, r);
            if constexpr (REFINE) {
                const double res = residual(rhs, (double)x0);
                T dz = fwd_solve<T, MR>((T)res, row, dinv, r);
                T dx = bwd_solve<T, MR>(dz, Lg, dinv, r);
                return (double)x0 + (double)dx;
            } else {
                return (double)x0;
            }
        };

        // ---------------- sigma^2 ------------------------------------------
        double bvec = 0.0, zk = 0.0, s2;
        if (REFINE || a.grad) {
            bvec = solve(kv);
            s2 = c - gsum<MR>(kv * bvec);
        } else {
            zk = (double)fwd_solve<T, MR>((T)kv, row, dinv, r);
            s2 = c - gsum<MR>(zk * zk);
        }
        const double sig = sqrt(s2);
        if (live && a.stdv && r == 0) a.stdv[p] = sig;

        // ---------------- outputs: mean / loglik / adjoint weights ---------
        double ll = 0.0, Gs = 0.0, at = 0.0;
        for (int o = 0; o < ny; ++o) {
            const double yv = vr ? a.y[(int64_t)jr * ny + o] : 0.0;
            double mu, av = 0.0;
            if (REFINE || a.grad) {
                av = solve(yv);
                mu = gsum<MR>(kv * av);
            } else {
                const double zy = (double)fwd_solve<T, MR>((T)yv, row, dinv, r);
                mu = gsum<MR>(zk * zy);
            }
            if (live && a.mean && r == 0) a.mean[p * ny + o] = mu;
            if (a.ys) {
                const double yo = pv ? a.ys[p * ny + o] : 0.0;
                const double res = yo - mu;
                ll += -0.5 * res * res / s2 - kHalfLog2Pi - log(sig);
                if (a.grad) {
                    Gs += -0.5 / s2 + 0.5 * res * res / (s2 * s2);
                    at = fma(res / s2, av, at);
                }
            }
        }
        if (live && a.pt && r == 0) a.pt[p * a.pt_stride] = ll;

        // ---------------- adjoint contraction -------------------------------
        if (a.grad && a.pt) {
            Bg[r] = bvec;
            Ag[r] = at;
            __syncthreads();
            double acc_w = 0.0, acc_C = 0.0, acc_s = 0.0, acc_l = 0.0;
            const double* xr = Xg + r * d;
            const double* xs = Xg + MR * d;
            // Matern with ARD: row r's stored E (read by this lane only from here
            // on) is replaced by the length-scale factor G for the per-feature pass.
            const bool g_pass = kp.n_ls > 1 && kp.fam != NNGP_COV_RBF;
            // G of the pair (row r, xo) from its E, and u into *uo
            auto pair_g = [&](const double* xo, double E, double* uo) -> double {
                double u = 0.0;
                if (kp.n_ls == 1) {
                    for (int f = 0; f < d; ++f) {
                        double t = xr[f] - xo[f];
                        u = fma(t, t, u);
                    }
                    u *= kp.h[0];
                } else {
                    for (int f = 0; f < d; ++f) {
                        double t = xr[f] - xo[f];
                        u = fma(t * t, kp.h[f], u);
                    }
                }
                *uo = u;
                return cov_g(kp.fam, u, E);
            };
            if (vr) {
                const double v = at - 2.0 * Gs * bvec;
                const double E0 = E0g[r];
                acc_C += v;
                acc_s += v * E0;
                if (kp.n_ls == 1 || g_pass) {
                    double u;
                    const double G0 = pair_g(xs, E0, &u);
                    if (kp.n_ls == 1) acc_l += v * G0 * 2.0 * u;
                    else E0g[r] = G0;
                }
                for (int l = 0; l < MR; ++l) {
                    if (!((vbits >> l) & 1ULL)) continue;
                    const double bl = Bg[l], al = Ag[l];
                    const double W = Gs * bvec * bl - 0.5 * (bvec * al + at * bl);
                    const double E = Eg[r * (MR + 1) + l];
                    if (l == r) acc_w += W;
                    acc_C += W;
                    acc_s += W * E;
                    if (kp.n_ls == 1 || g_pass) {
                        double u;
                        const double Gl = pair_g(Xg + l * d, E, &u);
                        if (kp.n_ls == 1) acc_l += W * Gl * 2.0 * u;
                        else Eg[r * (MR + 1) + l] = Gl;
                    }
                }
            }
            acc_w = gsum<MR>(acc_w);
            acc_C = gsum<MR>(acc_C);
            acc_s = gsum<MR>(acc_s);
            acc_l = gsum<MR>(acc_l);
            double* gp = a.pt + p * a.pt_stride + 1;
            if (live && r == 0) {
                for (int q = 0; q < a.n_theta; ++q) gp[q] = 0.0;
                if (kp.t_w >= 0) gp[kp.t_w] += kp.w * (acc_w + Gs);
                if (kp.t_C >= 0) gp[kp.t_C] += kp.C * (acc_C + Gs);
                if (kp.t_s >= 0) gp[kp.t_s] += kp.s * (acc_s + Gs);
                if (kp.t_l >= 0 && kp.n_ls == 1) gp[kp.t_l] += kp.s * acc_l;
            }
            if (kp.t_l >= 0 && kp.n_ls > 1) {
                for (int f = 0; f < d; ++f) {
                    double accf = 0.0;
                    if (vr) {
                        const double v = at - 2.0 * Gs * bvec;
                        double t0 = xr[f] - xs[f];
                        accf += v * E0g[r] * 2.0 * t0 * t0 * kp.h[f];
                        for (int l = 0; l < MR; ++l) {
                            if (!((vbits >> l) & 1ULL)) continue;
                            const double bl = Bg[l], al = Ag[l];
                            const double W = Gs * bvec * bl - 0.5 * (bvec * al + at * bl);
                            const double t = xr[f] - Xg[l * d + f];
                            accf += W * Eg[r * (MR + 1) + l] * 2.0 * t * t * kp.h[f];
                        }
                    }
                    accf = gsum<MR>(accf);
                    if (live && r == 0) gp[kp.t_l + f] += kp.s * accf;
                }
            }
        }
        __syncthreads();
    }
}

template <typename T, int MR>
size_t lds_bytes(int d) {
    constexpr int G = 64 / MR;
    size_t b = sizeof(double) * (size_t)G * MR * (MR + 1) + sizeof(double) * 3 * G * MR +
               sizeof(double) * (size_t)G * (MR + 1) * d + sizeof(T) * (size_t)G * MR * (MR + 1) +
               sizeof(T) * 2 * G * MR + sizeof(int) * G * MR;
    return (b + 15) & ~size_t(15);
}

template <typename T, int MR>
void launch_mr(const VecArgs& a, const RegExtra& ex, int64_t max_points, hipStream_t st) {
    constexpr int G = 64 / MR;
    int64_t blocks = (max_points + G - 1) / G;
    if (blocks < 1) blocks = 1;
    if (blocks > (1 << 20)) blocks = 1 << 20;
    const size_t lds = lds_bytes<T, MR>(a.d);
    hipLaunchKernelGGL((vecchia_reg_kernel<T, MR>), dim3((unsigned)blocks), dim3(64), lds, st, a, ex);
}

}  // namespace

int reg_mr(int m) { return m <= 16 ? 16 : (m <= 32 ? 32 : (m <= 64 ? 64 : 0)); }

size_t reg_lds_bytes(int m, int d, bool fp32) {
    switch (reg_mr(m)) {
        case 16: return fp32 ? lds_bytes<float, 16>(d) : lds_bytes<double, 16>(d);
        case 32: return fp32 ? lds_bytes<float, 32>(d) : lds_bytes<double, 32>(d);
        case 64: return fp32 ? lds_bytes<float, 64>(d) : lds_bytes<double, 64>(d);
        default: return (size_t)-1;
    }
}

int launch_reg(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points, hipStream_t st) {
    const int mr = reg_mr(a.m);
    if (fp32) {
        if (mr == 16) launch_mr<float, 16>(a, ex, max_points, st);
        else if (mr == 32) launch_mr<float, 32>(a, ex, max_points, st);
        else if (mr == 64) launch_mr<float, 64>(a, ex, max_points, st);
        else return -1;
    } else {
        if (mr == 16) launch_mr<double, 16>(a, ex, max_points, st);
        else if (mr == 32) launch_mr<double, 32>(a, ex, max_points, st);
        else if (mr == 64) launch_mr<double, 64>(a, ex, max_points, st);
        else return -1;
    }
    return 0;
}

}  // namespace nngp
