// Generic fused Vecchia point kernel (any m up to NNGP_MAX_NN, FP64), the
// materialising fill kernel of the plug-in surface, and the fixed-order
// reductions shared by all Vecchia kernels.
//
// Generic kernel layout: one thread per point; its packed lower-triangular
// factor and work vectors live in a caller workspace interleaved across threads
// (element e of thread t at ws[e * P + t]) so that neighbouring threads touch
// neighbouring addresses.  It is the path for m > 64; the register kernel in
// vecchia_reg.hip covers m <= 64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "vecchia.h"

namespace nngp {

namespace {

// k(a, b) without the white term; E (the stationary part) and u = sum_f h_f dx_f^2
// are returned for gradient use.
__device__ __forceinline__ double cov_entry(const KernParams& kp, const double* xa,
                                            const double* xb, int d, double* E_out,
                                            double* u_out) {
    if (kp.n_ls == 0) {
        *E_out = 0.0;
        *u_out = 0.0;
        return kp.C;
    }
    double u = 0.0;
    if (kp.n_ls == 1) {
        for (int f = 0; f < d; ++f) {
            double t = xa[f] - xb[f];
            u = fma(t, t, u);
        }
        u *= kp.h[0];
    } else {
        for (int f = 0; f < d; ++f) {
            double t = xa[f] - xb[f];
            u = fma(t * t, kp.h[f], u);
        }
    }
    double E = cov_e(kp.fam, u);
    *E_out = E;
    *u_out = u;
    return fma(kp.s, E, kp.C);
}

__global__ __launch_bounds__(256) void vecchia_generic_kernel(VecArgs a, int64_t p0, int64_t P,
                                                              double* wsL, double* wsv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p = p0 + t;
    if (t >= P || p >= a.npts) return;
    const int m = a.m, d = a.d, ny = a.ny;
    const KernParams& kp = a.kp;
    const int32_t* nnp = a.nn + p * m;
    const double* xsp = a.xs + p * d;
#define LP(i, j) wsL[((int64_t)(i) * ((i) + 1) / 2 + (j)) * P + t]
#define KV(i) wsv[((int64_t)(i)) * P + t]
#define BV(i) wsv[((int64_t)(m + (i))) * P + t]
#define TV(i) wsv[((int64_t)(2 * m + (i))) * P + t]
#define AT(i) wsv[((int64_t)(3 * m + (i))) * P + t]
    // neighbour slot i: a source row or -1 (invalid entries read as empty; the
    // fill's first read of each slot counts them in the header)
    auto slot = [&](int i) { return nbr_slot(nnp[i], a.n_src, nullptr); };
    // ---- fill ----
    double Ed, ud;
    for (int i = 0; i < m; ++i) {
        const int ji = nbr_slot(nnp[i], a.n_src, a.hdr);
        for (int j = 0; j <= i; ++j) {
            const int jj = slot(j);
            double v;
            if (ji < 0 || jj < 0) {
                v = (i == j) ? 1.0 : 0.0;
            } else {
                v = cov_entry(kp, a.x + (int64_t)ji * d, a.x + (int64_t)jj * d, d, &Ed, &ud);
                if (i == j) v += kp.w + ((a.alpha_cut && ji >= a.alpha_limit) ? 0.0 : a.alpha);
            }
            LP(i, j) = v;
        }
        KV(i) = (ji < 0) ? 0.0 : cov_entry(kp, xsp, a.x + (int64_t)ji * d, d, &Ed, &ud);
    }
    const double c = kp.s + kp.C + kp.w + (a.alpha_self ? a.alpha : 0.0);
    // ---- Cholesky (row by row) ----
    bool pd = true;
    for (int i = 0; i < m && pd; ++i) {
        for (int j = 0; j <= i; ++j) {
            double sum = LP(i, j);
            for (int k = 0; k < j; ++k) sum -= LP(i, k) * LP(j, k);
            if (i == j) {
                if (!(sum > 0.0)) {
                    pd = false;
                    break;
                }
                LP(i, i) = sqrt(sum);
            } else {
                LP(i, j) = sum / LP(j, j);
            }
        }
    }
    if (!pd) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.hdr->not_pd), 1ULL);
        if (a.pt)
            for (int q = 0; q < a.pt_stride; ++q) a.pt[p * a.pt_stride + q] = __builtin_nan("");
        if (a.mean)
            for (int o = 0; o < ny; ++o) a.mean[p * ny + o] = __builtin_nan("");
        if (a.stdv) a.stdv[p] = __builtin_nan("");
        return;
    }
    // ---- z_k = L^-1 k (in KV), sigma^2 ----
    double zz = 0.0;
    for (int i = 0; i < m; ++i) {
        double sum = KV(i);
        for (int k = 0; k < i; ++k) sum -= LP(i, k) * KV(k);
        double z = sum / LP(i, i);
        KV(i) = z;
        zz += z * z;
    }
    const double s2 = c - zz;
    const double sig = sqrt(s2);
    if (a.stdv) a.stdv[p] = sig;
    if (a.s2out) a.s2out[p] = s2;
    // b = L^-T z_k (gradient and conditioning weights)
    if (a.grad || a.wout) {
        for (int i = m - 1; i >= 0; --i) {
            double sum = KV(i);
            for (int k = i + 1; k < m; ++k) sum -= LP(k, i) * BV(k);
            BV(i) = sum / LP(i, i);
        }
        for (int i = 0; i < m; ++i) AT(i) = 0.0;
        if (a.wout)
            for (int i = 0; i < m; ++i) a.wout[p * m + i] = BV(i);
    }
    double ll = 0.0, Gs = 0.0;
    const double half_log_2pi = 0.91893853320467274178;
    for (int o = 0; o < ny; ++o) {
        for (int i = 0; i < m; ++i) {
            const int ji = slot(i);
            TV(i) = (ji < 0) ? 0.0 : a.y[(int64_t)ji * ny + o];
        }
        double mu = 0.0;
        for (int i = 0; i < m; ++i) {
            double sum = TV(i);
            for (int k = 0; k < i; ++k) sum -= LP(i, k) * TV(k);
            double z = sum / LP(i, i);
            TV(i) = z;
            mu += KV(i) * z;
        }
        if (a.mean) a.mean[p * ny + o] = mu;
        if (a.ys) {
            const double r = a.ys[p * ny + o] - mu;
            ll += -0.5 * r * r / s2 - half_log_2pi - log(sig);
            if (a.grad) {
                const double gmu = r / s2;
                Gs += -0.5 / s2 + 0.5 * r * r / (s2 * s2);
                for (int i = m - 1; i >= 0; --i) {
                    double sum = TV(i);
                    for (int k = i + 1; k < m; ++k) sum -= LP(k, i) * TV(k);
                    TV(i) = sum / LP(i, i);
                }
                for (int i = 0; i < m; ++i) AT(i) += gmu * TV(i);
            }
        }
    }
    if (a.pt) a.pt[p * a.pt_stride] = ll;
    if (!a.grad || !a.pt) return;
    // ---- adjoint contraction ----
    double acc_w = Gs, acc_C = Gs, acc_s = Gs, acc_l = 0.0;
    double acc_lf[NNGP_MAX_DIM];
    const bool ard = kp.n_ls > 1;
    if (ard)
        for (int f = 0; f < d; ++f) acc_lf[f] = 0.0;
    for (int j = 0; j < m; ++j) {
        const int jj = slot(j);
        if (jj < 0) continue;
        const double* xj = a.x + (int64_t)jj * d;
        const double bj = BV(j), aj = AT(j);
        // self - neighbour entry
        const double v = aj - 2.0 * Gs * bj;
        double E, u;
        cov_entry(kp, xsp, xj, d, &E, &u);
        acc_C += v;
        acc_s += v * E;
        double Gd = cov_g(kp.fam, u, E);
        if (kp.n_ls == 1) acc_l += v * Gd * 2.0 * u;
        if (ard)
            for (int f = 0; f < d; ++f) {
                double tt = xsp[f] - xj[f];
                acc_lf[f] += v * Gd * 2.0 * tt * tt * kp.h[f];
            }
        // neighbour - neighbour entries (l <= j; off-diagonal counted twice)
        for (int l = 0; l <= j; ++l) {
            const int jl = slot(l);
            if (jl < 0) continue;
            const double bl = BV(l), al = AT(l);
            double W = Gs * bj * bl - 0.5 * (bj * al + aj * bl);
            if (l != j) W *= 2.0;
            const double* xl = a.x + (int64_t)jl * d;
            cov_entry(kp, xj, xl, d, &E, &u);
            if (l == j) acc_w += W;
            acc_C += W;
            acc_s += W * E;
            Gd = cov_g(kp.fam, u, E);
            if (kp.n_ls == 1) acc_l += W * Gd * 2.0 * u;
            if (ard)
                for (int f = 0; f < d; ++f) {
                    double tt = xj[f] - xl[f];
                    acc_lf[f] += W * Gd * 2.0 * tt * tt * kp.h[f];
                }
        }
    }
    double* g = a.pt + p * a.pt_stride + 1;
    for (int q = 0; q < a.n_theta; ++q) g[q] = 0.0;
    if (kp.t_w >= 0) g[kp.t_w] += kp.w * acc_w;
    if (kp.t_C >= 0) g[kp.t_C] += kp.C * acc_C;
    if (kp.t_s >= 0) g[kp.t_s] += kp.s * acc_s;
    if (kp.t_l >= 0) {
        if (ard)
            for (int f = 0; f < d; ++f) g[kp.t_l + f] += kp.s * acc_lf[f];
        else
            g[kp.t_l] += kp.s * acc_l;
    }
#undef LP
#undef KV
#undef BV
#undef TV
#undef AT
}

// Materialised neighbour blocks: one thread per (point, row i).  A slot outside
// [-1, n_x) is never read through: its row and column are NaN (K and gradient).
__global__ __launch_bounds__(256) void fill_kernel(const double* x, int64_t n_x, int d,
                                                   const int32_t* nn, int64_t npts, int width,
                                                   KernParams kp, double* outK, double* outG) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= npts * width) return;
    const int64_t p = gid / width;
    const int i = (int)(gid % width);
    const int ri = nn[p * width + i];
    const int ji = nbr_slot(ri, n_x, nullptr);
    const int nt = kp.n_theta;
    for (int j = 0; j < width; ++j) {
        const int rj = nn[p * width + j];
        const int jj = nbr_slot(rj, n_x, nullptr);
        const int64_t e = (p * width + i) * width + j;
        double* g = outG ? outG + e * nt : nullptr;
        const bool bad = (ji < 0 && ri != -1) || (jj < 0 && rj != -1);
        if (g)
            for (int q = 0; q < nt; ++q) g[q] = bad ? __builtin_nan("") : 0.0;
        if (bad) {
            outK[e] = __builtin_nan("");
            continue;
        }
        if (ji < 0 || jj < 0) {
            outK[e] = (i == j) ? 1.0 : 0.0;
            continue;
        }
        double E, u;
        double v = cov_entry(kp, x + (int64_t)ji * d, x + (int64_t)jj * d, d, &E, &u);
        if (i == j) v += kp.w;
        outK[e] = v;
        if (!g) continue;
        if (kp.t_w >= 0 && i == j) g[kp.t_w] += kp.w;
        if (kp.t_C >= 0) g[kp.t_C] += kp.C;
        if (kp.t_s >= 0) g[kp.t_s] += kp.s * E;
        const double Gd = kp.n_ls > 0 ? cov_g(kp.fam, u, E) : 0.0;
        if (kp.t_l >= 0 && kp.n_ls == 1) g[kp.t_l] += kp.s * Gd * 2.0 * u;
        if (kp.t_l >= 0 && kp.n_ls > 1) {
            const double* xa = x + (int64_t)ji * d;
            const double* xb = x + (int64_t)jj * d;
            for (int f = 0; f < d; ++f) {
                double tt = xa[f] - xb[f];
                g[kp.t_l + f] += kp.s * Gd * 2.0 * tt * tt * kp.h[f];
            }
        }
    }
}

// Fixed-order two-level sum of `cols` interleaved columns over `rows` rows:
// stage 1 writes one partial per (block, column); stage 2 sums the partials.
__global__ __launch_bounds__(256) void reduce_stage1(const double* v, int64_t rows, int cols,
                                                     double* partial) {
    __shared__ double sh[256];
    const int64_t per = (rows + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per;
    const int64_t r1 = r0 + per < rows ? r0 + per : rows;
    for (int c = 0; c < cols; ++c) {
        double s = 0.0;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) s += v[r * cols + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * cols + c] = sh[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void reduce_stage2(const double* partial, int nblk, int cols,
                                                     double* out_ll, double* out_grad) {
    __shared__ double sh[256];
    for (int c = 0; c < cols; ++c) {
        double* out = c == 0 ? out_ll : (out_grad ? out_grad + (c - 1) : nullptr);
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(int64_t)b * cols + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0 && out) *out = sh[0];
        __syncthreads();
    }
}

}  // namespace

constexpr int kReduceBlocks = 256;

size_t generic_ws_doubles_per_point(int m) { return (size_t)m * (m + 1) / 2 + 5 * (size_t)m; }

int64_t generic_chunk(int64_t npts, int m) {
    const size_t per = generic_ws_doubles_per_point(m) * sizeof(double);
    int64_t P = (int64_t)((size_t)256 << 20) / (int64_t)per;
    P = P < 1024 ? 1024 : P;
    P = P > 65536 ? 65536 : P;
    return P < npts ? P : (npts > 0 ? npts : 1);
}

void launch_generic(const VecArgs& a, double* wsL, double* wsv, int64_t P, hipStream_t st) {
    for (int64_t p0 = 0; p0 < a.npts; p0 += P) {
        const int64_t cnt = (a.npts - p0) < P ? (a.npts - p0) : P;
        hipLaunchKernelGGL(vecchia_generic_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256),
                           0, st, a, p0, P, wsL, wsv);
    }
}

void launch_fill(const double* x, int64_t n_x, int d, const int32_t* nn, int64_t npts, int width,
                 const KernParams& kp, double* outK, double* outG, hipStream_t st) {
    const int64_t tot = npts * width;
    if (tot == 0) return;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, n_x, d,
                       nn, npts, width, kp, outK, outG);
}

size_t reduce_partial_doubles(int cols) { return (size_t)kReduceBlocks * cols; }

void launch_reduce(const double* v, int64_t rows, int cols, double* partial, double* out_ll,
                   double* out_grad, hipStream_t st) {
    hipLaunchKernelGGL(reduce_stage1, dim3(kReduceBlocks), dim3(256), 0, st, v, rows, cols, partial);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, st, partial, kReduceBlocks, cols,
                       out_ll, out_grad);
}

}  // namespace nngp
