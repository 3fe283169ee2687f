// Internal interface of the fused Vecchia kernels.
#pragma once
#include "common.h"

namespace nngp {

// Kernel parameters in device-friendly form (from nngp_kernel).
struct KernParams {
    double w, C, s;
    int n_ls;                   // 0, 1 or d
    int t_w, t_C, t_s, t_l;     // theta indices or -1
    int n_theta;
    int fam;                    // nngp_covariance of the stationary part
    double h[NNGP_MAX_DIM];     // 1 / (2 l_f^2)
};

// a * b + c with c in a scalar register pair: a Horner step whose coefficient the
// compiler would otherwise hold in a vector register and copy before every v_fmac
// (the two-address form).
__device__ __forceinline__ double fma_s(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// exp(-u) for u >= 0 (0 from u = 746 on, also for u = +inf).  Cody-Waite
// reduction -u = n ln2 + r, |r| <= ln2 / 2, with ln2 split into a head whose low 32
// bits are zero (n * head exact for |n| < 2^21) and a tail; e^r from its Taylor
// series to degree 13 (truncation below 6e-18 relative on that interval); then the
// scale by 2^n.  Measured against a long-double reference over [0, 746]: at most
// 0.88 ulp (the library exp: 0.51), in 20 instructions against the library's ~32
// (its Horner steps each copy a coefficient register before the fused multiply-add;
// fma_s keeps the coefficients in scalar registers).
__device__ __forceinline__ double exp_neg(double u) {
    const double x = fmax(-u, -746.0);
    const double n = rint(x * 1.4426950408889634074);
    const double r =
        fma(-n, 1.90821492927058770002e-10, fma(-n, 6.93147180369123816490e-01, x));
    double p = 1.0 / 6227020800.0;      // 1/13!
    p = fma_s(p, r, 1.0 / 479001600.0);   // 1/12!
    p = fma_s(p, r, 1.0 / 39916800.0);    // 1/11!
    p = fma_s(p, r, 1.0 / 3628800.0);     // 1/10!
    p = fma_s(p, r, 1.0 / 362880.0);      // 1/9!
    p = fma_s(p, r, 1.0 / 40320.0);       // 1/8!
    p = fma_s(p, r, 1.0 / 5040.0);        // 1/7!
    p = fma_s(p, r, 1.0 / 720.0);         // 1/6!
    p = fma_s(p, r, 1.0 / 120.0);         // 1/5!
    p = fma_s(p, r, 1.0 / 24.0);          // 1/4!
    p = fma_s(p, r, 1.0 / 6.0);           // 1/3!
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// Stationary part as a function of u = sum_f h_f dx_f^2 = r^2 / 2 (r: distance in
// length-scale units): E(u), and the factor G with dE/dlog l_f = G * 2 h_f dx_f^2
// (isotropic: dE/dlog l = G * 2 u).
//   RBF         E = exp(-u)                                G = E
//   Matern 1/2  E = exp(-r)                                G = E / r   (0 at r = 0)
//   Matern 3/2  E = (1 + a) exp(-a),           a = sqrt(3) r   G = 3 exp(-a)
//   Matern 5/2  E = (1 + a + a^2/3) exp(-a),   a = sqrt(5) r   G = 5/3 (1 + a) exp(-a)
// (dE/dr = -E, -a^2 exp(-a) / r and -a^2 (1 + a) exp(-a) / (3 r) respectively, and
// dr/dlog l_f = -2 h_f dx_f^2 / r.)  The r = 0 value of Matern 1/2 follows the
// reference, which divides by 1 where the distance is 0 (batched_kernels.py:1101-1104).
__device__ __forceinline__ double cov_e(int fam, double u) {
    if (fam == NNGP_COV_RBF) return exp_neg(u);
    if (fam == NNGP_COV_MATERN_12) return exp_neg(sqrt(2.0 * u));
    if (fam == NNGP_COV_MATERN_32) {
        const double a = sqrt(6.0 * u);
        return (1.0 + a) * exp_neg(a);
    }
    const double a = sqrt(10.0 * u);
    return fma(a, fma(a, 1.0 / 3.0, 1.0), 1.0) * exp_neg(a);
}

// G from u and E = cov_e(fam, u) (E may carry the rounding of a stored entry).
__device__ __forceinline__ double cov_g(int fam, double u, double E) {
    if (fam == NNGP_COV_RBF) return E;
    if (fam == NNGP_COV_MATERN_12) {
        const double r = sqrt(2.0 * u);
        return r > 0.0 ? E / r : 0.0;
    }
    if (fam == NNGP_COV_MATERN_32) {
        const double a = sqrt(6.0 * u);
        return 3.0 * E / (1.0 + a);
    }
    const double a = sqrt(10.0 * u);
    return (5.0 / 3.0) * (1.0 + a) * E / fma(a, fma(a, 1.0 / 3.0, 1.0), 1.0);
}

// One fused pass over points: loglik (ys != NULL) or predict (ys == NULL).
struct VecArgs {
    const double* x;       // neighbour coordinates source (n_src x d)
    const double* y;       // neighbour targets (n_src x ny)
    const double* xs;      // own coordinates (npts x d)
    const double* ys;      // own targets (npts x ny) for the log-likelihood, else NULL
    const int32_t* nn;     // npts x m neighbour indices into x, -1 padded
    int64_t n_src;         // rows of x and y (indices outside [-1, n_src) are invalid)
    int64_t npts;
    int d, ny, m;
    double alpha;
    int alpha_self;        // add alpha to the point's own variance (log-likelihood)
    KernParams kp;
    int n_theta;
    int grad;
    double* pt;            // per point: [loglik, grad[n_theta]] with stride pt_stride
    int pt_stride;
    double* mean;          // npts x ny (predict)
    double* stdv;          // npts (predict)
    WsHeader* hdr;
    // Generic kernel only (conditioning weights of a sequential prediction):
    int alpha_cut;         // 1: alpha only on neighbours with index < alpha_limit
    int64_t alpha_limit;
    double* wout;          // npts x m: K^-1 k, or NULL
    double* s2out;         // npts: K_00 - k.K^-1 k (unclamped), or NULL
};

size_t generic_ws_doubles_per_point(int m);
int64_t generic_chunk(int64_t npts, int m);
void launch_generic(const VecArgs& a, double* wsL, double* wsv, int64_t P, hipStream_t st);
void launch_fill(const double* x, int64_t n_x, int d, const int32_t* nn, int64_t npts, int width,
                 const KernParams& kp, double* outK, double* outG, hipStream_t st);
size_t reduce_partial_doubles(int cols);
void launch_reduce(const double* v, int64_t rows, int cols, double* partial, double* out_ll,
                   double* out_grad, hipStream_t st);

}  // namespace nngp
