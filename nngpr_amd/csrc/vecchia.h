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
    double h[NNGP_MAX_DIM];     // 1 / (2 l_f^2)
};

// One fused pass over points: loglik (ys != NULL) or predict (ys == NULL).
struct VecArgs {
    const double* x;       // neighbour coordinates source (n_src x d)
    const double* y;       // neighbour targets (n_src x ny)
    const double* xs;      // own coordinates (npts x d)
    const double* ys;      // own targets (npts x ny) for the log-likelihood, else NULL
    const int32_t* nn;     // npts x m neighbour indices into x, -1 padded
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
};

size_t generic_ws_doubles_per_point(int m);
int64_t generic_chunk(int64_t npts, int m);
void launch_generic(const VecArgs& a, double* wsL, double* wsv, int64_t P, hipStream_t st);
void launch_fill(const double* x, int d, const int32_t* nn, int64_t npts, int width,
                 const KernParams& kp, double* outK, double* outG, hipStream_t st);
size_t reduce_partial_doubles(int cols);
void launch_reduce(const double* v, int64_t rows, int cols, double* partial, double* out_ll,
                   double* out_grad, hipStream_t st);

}  // namespace nngp
