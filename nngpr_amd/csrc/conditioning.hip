// Sequential conditioning of a prediction: conditioning_method 'full' /
// 'full-double-nn', return_cov and sample_y (base_nngpr.py:616-761, 896-1030).
//
// Query i is conditioned on training rows and on earlier queries q < i (its
// neighbour list indexes [x_train; x_query]).  Everything that depends only on the
// coordinates is computed for all queries at once: with K the covariance of the
// neighbours (alpha on training neighbours only) and k their covariance with the
// query,
//   w_i = K^-1 k,  s2_i = K_00 - k.w_i,  mt_i = sum over training neighbours w y
// (the generic fused kernel, which writes w and s2 next to the mean).  What remains
// is linear in earlier queries' results and runs level by level over the
// dependency DAG (level of i = 1 + the largest level of its query neighbours,
// built on the host from the neighbour lists), one launch per level:
//   mean_i    = mt_i + sum_p w_p mean_{q_p}
//   std       var_i = max(0, s2_i) + w.C w, C over the query neighbours with
//             C_pp = var_{q_p} and C_pr = the stored covariance of the later of
//             q_p, q_r with the earlier one when that one is among its neighbours,
//             else 0; the covariances of i with its query neighbours, (C w)_p, are
//             stored for later queries (the reference's sparse partial covariance);
//   cov       f_i = sum_p w_p f_{q_p} + sqrt(max(0, s2_i)) e_i, i.e. rows of
//             M = (I - B)^-1 Lambda^1/2 (B: the query-neighbour weights), and the
//             joint covariance M M^T, which is the reference's dense recursion
//             partial_cov[i, :i] = w_Q partial_cov[Q, :i] in closed form;
//   samples   y_i = mt_i + sum_p w_p y_{q_p} + sigma_i z_i with sigma_i =
//             sqrt(s2_i) (0 where s2_i <= 0) and caller-provided normals z.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "conditioning.h"
#include "vecchia.h"
#include "vecchia_reg.h"

namespace nngp {

namespace {

struct CondLayout {
    WsHeader* hdr;
    double* xf;      // (nt + nq) x d
    double* yf;      // (nt + nq) x ny, query rows 0
    double* w;       // nq x m
    double* s2;      // nq
    double* mt;      // nq x ny
    int32_t* order;  // nq, queries by level
    double* wsL;     // generic-kernel scratch
    double* wsv;
    int64_t P;
    double* var;     // nq (std, cov)
    double* pc;      // nq x m (std)
    double* M;       // nq x nq (cov)
};

CondLayout carve_cond(Carver& c, int64_t nt, int64_t nq, int m, int d, int ny, int output) {
    CondLayout L{};
    L.hdr = c.take<WsHeader>(1);
    L.xf = c.take<double>((size_t)(nt + nq) * d);
    L.yf = c.take<double>((size_t)(nt + nq) * ny);
    L.w = c.take<double>((size_t)nq * m);
    L.s2 = c.take<double>((size_t)nq);
    L.mt = c.take<double>((size_t)nq * ny);
    L.order = c.take<int32_t>((size_t)nq);
    L.P = generic_chunk(nq, m);
    L.wsL = c.take<double>((size_t)L.P * ((size_t)m * (m + 1) / 2));
    L.wsv = c.take<double>((size_t)L.P * 5 * (size_t)m);
    if (output == COND_STD || output == COND_COV) L.var = c.take<double>((size_t)nq);
    if (output == COND_STD) L.pc = c.take<double>((size_t)nq * m);
    if (output == COND_COV) L.M = c.take<double>((size_t)nq * nq);
    return L;
}

// out = [a (na rows); b (nb rows, or zeros when b is NULL)], `cols` columns.
__global__ __launch_bounds__(256) void concat_rows_kernel(const double* a, int64_t na,
                                                          const double* b, int64_t nb, int cols,
                                                          double* out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (na + nb) * cols;
    if (e >= tot) return;
    const int64_t r = e / cols;
    out[e] = r < na ? a[e] : (b ? b[e - na * cols] : 0.0);
}

__global__ __launch_bounds__(256) void zero_kernel(double* p, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

struct LevelArgs {
    const int32_t* nn;
    int m;
    int64_t nt, nq;
    int ny;
    const double* w;
    const double* s2;
    const double* mt;
    const int32_t* order;
    double* mean;
    double* var;      // COND_STD: variances (else NULL)
    double* pc;
    double* M;
    double* var_cov;  // COND_COV: variances |M_i|^2
    int n_samples;
    const double* z;
    double* samples;
};

// Mean (and, with a.var, the sparse-covariance variance) of one level's queries:
// one wave per query.
__global__ __launch_bounds__(64) void cond_mean_std_kernel(LevelArgs a, int64_t lo) {
    const int64_t i = a.order[lo + blockIdx.x];
    const int lane = threadIdx.x;
    const int m = a.m;
    const int32_t* row = a.nn + i * m;
    const double* wi = a.w + i * m;
    const int64_t nt = a.nt;
    for (int o = lane; o < a.ny; o += 64) {
        double acc = a.mt[i * a.ny + o];
        for (int p = 0; p < m; ++p) {
            const int64_t j = row[p];
            if (j >= nt) acc = fma(wi[p], a.mean[(j - nt) * a.ny + o], acc);
        }
        a.mean[i * a.ny + o] = acc;
    }
    if (!a.var) return;
    double part = 0.0;
    for (int p = lane; p < m; p += 64) {
        const int64_t jp = row[p];
        if (jp < nt) {
            a.pc[i * m + p] = 0.0;
            continue;
        }
        const int64_t qp = jp - nt;
        double v = 0.0;
        for (int r = 0; r < m; ++r) {
            const int64_t jr = row[r];
            if (jr < nt) continue;
            const int64_t qr = jr - nt;
            double cpr = 0.0;
            if (r == p) {
                cpr = a.var[qp];
            } else {
                const int64_t hi = qp > qr ? qp : qr;
                const int64_t lo_q = qp > qr ? qr : qp;
                const int32_t* rh = a.nn + hi * m;
                for (int s = 0; s < m; ++s)
                    if ((int64_t)rh[s] == nt + lo_q) {
                        cpr = a.pc[hi * m + s];
                        break;
                    }
            }
            v = fma(cpr, wi[r], v);
        }
        a.pc[i * m + p] = v;
        part = fma(wi[p], v, part);
    }
    part = group_sum<64>(part);
    if (lane == 0) a.var[i] = fmax(0.0, fmax(0.0, a.s2[i]) + part);
}

// Row i of M (columns 0..i; the rest stays 0) and var_i = |M_i|^2; 256 threads
// per query.
__global__ __launch_bounds__(256) void cond_cov_row_kernel(LevelArgs a, int64_t lo) {
    __shared__ int64_t qs[NNGP_MAX_NN];
    __shared__ double ws[NNGP_MAX_NN];
    __shared__ double red[4];
    __shared__ int nq_s;
    const int64_t i = a.order[lo + blockIdx.x];
    const int tid = threadIdx.x;
    const int m = a.m;
    if (tid == 0) {
        int c = 0;
        for (int p = 0; p < m; ++p) {
            const int64_t j = a.nn[i * m + p];
            if (j >= a.nt) {
                qs[c] = j - a.nt;
                ws[c] = a.w[i * m + p];
                ++c;
            }
        }
        nq_s = c;
    }
    __syncthreads();
    const int cq = nq_s;
    const double sq = a.s2[i] > 0.0 ? sqrt(a.s2[i]) : 0.0;
    double ss = 0.0;
    double* Mi = a.M + i * a.nq;
    for (int64_t k = tid; k <= i; k += 256) {
        double acc = k == i ? sq : 0.0;
        for (int p = 0; p < cq; ++p) acc = fma(ws[p], a.M[qs[p] * a.nq + k], acc);
        Mi[k] = acc;
        ss = fma(acc, acc, ss);
    }
    ss = group_sum<64>(ss);
    if ((tid & 63) == 0) red[tid >> 6] = ss;
    __syncthreads();
    if (tid == 0) a.var_cov[i] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Samples of one level's queries; 256 threads per query over (output, sample).
__global__ __launch_bounds__(256) void cond_sample_kernel(LevelArgs a, int64_t lo) {
    const int64_t i = a.order[lo + blockIdx.x];
    const int m = a.m;
    const int32_t* row = a.nn + i * m;
    const double* wi = a.w + i * m;
    const double s2 = a.s2[i];
    const double sig = s2 > 0.0 ? sqrt(s2) : 0.0;
    const int ns = a.n_samples;
    const int64_t per = (int64_t)a.ny * ns;
    for (int64_t e = threadIdx.x; e < per; e += 256) {
        const int o = (int)(e / ns);
        double acc = a.mt[i * a.ny + o];
        for (int p = 0; p < m; ++p) {
            const int64_t j = row[p];
            if (j >= a.nt) acc = fma(wi[p], a.samples[(j - a.nt) * per + e], acc);
        }
        a.samples[i * per + e] = fma(sig, a.z[i * per + e], acc);
    }
}

// C = M M^T for lower-triangular M (n x n): one 64 x 64 tile of the lower
// triangle per block (written to both halves), 4 x 4 outputs per thread, 16-wide
// k steps through LDS, k limited to the tile's column range.
__global__ __launch_bounds__(256) void syrk_lower_kernel(const double* M, int64_t n, double* C) {
    __shared__ double As[16][65];
    __shared__ double Bs[16][65];
    const int64_t t = blockIdx.x;
    int64_t bi = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int64_t bj = t - bi * (bi + 1) / 2;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int64_t r0 = bi * 64, c0 = bj * 64;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    const int64_t kend = (c0 + 64 < n) ? c0 + 64 : n;
    for (int64_t k0 = 0; k0 < kend; k0 += 16) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int e = tid + l * 256;  // 0..1023: row e / 16, k e % 16
            const int rr = e >> 4, kk = e & 15;
            const int64_t gk = k0 + kk;
            const int64_t ga = r0 + rr, gb = c0 + rr;
            As[kk][rr] = (ga < n && gk < n) ? M[ga * n + gk] : 0.0;
            Bs[kk][rr] = (gb < n && gk < n) ? M[gb * n + gk] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[kk][ty * 4 + u];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Bs[kk][tx * 4 + v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t r = r0 + ty * 4 + u;
        if (r >= n) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t c = c0 + tx * 4 + v;
            if (c >= n || c > r) continue;
            C[r * n + c] = acc[u][v];
            C[c * n + r] = acc[u][v];
        }
    }
}

__global__ __launch_bounds__(256) void sqrt_kernel(const double* v, int64_t n, double* out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = sqrt(v[e]);
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

size_t cond_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d, int ny, int output) {
    Carver c(nullptr, 0);
    carve_cond(c, n_train, n_query, m, d, ny, output);
    return c.used + 256;
}

int cond_run(const CondProblem& P, void* ws, size_t ws_bytes, hipStream_t st) {
    const int64_t nt = P.n_train, nq = P.n_query;
    const int m = P.m, d = P.d, ny = P.ny;
    if (nq == 0) return NNGP_OK;
    Carver c(ws, ws_bytes);
    CondLayout L = carve_cond(c, nt, nq, m, d, ny, P.output);
    if (!c.ok() || !ws) return NNGP_ERR_WORKSPACE_TOO_SMALL;

    // ---- dependency levels (host), from one read of the neighbour lists; the
    // lists are validated before any kernel reads them
    std::vector<int32_t> h_nn((size_t)nq * m);
    NNGP_HIP_CHECK(hipMemcpyAsync(h_nn.data(), P.nn, h_nn.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, st));
    NNGP_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int32_t> level((size_t)nq, 0);
    int32_t n_levels = 0;
    for (int64_t i = 0; i < nq; ++i) {
        int32_t lv = 0;
        for (int p = 0; p < m; ++p) {
            const int64_t j = h_nn[(size_t)i * m + p];
            if (j < 0) continue;
            if (j >= nt + nq) return NNGP_ERR_INVALID_ARGUMENT;
            if (j >= nt) {
                const int64_t q = j - nt;
                if (q >= i) return NNGP_ERR_INVALID_ARGUMENT;  // only earlier queries
                lv = std::max(lv, level[(size_t)q] + 1);
            }
        }
        level[(size_t)i] = lv;
        n_levels = std::max(n_levels, lv + 1);
    }
    std::vector<int64_t> start((size_t)n_levels + 1, 0);
    for (int64_t i = 0; i < nq; ++i) ++start[(size_t)level[(size_t)i] + 1];
    for (int32_t l = 0; l < n_levels; ++l) start[(size_t)l + 1] += start[(size_t)l];
    std::vector<int32_t> order((size_t)nq);
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < nq; ++i) order[(size_t)fill[(size_t)level[(size_t)i]]++] = (int32_t)i;
    }
    NNGP_HIP_CHECK(hipMemcpyAsync(L.order, order.data(), order.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice, st));
    NNGP_HIP_CHECK(hipStreamSynchronize(st));  // `order` is host memory of this frame

    hipLaunchKernelGGL(zero_header_kernel, dim3(1), dim3(64), 0, st, L.hdr);
    hipLaunchKernelGGL(concat_rows_kernel, dim3(blocks_for((nt + nq) * d, 256)), dim3(256), 0, st,
                       P.x_train, nt, P.x_query, nq, d, L.xf);
    hipLaunchKernelGGL(concat_rows_kernel, dim3(blocks_for((nt + nq) * ny, 256)), dim3(256), 0, st,
                       P.y_train, nt, (const double*)nullptr, nq, ny, L.yf);

    // ---- per-query weights, conditional variances and training part of the mean
    VecArgs a{};
    a.x = L.xf;
    a.y = L.yf;
    a.xs = P.x_query;
    a.ys = nullptr;
    a.nn = P.nn;
    a.n_src = nt + nq;
    a.npts = nq;
    a.d = d;
    a.ny = ny;
    a.m = m;
    a.alpha = P.alpha;
    a.alpha_self = 0;
    a.kp = P.kp;
    a.n_theta = 0;
    a.grad = 0;
    a.mean = L.mt;
    a.hdr = L.hdr;
    a.alpha_cut = 1;
    a.alpha_limit = nt;
    a.wout = L.w;
    a.s2out = L.s2;
    // the segment kernel where it applies (m <= 64, d <= 32, LDS within the block
    // limit), else the generic kernel
    if (launch_seg_weights(a, st) != 0) launch_generic(a, L.wsL, L.wsv, L.P, st);
    NNGP_HIP_CHECK(hipGetLastError());

    // ---- level by level
    LevelArgs la{};
    la.nn = P.nn;
    la.m = m;
    la.nt = nt;
    la.nq = nq;
    la.ny = ny;
    la.w = L.w;
    la.s2 = L.s2;
    la.mt = L.mt;
    la.order = L.order;
    la.mean = P.out_mean;
    la.var = P.output == COND_STD ? L.var : nullptr;
    la.pc = L.pc;
    la.M = L.M;
    la.var_cov = P.output == COND_COV ? L.var : nullptr;
    la.n_samples = P.n_samples;
    la.z = P.normals;
    la.samples = P.out_samples;
    if (P.output == COND_COV)
        hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, st, L.M, nq * nq);
    for (int32_t l = 0; l < n_levels; ++l) {
        const int64_t lo = start[(size_t)l];
        const unsigned cnt = (unsigned)(start[(size_t)l + 1] - lo);
        if (P.output == COND_SAMPLE) {
            hipLaunchKernelGGL(cond_sample_kernel, dim3(cnt), dim3(256), 0, st, la, lo);
            continue;
        }
        hipLaunchKernelGGL(cond_mean_std_kernel, dim3(cnt), dim3(64), 0, st, la, lo);
        if (P.output == COND_COV)
            hipLaunchKernelGGL(cond_cov_row_kernel, dim3(cnt), dim3(256), 0, st, la, lo);
    }
    NNGP_HIP_CHECK(hipGetLastError());
    if (P.output == COND_COV) {
        const int64_t nb = (nq + 63) / 64;
        hipLaunchKernelGGL(syrk_lower_kernel, dim3((unsigned)(nb * (nb + 1) / 2)), dim3(256), 0, st,
                           L.M, nq, P.out_cov);
    }
    if ((P.output == COND_STD || P.output == COND_COV) && P.out_std)
        hipLaunchKernelGGL(sqrt_kernel, dim3(blocks_for(nq, 256)), dim3(256), 0, st, L.var, nq,
                           P.out_std);
    NNGP_HIP_CHECK(hipGetLastError());

    // ---- a factor that failed is reported (its outputs are not meaningful)
    WsHeader h_hdr;
    NNGP_HIP_CHECK(hipMemcpyAsync(&h_hdr, L.hdr, sizeof(WsHeader), hipMemcpyDeviceToHost, st));
    NNGP_HIP_CHECK(hipStreamSynchronize(st));
    if (h_hdr.not_pd > 0) return NNGP_ERR_NOT_POSITIVE_DEFINITE;
    return NNGP_OK;
}

}  // namespace nngp
