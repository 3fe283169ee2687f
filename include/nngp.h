/*
 * nngp.h -- C ABI of the MI355X (gfx950) NNGP / Vecchia engine.
 *
 * Plain C types only: device pointers are raw pointers to HIP device memory
 * owned by the caller, sizes are int64_t/size_t, the stream is a hipStream_t
 * passed as void*.  Every call is asynchronous on that stream unless stated.
 * No function allocates or frees caller memory; scratch space comes from the
 * caller through (workspace, workspace_bytes), sized by the *_workspace_bytes
 * query of the same call.  Functions return an nngp_status (0 = success).
 *
 * These entry points stand in for the reference's accelerator plug-in surface,
 * nngpr/cuda_kernel_accelerators.py (CudaKernelAccelerator / RbfAccelerator) and
 * the vendor factor/solve calls of its torch/cupy workers; each declaration cites
 * the reference interface it replaces.
 */
#ifndef NNGP_H
#define NNGP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNGP_ABI_VERSION 2
#define NNGP_MAX_DIM 64   /* maximum number of input features */
#define NNGP_MAX_NN 1024  /* maximum neighbours per point (fill, fused and sequential entry points) */
#define NNGP_MAX_K 256    /* maximum k of a kNN search (nngp_knn, nngp_model_create) */

typedef enum {
    NNGP_OK = 0,
    NNGP_ERR_INVALID_ARGUMENT = -1,
    NNGP_ERR_UNSUPPORTED = -2,           /* caller maps to NotImplementedError */
    NNGP_ERR_NOT_POSITIVE_DEFINITE = -3, /* caller maps to LinAlgError */
    NNGP_ERR_WORKSPACE_TOO_SMALL = -4,
    NNGP_ERR_HIP = -5,
    NNGP_ERR_KNN_OVERFLOW = -6
} nngp_status;

typedef enum {
    NNGP_PRECISION_FP64 = 0,         /* plain FP64 form */
    NNGP_PRECISION_FP32_REFINED = 1  /* FP32 screening/factor + FP64 refinement */
} nngp_precision;

/*
 * Stationary part of the covariance as a function of the distance in length-scale
 * units r = sqrt(sum_f (x_if - x_jf)^2 / l_f^2):
 *   RBF         exp(-r^2 / 2)
 *   Matern 1/2  exp(-r)
 *   Matern 3/2  (1 + a) exp(-a),             a = sqrt(3) r
 *   Matern 5/2  (1 + a + a^2 / 3) exp(-a),   a = sqrt(5) r
 * (batched_kernels.py:844-920 RBF, :1003-1131 Matern; Matern with nu = inf is the
 * RBF).
 */
typedef enum {
    NNGP_COV_RBF = 0,
    NNGP_COV_MATERN_12 = 1,
    NNGP_COV_MATERN_32 = 2,
    NNGP_COV_MATERN_52 = 3
} nngp_covariance;

/*
 * Decomposed covariance  k(x_i, x_j) = s * E(r_ij) + C + w * [i == j (by position)]
 * with E the stationary part above, i.e. WhiteKernel(w) + ConstantKernel(C) +
 * ConstantKernel(s) * RBF(l) or Matern(l, nu), the form the reference reduces a
 * kernel tree to (cuda_kernel_accelerators.py:520-590).  theta_* give the index of
 * each parameter's log in kernel.theta, or -1 when the parameter is fixed or
 * absent (cuda_kernel_accelerators.py:592-649); ARD length scales occupy
 * theta_length_scale .. theta_length_scale + n_length_scale - 1.
 */
typedef struct {
    double noise_level;     /* w */
    double constant_value;  /* C */
    double rbf_level;       /* s */
    int32_t n_length_scale; /* 0: no RBF part, 1: isotropic, d: ARD */
    int32_t theta_noise_level;
    int32_t theta_constant_value;
    int32_t theta_rbf_level;
    int32_t theta_length_scale;
    int32_t n_theta;        /* length of kernel.theta */
    double length_scale[NNGP_MAX_DIM];
    int32_t covariance;     /* nngp_covariance of the stationary part */
} nngp_kernel;

/* Human-readable text of a status code (static storage). */
const char* nngp_status_string(int status);

/* Text of the last HIP runtime error seen by this thread's calls (static). */
const char* nngp_last_error_detail(void);

int nngp_abi_version(void);

/*
 * k-nearest-neighbour search.
 * Replaces RbfAccelerator.find_nearest_neighbors + the calc_nn kernel
 * (cuda_kernel_accelerators.py:196-381, 663-745) and NngprWorker.find_nearest_neighbors
 * (base_nngpr.py:1613-1686).
 *   ref (n_ref x d), query (n_query x d): FP64 row-major device arrays.
 *   scale (host, n_scale doubles): coordinates are divided by it before the
 *     search (kernel-space search with RBF length scales); n_scale 0 = raw.
 *   k (1..NNGP_MAX_K) neighbours per query; has_cut != 0 limits query i to refs [0, i + ref_cut)
 *     (the reference's ref_cut_index; None <=> has_cut == 0).
 *   out_idx (n_query x k int32, device): ascending (key, index) order, -1 padded.
 * Key = FP64 squared distance of scaled coordinates, accumulated with FMA over
 * features in order; ties keep the lower index.  Both precisions return the
 * same indices: FP32_REFINED screens in FP32 and re-ranks a candidate margin in
 * FP64, certifying every row (uncertified rows are redone exactly); the grid
 * search (see nngp_knn_set_method) computes FP64 keys in both.
 */
size_t nngp_knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision);
int nngp_knn(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
             const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut,
             int precision, int32_t* out_idx, void* workspace, size_t workspace_bytes,
             void* stream);

/*
 * Exact kNN lists at a new length scale from exact lists at an old one, for
 * optimizer loops.  The reference searches again at every log-likelihood
 * evaluation of fit (base_nngpr.py:530-548, RbfAccelerator.find_nearest_neighbors
 * with the new length scales, cuda_kernel_accelerators.py:196-381).
 *   cand (n_query x K int32, device): the output of nngp_knn with k = K
 *     (k <= K <= 72) at scale_old for the same ref, query and cut.
 *   state (device, nngp_knn_rescale_state_bytes): nngp_knn_rescale_init fills it
 *     once per cand (per row: the K-th candidate key and whether the first k + 1
 *     keys are separated beyond rounding); nngp_knn_rescale reads it and also
 *     uses it as scratch (calls sharing a state must not overlap).
 *   nngp_knn_rescale writes to out_idx (n_query x k) exactly the lists
 *     nngp_knn(k, scale_new) returns: under an isotropic change rows marked
 *     stable keep their lists; other rows re-rank their K candidates at the new
 *     scale and are certified by a lower bound on every other ref's key; rows
 *     that fail are recomputed exactly (their count: nngp_read_stats(state)).
 * scale_old / scale_new: host arrays of n_scale (0, 1 or d) positive values.
 */
size_t nngp_knn_rescale_state_bytes(int64_t n_ref, int64_t n_query, int d, int k, int K);
int nngp_knn_rescale_init(const double* ref, int64_t n_ref, const double* query, int64_t n_query,
                          int d, const double* scale, int n_scale, const int32_t* cand, int K,
                          int k, void* state, size_t state_bytes, void* stream);
int nngp_knn_rescale(const double* ref, int64_t n_ref, const double* query, int64_t n_query,
                     int d, const double* scale_old, const double* scale_new, int n_scale,
                     int has_cut, int64_t ref_cut, const int32_t* cand, int K, int k,
                     int32_t* out_idx, void* state, size_t state_bytes, void* stream);

/*
 * Search method of nngp_knn and of the serving model's query search,
 * process-wide, read when the refs are prepared (at each nngp_knn call, at
 * nngp_model_create).  AUTO: a uniform-grid exact search for d <= 3 and k <= 64
 * (the lowest Vecchia rows scanned brute force), the brute-force scans
 * otherwise.  BRUTE_FORCE: always the scans.  In both, the FP32_REFINED screen
 * of d >= 4 runs on the matrix cores (bf16 hi/lo split products, FP32
 * accumulation, its own certified error bound).  BRUTE_FORCE_VALU: the scans,
 * with the FP32 VALU screen for every d.  The grid search runs one query per
 * lane (64 per wave) above 100,000 queries and one query per wave below; GRID_LANES
 * and GRID_WAVES are AUTO with one of the two for every query count.  Every
 * method returns the same indices.  No reference counterpart (the reference has
 * the brute-force scan only, cuda_kernel_accelerators.py:663-745).
 */
typedef enum {
    NNGP_KNN_AUTO = 0,
    NNGP_KNN_BRUTE_FORCE = 1,
    NNGP_KNN_BRUTE_FORCE_VALU = 2,
    NNGP_KNN_GRID_LANES = 3,
    NNGP_KNN_GRID_WAVES = 4
} nngp_knn_method;
int nngp_knn_set_method(int method);
int nngp_knn_get_method(void);

/*
 * Materialised neighbour covariance blocks (and theta gradients).
 * Replaces RbfAccelerator.fill_nn_kernel + the fill_nn_kernel kernel
 * (cuda_kernel_accelerators.py:383-518, 748-864) and NngprWorker.fill_nn_kernel
 * (base_nngpr.py:1781-1826).
 *   x (n_x x d FP64), nn (n_points x width int32, -1 = empty slot).
 *   out_k (n_points x width x width FP64, row-major), out_grad (n_points x width
 *   x width x n_theta FP64, or NULL).
 * The fused entry points below never materialise these blocks; this one exists
 * for the plug-in surface and for tests.  A slot outside [-1, n_x) is never
 * read through: its row and column of out_k (and of out_grad) are NaN.
 */
int nngp_fill_nn_kernel(const double* x, int64_t n_x, int d, const int32_t* nn,
                        int64_t n_points, int width, const nngp_kernel* kernel,
                        double* out_k, double* out_grad, void* stream);

/*
 * Vecchia log-likelihood (and gradient w.r.t. log-hyperparameters) of training
 * rows [row0, row1).  Replaces the per-batch body of Nngpr._log_likelihood_worker
 * (base_nngpr.py:549-614): fill -> +alpha -> Cholesky -> mu, sigma -> l_i, and the
 * gradient, fused into one pass per point that never stores K or K_gradient.
 *   x (n x d), y (n x ny) FP64; nn (n x m int32): row i's neighbours among rows
 *   < i (as returned by nngp_knn with has_cut = 1, ref_cut = 0), -1 padded.
 *   out_loglik (device, 1 double): sum over rows and outputs.
 *   out_grad (device, n_theta doubles, or NULL when eval_gradient == 0).
 *   out_point_loglik (device, row1 - row0 doubles, or NULL).
 * Sums are reduced in a fixed order: results are bitwise reproducible.
 * Neighbour indices outside [-1, n) are read as empty slots (no gather leaves x
 * and y) and counted; nngp_read_stats on the workspace then returns
 * NNGP_ERR_INVALID_ARGUMENT.  The same holds for nngp_predict with n_train.
 */
size_t nngp_vecchia_workspace_bytes(int64_t n_rows, int m, int d, int ny, int n_theta,
                                    int precision, int eval_gradient);
int nngp_vecchia_loglik(const double* x, const double* y, int64_t n, int d, int ny,
                        const int32_t* nn, int m, int64_t row0, int64_t row1,
                        const nngp_kernel* kernel, double alpha, int precision,
                        int eval_gradient, double* out_loglik, double* out_grad,
                        double* out_point_loglik, void* workspace, size_t workspace_bytes,
                        void* stream);

/*
 * nngp_vecchia_loglik without a workspace: it allocates nothing (no hipMalloc,
 * no caller buffer) and may be called from code under HIP graph capture or from
 * loops that must not allocate.  No counterpart in the reference; it is for
 * callers that cannot allocate.  Same arguments and results as
 * nngp_vecchia_loglik minus (workspace, workspace_bytes); its scratch is a
 * fixed NNGP_NOALLOC_SCRATCH_BYTES array in device memory, one per GPU, held
 * from the first call on that GPU until the process ends.  Calls whose scratch
 * need (nngp_vecchia_workspace_bytes of the same arguments: about
 * (row1 - row0) * (1 + n_theta) * 8 bytes, plus (row1 - row0) * 8 in the FP32
 * form, for m <= 64) exceeds it return NNGP_ERR_WORKSPACE_TOO_SMALL (detail in
 * nngp_last_error_detail) before any launch: 434,874 rows with 3 theta fit in
 * either form.  Every call on a GPU shares that GPU's scratch, so calls that
 * overlap on one GPU (other streams, other threads) race.  A block that cannot
 * be factored or a neighbour index outside [-1, n) makes out_loglik (and
 * out_grad) NaN.
 */
#define NNGP_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int nngp_vecchia_loglik_noalloc(const double* x, const double* y, int64_t n, int d, int ny,
                                const int32_t* nn, int m, int64_t row0, int64_t row1,
                                const nngp_kernel* kernel, double alpha, int precision,
                                int eval_gradient, double* out_loglik, double* out_grad,
                                double* out_point_loglik, void* stream);

/*
 * Only-train prediction.  Replaces Nngpr._predict_worker (base_nngpr.py:1101-1127):
 * per query, K over [query, its m train neighbours]; alpha on the train
 * neighbours only; mean = K_xn K^-1 y_N, std = sqrt(K_00 - K_xn K^-1 K_nx).
 *   x_train (n_train x d), y_train (n_train x ny), x_query (n_query x d) FP64;
 *   nn (n_query x m int32) indices into x_train, -1 padded.
 *   out_mean (n_query x ny), out_std (n_query, or NULL).
 */
size_t nngp_predict_workspace_bytes(int64_t n_query, int m, int d, int ny, int precision);
int nngp_predict(const double* x_train, const double* y_train, int64_t n_train, int d,
                 int ny, const double* x_query, int64_t n_query, const int32_t* nn, int m,
                 const nngp_kernel* kernel, double alpha, int precision, double* out_mean,
                 double* out_std, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Sequential conditioning: prediction where query i is conditioned on training rows
 * and on earlier queries (conditioning_method 'full' / 'full-double-nn') and the
 * joint predictive covariance (return_cov).  Replaces the sequential branch of
 * Nngpr.predict (base_nngpr.py:896-1030).
 *   x_train (n_train x d), y_train (n_train x ny): FP64 device arrays (NULL when
 *     n_train == 0); x_query (n_query x d).
 *   nn (n_query x m int32, device): row i indexes the concatenation
 *     [x_train; x_query]: j < n_train a training row, j = n_train + q an earlier
 *     query q < i, -1 an empty slot (anything else: NNGP_ERR_INVALID_ARGUMENT).
 *   alpha is added to the variance of training neighbours only (:948-950).
 *   output NNGP_COND_MEAN: out_mean (n_query x ny) only.
 *          NNGP_COND_STD: out_mean and out_std (n_query), the standard deviation
 *            through the reference's sparse partial covariance (:918-1008): the
 *            covariance of two query neighbours is the one stored when the later
 *            had the earlier among its neighbours, else 0.
 *          NNGP_COND_COV: out_mean, out_cov (n_query x n_query, row-major, the
 *            joint covariance of the dense recursion, :933-1001) and out_std
 *            (its diagonal's square root, or NULL).
 * Outputs are for the normalised targets (the caller undoes normalize_y).  The
 * call reads nn on the host (it synchronises `stream`) to order the queries by
 * dependency level; every query of a level is computed at once, one level after
 * the other.  FP64 only, as the reference's sequential branch.
 */
typedef enum {
    NNGP_COND_MEAN = 0,
    NNGP_COND_STD = 1,
    NNGP_COND_COV = 2
} nngp_cond_output;
size_t nngp_predict_sequential_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d,
                                               int ny, int output);
int nngp_predict_sequential(const double* x_train, const double* y_train, int64_t n_train,
                            const double* x_query, int64_t n_query, int d, int ny,
                            const int32_t* nn, int m, const nngp_kernel* kernel, double alpha,
                            int output, double* out_mean, double* out_std, double* out_cov,
                            void* workspace, size_t workspace_bytes, void* stream);

/*
 * Samples of the sequential process.  Replaces the sequential branch of
 * Nngpr.sample_y (base_nngpr.py:680-752): y_i = K_xn K^-1 y_N + sigma_i z_i, with
 * y_N the training targets and the samples already drawn at earlier query
 * neighbours (sample by sample), sigma_i = sqrt(K_00 - K_xn K^-1 K_nx) (0 where
 * that is not positive).  Same inputs as nngp_predict_sequential; normals
 * (n_query x ny x n_samples, device) are the caller's standard normal draws z and
 * out_samples has the same shape.  Synchronises `stream`.
 */
size_t nngp_sample_sequential_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d,
                                              int ny);
int nngp_sample_sequential(const double* x_train, const double* y_train, int64_t n_train,
                           const double* x_query, int64_t n_query, int d, int ny,
                           const int32_t* nn, int m, const nngp_kernel* kernel, double alpha,
                           int n_samples, const double* normals, double* out_samples,
                           void* workspace, size_t workspace_bytes, void* stream);

/*
 * Resident prediction model for serving.  No single reference counterpart: it
 * keeps the state the reference rebuilds on every predict call
 * (_find_nn_indices_for_train_and_eval + _predict_worker, base_nngpr.py:1370-1464,
 * 1032-1135) resident on the device: the training set, the kNN ref-side prepared
 * coordinates/norms and the kernel parameters.  nngp_model_create copies
 * x_train/y_train (device pointers), allocates the model's own device memory and
 * synchronises `stream`; nngp_model_destroy frees it.  nngp_model_predict is
 * nngp_knn (query side only, no cut, k = min(num_nn, n_train)) followed by
 * nngp_predict, entirely on `stream`, with no allocation, no host
 * synchronisation and no memset nodes: it may be captured into a HIP graph.
 * Concurrent predicts on one model need distinct workspaces.
 */
typedef struct nngp_model nngp_model;
int nngp_model_create(const double* x_train, const double* y_train, int64_t n_train, int d, int ny,
                      const nngp_kernel* kernel, double alpha, int num_nn, const double* scale,
                      int n_scale, int precision, nngp_model** out_model, void* stream);
int nngp_model_destroy(nngp_model* model);
size_t nngp_model_predict_workspace_bytes(const nngp_model* model, int64_t max_query);
int nngp_model_predict(const nngp_model* model, const double* x_query, int64_t n_query,
                       double* out_mean, double* out_std, void* workspace, size_t workspace_bytes,
                       void* stream);

/*
 * Counters of the last nngp_knn / nngp_vecchia_loglik / nngp_predict call that
 * used this workspace: rows whose kNN certificate failed, or whose uniform-grid
 * search (lists of 32 or 64) met a key equal to its k-th, and were recomputed
 * exactly, points re-factored in FP64 after the FP32 pivot check.  Synchronises
 * the stream.  Returns NNGP_ERR_INVALID_ARGUMENT when that call met neighbour
 * indices outside [-1, n), else NNGP_ERR_NOT_POSITIVE_DEFINITE when a block
 * could not be factored, else NNGP_OK.  No reference counterpart (diagnostics of
 * the new precision mode).
 */
typedef struct {
    int64_t knn_rows_recomputed;
    int64_t points_refactored_fp64;
} nngp_stats;
int nngp_read_stats(const void* workspace, nngp_stats* out, void* stream);

/*
 * Kernel timing with HIP events recorded on the launch stream around the main
 * kernels: 0 = kNN brute-force scan, 1 = kNN re-rank, 2 = fused log-likelihood
 * pass, 3 = fused prediction pass, 4 = kNN grid search, 5 = kNN grid build (all
 * its kernels).  Off by default; nngp_timing_read synchronises on
 * the recorded events and returns the summed milliseconds and launch count.
 * No reference counterpart (measurement support).
 */
int nngp_timing_enable(int on);
int nngp_timing_read(int kernel, int reset, double* total_ms, int64_t* launches);
/*
 * Work measured inside a timed kernel while timing is on, summed over its
 * launches: for the kNN grid search (4) the (query, ref) pairs whose key was
 * computed; 0 for the others.  Synchronises the device.
 */
int nngp_timing_work(int kernel, int reset, int64_t* units);

#ifdef __cplusplus
}
#endif

#endif /* NNGP_H */
