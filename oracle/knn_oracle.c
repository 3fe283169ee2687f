/*
 * CPU restatement of the reference's brute-force nearest-neighbour search.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP kNN kernels are
 * checked against; the product never links or calls it.
 *
 * Semantics restated (not translated) from the reference's `calc_nn` CUDA kernel
 * and its host wrapper:
 *   - coordinates are divided by the RBF length scale(s) first (kernel-space
 *     search; cuda_kernel_accelerators.py:235-239), or used raw (input-space);
 *   - the key of a (query, ref) pair is the squared Euclidean distance of the
 *     scaled coordinates, accumulated feature by feature as
 *     key = fma(t_f, t_f, key) with t_f = ref_f - query_f
 *     (cuda_kernel_accelerators.py:711-714, compiled with FMA contraction);
 *   - each query keeps the k smallest keys in ascending order; a candidate is
 *     admitted only when key < current k-th key (strict, :725) and is placed
 *     after entries with an equal key (:728), so among equal keys the lower
 *     ref index comes first;
 *   - query i only searches refs [0, i + ref_cut) when a cut is given
 *     (MAX_SEARCH_POINT, :358-362), else all refs; slots that stay empty are -1
 *     (:362, :740-742).
 * The result is therefore the first k entries of the refs sorted by (key, index).
 *
 * Built with -ffp-contract=off so that the only fused operations are the
 * explicit fma() calls below.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Scale coordinates: u = x / ls (per feature when n_ls == d, uniform when
 * n_ls == 1, untouched when n_ls == 0). */
static void scale_points(const double *x, int64_t n, int d, const double *ls,
                         int n_ls, double *u) {
    for (int64_t i = 0; i < n; ++i) {
        for (int f = 0; f < d; ++f) {
            double v = x[i * d + f];
            if (n_ls == 1)
                v = v / ls[0];
            else if (n_ls == d)
                v = v / ls[f];
            u[i * d + f] = v;
        }
    }
}

double oracle_knn_key(const double *ur, const double *uq, int d) {
    double key = 0.0;
    for (int f = 0; f < d; ++f) {
        double t = ur[f] - uq[f];
        key = fma(t, t, key);
    }
    return key;
}

/* out: n_query x k int32, row-major. has_cut != 0 enables the Vecchia cut. */
int oracle_knn(const double *ref, int64_t n_ref, const double *query,
               int64_t n_query, int d, const double *ls, int n_ls, int k,
               int has_cut, int64_t cut, int32_t *out, double *out_keys) {
    if (k <= 0 || d <= 0) return -1;
    double *ur = (double *)malloc(sizeof(double) * (size_t)(n_ref > 0 ? n_ref : 1) * d);
    double *uq = (double *)malloc(sizeof(double) * (size_t)(n_query > 0 ? n_query : 1) * d);
    if (!ur || !uq) { free(ur); free(uq); return -2; }
    scale_points(ref, n_ref, d, ls, n_ls, ur);
    scale_points(query, n_query, d, ls, n_ls, uq);

#pragma omp parallel
    {
        double *keys = (double *)malloc(sizeof(double) * (size_t)k);
        int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < n_query; ++q) {
            int64_t lim = n_ref;
            if (has_cut) {
                lim = q + cut;
                if (lim < 0) lim = 0;
                if (lim > n_ref) lim = n_ref;
            }
            for (int j = 0; j < k; ++j) { keys[j] = INFINITY; idx[j] = -1; }
            double kth = INFINITY;
            const double *pq = uq + q * d;
            for (int64_t r = 0; r < lim; ++r) {
                double key = oracle_knn_key(ur + r * d, pq, d);
                if (key < kth) {
                    int ii = k - 1;
                    while (ii > 0 && !(keys[ii - 1] <= key)) {
                        keys[ii] = keys[ii - 1];
                        idx[ii] = idx[ii - 1];
                        --ii;
                    }
                    keys[ii] = key;
                    idx[ii] = r;
                    kth = keys[k - 1];
                }
            }
            for (int j = 0; j < k; ++j) {
                out[q * k + j] = (int32_t)idx[j];
                if (out_keys) out_keys[q * k + j] = keys[j];
            }
        }
        free(keys);
        free(idx);
    }
    free(ur);
    free(uq);
    return 0;
}
