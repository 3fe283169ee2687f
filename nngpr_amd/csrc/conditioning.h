// Internal interface of the sequential-conditioning prediction (conditioning.hip).
#pragma once
#include "vecchia.h"

namespace nngp {

// Outputs of a sequential prediction.
enum CondOutput { COND_MEAN = 0, COND_STD = 1, COND_COV = 2, COND_SAMPLE = 3 };

struct CondProblem {
    const double* x_train;   // n_train x d (may be NULL when n_train == 0)
    const double* y_train;   // n_train x ny
    int64_t n_train;
    const double* x_query;   // n_query x d
    int64_t n_query;
    int d, ny, m;
    const int32_t* nn;       // n_query x m, indices into [x_train; x_query]
    KernParams kp;
    double alpha;
    int output;              // CondOutput
    int n_samples;           // COND_SAMPLE
    const double* normals;   // COND_SAMPLE: n_query x ny x n_samples standard normals
    double* out_mean;        // n_query x ny (not COND_SAMPLE)
    double* out_std;         // n_query (COND_STD, COND_COV; may be NULL)
    double* out_cov;         // n_query x n_query (COND_COV)
    double* out_samples;     // n_query x ny x n_samples (COND_SAMPLE)
};

size_t cond_workspace_bytes(int64_t n_train, int64_t n_query, int m, int d, int ny, int output);
// Runs the whole prediction.  Synchronises the stream: the neighbour lists are
// read on the host (validated, and the dependency levels built) before any kernel
// runs, the level order is copied back, and the factor status is read at the end.
int cond_run(const CondProblem& P, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace nngp
