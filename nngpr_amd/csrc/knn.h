// Internal interface of the kNN module (knn.hip).
#pragma once
#include "common.h"

namespace nngp {

size_t knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision,
                           bool same);
int knn_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
            const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut, int precision,
            int32_t* out_idx, void* ws, size_t ws_bytes, hipStream_t st, bool reset_stats);

}  // namespace nngp
