// Internal interface of the kNN module (knn.hip).
#pragma once
#include "common.h"

namespace nngp {

// Ref-side state of the kNN search: scaled coordinates (AoS), blocked scan
// layout (FP32 centred + norms for the screen, or FP64), centre and max radius.
struct KnnRefs {
    const double* u64;
    const double* s64;
    const float* s32;
    const long long* mm;
    const unsigned int* smax;
    int64_t n;
    int d;
    int precision;
    int n_scale;
    double scale[NNGP_MAX_DIM];
};

size_t knn_refs_bytes(int64_t n_ref, int d, int precision);
int knn_prepare_refs(const double* ref, int64_t n_ref, int d, const double* scale, int n_scale,
                     int precision, void* mem, size_t bytes, KnnRefs* out, hipStream_t st);
size_t knn_search_bytes(int64_t n_query, int d, int k, int precision, bool same);
// query == NULL searches the ref set itself (Vecchia self-search).
int knn_search(const KnnRefs& R, const double* query, int64_t n_query, int k, int has_cut,
               int64_t ref_cut, int32_t* out_idx, void* ws, size_t ws_bytes, WsHeader* hdr,
               hipStream_t st);

size_t knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision,
                           bool same);
int knn_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
            const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut, int precision,
            int32_t* out_idx, void* ws, size_t ws_bytes, hipStream_t st, bool reset_stats);

}  // namespace nngp
