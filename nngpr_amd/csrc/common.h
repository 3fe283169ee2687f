// Shared device helpers for the NNGP HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nngp.h"

#define NNGP_WAVE 64

namespace nngp {

// Error-checking helper for host code: returns NNGP_ERR_HIP from the enclosing
// function on any HIP runtime error.
#define NNGP_HIP_CHECK(expr)                                 \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) {                              \
            nngp::set_last_hip_error(_e, #expr, __LINE__);   \
            return NNGP_ERR_HIP;                             \
        }                                                    \
    } while (0)

void set_last_hip_error(hipError_t e, const char* what, int line);

// Optional per-kernel timing with HIP events on the launch stream (off by default).
enum TimedKernel {
    TK_KNN_SCAN = 0,
    TK_KNN_RERANK = 1,
    TK_LOGLIK = 2,
    TK_PREDICT = 3,
    TK_KNN_GRID = 4,
    TK_KNN_GRID_BUILD = 5,
    TK_COUNT = 6
};
void timing_begin(int id, hipStream_t st);
void timing_end(int id, hipStream_t st);
// Device counter of measured work units of a timed kernel (the grid search: the
// (query, ref) pairs it keyed), or NULL when timing is off.
unsigned long long* timing_work_counter(int id);

template <typename T>
__device__ __forceinline__ T dev_inf();
template <>
__device__ __forceinline__ float dev_inf<float>() { return __builtin_inff(); }
template <>
__device__ __forceinline__ double dev_inf<double>() { return __builtin_inf(); }

// Sum over the lanes [base, base + width) of a wave; every lane of the group gets
// the total.  width is a power of two <= 64.
template <int WIDTH>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = WIDTH / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WIDTH);
    return v;
}

// Header at offset 0 of every workspace: counters of the last call.
struct WsHeader {
    int64_t knn_rows_recomputed;
    int64_t points_refactored_fp64;
    int64_t not_pd;
    int64_t bad_index;  // neighbour slots outside [-1, n_src), read as empty
    int64_t reserved[28];
};

// A neighbour slot as the fused kernels use it: a row of the source set
// [0, n_src) or -1 (empty).  Any other value is counted in hdr->bad_index (when
// hdr is given) and read as an empty slot, so no gather leaves the source rows.
__device__ __forceinline__ int nbr_slot(int j, int64_t n_src, WsHeader* hdr) {
    if (j >= 0 && (int64_t)j < n_src) return j;
    if (j != -1 && hdr) atomicAdd(reinterpret_cast<unsigned long long*>(&hdr->bad_index), 1ULL);
    return -1;
}

static __global__ void zero_header_kernel(WsHeader* h) {
    int i = threadIdx.x;
    if (i < (int)(sizeof(WsHeader) / sizeof(int64_t))) reinterpret_cast<int64_t*>(h)[i] = 0;
}

// Workspace carving: bump allocator over a caller-provided device buffer.
struct Carver {
    char* base;
    size_t cap;
    size_t used;
    __host__ Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c), used(0) {}
    template <typename T>
    __host__ T* take(size_t count) {
        size_t off = (used + 255) & ~size_t(255);
        used = off + count * sizeof(T);
        if (base == nullptr) return nullptr;
        return reinterpret_cast<T*>(base + off);
    }
    __host__ bool ok() const { return used <= cap; }
};

}  // namespace nngp
