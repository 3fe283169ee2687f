// Internal interface of the register-resident fused Vecchia kernel (m <= 64).
#pragma once
#include "vecchia.h"

namespace nngp {

struct RegExtra {
    const int64_t* plist;                 // optional indirection: slot -> point
    const unsigned long long* pcount;     // optional device count of slots
    int64_t* flag_list;                   // FP32 form: points to redo in FP64
    unsigned long long* flag_count;
};

int reg_mr(int m);
size_t reg_lds_bytes(int m, int d, bool fp32);
int launch_reg(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points, hipStream_t st);

// Segment kernel (m <= 64, d <= 32): MR / 2 lanes per block of MR rows, two rows
// per lane, factor in registers.  fp32: FP32 factor + FP64 refinement, flagged
// blocks go to ex.flag_list; else the FP64 form.  launch_seg dispatches on the
// covariance family; launch_seg_matern runs the Matern instances.
bool seg_supported(int m, int d);
int launch_seg(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points,
               hipStream_t st);
int launch_seg_matern(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points,
                      hipStream_t st);
// FP64 conditioning weights of a sequential prediction (a.wout, a.s2out, alpha
// only on neighbours below a.alpha_limit) for every point of a.
int launch_seg_weights(const VecArgs& a, hipStream_t st);

}  // namespace nngp
