// RBF instances of the fused Vecchia segment kernel (vecchia_seg_impl.h) and the
// dispatch over covariance families.
#include "vecchia_seg_impl.h"

namespace nngp {

bool seg_supported(int m, int d) { return m >= 1 && m <= 64 && d >= 1 && d <= kSegMaxDim; }

// fp32: FP32 factor + FP64 refinement form (flagged blocks are queued in
// ex.flag_list); else the FP64 form.
int launch_seg(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points,
               hipStream_t st) {
    if (!seg_supported(a.m, a.d)) return -1;
    if (a.kp.fam != NNGP_COV_RBF && a.kp.n_ls > 0) return launch_seg_matern(a, ex, fp32, max_points, st);
    launch_seg_family<false>(a, ex, fp32, max_points, st);
    return 0;
}

}  // namespace nngp
