// Matern instances of the fused Vecchia segment kernel (vecchia_seg_impl.h); the
// family (nu = 1/2, 3/2, 5/2) is read from the kernel parameters at run time.
#include "vecchia_seg_impl.h"

namespace nngp {

int launch_seg_matern(const VecArgs& a, const RegExtra& ex, bool fp32, int64_t max_points,
                      hipStream_t st) {
    if (!seg_supported(a.m, a.d)) return -1;
    launch_seg_family<true>(a, ex, fp32, max_points, st);
    return 0;
}

}  // namespace nngp
