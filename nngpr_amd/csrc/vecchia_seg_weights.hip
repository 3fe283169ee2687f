// Conditioning-weight instances of the fused Vecchia segment kernel
// (vecchia_seg_impl.h): the per-query pass of a sequential prediction (FP64,
// RBF or Matern), writing w = K^-1 k and s2 = K_00 - k.w next to the mean.
#include "vecchia_seg_impl.h"

namespace nngp {

namespace {

// LDS of one block of the weights kernel that serves (m, d).
size_t weights_lds_bytes(int m, int d) {
    auto by_d = [&](auto mr) -> size_t {
        constexpr int MR = decltype(mr)::value;
        return d <= 3 ? seg_lds<double, MR, 3>()
                      : (d <= 8 ? seg_lds<double, MR, 8>() : seg_lds<double, MR, kSegMaxDim>());
    };
    return m <= 16 ? by_d(std::integral_constant<int, 16>{})
                   : (m <= 32 ? by_d(std::integral_constant<int, 32>{})
                              : by_d(std::integral_constant<int, 64>{}));
}

}  // namespace

// Returns -1 (nothing launched) when the segment kernel does not serve (m, d) or
// its LDS exceeds the device's per-block limit.
int launch_seg_weights(const VecArgs& a, hipStream_t st) {
    if (!seg_supported(a.m, a.d) || !a.wout || !a.s2out) return -1;
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        weights_lds_bytes(a.m, a.d) > (size_t)lds)
        return -1;
    if (a.npts == 0) return 0;
    if (a.kp.fam != NNGP_COV_RBF && a.kp.n_ls > 0) launch_seg_weights_family<true>(a, st);
    else launch_seg_weights_family<false>(a, st);
    return 0;
}

}  // namespace nngp
