// Launchers of the forward kernels for float fields (storage formats 0, 1, 2).
#define TIDE_TM2D_LAUNCH_DEFINITIONS
#include "tm2d_launch.h"

namespace tide_hip {
namespace tm2d {

template void launch_fwd<float, 0>(int, bool, dim3, dim3, hipStream_t, const FwdArgs<float> &);
template void launch_fwd2<float, 0>(int, int, int64_t, dim3, hipStream_t, bool, const Fwd2Args<float> &);
template void launch_fwd<float, 1>(int, bool, dim3, dim3, hipStream_t, const FwdArgs<float> &);
template void launch_fwd2<float, 1>(int, int, int64_t, dim3, hipStream_t, bool, const Fwd2Args<float> &);
template void launch_fwd<float, 2>(int, bool, dim3, dim3, hipStream_t, const FwdArgs<float> &);
template void launch_fwd2<float, 2>(int, int, int64_t, dim3, hipStream_t, bool, const Fwd2Args<float> &);

}  // namespace tm2d
}  // namespace tide_hip
