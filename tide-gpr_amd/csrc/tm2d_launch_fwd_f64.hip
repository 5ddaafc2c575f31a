// Launchers of the forward kernels for double fields (storage formats 0, 1).
#define TIDE_TM2D_LAUNCH_DEFINITIONS
#include "tm2d_launch.h"

namespace tide_hip {
namespace tm2d {

template void launch_fwd<double, 0>(int, bool, dim3, dim3, hipStream_t, const FwdArgs<double> &);
template void launch_fwd2<double, 0>(int, int, int64_t, dim3, hipStream_t, bool, const Fwd2Args<double> &);
template void launch_fwd<double, 1>(int, bool, dim3, dim3, hipStream_t, const FwdArgs<double> &);
template void launch_fwd2<double, 1>(int, int, int64_t, dim3, hipStream_t, bool, const Fwd2Args<double> &);

}  // namespace tm2d
}  // namespace tide_hip
