// Launchers of the K-step adj kernels for double fields.
#define TIDE_TM2D_LAUNCH_K_DEFINITIONS
#include "tm2d_launch_k.h"

namespace tide_hip {
namespace tm2d {

template void launch_adjk<double, 0>(int, int, int, int64_t, dim3, hipStream_t, const AdjKArgs<double> &);
template void launch_adjk<double, 1>(int, int, int, int64_t, dim3, hipStream_t, const AdjKArgs<double> &);

}  // namespace tm2d
}  // namespace tide_hip
