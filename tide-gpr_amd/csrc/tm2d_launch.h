// Kernel launchers of the 2-D TM propagator: the launch shape (rows per
// wavefront, border tile height, single- or two-step tile set) selects the
// kernel instantiation.  The definitions are compiled in one translation unit
// per pass and field type (tm2d_launch_{fwd,adj}_{f32,f64}.hip) so the kernel
// families build in parallel; every other user sees the declarations only.
#pragma once

#include "tm2d_device.h"
#include "tm2d_fused.h"

namespace tide_hip {
namespace tm2d {

template <typename T, int FMT>
void launch_fwd(int rows, bool loop, dim3 grid, dim3 block, hipStream_t stream,
                const FwdArgs<T> &a);
template <typename T, int FMT>
void launch_fwd2(int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 bool split, const Fwd2Args<T> &a);
template <typename T, int FMT>
void launch_adj(int rows, bool loop, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
                const AdjArgs<T> &a);
template <typename T, int FMT>
void launch_adj2(int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 bool split, const Adj2Args<T> &a);

#ifdef TIDE_TM2D_LAUNCH_DEFINITIONS

template <typename T, int FMT>
void launch_fwd(int rows, bool loop, dim3 grid, dim3 block, hipStream_t stream,
                const FwdArgs<T> &a) {
  if (loop) {
    switch (rows) {
      case 2: fwd_loop<T, 2, FMT><<<grid, block, 0, stream>>>(a); break;
      case 4: fwd_loop<T, 4, FMT><<<grid, block, 0, stream>>>(a); break;
      case 16: fwd_loop<T, 16, FMT><<<grid, block, 0, stream>>>(a); break;
      default: fwd_loop<T, 8, FMT><<<grid, block, 0, stream>>>(a); break;
    }
    return;
  }
  switch (rows) {
    case 2: fwd_step<T, 2, FMT><<<grid, block, 0, stream>>>(a); break;
    case 4: fwd_step<T, 4, FMT><<<grid, block, 0, stream>>>(a); break;
    case 16: fwd_step<T, 16, FMT><<<grid, block, 0, stream>>>(a); break;
    default: fwd_step<T, 8, FMT><<<grid, block, 0, stream>>>(a); break;
  }
}

// Two-step launches: interior tiles of height TY, border tiles of height TYB,
// either split (the interior, the border tiles beside the interior rows and
// those above and below it: three kernels with their own register budgets) or
// in one kernel that takes the border tiles first.
template <typename T, int TY, int TYB, int FMT>
void launch_fwd2_parts(int64_t groups, dim3 block, hipStream_t stream, bool split,
                       const Fwd2Args<T> &a) {
  if (!split) {
    const int64_t tiles = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles) +
                          split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBorderTiles);
    fwd2_step<T, TY, FMT, kOrderedTiles, TYB><<<(unsigned)(tiles * groups), block, 0, stream>>>(a);
    return;
  }
  const int64_t inner = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles);
  const int64_t sides = split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kSideTiles);
  const int64_t bands = split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBandTiles);
  if (inner > 0) fwd2_step<T, TY, FMT, kInnerTiles><<<(unsigned)(inner * groups), block, 0, stream>>>(a);
  if (sides > 0) fwd2_step<T, TYB, FMT, kSideTiles><<<(unsigned)(sides * groups), block, 0, stream>>>(a);
  if (bands > 0) fwd2_step<T, TYB, FMT, kBandTiles><<<(unsigned)(bands * groups), block, 0, stream>>>(a);
}

// border_rows 0 gives the border tiles the interior height.
template <typename T, int FMT>
void launch_fwd2(int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 bool split, const Fwd2Args<T> &a) {
  const int b = border_rows == 0 ? rows : border_rows;
  if (rows == 2) {
    return b == 1 ? launch_fwd2_parts<T, 2, 1, FMT>(groups, block, stream, split, a)
                  : launch_fwd2_parts<T, 2, 2, FMT>(groups, block, stream, split, a);
  }
  if constexpr (sizeof(T) == 4) {
    if (rows == 8)
      return b == 1   ? launch_fwd2_parts<T, 8, 1, FMT>(groups, block, stream, split, a)
             : b == 2 ? launch_fwd2_parts<T, 8, 2, FMT>(groups, block, stream, split, a)
                      : launch_fwd2_parts<T, 8, 8, FMT>(groups, block, stream, split, a);
  }
  b == 1   ? launch_fwd2_parts<T, 4, 1, FMT>(groups, block, stream, split, a)
  : b == 2 ? launch_fwd2_parts<T, 4, 2, FMT>(groups, block, stream, split, a)
           : launch_fwd2_parts<T, 4, 4, FMT>(groups, block, stream, split, a);
}

template <typename T, int TY, int TYB, int FMT>
void launch_adj2_parts(int64_t groups, dim3 block, hipStream_t stream, bool split,
                       const Adj2Args<T> &a) {
  const size_t lds = 2 * (size_t)block.y * TY * 64 * sizeof(double);
  const size_t lds_b = 2 * (size_t)block.y * TYB * 64 * sizeof(double);
  if (!split) {
    const int64_t tiles = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles) +
                          split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBorderTiles);
    adj2_step<T, TY, FMT, kOrderedTiles, TYB><<<(unsigned)(tiles * groups), block, lds, stream>>>(a);
    return;
  }
  const int64_t inner = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles);
  const int64_t sides = split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kSideTiles);
  const int64_t bands = split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBandTiles);
  if (inner > 0) adj2_step<T, TY, FMT, kInnerTiles><<<(unsigned)(inner * groups), block, lds, stream>>>(a);
  if (sides > 0)
    adj2_step<T, TYB, FMT, kSideTiles><<<(unsigned)(sides * groups), block, lds_b, stream>>>(a);
  if (bands > 0)
    adj2_step<T, TYB, FMT, kBandTiles><<<(unsigned)(bands * groups), block, lds_b, stream>>>(a);
}

template <typename T, int FMT>
void launch_adj2(int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 bool split, const Adj2Args<T> &a) {
  const int b = border_rows == 0 ? rows : border_rows;
  if (rows == 2)
    return b == 1 ? launch_adj2_parts<T, 2, 1, FMT>(groups, block, stream, split, a)
                  : launch_adj2_parts<T, 2, 2, FMT>(groups, block, stream, split, a);
  b == 1   ? launch_adj2_parts<T, 4, 1, FMT>(groups, block, stream, split, a)
  : b == 2 ? launch_adj2_parts<T, 4, 2, FMT>(groups, block, stream, split, a)
           : launch_adj2_parts<T, 4, 4, FMT>(groups, block, stream, split, a);
}

template <typename T, int FMT>
void launch_adj(int rows, bool loop, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
                const AdjArgs<T> &a) {
  if (loop) {
    switch (rows) {
      case 2: adj_loop<T, 2, FMT><<<grid, block, 0, stream>>>(a); break;
      case 8: adj_loop<T, 8, FMT><<<grid, block, 0, stream>>>(a); break;
      default: adj_loop<T, 4, FMT><<<grid, block, 0, stream>>>(a); break;
    }
    return;
  }
  switch (rows) {
    case 2: adj_step<T, 2, FMT><<<grid, block, lds, stream>>>(a); break;
    case 8: adj_step<T, 8, FMT><<<grid, block, lds, stream>>>(a); break;
    default: adj_step<T, 4, FMT><<<grid, block, lds, stream>>>(a); break;
  }
}

#endif  // TIDE_TM2D_LAUNCH_DEFINITIONS

}  // namespace tm2d
}  // namespace tide_hip
