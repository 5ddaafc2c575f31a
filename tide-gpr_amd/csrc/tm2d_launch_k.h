// Launchers of the K-step (K = 3, 4) 2-D TM kernels of tm2d_blockk.h.  As
// in tm2d_launch.h, the definitions compile in one translation unit per
// pass and field type (tm2d_launch_k_{fwd,adj}_{f32,f64}.hip).
#pragma once

#include "tm2d_blockk.h"

namespace tide_hip {
namespace tm2d {

// One ordered-tile launch: interior tiles of `rows` rows, border tiles of
// `border_rows` (0: the interior height).
template <typename T, int FMT>
void launch_fwdk(int k, int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 const FwdKArgs<T> &a);
template <typename T, int FMT>
void launch_adjk(int k, int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 const AdjKArgs<T> &a);

#ifdef TIDE_TM2D_LAUNCH_K_DEFINITIONS

template <typename T, int TY, int TYB, int K, int FMT>
void fwdk_parts(int64_t groups, dim3 block, hipStream_t stream, const FwdKArgs<T> &a) {
  const int64_t tiles = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles) +
                        split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBorderTiles);
  fwdk_step<T, TY, K, FMT, kOrderedTiles, TYB><<<(unsigned)(tiles * groups), block, 0, stream>>>(a);
}

template <typename T, int FMT>
void launch_fwdk(int k, int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 const FwdKArgs<T> &a) {
  const bool b2 = border_rows != 0 && border_rows < rows;
  if constexpr (sizeof(T) == 4) {
    if (rows >= 8) {
      if (k == 3) return b2 ? fwdk_parts<T, 8, 2, 3, FMT>(groups, block, stream, a)
                            : fwdk_parts<T, 8, 8, 3, FMT>(groups, block, stream, a);
      return b2 ? fwdk_parts<T, 8, 2, 4, FMT>(groups, block, stream, a)
                : fwdk_parts<T, 8, 8, 4, FMT>(groups, block, stream, a);
    }
  }
  if (rows >= 4) {
    if (k == 3) return b2 ? fwdk_parts<T, 4, 2, 3, FMT>(groups, block, stream, a)
                          : fwdk_parts<T, 4, 4, 3, FMT>(groups, block, stream, a);
    return b2 ? fwdk_parts<T, 4, 2, 4, FMT>(groups, block, stream, a)
              : fwdk_parts<T, 4, 4, 4, FMT>(groups, block, stream, a);
  }
  if (k == 3) return fwdk_parts<T, 2, 2, 3, FMT>(groups, block, stream, a);
  fwdk_parts<T, 2, 2, 4, FMT>(groups, block, stream, a);
}

template <typename T, int TY, int TYB, int K, int FMT>
void adjk_parts(int64_t groups, dim3 block, hipStream_t stream, const AdjKArgs<T> &a) {
  const size_t lds = 2 * (size_t)block.y * TY * 64 * sizeof(double);
  const int64_t tiles = split_tiles<TY>(a.split, a.pts.n_xt, a.g.ny, kInnerTiles) +
                        split_tiles<TYB>(a.split, a.pts.n_xt, a.g.ny, kBorderTiles);
  const unsigned grid = (unsigned)(tiles * groups);
  if constexpr (sizeof(T) == 4) {
    const int waves = block.x * block.y <= 256 ? g_tuning.adj_waves : 0;
    if (waves == 4)
      return adjk_step_w<T, TY, K, FMT, kOrderedTiles, TYB, 4><<<grid, block, lds, stream>>>(a);
  }
  adjk_step<T, TY, K, FMT, kOrderedTiles, TYB><<<grid, block, lds, stream>>>(a);
}

template <typename T, int FMT>
void launch_adjk(int k, int rows, int border_rows, int64_t groups, dim3 block, hipStream_t stream,
                 const AdjKArgs<T> &a) {
  const bool b2 = border_rows != 0 && border_rows < rows;
  if (rows >= 4) {
    if (k == 3) return b2 ? adjk_parts<T, 4, 2, 3, FMT>(groups, block, stream, a)
                          : adjk_parts<T, 4, 4, 3, FMT>(groups, block, stream, a);
    return b2 ? adjk_parts<T, 4, 2, 4, FMT>(groups, block, stream, a)
              : adjk_parts<T, 4, 4, 4, FMT>(groups, block, stream, a);
  }
  if (k == 3) return adjk_parts<T, 2, 2, 3, FMT>(groups, block, stream, a);
  adjk_parts<T, 2, 2, 4, FMT>(groups, block, stream, a);
}

#endif  // TIDE_TM2D_LAUNCH_K_DEFINITIONS

}  // namespace tm2d
}  // namespace tide_hip
