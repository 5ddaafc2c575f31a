// Per-problem descriptors of the multi-problem 2-D TM launches (batched
// plans and the serving engine): one BatchItem per problem, a flat tile
// index decoded into (tile column, tile row, shot group) of that problem,
// and the step arguments of its forward step k / adjoint iteration k.
#pragma once

#include "tm2d_device.h"

namespace tide_hip {
namespace tm2d {

template <typename T> struct BatchItem {
  Geom<T> g;
  Points pts;
  const T *ca, *cb, *cq;
  T *fwd[2][5];  // [parity]: ey, hx, hz, m_ey_x, m_ey_z (parity 0: caller's arrays)
  T *mhzx, *mhxz;
  const T *f;
  T *r;
  void *ey_store, *curl_store;
  int64_t store_step_bytes;
  T *lam[2][7];  // [parity]: ley, lhx, lhz, m_hz_x, m_hx_z, m_ey_x, m_ey_z
  const T *grad_r;
  T *grad_f;
  double *gca, *gcb;           // what the kernels add to (group slices when n_groups > 1)
  int64_t group_stride;        // elements between group slices (0: one group)
  int32_t n_shots, n_src, n_rec, n_yt, nt, interval, n_groups, tile0;
};

__device__ __forceinline__ TileCoord batch_tile(int local, int n_xt, int n_yt) {
  const int n_tiles = n_xt * n_yt;
  TileCoord tc;
  tc.grp = local / n_tiles;
  const int tile = local - tc.grp * n_tiles;
  tc.yt = tile / n_xt;
  tc.xt = tile - tc.yt * n_xt;
  return tc;
}

template <typename T>
__device__ __forceinline__ FwdArgs<T> item_fwd_args(const BatchItem<T> &it, int k) {
  const int p = k & 1;
  FwdArgs<T> a;
  a.g = it.g;
  a.pts = it.pts;
  a.ca = it.ca;
  a.cb = it.cb;
  a.cq = it.cq;
  a.ey_i = it.fwd[p][0];
  a.hx_i = it.fwd[p][1];
  a.hz_i = it.fwd[p][2];
  a.mex_i = it.fwd[p][3];
  a.mez_i = it.fwd[p][4];
  a.ey_o = it.fwd[p ^ 1][0];
  a.hx_o = it.fwd[p ^ 1][1];
  a.hz_o = it.fwd[p ^ 1][2];
  a.mex_o = it.fwd[p ^ 1][3];
  a.mez_o = it.fwd[p ^ 1][4];
  a.mhzx = it.mhzx;
  a.mhxz = it.mhxz;
  a.f_t = it.n_src > 0 ? it.f + (int64_t)k * it.n_shots * it.n_src : nullptr;
  a.r_t = it.n_rec > 0 ? it.r + (int64_t)k * it.n_shots * it.n_rec : nullptr;
  const bool st = (k % it.interval) == 0;
  const int64_t off = (int64_t)(k / it.interval) * it.store_step_bytes;
  a.ey_store = st && it.ey_store ? (char *)it.ey_store + off : nullptr;
  a.curl_store = st && it.curl_store ? (char *)it.curl_store + off : nullptr;
  a.n_shots = it.n_shots;
  a.n_yt = it.n_yt;
  return a;
}

// Adjoint iteration k applies the transpose of forward step t = nt-1-k.
template <typename T>
__device__ __forceinline__ AdjArgs<T> item_adj_args(const BatchItem<T> &it, int k) {
  const int p = k & 1;
  const int t = it.nt - 1 - k;
  AdjArgs<T> a;
  a.g = it.g;
  a.pts = it.pts;
  a.ca = it.ca;
  a.cb = it.cb;
  a.cq = it.cq;
  a.ley_i = it.lam[p][0];
  a.lhx_i = it.lam[p][1];
  a.lhz_i = it.lam[p][2];
  a.mx_i = it.lam[p][3];
  a.my_i = it.lam[p][4];
  a.nx_i = it.lam[p][5];
  a.nz_i = it.lam[p][6];
  a.ley_o = it.lam[p ^ 1][0];
  a.lhx_o = it.lam[p ^ 1][1];
  a.lhz_o = it.lam[p ^ 1][2];
  a.mx_o = it.lam[p ^ 1][3];
  a.my_o = it.lam[p ^ 1][4];
  a.nx_o = it.lam[p ^ 1][5];
  a.nz_o = it.lam[p ^ 1][6];
  const bool g_step = (it.gca != nullptr || it.gcb != nullptr) && (t % it.interval) == 0;
  const int64_t off = (int64_t)(t / it.interval) * it.store_step_bytes;
  a.ey_store = g_step && it.gca ? (const char *)it.ey_store + off : nullptr;
  a.curl_store = g_step && it.gcb ? (const char *)it.curl_store + off : nullptr;
  a.gca = g_step ? it.gca : nullptr;
  a.gcb = g_step ? it.gcb : nullptr;
  a.per_shot_grad = 0;
  a.step_scale = (double)it.interval;
  a.grad_r_t = it.n_rec > 0 ? it.grad_r + (int64_t)t * it.n_shots * it.n_rec : nullptr;
  a.grad_f_t = (it.n_src > 0 && it.grad_f) ? it.grad_f + (int64_t)t * it.n_shots * it.n_src
                                           : nullptr;
  a.n_shots = it.n_shots;
  a.n_yt = it.n_yt;
  a.group_stride = it.group_stride;
  return a;
}

}  // namespace tm2d
}  // namespace tide_hip
