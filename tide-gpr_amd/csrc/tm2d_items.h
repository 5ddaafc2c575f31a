// Per-problem descriptors of the multi-problem 2-D TM launches (batched
// plans and the serving engine): one BatchItem per problem, a flat tile
// index decoded into (tile column, tile row, shot group) of that problem,
// and the step arguments of its forward step k / adjoint iteration k.
#pragma once

#include "tm2d_fused.h"

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
  // Two-step launches (tm2d_fused.h): point table with halo copies, interior
  // rectangle of the two-step tiles, second copies of the E-phase memories
  // (the two-step kernels ping-pong them), tile rows and first flat tile.
  Points pts2;
  TileSplit split2;
  T *mhzx_b, *mhxz_b;
  int32_t n_yt2, tile0_2;
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

// Forward step k reading the state buffers of parity `par` (writing par^1)
// and updating the E-phase memories of copy `epar` in place.
template <typename T>
__device__ __forceinline__ FwdArgs<T> item_fwd_args(const BatchItem<T> &it, int k, int par, int epar) {
  const int p = par;
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
  a.mhzx = epar ? it.mhzx_b : it.mhzx;
  a.mhxz = epar ? it.mhxz_b : it.mhxz;
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

// One step per launch: step k reads parity k & 1.
template <typename T>
__device__ __forceinline__ FwdArgs<T> item_fwd_args(const BatchItem<T> &it, int k) {
  return item_fwd_args(it, k, k & 1, 0);
}

// Snapshot slot of step s in a snapshot buffer (null: none kept at step s).
template <typename T, typename V>
__device__ __forceinline__ V *item_snap(const BatchItem<T> &it, V *buf, int s) {
  if (buf == nullptr || (s % it.interval) != 0) return nullptr;
  using B = typename std::conditional<std::is_const<V>::value, const char, char>::type;
  return (V *)((B *)buf + (int64_t)(s / it.interval) * it.store_step_bytes);
}

// Forward steps k and k+1 in one launch, reading parity `par` (fields and
// both memory pairs) and writing par^1.
template <typename T>
__device__ __forceinline__ Fwd2Args<T> item_fwd2_args(const BatchItem<T> &it, int k, int par) {
  Fwd2Args<T> a;
  a.g = it.g;
  a.pts = it.pts2;
  a.ca = it.ca;
  a.cb = it.cb;
  a.cq = it.cq;
  a.ey_i = it.fwd[par][0];
  a.hx_i = it.fwd[par][1];
  a.hz_i = it.fwd[par][2];
  a.mex_i = it.fwd[par][3];
  a.mez_i = it.fwd[par][4];
  a.ey_o = it.fwd[par ^ 1][0];
  a.hx_o = it.fwd[par ^ 1][1];
  a.hz_o = it.fwd[par ^ 1][2];
  a.mex_o = it.fwd[par ^ 1][3];
  a.mez_o = it.fwd[par ^ 1][4];
  T *const e0[2] = {it.mhzx, it.mhzx_b}, *const e1[2] = {it.mhxz, it.mhxz_b};
  a.mhzx_i = e0[par];
  a.mhxz_i = e1[par];
  a.mhzx_o = e0[par ^ 1];
  a.mhxz_o = e1[par ^ 1];
  const int64_t fs = (int64_t)it.n_shots * it.n_src, rs = (int64_t)it.n_shots * it.n_rec;
  a.f0 = it.n_src > 0 ? it.f + k * fs : nullptr;
  a.f1 = it.n_src > 0 ? it.f + (k + 1) * fs : nullptr;
  a.r0 = it.n_rec > 0 ? it.r + k * rs : nullptr;
  a.r1 = it.n_rec > 0 ? it.r + (k + 1) * rs : nullptr;
  a.ey_store0 = item_snap(it, it.ey_store, k);
  a.curl_store0 = item_snap(it, it.curl_store, k);
  a.ey_store1 = item_snap(it, it.ey_store, k + 1);
  a.curl_store1 = item_snap(it, it.curl_store, k + 1);
  a.n_shots = it.n_shots;
  a.n_yt = it.n_yt2;
  a.split = it.split2;
  return a;
}

// Adjoint iteration k applies the transpose of forward step t = nt-1-k,
// reading the adjoint state of parity `par` and writing par^1.
template <typename T>
__device__ __forceinline__ AdjArgs<T> item_adj_args(const BatchItem<T> &it, int k, int par) {
  const int p = par;
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

template <typename T>
__device__ __forceinline__ AdjArgs<T> item_adj_args(const BatchItem<T> &it, int k) {
  return item_adj_args(it, k, k & 1);
}

// Adjoint iterations k and k+1 (transposes of steps t = nt-1-k and t-1) in
// one launch, reading parity `par` and writing par^1.
template <typename T>
__device__ __forceinline__ Adj2Args<T> item_adj2_args(const BatchItem<T> &it, int k, int par) {
  const int t = it.nt - 1 - k;
  Adj2Args<T> a;
  a.g = it.g;
  a.pts = it.pts2;
  a.ca = it.ca;
  a.cb = it.cb;
  a.cq = it.cq;
  a.ley_i = it.lam[par][0];
  a.lhx_i = it.lam[par][1];
  a.lhz_i = it.lam[par][2];
  a.mx_i = it.lam[par][3];
  a.my_i = it.lam[par][4];
  a.nx_i = it.lam[par][5];
  a.nz_i = it.lam[par][6];
  a.ley_o = it.lam[par ^ 1][0];
  a.lhx_o = it.lam[par ^ 1][1];
  a.lhz_o = it.lam[par ^ 1][2];
  a.mx_o = it.lam[par ^ 1][3];
  a.my_o = it.lam[par ^ 1][4];
  a.nx_o = it.lam[par ^ 1][5];
  a.nz_o = it.lam[par ^ 1][6];
  const void *const es = it.gca ? it.ey_store : nullptr;
  const void *const cs = it.gcb ? it.curl_store : nullptr;
  a.ey_store0 = item_snap(it, es, t);
  a.curl_store0 = item_snap(it, cs, t);
  a.ey_store1 = item_snap(it, es, t - 1);
  a.curl_store1 = item_snap(it, cs, t - 1);
  a.gca = it.gca;
  a.gcb = it.gcb;
  a.per_shot_grad = 0;
  a.step_scale = (double)it.interval;
  const int64_t fs = (int64_t)it.n_shots * it.n_src, rs = (int64_t)it.n_shots * it.n_rec;
  a.grad_r0 = it.n_rec > 0 ? it.grad_r + t * rs : nullptr;
  a.grad_r1 = it.n_rec > 0 ? it.grad_r + (t - 1) * rs : nullptr;
  a.grad_f0 = (it.n_src > 0 && it.grad_f) ? it.grad_f + t * fs : nullptr;
  a.grad_f1 = (it.n_src > 0 && it.grad_f) ? it.grad_f + (t - 1) * fs : nullptr;
  a.n_shots = it.n_shots;
  a.n_yt = it.n_yt2;
  a.group_stride = it.group_stride;
  a.split = it.split2;
  return a;
}

}  // namespace tm2d
}  // namespace tide_hip
