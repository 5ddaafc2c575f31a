// Device code of the 2-D TM step kernels (tile functions and the
// single-problem kernels), shared by the single-problem driver (tm2d.hip)
// and the batched plan (tm2d_batch.hip).  See tm2d.hip for the design.
#pragma once

#include "tide_common.h"

#include <type_traits>

namespace tide_hip {
namespace tm2d {

constexpr int kOwn = 62;  // columns owned by a wavefront (64 lanes - 2 halo)


// CPML profiles packed per axis: py = [a, b, k, ah, bh, kh] x ny and
// px = [a, b, k, ah, bh, kh] x nx.
enum { kA = 0, kB = 1, kK = 2, kAh = 3, kBh = 4, kKh = 5 };

template <typename T> struct Geom {
  const T *py, *px;
  T rdy, rdx;
  int32_t ny, nx, numel;
  int32_t pml_y0, pml_y1, pml_x0, pml_x1, pml_y1h, pml_x1h;
  int32_t ca_b, cb_b, cq_b;  // coefficient is per shot
  __device__ __forceinline__ T Y(int k, int y) const { return py[k * ny + y]; }
  __device__ __forceinline__ T X(int k, int x) const { return px[k * nx + x]; }
};

struct Points {
  const int32_t *row_slot;
  const int32_t *seg;
  const int2 *ent;
  int32_t n_xt;
};

template <typename T> struct FwdArgs {
  Geom<T> g;
  Points pts;
  const T *ca, *cb, *cq;
  const T *ey_i, *hx_i, *hz_i;
  T *ey_o, *hx_o, *hz_o;
  const T *mex_i, *mez_i;  // m_Ey_x, m_Ey_z: H-phase memories (ping-pong)
  T *mex_o, *mez_o;
  T *mhzx, *mhxz;          // m_Hz_x, m_Hx_z: E-phase memories (in place)
  void *ey_store, *curl_store;
  const T *f_t;
  T *r_t;
  int32_t n_shots, n_yt;
};

template <typename T> struct AdjArgs {
  Geom<T> g;
  Points pts;
  const T *ca, *cb, *cq;
  const T *ley_i, *lhx_i, *lhz_i;
  T *ley_o, *lhx_o, *lhz_o;
  // adjoint CPML memories: mx = m_lambda_hz_x, my = m_lambda_hx_z,
  // nx = m_lambda_ey_x, nz = m_lambda_ey_z (all ping-pong)
  const T *mx_i, *my_i, *nx_i, *nz_i;
  T *mx_o, *my_o, *nx_o, *nz_o;
  const void *ey_store, *curl_store;
  double *gca, *gcb;
  int32_t per_shot_grad;  // gradient arrays are [n_shots, ny, nx]
  double step_scale;
  const T *grad_r_t;
  T *grad_f_t;
  int32_t n_shots, n_yt;
  // With several shot groups, group g adds its sums to gca/gcb + g*group_stride
  // (separate FP64 slices, summed in group order after the time loop), so
  // no two workgroups update the same cell: no atomics, bitwise repeatable.
  int64_t group_stride;
};

// Snapshots are written once and read once, much later: non-temporal stores
// keep them from evicting the fields, which the next step reads again.
template <typename T, int FMT>
__device__ __forceinline__ void put_store(void *base, int64_t i, T v) {
  if constexpr (FMT == 2) {
    __builtin_nontemporal_store(f32_to_bf16((float)v), (uint16_t *)base + i);
  } else {
    __builtin_nontemporal_store(v, (T *)base + i);
  }
}

// The adjoint reads each snapshot once: non-temporal loads keep the adjoint
// fields cached.
template <typename T, int FMT>
__device__ __forceinline__ double get_store(const void *base, int64_t i) {
  if constexpr (FMT == 2) {
    return (double)bf16_to_f32(__builtin_nontemporal_load((const uint16_t *)base + i));
  } else {
    return (double)__builtin_nontemporal_load((const T *)base + i);
  }
}

struct TileCoord {
  int xt, yt, grp;
};

__device__ __forceinline__ TileCoord decode_tile(int n_xt, int n_yt, int n_groups) {
  const int64_t n_tiles = (int64_t)n_xt * n_yt;
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, n_tiles * n_groups);
  TileCoord c;
  c.grp = (int)(flat / n_tiles);
  const int64_t tile = flat - (int64_t)c.grp * n_tiles;
  c.yt = (int)(tile / n_xt);
  c.xt = (int)(tile - (int64_t)c.yt * n_xt);
  return c;
}

// The six y profiles (kA .. kKh) of up to 64 consecutive rows, held across
// the lanes of a wavefront (lane k: row y_first + k, clamped into the grid).
// A tile loads them with one vector load per profile up front and reads a
// row's value back with a cross-lane read, instead of a dependent scalar load
// (and wait) at every use.
template <typename T> struct LaneRows {
  T v[6];
  int y_first;
  __device__ __forceinline__ void load(const Geom<T> &g, int y_first_, int lane) {
    y_first = y_first_;
    const int yc = min(max(y_first_ + lane, 0), g.ny - 1);
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = g.Y(k, yc);
  }
  __device__ __forceinline__ T operator()(int k, int y) const { return lane_read(v[k], y - y_first); }
  // Entries (pa, pb, pk) = (ka, kb, kk) of row y, or (0, 0, 1) when `on`
  // (wave-uniform) is false.  These are cross-lane reads: call this where
  // every lane of the wavefront is active.  In a lane-divergent branch the
  // lane read may be inactive, and a register that the compiler spills and
  // reloads there keeps its active lanes only.
  __device__ __forceinline__ void row3(int ka, int kb, int kk, int y, bool on, T &pa, T &pb,
                                       T &pk) const {
    pa = (T)0;
    pb = (T)0;
    pk = (T)1;
    if (on) {
      pa = (*this)(ka, y);
      pb = (*this)(kb, y);
      pk = (*this)(kk, y);
    }
  }
};

// Row loader for one wavefront.  Interior tiles (EDGE false) need no bounds
// checks: every row and column they touch exists.  Edge tiles read a
// clamped column and zero it outside the grid.
template <typename T, bool EDGE> struct RowLoader {
  int x, xc, nx, ny;
  bool xv;
  __device__ __forceinline__ RowLoader(int x_, int nx_, int ny_) : x(x_), nx(nx_), ny(ny_) {
    xv = x_ >= 0 && x_ < nx_;
    xc = min(max(x_, 0), nx_ - 1);
  }
  __device__ __forceinline__ T operator()(const T *__restrict__ p, int y) const {
    if constexpr (EDGE) {
      // Branch-free (clamped row and column, then a select): a branch around
      // the load would make the wavefront wait for it before the next load.
      const int yc = min(max(y, 0), ny - 1);
      const T v = *elem32(p, (uint32_t)(yc * nx + xc));
      return keep_or_zero(v, xv && y == yc);
    } else {
      return *elem32(p, (uint32_t)(y * nx + x));
    }
  }
};

// ----------------------------------------------------------------- forward
// FMT: 0 = no snapshot, 1 = T payload, 2 = BF16 payload.
template <typename T, int TY, int FMT, bool EDGE>
__device__ __forceinline__ void fwd_tile(const FwdArgs<T> &a, int s, int xt, int yt) {
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * kOwn - 1 + lane;
  const int y0 = yt * TY;
  const bool own = lane >= 1 && lane <= kOwn && x < nx;
  // Update domains for stencil 2 (FD_PAD = 1); all true inside interior tiles.
  const bool e_col = !EDGE || (x >= 1 && x < nx);       // Ey and Hx columns
  const bool hz_col = !EDGE || (x >= 1 && x < nx - 1);  // Hz columns
  const int x0 = xt * kOwn - 1;
  const bool tile_pml_x = x0 < g.pml_x0 || x0 + 63 >= g.pml_x1h;
  const bool tile_pml_y = y0 - 1 < g.pml_y0 || y0 + TY - 1 >= g.pml_y1h;
  const bool pml_xe = tile_pml_x && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = tile_pml_x && (x < g.pml_x0 || x >= g.pml_x1h);

  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ey_i = a.ey_i + base;
  const T *__restrict__ hx_i = a.hx_i + base;
  const T *__restrict__ hz_i = a.hz_i + base;
  const T *__restrict__ mex_i = a.mex_i + base;
  const T *__restrict__ mez_i = a.mez_i + base;
  T *__restrict__ ey_o = a.ey_o + base;
  T *__restrict__ hx_o = a.hx_o + base;
  T *__restrict__ hz_o = a.hz_o + base;
  T *__restrict__ mex_o = a.mex_o + base;
  T *__restrict__ mez_o = a.mez_o + base;
  T *__restrict__ mhzx = a.mhzx + base;
  T *__restrict__ mhxz = a.mhxz + base;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, EDGE> L(x, nx, ny);

  // ---- loads: Ey rows y0-1..y0+TY, Hx and cq rows y0-1..y0+TY-1, the rest
  // rows y0..y0+TY-1.
  T ey[TY + 2], hx[TY + 1], cqv[TY + 1], hz[TY], cav[TY], cbv[TY];
#pragma unroll
  for (int k = 0; k < TY + 2; ++k) ey[k] = L(ey_i, y0 - 1 + k);
#pragma unroll
  for (int k = 0; k < TY + 1; ++k) {
    hx[k] = L(hx_i, y0 - 1 + k);
    cqv[k] = L(cq, y0 - 1 + k);
  }
#pragma unroll
  for (int k = 0; k < TY; ++k) {
    hz[k] = L(hz_i, y0 + k);
    cav[k] = L(ca, y0 + k);
    cbv[k] = L(cb, y0 + k);
  }

  // Hx^{n+1/2}(y0-1, x): recomputed here, written by the tile above.
  T hx_up = hx[0];
  {
    const int y = y0 - 1;
    if (e_col && (!EDGE || (y >= 1 && y < ny - 1))) {
      T d = (ey[1] - ey[0]) * g.rdy;
      if (tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1h)) {
        const T m = g.Y(kBh, y) * mez_i[y * nx + x] + g.Y(kAh, y) * d;
        d = d / g.Y(kKh, y) + m;
      }
      hx_up = hx_up - cqv[0] * d;
    }
  }

#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int y = y0 + r;
    if (EDGE && y >= ny) break;  // uniform
    const int j = y * nx + x;
    const int64_t i = base + j;
    const bool e_row = !EDGE || y >= 1;
    const bool hx_row = !EDGE || (y >= 1 && y < ny - 1);
    const bool pml_ye = tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1);
    const bool pml_yh = tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1h);
    const T ey_c = ey[r + 1];
    const T ey_r = shfl_dn1(ey_c);  // Ey(y, x+1); unused at lane 63

    // ---- H half-step (every lane: lanes 0 and 63 are halo copies).
    T hx_n = hx[r + 1];
    if (e_col && hx_row) {
      T d = (ey[r + 2] - ey_c) * g.rdy;
      if (pml_yh) {
        const T m = g.Y(kBh, y) * mez_i[j] + g.Y(kAh, y) * d;
        if (own) mez_o[j] = m;
        d = d / g.Y(kKh, y) + m;
      }
      hx_n = hx_n - cqv[r + 1] * d;
    }
    T hz_n = hz[r];
    if (hz_col && e_row) {
      T d = (ey_r - ey_c) * g.rdx;
      if (pml_xh) {
        const T m = g.X(kBh, x) * mex_i[j] + g.X(kAh, x) * d;
        if (own) mex_o[j] = m;
        d = d / g.X(kKh, x) + m;
      }
      hz_n = hz_n + cqv[r + 1] * d;
    }
    const T hz_l = shfl_up1(hz_n);  // Hz(y, x-1); unused at lane 0

    // ---- E step for owned cells.
    T ey_n = ey_c;
    if (own && e_col && e_row) {
      T dhz = (hz_n - hz_l) * g.rdx;
      T dhx = (hx_n - hx_up) * g.rdy;
      if (pml_xe) {
        const T m = g.X(kB, x) * mhzx[j] + g.X(kA, x) * dhz;
        mhzx[j] = m;
        dhz = dhz / g.X(kK, x) + m;
      }
      if (pml_ye) {
        const T m = g.Y(kB, y) * mhxz[j] + g.Y(kA, y) * dhx;
        mhxz[j] = m;
        dhx = dhx / g.Y(kK, y) + m;
      }
      const T curl = dhz - dhx;
      if constexpr (FMT != 0) {
        if (a.ey_store) put_store<T, FMT>(a.ey_store, i, ey_c);
        if (a.curl_store) put_store<T, FMT>(a.curl_store, i, curl);
      }
      ey_n = cav[r] * ey_c + cbv[r] * curl;
    }

    // ---- sources, then receivers, of this shot in this row segment.
    const int slot = a.pts.row_slot[y];
    if (slot >= 0) {
      const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + xt;
      const int e1 = a.pts.seg[si + 1];
      for (int e = a.pts.seg[si]; e < e1; ++e) {
        const int2 en = a.pts.ent[e];
        if ((en.x & 0xff) == lane) {
          if ((en.x >> 8) == 0)
            ey_n += a.f_t[en.y];
          else
            a.r_t[en.y] = ey_n;
        }
      }
    }

    if (own) {
      hx_o[j] = hx_n;
      hz_o[j] = hz_n;
      ey_o[j] = ey_n;
    }
    hx_up = hx_n;
  }
}

template <typename T, int TY>
__device__ __forceinline__ bool interior_tile(const Geom<T> &g, int xt, int yt) {
  const int x0 = xt * kOwn - 1, y0 = yt * TY;
  return x0 >= 1 && x0 + 63 <= g.nx - 2 && y0 >= 2 && y0 + TY + 1 <= g.ny - 1;
}

// One forward workgroup: wavefront threadIdx.y handles shot grp*G + y.
template <typename T, int TY, int FMT>
__device__ __forceinline__ void fwd_workgroup(const FwdArgs<T> &a, const TileCoord &tc) {
  const int G = (int)blockDim.y;
  const int s = tc.grp * G + (int)threadIdx.y;
  if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
  if (interior_tile<T, TY>(a.g, tc.xt, tc.yt))
    fwd_tile<T, TY, FMT, false>(a, s, tc.xt, tc.yt);
  else
    fwd_tile<T, TY, FMT, true>(a, s, tc.xt, tc.yt);
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void fwd_step(FwdArgs<T> a) {
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  fwd_workgroup<T, TY, FMT>(a, decode_tile(a.pts.n_xt, a.n_yt, n_groups));
}

// ----------------------------------------------------------------- adjoint
// One reverse iteration: transpose of the E step, transpose of the H
// half-step, receiver injection, source read-out, coefficient gradient.
// SINK 0: the gradient terms of shot s go to the LDS slots red_*[r*64+lane]
// (summed over the workgroup's shots afterwards); SINK 1: they are added to
// the per-thread accumulators red_*[r] (the caller walks the shots).
template <typename T, int TY, int FMT, bool EDGE, int SINK = 0>
__device__ __forceinline__ void adj_tile(const AdjArgs<T> &a, int s, int xt, int yt,
                                         double *red_ca, double *red_cb, bool reduce) {
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * kOwn - 1 + lane;
  const int y0 = yt * TY;
  const bool own = lane >= 1 && lane <= kOwn && x < nx;
  const bool e_col = !EDGE || (x >= 1 && x < nx);
  const int x0 = xt * kOwn - 1;
  const bool tile_pml_x = x0 - 1 < g.pml_x0 || x0 + 64 >= g.pml_x1h;
  const bool tile_pml_y = y0 - 1 < g.pml_y0 || y0 + TY >= g.pml_y1h;
  const bool pml_xe = tile_pml_x && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = tile_pml_x && (x < g.pml_x0 || x >= g.pml_x1h);
  const bool want_grad = (a.gca != nullptr || a.gcb != nullptr) && FMT != 0;

  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ley_i = a.ley_i + base;
  const T *__restrict__ lhx_i = a.lhx_i + base;
  const T *__restrict__ lhz_i = a.lhz_i + base;
  const T *__restrict__ mx_i = a.mx_i + base;
  const T *__restrict__ my_i = a.my_i + base;
  const T *__restrict__ nx_i = a.nx_i + base;
  const T *__restrict__ nz_i = a.nz_i + base;
  T *__restrict__ ley_o = a.ley_o + base;
  T *__restrict__ lhx_o = a.lhx_o + base;
  T *__restrict__ lhz_o = a.lhz_o + base;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, EDGE> L(x, nx, ny);

  // ---- loads.  lam/cbv rows y0-1..y0+TY; lhx/cqv rows y0-1..y0+TY-1; the
  // rest rows y0..y0+TY-1.
  T lam[TY + 2], cbv[TY + 2], lhx[TY + 1], cqv[TY + 1], lhz[TY], cav[TY];
  float se[TY], sc[TY];  // snapshot values (exact in fp32 for fp32/bf16 payloads)
#pragma unroll
  for (int k = 0; k < TY + 2; ++k) {
    lam[k] = L(ley_i, y0 - 1 + k);
    cbv[k] = L(cb, y0 - 1 + k);
  }
#pragma unroll
  for (int k = 0; k < TY + 1; ++k) {
    lhx[k] = L(lhx_i, y0 - 1 + k);
    cqv[k] = L(cq, y0 - 1 + k);
  }
#pragma unroll
  for (int k = 0; k < TY; ++k) {
    const int y = y0 + k;
    lhz[k] = L(lhz_i, y);
    cav[k] = L(ca, y);
    se[k] = 0.0f;
    sc[k] = 0.0f;
    if constexpr (sizeof(T) == 4) {
      if (want_grad && own && (!EDGE || (y >= 1 && y < ny && x >= 1))) {
        const int64_t i = base + y * nx + x;
        if (a.gca) se[k] = (float)get_store<T, FMT>(a.ey_store, i);
        if (a.gcb) sc[k] = (float)get_store<T, FMT>(a.curl_store, i);
      }
    }
  }

  // Adjoint of dHz/dx at (y, x) from lambda_Ey and cb there.
  auto raw_x = [&](int y, T lm, T cbx) -> T {
    if (EDGE && (y < 1 || y >= ny || x < 1 || x >= nx)) return (T)0;
    const T q = cbx * lm;
    if (!pml_xe) return q;
    const T mn = g.X(kB, x) * mx_i[y * nx + x] + q;
    return g.X(kA, x) * mn + q / g.X(kK, x);
  };
  // Negated adjoint of dHx/dz at (y, x): q off the strips, q/k - a*mn in
  // them.  The reference's term is -q resp. a*mn - q/k; IEEE subtraction
  // is sign-symmetric, so carrying the negation and swapping the operands
  // of the differences below gives bit-identical results.  (Returning -q
  // from a branch of this lambda was miscompiled by the gfx950 backend of
  // ROCm 7.2: inside tiles that touch a y strip, the interior-row path
  // skipped the negation.)
  auto nraw_y = [&](int y, T lm, T cbx) -> T {
    if (EDGE && (y < 1 || y >= ny || !e_col)) return (T)0;
    const T q = cbx * lm;
    if (!(tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1))) return q;
    const T mn = g.Y(kB, y) * my_i[y * nx + x] - q;
    return q / g.Y(kK, y) - g.Y(kA, y) * mn;
  };
  // Adjoint of the Hx update's dEy/dz at half row y.
  auto raw_z = [&](int y, T lhxv, T cqx) -> T {
    if (EDGE && (y < 1 || y >= ny - 1 || !e_col)) return (T)0;
    const T adj = -cqx * lhxv;
    if (!(tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1h))) return adj;
    const T mn = g.Y(kBh, y) * nz_i[y * nx + x] + adj;
    return g.Y(kAh, y) * mn + adj / g.Y(kKh, y);
  };
  // Adjoint of the Hz update's dEy/dx at half column x.
  auto raw_xh = [&](int y, T lhzv, T cqx) -> T {
    if (EDGE && (y < 1 || y >= ny || x < 1 || x >= nx - 1)) return (T)0;
    const T adj = cqx * lhzv;
    if (!pml_xh) return adj;
    const T mn = g.X(kBh, x) * nx_i[y * nx + x] + adj;
    return g.X(kAh, x) * mn + adj / g.X(kKh, x);
  };

  T nrawy_c = nraw_y(y0, lam[1], cbv[1]);
  T rawz_up = (T)0;
  if (!EDGE || y0 >= 1) {
    const int y = y0 - 1;
    T v = lhx[0];
    if (e_col && (!EDGE || y >= 1)) v = v + (nrawy_c - nraw_y(y, lam[0], cbv[0])) * g.rdy;
    rawz_up = raw_z(y, v, cqv[0]);
  }

#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int y = y0 + r;
    if (EDGE && y >= ny) break;
    const int j = y * nx + x;
    const bool in_e = e_col && (!EDGE || y >= 1);
    const T ley_c = lam[r + 1];

    // ---- transpose of the E step.
    const T rawx_c = raw_x(y, ley_c, cbv[r + 1]);
    const T rawx_r = shfl_dn1(rawx_c);  // unused at lane 63
    T lhz_n = lhz[r];
    if (in_e) lhz_n = lhz_n + (rawx_c - rawx_r) * g.rdx;
    const T nrawy_dn = nraw_y(y + 1, lam[r + 2], cbv[r + 2]);
    T lhx_n = lhx[r + 1];
    if (in_e) lhx_n = lhx_n + (nrawy_dn - nrawy_c) * g.rdy;
    if (own && in_e) {
      const T q = cbv[r + 1] * ley_c;
      if (pml_xe) a.mx_o[base + j] = g.X(kB, x) * mx_i[j] + q;
      if (tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1))
        a.my_o[base + j] = g.Y(kB, y) * my_i[j] - q;
    }

    // ---- transpose of the H half-step.
    const T rawz_c = raw_z(y, lhx_n, cqv[r + 1]);
    const T rawxh_c = raw_xh(y, lhz_n, cqv[r + 1]);
    const T rawxh_l = shfl_up1(rawxh_c);  // unused at lane 0
    T ley_n = ley_c;
    if (own && in_e) {
      const T cqx = cqv[r + 1];
      const T adj_y = -cqx * lhx_n;
      const T adj_x = cqx * lhz_n;
      if ((!EDGE || y < ny - 1) && tile_pml_y && (y < g.pml_y0 || y >= g.pml_y1h))
        a.nz_o[base + j] = g.Y(kBh, y) * nz_i[j] + adj_y;
      if ((!EDGE || x < nx - 1) && pml_xh) a.nx_o[base + j] = g.X(kBh, x) * nx_i[j] + adj_x;
      const T t1 = cav[r] * ley_c;
      const T t2 = (rawz_up - rawz_c) * g.rdy + (rawxh_l - rawxh_c) * g.rdx;
      ley_n = t1 + t2;
    }

    // ---- receivers inject grad_r[t]; sources read grad_f[t].
    const int slot = a.pts.row_slot[y];
    if (slot >= 0) {
      const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + xt;
      const int e0 = a.pts.seg[si], e1 = a.pts.seg[si + 1];
      for (int e = e0; e < e1; ++e) {
        const int2 en = a.pts.ent[e];
        if ((en.x >> 8) == 1 && (en.x & 0xff) == lane) ley_n += a.grad_r_t[en.y];
      }
      if (a.grad_f_t != nullptr) {
        for (int e = e0; e < e1; ++e) {
          const int2 en = a.pts.ent[e];
          if ((en.x >> 8) == 0 && (en.x & 0xff) == lane) a.grad_f_t[en.y] = ley_n;
        }
      }
    }

    // ---- coefficient gradients against the stored forward state.
    if (want_grad && own && in_e) {
      const double lm = (double)ley_n;
      double ve, vc;
      if constexpr (sizeof(T) == 4) {
        ve = (double)se[r];
        vc = (double)sc[r];
      } else {
        ve = a.gca ? get_store<T, FMT>(a.ey_store, base + j) : 0.0;
        vc = a.gcb ? get_store<T, FMT>(a.curl_store, base + j) : 0.0;
      }
      const double pa = lm * ve * a.step_scale;
      const double pb = lm * vc * a.step_scale;
      if (reduce) {
        if constexpr (SINK == 0) {
          red_ca[r * 64 + lane] = pa;
          red_cb[r * 64 + lane] = pb;
        } else {
          red_ca[r] += pa;
          red_cb[r] += pb;
        }
      } else {
        const int64_t gi = (a.per_shot_grad ? base : 0) + j;
        if (a.gca) a.gca[gi] += pa;
        if (a.gcb) a.gcb[gi] += pb;
      }
    }

    if (own) {
      ley_o[j] = ley_n;
      lhx_o[j] = lhx_n;
      lhz_o[j] = lhz_n;
    }
    nrawy_c = nrawy_dn;
    rawz_up = rawz_c;
  }
}

// The per-shot FP64 gradient terms of a tile are summed across its
// wavefronts (shots) through LDS in a fixed order, bitwise reproducibly,
// and added once per cell.
// One adjoint workgroup; `red` is the dynamic LDS block [2][G][TY][64].
template <typename T, int TY, int FMT>
__device__ __forceinline__ void adj_workgroup(const AdjArgs<T> &a, const TileCoord &tc,
                                              double *red) {
  const int lane = threadIdx.x;
  const int w = (int)threadIdx.y;
  const int G = (int)blockDim.y;
  const int s = tc.grp * G + w;
  const bool want_grad = (a.gca != nullptr || a.gcb != nullptr) && FMT != 0;
  const bool reduce = want_grad && !a.per_shot_grad;
  double *red_ca = red + (int64_t)w * TY * 64;
  double *red_cb = red + ((int64_t)G + w) * TY * 64;
  if (reduce) {
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      red_ca[r * 64 + lane] = 0.0;
      red_cb[r * 64 + lane] = 0.0;
    }
  }
  if (s < a.n_shots) {  // inactive wavefronts still join the barrier below
    if (interior_tile<T, TY>(a.g, tc.xt, tc.yt))
      adj_tile<T, TY, FMT, false>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    else
      adj_tile<T, TY, FMT, true>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
  }
  if (!reduce) return;  // uniform over the workgroup
  __syncthreads();
  const int x = tc.xt * kOwn - 1 + lane;
  const int y0 = tc.yt * TY;
  if (lane < 1 || lane > kOwn || x < 1 || x >= a.g.nx) return;
  for (int pass = 0; pass < 2; ++pass) {
    double *out = pass == 0 ? a.gca : a.gcb;
    if (out == nullptr) continue;
    out += tc.grp * a.group_stride;
    const double *slots = red + (int64_t)pass * G * TY * 64;
    for (int r = w; r < TY; r += G) {
      const int y = y0 + r;
      if (y >= 1 && y < a.g.ny) {
        double sum = 0.0;
        for (int k = 0; k < G; ++k) sum += slots[(k * TY + r) * 64 + lane];
        const int64_t j = (int64_t)y * a.g.nx + x;
        out[j] += sum;
      }
    }
  }
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void adj_step(AdjArgs<T> a) {
  extern __shared__ double red[];
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  adj_workgroup<T, TY, FMT>(a, decode_tile(a.pts.n_xt, a.n_yt, n_groups), red);
}

// ------------------------------------------------------ shot-looping kernels
// One wavefront per tile walking every shot of the call: workgroup b takes
// tiles b*W .. b*W+W-1 (W = blockDim.y), no LDS and no barrier.  The
// adjoint sums each cell's FP64 gradient terms over the shots in registers
// (shot order) and adds them once per cell and step.
__device__ __forceinline__ int loop_tile(int n_tiles) {
  const int64_t blk = xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  const int64_t tile = blk * blockDim.y + threadIdx.y;
  return tile < n_tiles ? (int)tile : -1;
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(256) void fwd_loop(FwdArgs<T> a) {
  const int n_xt = a.pts.n_xt;
  const int tile = loop_tile(n_xt * a.n_yt);
  if (tile < 0) return;
  const int yt = tile / n_xt, xt = tile - yt * n_xt;
  const bool interior = interior_tile<T, TY>(a.g, xt, yt);
  for (int s = 0; s < a.n_shots; ++s) {
    if (interior)
      fwd_tile<T, TY, FMT, false>(a, s, xt, yt);
    else
      fwd_tile<T, TY, FMT, true>(a, s, xt, yt);
  }
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(256) void adj_loop(AdjArgs<T> a) {
  const int n_xt = a.pts.n_xt;
  const int tile = loop_tile(n_xt * a.n_yt);
  if (tile < 0) return;
  const int yt = tile / n_xt, xt = tile - yt * n_xt;
  const bool want_grad = (a.gca != nullptr || a.gcb != nullptr) && FMT != 0;
  const bool reduce = want_grad && !a.per_shot_grad;
  double acc_ca[TY], acc_cb[TY];
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    acc_ca[r] = 0.0;
    acc_cb[r] = 0.0;
  }
  const bool interior = interior_tile<T, TY>(a.g, xt, yt);
  for (int s = 0; s < a.n_shots; ++s) {
    if (interior)
      adj_tile<T, TY, FMT, false, 1>(a, s, xt, yt, acc_ca, acc_cb, reduce);
    else
      adj_tile<T, TY, FMT, true, 1>(a, s, xt, yt, acc_ca, acc_cb, reduce);
  }
  if (!reduce) return;
  const int lane = threadIdx.x;
  const int x = xt * kOwn - 1 + lane;
  if (lane < 1 || lane > kOwn || x < 1 || x >= a.g.nx) return;
  const int y0 = yt * TY;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int y = y0 + r;
    if (y >= 1 && y < a.g.ny) {
      const int64_t j = (int64_t)y * a.g.nx + x;
      if (a.gca) a.gca[j] += acc_ca[r];
      if (a.gcb) a.gcb[j] += acc_cb[r];
    }
  }
}

// out[i] += sum over g of acc[g*n + i], summed in group order.
static __global__ void fold_groups(double *__restrict__ out, const double *__restrict__ acc, int64_t n,
                            int groups) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sum = acc[i];
  for (int g = 1; g < groups; ++g) sum += acc[(int64_t)g * n + i];
  out[i] += sum;
}

// Adds an FP64 accumulator into the caller's gradient array (reference
// combine_grad semantics: grad += sum over shots).
template <typename T>
__global__ void fold_grad(T *__restrict__ out, const double *__restrict__ acc, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)((double)out[i] + acc[i]);
}

constexpr int kMaxGroup = 8;  // shots (wavefronts) per workgroup

// Launch shape: rows marched per wavefront and shots (wavefronts) per
// workgroup for each kernel.  Process-wide (defined in tm2d.hip); set by the
// autotuner through tide_tm2d_set_tuning().  fp64 fields use half the rows
// (twice the registers per value).
// fwd_loop / adj_loop select the shot-looping kernels; their group is then
// the number of tiles (wavefronts) per workgroup.
// fwd_steps / adj_steps (1 to 4) are the time steps per launch: 2 selects
// the two-step kernels (tm2d_fused.h) for pairs of steps, 3 and 4 the K-step
// kernels (tm2d_blockk.h); fwd_split / adj_split run the two-step kernels'
// interior and border tiles as two kernels.  On a 1065^2 grid with 4 or 7
// shots per call the forward is fastest at 2 steps per launch (20.8 us per
// step against 22+ at 3 or 4) and the adjoint at 3 (32.2 against 33.6 at
// 2 and 40.5 at 4, at the compiler's register count).
struct Tuning {
  int fwd_rows = 4, fwd_group = 1, adj_rows = 4, adj_group = 4;
  int fwd_loop = 0, adj_loop = 0;
  int fwd_steps = 2, adj_steps = 3;
  int fwd_split = 0, adj_split = 0;
  // Border tile height of the two-step kernels: 1, 2 or 0 (the interior
  // height).  The forward is fastest with interior-height border tiles, the
  // adjoint (more live state per border tile) with 2-row ones.
  int fwd_border_rows = 0, adj_border_rows = 2;
  // fp32 K-step adjoint held to four waves per SIMD (128 VGPRs, about 19
  // spilled dwords per lane at K = 3) when 4, the compiler's own register
  // count (148 VGPRs, three waves) when 0.  Four waves hide more latency:
  // 29.3 against 31.7 us per step on a 1065^2 grid with 4 shots, 54.7
  // against 58.5 with 7; five waves (96 VGPRs) spill too much: 44.5.  fp64
  // kernels (229+ VGPRs) always use the compiler's count.
  int adj_waves = 4;
};
extern Tuning g_tuning;

template <typename T> inline int rows_for(int rows) {
  if (sizeof(T) == 8) rows = rows / 2 > 2 ? rows / 2 : 2;
  return rows;
}

// Shots per adjoint workgroup: the fewest groups of at most `max_group`
// shots, balanced (6 shots at 4 per group run as 3 + 3, not 4 + 2).  On a
// 1065^2 grid the 3 + 3 split runs the two-step adjoint 12% faster than
// 4 + 2 and than one group of 6; 7 shots stay 4 + 3.
inline int balanced_group(int n_shots, int max_group) {
  const int g = std::max(1, std::min(n_shots, max_group));
  const int n_groups = (n_shots + g - 1) / g;
  return (n_shots + n_groups - 1) / n_groups;
}

}  // namespace tm2d
}  // namespace tide_hip
