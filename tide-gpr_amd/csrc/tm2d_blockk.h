// K-step (temporally blocked) 2-D TM forward tiles, K = 3 or 4.
//
// The two-step tiles of tm2d_fused.h generalised to K steps per launch: a
// wavefront owns 64 - 2K columns (lanes K .. 63-K) of TY rows of one shot.
// It loads the step-t state of its rows and columns plus a halo of K, and
// step j (1..K) updates the cells whose dependence cone still lies inside
// what it holds: the row range shrinks by one row per side and the valid
// lanes by one per side each step, so after K steps exactly the owned cells
// are current.  HBM sees the wavefield state once per K steps (24/K bytes
// per cell-update) instead of once per two; snapshots are written for every
// step as before.  Every cell update is the single-step expression in the
// same operation order, so K-step results are bit-identical to K single
// steps.  The register arrays are updated in place, one step after the
// other (H rows first, then Hz and Ey row by row), with every index a
// compile-time constant after unrolling.
#pragma once

#include "tm2d_fused.h"

namespace tide_hip {
namespace tm2d {

constexpr int kMaxBlockSteps = 4;

template <typename T> struct FwdKArgs {
  Geom<T> g;
  Points pts;  // built with build_point_table_halo(tile_w 64-2K, lane offset K)
  const T *ca, *cb, *cq;
  const T *ey_i, *hx_i, *hz_i;
  T *ey_o, *hx_o, *hz_o;
  const T *mex_i, *mez_i, *mhzx_i, *mhxz_i;  // m_ey_x m_ey_z m_hz_x m_hx_z, step t
  T *mex_o, *mez_o, *mhzx_o, *mhxz_o;        // step t+K
  const T *f[kMaxBlockSteps];                // source terms of steps t .. t+K-1
  T *r[kMaxBlockSteps];                      // receiver rows of steps t .. t+K-1
  void *ey_store[kMaxBlockSteps], *curl_store[kMaxBlockSteps];  // null: none
  int32_t n_shots, n_yt;
  TileSplit split;
};

// Interior rectangle of K-step tiles (tile width 64 - 2K, halo K): a tile is
// interior when its columns x0-K .. x0+63-K and rows y0-K .. y0+TY+K-1 lie
// in [1, n-2] and outside the CPML strips (H-phase bounds).
template <typename T>
inline TileSplit make_split_k(const Geom<T> &g, int n_xt, int n_yt, int ty, int k) {
  const int own = 64 - 2 * k;
  auto col_ok = [&](int xt) {
    const int x0 = xt * own - k;
    return x0 >= 1 && x0 + 63 <= g.nx - 2 && x0 >= g.pml_x0 && x0 + 63 < g.pml_x1h;
  };
  auto row_ok = [&](int yt) {
    const int y0 = yt * ty;
    return y0 - k >= 1 && y0 + ty + k - 1 <= g.ny - 2 && y0 - k >= g.pml_y0 &&
           y0 + ty + k - 2 < g.pml_y1h;
  };
  TileSplit sp{0, 0, 0, 0};
  int xa = 0;
  while (xa < n_xt && !col_ok(xa)) ++xa;
  int xb = xa;
  while (xb < n_xt && col_ok(xb)) ++xb;
  int ya = 0;
  while (ya < n_yt && !row_ok(ya)) ++ya;
  int yb = ya;
  while (yb < n_yt && row_ok(yb)) ++yb;
  if (xb > xa && yb > ya) sp = TileSplit{xa, xb, ya * ty, yb * ty};
  return sp;
}

// K steps of TY rows of one shot.  EDGE bit 0: x checks, bit 1: y checks.
template <typename T, int TY, int K, int FMT, int EDGE>
__device__ __forceinline__ void fwdk_tile(const FwdKArgs<T> &a, int s, int xt, int yt) {
  constexpr bool GX = (EDGE & 1) != 0, GY = (EDGE & 2) != 0, GEN = EDGE != 0;
  constexpr int OWN = 64 - 2 * K;
  // rows: Ey index k <-> row yb+k (NE); Hx and cq index k <-> row yb+k (NX);
  // Hz, ca, cb and the E-phase memories index k <-> row yb+1+k (NZ)
  constexpr int NE = TY + 2 * K, NX = NE - 1, NZ = NE - 2;
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * OWN - K + lane;
  const int y0 = yt * TY;
  const int yb = y0 - K;
  const bool own = lane >= K && lane <= 63 - K && x < nx;
  const bool e_col = !GX || (x >= 1 && x < nx);
  const bool hz_col = !GX || (x >= 1 && x < nx - 1);
  const bool pml_xe = GX && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = GX && (x < g.pml_x0 || x >= g.pml_x1h);
  T xb = 0, xa = 0, xk = 1, xbh = 0, xah = 0, xkh = 1;
  if constexpr (GX) {
    const int xc = min(max(x, 0), nx - 1);
    xb = g.X(kB, xc), xa = g.X(kA, xc), xk = g.X(kK, xc);
    xbh = g.X(kBh, xc), xah = g.X(kAh, xc), xkh = g.X(kKh, xc);
  }
  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, GEN> L(x, nx, ny);
  int slot[NZ];  // Ey update rows yb+1 .. yb+NZ
  tile_slots<NZ, GY>(a.pts, yb + 1, ny, slot);
  LaneRows<T> yp;  // y profiles of rows yb .. yb+63
  if constexpr (GY) yp.load(g, yb, lane);

  T ey[NE], hx[NX], cqv[NX], hz[NZ], cav[NZ], cbv[NZ];
  {
    const T *__restrict__ ey_i = a.ey_i + base;
    const T *__restrict__ hx_i = a.hx_i + base;
    const T *__restrict__ hz_i = a.hz_i + base;
#pragma unroll
    for (int k = 0; k < NE; ++k) ey[k] = L(ey_i, yb + k);
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      hx[k] = L(hx_i, yb + k);
      cqv[k] = L(cq, yb + k);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      hz[k] = L(hz_i, yb + 1 + k);
      cav[k] = L(ca, yb + 1 + k);
      cbv[k] = L(cb, yb + 1 + k);
    }
  }
  // CPML memories (GEN only, zero outside their strips): m_ey_z at the Hx
  // rows, m_ey_x at the Hz rows, m_hz_x and m_hx_z at the Ey update rows.
  T mez[NX], mex[NZ], mhzx[NZ], mhxz[NZ];
  if constexpr (GY) {
    const T *__restrict__ mez_i = a.mez_i + base;
    const T *__restrict__ mhxz_i = a.mhxz_i + base;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int y = yb + k;
      mez[k] = keep_or_zero(L(mez_i, y), y < g.pml_y0 || y >= g.pml_y1h);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int y = yb + 1 + k;
      mhxz[k] = keep_or_zero(L(mhxz_i, y), y < g.pml_y0 || y >= g.pml_y1);
    }
  }
  if constexpr (GX) {
    const T *__restrict__ mex_i = a.mex_i + base;
    const T *__restrict__ mhzx_i = a.mhzx_i + base;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int y = yb + 1 + k;
      mex[k] = keep_or_zero(L(mex_i, y), pml_xh);
      mhzx[k] = keep_or_zero(L(mhzx_i, y), pml_xe);
    }
  }

#pragma unroll
  for (int j = 0; j < K; ++j) {  // step t+j; the valid rows shrink by j per side
    // Hx^{t+j+1/2}, rows yb+k for k in [j, NX-j)
#pragma unroll
    for (int k = j; k < NX - j; ++k) {
      const int y = yb + k;
      const bool pyh = GY && (y < g.pml_y0 || y >= g.pml_y1h);
      T pa, pb, pk;  // read with every lane active
      yp.row3(kAh, kBh, kKh, y, pyh, pa, pb, pk);
      if (e_col && (!GY || (y >= 1 && y < ny - 1))) {
        T d = (ey[k + 1] - ey[k]) * g.rdy;
        if (pyh) {
          const T m = pb * mez[k] + pa * d;
          mez[k] = m;
          d = d / pk + m;
        }
        hx[k] = hx[k] - cqv[k] * d;
      }
    }
    void *const es = snap_plane<T, FMT>(a.ey_store[j], base);
    void *const cs = snap_plane<T, FMT>(a.curl_store[j], base);
    // Hz^{t+j+1/2} and Ey^{t+j+1}, rows yb+1+k for k in [j, NZ-j)
#pragma unroll
    for (int k = j; k < NZ - j; ++k) {
      const int y = yb + 1 + k;
      const bool row = !GY || (y >= 1 && y < ny);
      const T ey_c = ey[k + 1];
      const T ey_r = shfl_dn1(ey_c);  // unused at lane 63
      if (hz_col && row) {
        T d = (ey_r - ey_c) * g.rdx;
        if (pml_xh) {
          const T m = xbh * mex[k] + xah * d;
          mex[k] = m;
          d = d / xkh + m;
        }
        hz[k] = hz[k] + cqv[k + 1] * d;
      }
      const T hz_l = shfl_up1(hz[k]);  // unused at lane 0
      T ey_n = ey_c;
      const bool pye = GY && (y < g.pml_y0 || y >= g.pml_y1);
      T pa, pb, pk;  // read with every lane active
      yp.row3(kA, kB, kK, y, pye, pa, pb, pk);
      if (e_col && row) {
        T dhz = (hz[k] - hz_l) * g.rdx;
        T dhx = (hx[k + 1] - hx[k]) * g.rdy;
        if (pml_xe) {
          const T m = xb * mhzx[k] + xa * dhz;
          mhzx[k] = m;
          dhz = dhz / xk + m;
        }
        if (pye) {
          const T m = pb * mhxz[k] + pa * dhx;
          mhxz[k] = m;
          dhx = dhx / pk + m;
        }
        const T curl = dhz - dhx;
        if constexpr (FMT != 0) {
          if (own && k >= K - 1 && k < K - 1 + TY && (!GY || y < ny)) {
            const uint32_t i = (uint32_t)(y * nx + x);
            if (es) put_store32<T, FMT>(es, i, ey_c);
            if (cs) put_store32<T, FMT>(cs, i, curl);
          }
        }
        ey_n = cav[k] * ey_c + cbv[k] * curl;
      }
      fused_points<T>(a.pts, slot[k], a.n_shots, s, xt, lane, a.f[j], a.r[j],
                      own && k >= K - 1 && k < K - 1 + TY, ey_n);
      ey[k + 1] = ey_n;
    }
  }

  // owned rows y0 .. y0+TY-1: Ey index K+r, Hx index K+r, Hz index K-1+r
  T *__restrict__ ey_o = a.ey_o + base;
  T *__restrict__ hx_o = a.hx_o + base;
  T *__restrict__ hz_o = a.hz_o + base;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int y = y0 + r;
    if (GY && y >= ny) break;  // uniform
    if (!own) continue;
    const uint32_t o = (uint32_t)(y * nx + x);
    *elem32(ey_o, o) = ey[K + r];
    *elem32(hx_o, o) = hx[K + r];
    *elem32(hz_o, o) = hz[K - 1 + r];
    if constexpr (GY) {
      if (y < g.pml_y0 || y >= g.pml_y1h) *elem32(a.mez_o + base, o) = mez[K + r];
      if (y < g.pml_y0 || y >= g.pml_y1) *elem32(a.mhxz_o + base, o) = mhxz[K - 1 + r];
    }
    if constexpr (GX) {
      if (pml_xh) *elem32(a.mex_o + base, o) = mex[K - 1 + r];
      if (pml_xe) *elem32(a.mhzx_o + base, o) = mhzx[K - 1 + r];
    }
  }
}

template <typename T, int TY, int K, int FMT, int SET, int TYB = TY>
__global__ __launch_bounds__(512) void fwdk_step(FwdKArgs<T> a) {
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  int edge = 3;
  const TileCoord tc = pick_tile<SET, TY, TYB>(a.split, a.pts.n_xt, a.n_yt, a.g.ny, n_groups, &edge);
  const int s = tc.grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
  if constexpr (SET == kOrderedTiles) {
    if (edge == 0)
      fwdk_tile<T, TY, K, FMT, 0>(a, s, tc.xt, tc.yt);
    else if (edge == 1)
      fwdk_tile<T, TYB, K, FMT, 1>(a, s, tc.xt, tc.yt);
    else
      fwdk_tile<T, TYB, K, FMT, 3>(a, s, tc.xt, tc.yt);
  } else {
    if (edge == 0)
      fwdk_tile<T, TY, K, FMT, 0>(a, s, tc.xt, tc.yt);
    else
      fwdk_tile<T, TY, K, FMT, 3>(a, s, tc.xt, tc.yt);
  }
}

// ----------------------------------------------------------------- adjoint
template <typename T> struct AdjKArgs {
  Geom<T> g;
  Points pts;  // built with build_point_table_halo(tile_w 64-2K, lane offset K)
  const T *ca, *cb, *cq;
  const T *ley_i, *lhx_i, *lhz_i;
  T *ley_o, *lhx_o, *lhz_o;
  // adjoint CPML memories (mx: m_hz_x, my: m_hx_z, nx: m_ey_x, nz: m_ey_z)
  const T *mx_i, *my_i, *nx_i, *nz_i;
  T *mx_o, *my_o, *nx_o, *nz_o;
  // Iteration j is the transpose of forward step t-j.
  const void *ey_store[kMaxBlockSteps], *curl_store[kMaxBlockSteps];  // null: no term
  double *gca, *gcb;
  int32_t per_shot_grad;
  double step_scale;
  const T *grad_r[kMaxBlockSteps];  // receiver cotangents of steps t .. t-K+1
  T *grad_f[kMaxBlockSteps];        // source-term gradients of steps t .. t-K+1
  int32_t n_shots, n_yt;
  int64_t group_stride;
  TileSplit split;
};

// K reverse iterations of TY rows of one shot: the two-step adjoint tile
// (adj2_tile) with the row ranges shrinking by one per side each iteration.
// When `reduce`, each iteration's FP64 gradient terms of the owned cells are
// summed into one LDS slot per cell (red_ca / red_cb, [TY][64]); otherwise
// they go straight to the per-shot gradient arrays, iteration by iteration.
template <typename T, int TY, int K, int FMT, int EDGE>
__device__ __forceinline__ void adjk_tile(const AdjKArgs<T> &a, int s, int xt, int yt,
                                          double *red_ca, double *red_cb, bool reduce) {
  constexpr bool GX = (EDGE & 1) != 0, GY = (EDGE & 2) != 0, GEN = EDGE != 0;
  constexpr int OWN = 64 - 2 * K;
  // lambda_Ey, cb and my index k <-> row yb+k (NE); lambda_Hx, cq and nz
  // index k <-> row yb+k (NX); lambda_Hz, ca, mx and nx index k <-> row
  // yb+1+k (NZ)
  constexpr int NE = TY + 2 * K, NX = NE - 1, NZ = NE - 2;
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * OWN - K + lane;
  const int y0 = yt * TY;
  const int yb = y0 - K;
  const bool own = lane >= K && lane <= 63 - K && x < nx;
  const bool e_col = !GX || (x >= 1 && x < nx);
  const bool pml_xe = GX && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = GX && (x < g.pml_x0 || x >= g.pml_x1h);
  T xb = 0, xa = 0, xk = 1, xbh = 0, xah = 0, xkh = 1;
  if constexpr (GX) {
    const int xc = min(max(x, 0), nx - 1);
    xb = g.X(kB, xc), xa = g.X(kA, xc), xk = g.X(kK, xc);
    xbh = g.X(kBh, xc), xah = g.X(kAh, xc), xkh = g.X(kKh, xc);
  }
  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, GEN> L(x, nx, ny);
  int slot[NZ];  // rows yb+1 .. yb+NZ
  tile_slots<NZ, GY>(a.pts, yb + 1, ny, slot);
  LaneRows<T> yp;  // y profiles of rows yb .. yb+63
  if constexpr (GY) yp.load(g, yb, lane);

  T ley[NE], cbv[NE], lhx[NX], cqv[NX], lhz[NZ], cav[NZ];
  {
    const T *__restrict__ ley_i = a.ley_i + base;
    const T *__restrict__ lhx_i = a.lhx_i + base;
    const T *__restrict__ lhz_i = a.lhz_i + base;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      ley[k] = L(ley_i, yb + k);
      cbv[k] = L(cb, yb + k);
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      lhx[k] = L(lhx_i, yb + k);
      cqv[k] = L(cq, yb + k);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      lhz[k] = L(lhz_i, yb + 1 + k);
      cav[k] = L(ca, yb + 1 + k);
    }
  }
  // Snapshots of the owned rows for every iteration, loaded with the rows
  // above and without per-lane or per-buffer branches (an absent buffer
  // reads this shot's lambda_Ey plane, a lane outside the grid its clamped
  // column; keep_or_zero clears those values).
  bool grad_on[K];
  float se[K][TY], sc[K][TY];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    grad_on[j] = FMT != 0 && (a.ey_store[j] != nullptr || a.curl_store[j] != nullptr);
    const void *const es = snap_plane<T, FMT>(a.ey_store[j], base);
    const void *const cs = snap_plane<T, FMT>(a.curl_store[j], base);
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      se[j][r] = sc[j][r] = 0.0f;
      if constexpr (sizeof(T) == 4 && FMT != 0) {
        const int y = y0 + r;
        const void *const dflt = a.ley_i + base;
        const int xc = GX ? min(max(x, 0), nx - 1) : x;
        const int yc = GY ? min(y, ny - 1) : y;
        const uint32_t i = (uint32_t)(yc * nx + xc);
        const float v0 = (float)get_store32<T, FMT>(es ? es : dflt, i);
        const float v1 = (float)get_store32<T, FMT>(cs ? cs : dflt, i);
        const bool in = (!GX || (x >= 1 && x < nx)) && (!GY || y < ny);
        se[j][r] = keep_or_zero(v0, grad_on[j] && es && in);
        sc[j][r] = keep_or_zero(v1, grad_on[j] && cs && in);
      }
    }
  }
  // CPML memories (GEN only, zero outside their strips)
  T my[NE], nz[NX], mx[NZ], nxm[NZ];
  if constexpr (GY) {
    const T *__restrict__ my_i = a.my_i + base;
    const T *__restrict__ nz_i = a.nz_i + base;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int y = yb + k;
      my[k] = keep_or_zero(L(my_i, y), y < g.pml_y0 || y >= g.pml_y1);
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int y = yb + k;
      nz[k] = keep_or_zero(L(nz_i, y), y < g.pml_y0 || y >= g.pml_y1h);
    }
  }
  if constexpr (GX) {
    const T *__restrict__ mx_i = a.mx_i + base;
    const T *__restrict__ nx_i = a.nx_i + base;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      mx[k] = keep_or_zero(L(mx_i, yb + 1 + k), pml_xe);
      nxm[k] = keep_or_zero(L(nx_i, yb + 1 + k), pml_xh);
    }
  }

  // The single-step kernel's terms, each also advancing its CPML memory.
  auto raw_x = [&](int y, T lm, T cbx, T &m) -> T {  // adjoint of dHz/dx
    if ((GY && (y < 1 || y >= ny)) || (GX && (x < 1 || x >= nx))) return (T)0;
    const T q = cbx * lm;
    if (!pml_xe) return q;
    const T mn = xb * m + q;
    m = mn;
    return xa * mn + q / xk;
  };
  // nraw_y and raw_z read the row's y profile before any lane leaves (the
  // cross-lane reads need every lane active); call them with all lanes active.
  auto nraw_y = [&](int y, T lm, T cbx, T &m) -> T {  // negated adjoint of dHx/dz
    const bool pye = GY && (y < g.pml_y0 || y >= g.pml_y1);
    T pa, pb, pk;
    yp.row3(kA, kB, kK, y, pye, pa, pb, pk);
    if ((GY && (y < 1 || y >= ny)) || !e_col) return (T)0;
    const T q = cbx * lm;
    if (!pye) return q;
    const T mn = pb * m - q;
    m = mn;
    return q / pk - pa * mn;
  };
  auto raw_z = [&](int y, T lhxv, T cqx, T &m) -> T {  // adjoint of dEy/dz (Hx update)
    const bool pyh = GY && (y < g.pml_y0 || y >= g.pml_y1h);
    T pa, pb, pk;
    yp.row3(kAh, kBh, kKh, y, pyh, pa, pb, pk);
    if ((GY && (y < 1 || y >= ny - 1)) || !e_col) return (T)0;
    const T adj = -cqx * lhxv;
    if (!pyh) return adj;
    const T mn = pb * m + adj;
    m = mn;
    return pa * mn + adj / pk;
  };
  auto raw_xh = [&](int y, T lhzv, T cqx, T &m) -> T {  // adjoint of dEy/dx (Hz update)
    if ((GY && (y < 1 || y >= ny)) || (GX && (x < 1 || x >= nx - 1))) return (T)0;
    const T adj = cqx * lhzv;
    if (!pml_xh) return adj;
    const T mn = xbh * m + adj;
    m = mn;
    return xah * mn + adj / xkh;
  };
  auto in_e = [&](int y) -> bool { return e_col && (!GY || (y >= 1 && y < ny)); };

#pragma unroll
  for (int j = 0; j < K; ++j) {  // iteration j: transpose of forward step t-j
    T nr[NE], rz[NX];
#pragma unroll
    for (int k = j; k < NE - j; ++k) nr[k] = nraw_y(yb + k, ley[k], cbv[k], my[k]);
#pragma unroll
    for (int k = j; k < NX - j; ++k) {
      const int y = yb + k;
      if (in_e(y)) lhx[k] = lhx[k] + (nr[k + 1] - nr[k]) * g.rdy;
      rz[k] = raw_z(y, lhx[k], cqv[k], nz[k]);
    }
    const void *const es = snap_plane<T, FMT>(a.ey_store[j], base);
    const void *const cs = snap_plane<T, FMT>(a.curl_store[j], base);
#pragma unroll
    for (int k = j; k < NZ - j; ++k) {
      const int y = yb + 1 + k;
      const T ley_c = ley[k + 1];
      const T rx = raw_x(y, ley_c, cbv[k + 1], mx[k]);
      const T rx_r = shfl_dn1(rx);  // unused at lane 63
      if (in_e(y)) lhz[k] = lhz[k] + (rx - rx_r) * g.rdx;
      const T rxh = raw_xh(y, lhz[k], cqv[k + 1], nxm[k]);
      const T rxh_l = shfl_up1(rxh);  // unused at lane 0
      T ley_n = ley_c;
      if (in_e(y)) {
        const T t1 = cav[k] * ley_c;
        const T t2 = (rz[k] - rz[k + 1]) * g.rdy + (rxh_l - rxh) * g.rdx;
        ley_n = t1 + t2;
      }
      const bool owned_row = k >= K - 1 && k < K - 1 + TY;
      fused_adj_points<T>(a.pts, slot[k], a.n_shots, s, xt, lane, a.grad_r[j], a.grad_f[j],
                          own && owned_row, ley_n);
      if (owned_row && grad_on[j] && own && in_e(y)) {
        const int r = k - (K - 1);
        const double lm = (double)ley_n;
        double ve, vc;
        if constexpr (sizeof(T) == 4) {
          ve = (double)se[j][r];
          vc = (double)sc[j][r];
        } else {
          const uint32_t i = (uint32_t)(y * nx + x);
          ve = es ? get_store32<T, FMT>(es, i) : 0.0;
          vc = cs ? get_store32<T, FMT>(cs, i) : 0.0;
        }
        const double pa = lm * ve * a.step_scale;
        const double pb = lm * vc * a.step_scale;
        if (reduce) {
          double *const ra = red_ca + r * 64 + lane;
          double *const rb = red_cb + r * 64 + lane;
          *ra = j == 0 ? pa : *ra + pa;
          *rb = j == 0 ? pb : *rb + pb;
        } else {
          const int64_t gb = a.per_shot_grad ? base : 0;
          const uint32_t gi = (uint32_t)(y * nx + x);
          if (a.gca && es) *elem32(a.gca + gb, gi) += pa;
          if (a.gcb && cs) *elem32(a.gcb + gb, gi) += pb;
        }
      }
      ley[k + 1] = ley_n;
    }
  }

  // owned rows y0 .. y0+TY-1: lambda_Ey / lambda_Hx / my / nz index K+r,
  // lambda_Hz / mx / nx index K-1+r
  T *__restrict__ ley_o = a.ley_o + base;
  T *__restrict__ lhx_o = a.lhx_o + base;
  T *__restrict__ lhz_o = a.lhz_o + base;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int y = y0 + r;
    if (GY && y >= ny) break;  // uniform
    if (!own) continue;
    const uint32_t o = (uint32_t)(y * nx + x);
    *elem32(ley_o, o) = ley[K + r];
    *elem32(lhx_o, o) = lhx[K + r];
    *elem32(lhz_o, o) = lhz[K - 1 + r];
    if constexpr (GEN) {
      if (in_e(y)) {
        if constexpr (GX) {
          if (pml_xe) *elem32(a.mx_o + base, o) = mx[K - 1 + r];
          if (x < nx - 1 && pml_xh) *elem32(a.nx_o + base, o) = nxm[K - 1 + r];
        }
        if constexpr (GY) {
          if (y < g.pml_y0 || y >= g.pml_y1) *elem32(a.my_o + base, o) = my[K + r];
          if (y < ny - 1 && (y < g.pml_y0 || y >= g.pml_y1h)) *elem32(a.nz_o + base, o) = nz[K + r];
        }
      }
    }
  }
}

// One K-step adjoint workgroup; `red` is the dynamic LDS block
// [2 coefficients][G][TY][64].  Each slot holds one shot's terms of all K
// iterations; the shots' slots are summed in shot order and added to the
// cell once per launch (FP64).  ALLOW as in adj2_workgroup.
template <typename T, int TY, int K, int FMT, int ALLOW>
__device__ __forceinline__ void adjk_workgroup(const AdjKArgs<T> &a, const TileCoord &tc,
                                               int edge, double *red) {
  constexpr int OWN = 64 - 2 * K;
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  const int G = (int)blockDim.y;
  const int s = tc.grp * G + w;
  const bool any_grad = FMT != 0 && (a.gca != nullptr || a.gcb != nullptr);
  const bool reduce = any_grad && !a.per_shot_grad;
  const int64_t slab = (int64_t)G * TY * 64;
  double *red_ca = red + (int64_t)w * TY * 64;
  double *red_cb = red_ca + slab;
  if (reduce) {
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      red_ca[r * 64 + lane] = 0.0;
      red_cb[r * 64 + lane] = 0.0;
    }
  }
  const int x = tc.xt * OWN - K + lane;
  const int y0 = tc.yt * TY;
  const bool red_lane = lane >= K && lane <= 63 - K && x >= 1 && x < a.g.nx;
  bool any_ey = false, any_curl = false;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    any_ey = any_ey || a.ey_store[j] != nullptr;
    any_curl = any_curl || a.curl_store[j] != nullptr;
  }
  double pre[2] = {0.0, 0.0};
  if (reduce && red_lane && w < TY && y0 + w >= 1 && y0 + w < a.g.ny) {
    const uint32_t j = (uint32_t)((y0 + w) * a.g.nx + x);
    if (a.gca && any_ey) pre[0] = *elem32(a.gca + tc.grp * a.group_stride, j);
    if (a.gcb && any_curl) pre[1] = *elem32(a.gcb + tc.grp * a.group_stride, j);
  }
  if (s < a.n_shots) {  // inactive wavefronts still join the barrier below
    constexpr bool a0 = ALLOW & 1, a1 = ALLOW & 2, a3 = ALLOW & 8;
    if (a0 && (!(a1 || a3) || edge == 0)) {
      if constexpr (a0) adjk_tile<T, TY, K, FMT, 0>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    } else if (a1 && (!a3 || edge == 1)) {
      if constexpr (a1) adjk_tile<T, TY, K, FMT, 1>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    } else {
      if constexpr (a3) adjk_tile<T, TY, K, FMT, 3>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    }
  }
  if (!reduce) return;  // uniform over the workgroup
  __syncthreads();
  if (!red_lane) return;
  for (int pass = 0; pass < 2; ++pass) {
    double *out = pass == 0 ? a.gca : a.gcb;
    if (out == nullptr || !(pass == 0 ? any_ey : any_curl)) continue;
    out += tc.grp * a.group_stride;
    const double *slots = red + (int64_t)pass * slab;
    for (int r = w; r < TY; r += G) {
      const int y = y0 + r;
      if (y >= 1 && y < a.g.ny) {
        const uint32_t j = (uint32_t)(y * a.g.nx + x);
        double sum = 0.0;
        for (int k = 0; k < G; ++k) sum += slots[(k * TY + r) * 64 + lane];
        *elem32(out, j) = (r == w ? pre[pass] : *elem32(out, j)) + sum;
      }
    }
  }
}

template <typename T, int TY, int K, int FMT, int SET, int TYB>
__device__ __forceinline__ void adjk_body(const AdjKArgs<T> &a) {
  extern __shared__ double red[];
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  int edge = 3;
  const TileCoord tc = pick_tile<SET, TY, TYB>(a.split, a.pts.n_xt, a.n_yt, a.g.ny, n_groups, &edge);
  if constexpr (SET == kOrderedTiles && TYB == TY) {
    adjk_workgroup<T, TY, K, FMT, 1 | 2 | 8>(a, tc, edge, red);
  } else if constexpr (SET == kOrderedTiles) {
    if (edge == 0)  // uniform over the workgroup
      adjk_workgroup<T, TY, K, FMT, 1>(a, tc, edge, red);
    else
      adjk_workgroup<T, TYB, K, FMT, 2 | 8>(a, tc, edge, red);
  } else {
    adjk_workgroup<T, TY, K, FMT, 1 | 8>(a, tc, edge, red);
  }
}

template <typename T, int TY, int K, int FMT, int SET, int TYB = TY>
__global__ __launch_bounds__(512) void adjk_step(AdjKArgs<T> a) {
  adjk_body<T, TY, K, FMT, SET, TYB>(a);
}

// The same kernel held to W waves per SIMD (512 / W VGPRs; workgroups of at
// most 256 threads), spilling what does not fit.
template <typename T, int TY, int K, int FMT, int SET, int TYB, int W>
__global__ __launch_bounds__(256, W) void adjk_step_w(AdjKArgs<T> a) {
  adjk_body<T, TY, K, FMT, SET, TYB>(a);
}

}  // namespace tm2d
}  // namespace tide_hip
