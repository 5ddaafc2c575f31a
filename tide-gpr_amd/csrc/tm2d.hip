// 2-D TM Maxwell FDTD propagator for MI355X (gfx950).
//
// One fused kernel per forward time step and one per adjoint time step.
//
// Layout: every field is [n_shots, ny, nx] fp32 (or fp64), x fastest.  A
// workgroup is one tile of 62 columns x TY rows; each of its wavefronts
// handles one shot of that tile.  The 64 lanes of a wavefront cover columns
// x0-1 .. x0+62: lanes 1..62 own their cells, lanes 0 and 63 only supply the
// x-halo, so every lane runs the same instruction stream and x-neighbours
// come from one cross-lane shuffle per direction.  A wavefront loads all its
// rows before computing (many loads in flight per lane) and marches down the
// rows carrying the previous row's values in registers; the coefficients
// are read by every wavefront of the tile from the same cache lines.  Fields
// ping-pong between two buffers per step (the caller's arrays and a scratch
// copy), which removes every inter-workgroup read/write race of fusing the H
// and E half-steps.  Sources, receivers, the adjoint injection and the FP64
// coefficient-gradient accumulation happen in the thread owning the cell.
//
// Semantics follow the reference propagator (paths relative to src/tide/):
//   H update          csrc/tm2d/maxwell_tm_core.cuh:448-502
//   E update/storage  csrc/tm2d/maxwell_tm_core.cuh:505-599
//   point ops         csrc/tm2d/maxwell.cu:89-161
//   exact adjoint     csrc/tm2d/maxwell_tm_cuda_inst.cu:922-1254
//   reverse loop      csrc/tm2d/maxwell_tm_cuda_inst.cu:2540-2731
// with the same floating-point operation order (no FMA contraction), so
// results match the reference CPU arithmetic.  Stencil order 2 (the
// reference default) is the one built here.

#include "tide_common.h"
#include "tide_hip.h"

#include <algorithm>
#include <cstring>

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
  int32_t use_atomics;
};

template <typename T, int FMT>
__device__ __forceinline__ void put_store(void *base, int64_t i, T v) {
  if constexpr (FMT == 2) {
    ((uint16_t *)base)[i] = f32_to_bf16((float)v);
  } else {
    ((T *)base)[i] = v;
  }
}

template <typename T, int FMT>
__device__ __forceinline__ double get_store(const void *base, int64_t i) {
  if constexpr (FMT == 2) {
    return (double)bf16_to_f32(((const uint16_t *)base)[i]);
  } else {
    return (double)((const T *)base)[i];
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
      if (y < 0 || y >= ny) return (T)0;  // uniform
      const T v = p[y * nx + xc];
      return xv ? v : (T)0;
    } else {
      return p[y * nx + x];
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
    const T ey_r = __shfl_down(ey_c, 1, 64);  // Ey(y, x+1); unused at lane 63

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
    const T hz_l = __shfl_up(hz_n, 1, 64);  // Hz(y, x-1); unused at lane 0

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

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void fwd_step(FwdArgs<T> a) {
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const TileCoord tc = decode_tile(a.pts.n_xt, a.n_yt, n_groups);
  const int s = tc.grp * G + (int)threadIdx.y;
  if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
  if (interior_tile<T, TY>(a.g, tc.xt, tc.yt))
    fwd_tile<T, TY, FMT, false>(a, s, tc.xt, tc.yt);
  else
    fwd_tile<T, TY, FMT, true>(a, s, tc.xt, tc.yt);
}

// ----------------------------------------------------------------- adjoint
// One reverse iteration: transpose of the E step, transpose of the H
// half-step, receiver injection, source read-out, coefficient gradient.
template <typename T, int TY, int FMT, bool EDGE>
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
    const T rawx_r = __shfl_down(rawx_c, 1, 64);  // unused at lane 63
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
    const T rawxh_l = __shfl_up(rawxh_c, 1, 64);  // unused at lane 0
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
        red_ca[r * 64 + lane] = pa;
        red_cb[r * 64 + lane] = pb;
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
template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void adj_step(AdjArgs<T> a) {
  extern __shared__ double red[];  // [2][G][TY][64] when reducing
  const int lane = threadIdx.x;
  const int w = (int)threadIdx.y;
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const TileCoord tc = decode_tile(a.pts.n_xt, a.n_yt, n_groups);
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
    const double *slots = red + (int64_t)pass * G * TY * 64;
    for (int r = w; r < TY; r += G) {
      const int y = y0 + r;
      if (y >= 1 && y < a.g.ny) {
        double sum = 0.0;
        for (int k = 0; k < G; ++k) sum += slots[(k * TY + r) * 64 + lane];
        const int64_t j = (int64_t)y * a.g.nx + x;
        if (a.use_atomics)
          atomicAdd(&out[j], sum);
        else
          out[j] += sum;
      }
    }
  }
}

// Adds an FP64 accumulator into the caller's gradient array (reference
// combine_grad semantics: grad += sum over shots).
template <typename T>
__global__ void fold_grad(T *__restrict__ out, const double *__restrict__ acc, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)((double)out[i] + acc[i]);
}

// ------------------------------------------------------------------ driver

template <typename T> struct Problem {
  Geom<T> g;
  const T *ca, *cb, *cq;
  const T *prof[12];  // ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
  int32_t n_shots;
  int64_t n_src, n_rec;
};

constexpr int kMaxGroup = 8;  // shots (wavefronts) per workgroup

// Launch shape: rows marched per wavefront and shots (wavefronts) per
// workgroup for each kernel.  Process-wide; set by the autotuner through
// tide_tm2d_set_tuning().  fp64 fields use half the rows (twice the
// registers per value).
struct Tuning {
  int fwd_rows = 8, fwd_group = 8, adj_rows = 4, adj_group = 8;
};
Tuning g_tuning;

template <typename T> int rows_for(int rows) {
  if (sizeof(T) == 8) rows = std::max(rows / 2, 2);
  return rows;
}

template <typename T, int FMT>
void launch_fwd(int rows, dim3 grid, dim3 block, hipStream_t stream, const FwdArgs<T> &a) {
  switch (rows) {
    case 2: fwd_step<T, 2, FMT><<<grid, block, 0, stream>>>(a); break;
    case 4: fwd_step<T, 4, FMT><<<grid, block, 0, stream>>>(a); break;
    case 16: fwd_step<T, 16, FMT><<<grid, block, 0, stream>>>(a); break;
    default: fwd_step<T, 8, FMT><<<grid, block, 0, stream>>>(a); break;
  }
}

template <typename T, int FMT>
void launch_adj(int rows, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
                const AdjArgs<T> &a) {
  switch (rows) {
    case 2: adj_step<T, 2, FMT><<<grid, block, lds, stream>>>(a); break;
    case 8: adj_step<T, 8, FMT><<<grid, block, lds, stream>>>(a); break;
    default: adj_step<T, 4, FMT><<<grid, block, lds, stream>>>(a); break;
  }
}

inline int check_geometry(int64_t n_shots, int64_t ny, int64_t nx) {
  if (n_shots <= 0 || ny < 3 || nx < 3) {
    set_error("n_shots must be positive and the padded grid at least 3x3");
    return kInvalidArgument;
  }
  if (ny * nx >= (int64_t)1 << 29 || n_shots > (1 << 30)) {
    set_error("a shot plane must hold fewer than 2^29 cells");
    return kInvalidArgument;
  }
  return kOk;
}

// Reads the flat source/receiver indices (device or host memory) to the host.
inline int fetch_indices(const int64_t *p, int64_t n, std::vector<int64_t> &out,
                         hipStream_t stream) {
  out.assign((size_t)std::max<int64_t>(n, 0), 0);
  if (n <= 0 || p == nullptr) return kOk;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e == hipSuccess && attr.type == hipMemoryTypeDevice) {
    TIDE_HIP_TRY(hipMemcpyAsync(out.data(), p, (size_t)n * sizeof(int64_t),
                                hipMemcpyDeviceToHost, stream));
    TIDE_HIP_TRY(hipStreamSynchronize(stream));
  } else {
    (void)hipGetLastError();
    std::memcpy(out.data(), p, (size_t)n * sizeof(int64_t));
  }
  return kOk;
}

struct Scratch {
  std::vector<void *> bufs;
  hipStream_t stream;
  explicit Scratch(hipStream_t s) : stream(s) {}
  ~Scratch() {
    for (void *p : bufs) (void)hipFreeAsync(p, stream);
  }
  template <typename U> int alloc(U **out, size_t count) {
    void *p = nullptr;
    TIDE_HIP_TRY(hipMallocAsync(&p, std::max<size_t>(count, 1) * sizeof(U), stream));
    bufs.push_back(p);
    *out = (U *)p;
    return kOk;
  }
};

#define TIDE_CHECK(expr)                    \
  do {                                      \
    const int _rc = (expr);                 \
    if (_rc != ::tide_hip::kOk) return _rc; \
  } while (0)

// Packs the 12 CPML profiles into [6][ny] + [6][nx] device memory.
template <typename T>
int pack_profiles(Problem<T> &p, Scratch &scr, hipStream_t stream) {
  const int ny = p.g.ny, nx = p.g.nx;
  T *buf;
  TIDE_CHECK(scr.alloc(&buf, (size_t)6 * (ny + nx)));
  // order in prof[]: ay0 by1 ayh2 byh3 ax4 bx5 axh6 bxh7 ky8 kyh9 kx10 kxh11
  const int ysrc[6] = {0, 1, 8, 2, 3, 9};    // a, b, k, ah, bh, kh
  const int xsrc[6] = {4, 5, 10, 6, 7, 11};
  for (int k = 0; k < 6; ++k) {
    TIDE_HIP_TRY(hipMemcpyAsync(buf + (size_t)k * ny, p.prof[ysrc[k]], (size_t)ny * sizeof(T),
                                hipMemcpyDefault, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(buf + (size_t)6 * ny + (size_t)k * nx, p.prof[xsrc[k]],
                                (size_t)nx * sizeof(T), hipMemcpyDefault, stream));
  }
  p.g.py = buf;
  p.g.px = buf + (size_t)6 * ny;
  return kOk;
}

template <typename T>
int upload_points(const Problem<T> &p, const int64_t *src, const int64_t *rec,
                  DevicePointTable &dpt, hipStream_t stream) {
  std::vector<int64_t> hsrc, hrec;
  TIDE_CHECK(fetch_indices(src, (int64_t)p.n_shots * p.n_src, hsrc, stream));
  TIDE_CHECK(fetch_indices(rec, (int64_t)p.n_shots * p.n_rec, hrec, stream));
  PointTable table;
  if (!build_point_table(hsrc.data(), p.n_src, hrec.data(), p.n_rec, p.n_shots, p.g.ny,
                         p.g.nx, kOwn, 1, table))
    return kInvalidArgument;
  return dpt.upload(table, stream);
}

template <typename T>
int forward_impl(Problem<T> p, T *ey, T *hx, T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z,
                 T *m_hz_x, const T *f, T *r, const int64_t *src, const int64_t *rec,
                 int64_t nt, int64_t start_t, int64_t step_ratio, void *ey_store,
                 void *curl_store, int fmt, size_t store_step_bytes, hipStream_t stream) {
  TIDE_CHECK(check_geometry(p.n_shots, p.g.ny, p.g.nx));
  if (nt <= 0) return kOk;
  if (step_ratio <= 0) {
    set_error("step_ratio must be positive");
    return kInvalidArgument;
  }
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  const int64_t n_shots = p.n_shots, total = n_shots * p.g.numel;
  DevicePointTable dpt;
  TIDE_CHECK(upload_points(p, src, rec, dpt, stream));
  Scratch scr(stream);
  TIDE_CHECK(pack_profiles(p, scr, stream));
  T *ey2, *hx2, *hz2, *mex2, *mez2;
  TIDE_CHECK(scr.alloc(&ey2, (size_t)total));
  TIDE_CHECK(scr.alloc(&hx2, (size_t)total));
  TIDE_CHECK(scr.alloc(&hz2, (size_t)total));
  TIDE_CHECK(scr.alloc(&mex2, (size_t)total));
  TIDE_CHECK(scr.alloc(&mez2, (size_t)total));
  const size_t bytes = (size_t)total * sizeof(T);
  // H-phase memories are only written inside the strips; both copies start
  // equal so the untouched cells agree whichever buffer ends up current.
  TIDE_HIP_TRY(hipMemcpyAsync(mex2, m_ey_x, bytes, hipMemcpyDeviceToDevice, stream));
  TIDE_HIP_TRY(hipMemcpyAsync(mez2, m_ey_z, bytes, hipMemcpyDeviceToDevice, stream));

  FwdArgs<T> a;
  a.g = p.g;
  a.pts = Points{dpt.row_slot, dpt.seg, dpt.ent, dpt.n_xt};
  a.ca = p.ca;
  a.cb = p.cb;
  a.cq = p.cq;
  a.mhzx = m_hz_x;
  a.mhxz = m_hx_z;
  a.n_shots = p.n_shots;
  const int rows = rows_for<T>(g_tuning.fwd_rows);
  const int group = std::min<int>(p.n_shots, g_tuning.fwd_group);
  a.n_yt = (p.g.ny + rows - 1) / rows;
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * a.n_yt * n_groups));

  T *cur[5] = {ey, hx, hz, m_ey_x, m_ey_z};
  T *nxt[5] = {ey2, hx2, hz2, mex2, mez2};
  const bool storing = ey_store != nullptr || curl_store != nullptr;
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    a.ey_i = cur[0]; a.hx_i = cur[1]; a.hz_i = cur[2];
    a.mex_i = cur[3]; a.mez_i = cur[4];
    a.ey_o = nxt[0]; a.hx_o = nxt[1]; a.hz_o = nxt[2];
    a.mex_o = nxt[3]; a.mez_o = nxt[4];
    a.f_t = p.n_src > 0 ? f + t * n_shots * p.n_src : nullptr;
    a.r_t = p.n_rec > 0 ? r + t * n_shots * p.n_rec : nullptr;
    const bool st = storing && (t % step_ratio) == 0;
    const size_t off = (size_t)(t / step_ratio) * store_step_bytes;
    a.ey_store = st && ey_store ? (char *)ey_store + off : nullptr;
    a.curl_store = st && curl_store ? (char *)curl_store + off : nullptr;
    if (!st) {
      launch_fwd<T, 0>(rows, grid, block, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_fwd<T, 2>(rows, grid, block, stream, a);
    } else {
      launch_fwd<T, 1>(rows, grid, block, stream, a);
    }
    for (int k = 0; k < 5; ++k) std::swap(cur[k], nxt[k]);
  }
  TIDE_HIP_TRY(hipGetLastError());
  if (cur[0] != ey) {
    TIDE_HIP_TRY(hipMemcpyAsync(ey, cur[0], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(hx, cur[1], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(hz, cur[2], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(m_ey_x, cur[3], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(m_ey_z, cur[4], bytes, hipMemcpyDeviceToDevice, stream));
  }
  dpt.release(stream);
  return kOk;
}

// gca/gcb: FP64 accumulators ([ny,nx] summed over shots, or [n_shots,ny,nx]
// when the coefficient is per shot), added to; null when not required.
template <typename T>
int backward_impl(Problem<T> p, const T *grad_r, T *l_ey, T *l_hx, T *l_hz, T *ml_ey_x,
                  T *ml_ey_z, T *ml_hx_z, T *ml_hz_x, const void *ey_store,
                  const void *curl_store, int fmt, size_t store_step_bytes, T *grad_f,
                  double *gca, double *gcb, const int64_t *src, const int64_t *rec,
                  int64_t nt, int64_t start_t, int64_t step_ratio, hipStream_t stream) {
  TIDE_CHECK(check_geometry(p.n_shots, p.g.ny, p.g.nx));
  if (nt <= 0) return kOk;
  if (step_ratio <= 0) {
    set_error("step_ratio must be positive");
    return kInvalidArgument;
  }
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  if ((gca && !ey_store) || (gcb && !curl_store)) {
    set_error("coefficient gradients require the stored forward snapshots");
    return kInvalidArgument;
  }
  const int64_t n_shots = p.n_shots, total = n_shots * p.g.numel;
  DevicePointTable dpt;
  TIDE_CHECK(upload_points(p, src, rec, dpt, stream));
  Scratch scr(stream);
  TIDE_CHECK(pack_profiles(p, scr, stream));
  T *ley2, *lhx2, *lhz2, *m2[4];
  TIDE_CHECK(scr.alloc(&ley2, (size_t)total));
  TIDE_CHECK(scr.alloc(&lhx2, (size_t)total));
  TIDE_CHECK(scr.alloc(&lhz2, (size_t)total));
  for (int k = 0; k < 4; ++k) {
    TIDE_CHECK(scr.alloc(&m2[k], (size_t)total));
    TIDE_HIP_TRY(hipMemsetAsync(m2[k], 0, (size_t)total * sizeof(T), stream));
  }
  const size_t bytes = (size_t)total * sizeof(T);

  AdjArgs<T> a;
  a.g = p.g;
  a.pts = Points{dpt.row_slot, dpt.seg, dpt.ent, dpt.n_xt};
  a.ca = p.ca;
  a.cb = p.cb;
  a.cq = p.cq;
  a.per_shot_grad = ((gca && p.g.ca_b) || (gcb && p.g.cb_b)) ? 1 : 0;
  a.step_scale = (double)step_ratio;
  a.n_shots = p.n_shots;
  const int rows = rows_for<T>(g_tuning.adj_rows);
  const int group = std::min<int>(p.n_shots, g_tuning.adj_group);
  a.n_yt = (p.g.ny + rows - 1) / rows;
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  // More shots than one workgroup holds: the per-group sums meet in FP64
  // atomics (order-dependent in the last bit); otherwise one plain add.
  a.use_atomics = n_groups > 1 ? 1 : 0;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * a.n_yt * n_groups));
  const size_t lds = 2 * (size_t)group * rows * 64 * sizeof(double);

  // order: ley, lhx, lhz, mx (hz_x), my (hx_z), nx (ey_x), nz (ey_z)
  T *cur[7] = {l_ey, l_hx, l_hz, ml_hz_x, ml_hx_z, ml_ey_x, ml_ey_z};
  T *nxt[7] = {ley2, lhx2, lhz2, m2[0], m2[1], m2[2], m2[3]};
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
    a.ley_i = cur[0]; a.lhx_i = cur[1]; a.lhz_i = cur[2];
    a.mx_i = cur[3]; a.my_i = cur[4]; a.nx_i = cur[5]; a.nz_i = cur[6];
    a.ley_o = nxt[0]; a.lhx_o = nxt[1]; a.lhz_o = nxt[2];
    a.mx_o = nxt[3]; a.my_o = nxt[4]; a.nx_o = nxt[5]; a.nz_o = nxt[6];
    a.grad_r_t = p.n_rec > 0 ? grad_r + t * n_shots * p.n_rec : nullptr;
    a.grad_f_t = (p.n_src > 0 && grad_f) ? grad_f + t * n_shots * p.n_src : nullptr;
    const bool g_step = (gca || gcb) && (t % step_ratio) == 0;
    const size_t off = (size_t)(t / step_ratio) * store_step_bytes;
    a.ey_store = g_step && gca ? (const char *)ey_store + off : nullptr;
    a.curl_store = g_step && gcb ? (const char *)curl_store + off : nullptr;
    a.gca = g_step ? gca : nullptr;
    a.gcb = g_step ? gcb : nullptr;
    if (!g_step) {
      launch_adj<T, 0>(rows, grid, block, lds, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_adj<T, 2>(rows, grid, block, lds, stream, a);
    } else {
      launch_adj<T, 1>(rows, grid, block, lds, stream, a);
    }
    for (int k = 0; k < 7; ++k) std::swap(cur[k], nxt[k]);
  }
  TIDE_HIP_TRY(hipGetLastError());
  T *orig[7] = {l_ey, l_hx, l_hz, ml_hz_x, ml_hx_z, ml_ey_x, ml_ey_z};
  for (int k = 0; k < 7; ++k)
    if (cur[k] != orig[k])
      TIDE_HIP_TRY(hipMemcpyAsync(orig[k], cur[k], bytes, hipMemcpyDeviceToDevice, stream));
  dpt.release(stream);
  return kOk;
}

template <typename T>
int fold_into(T *out, const double *acc, int64_t n, hipStream_t stream) {
  if (n <= 0) return kOk;
  const int threads = 256;
  fold_grad<T><<<(unsigned)((n + threads - 1) / threads), threads, 0, stream>>>(out, acc, n);
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace tm2d
}  // namespace tide_hip

// ===================================================================== C ABI
// Reference-compatible entry points.  Argument order and meaning follow the
// reference's ctypes specs (src/tide/backend_utils.py:256-308); the return
// value is a status code (0 = success) instead of void, readable through
// tide_last_error().  Storage modes: 0 device, 3 none; host-staged modes 1/2
// are reported as unsupported.

using namespace tide_hip;
using namespace tide_hip::tm2d;

namespace {

constexpr int64_t kStorageDevice = 0;
constexpr int64_t kStorageNone = 3;
constexpr int64_t kFormatBf16 = 1;

template <typename T>
Problem<T> make_problem(const T *ca, const T *cb, const T *cq, const T *const prof[12],
                        T rdy, T rdx, int64_t n_shots, int64_t ny, int64_t nx,
                        int64_t n_src, int64_t n_rec, bool ca_b, bool cb_b, bool cq_b,
                        int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1) {
  Problem<T> p;
  Geom<T> &g = p.g;
  g.py = g.px = nullptr;
  g.rdy = rdy;
  g.rdx = rdx;
  g.ny = (int32_t)ny;
  g.nx = (int32_t)nx;
  g.numel = (int32_t)(ny * nx);
  g.pml_y0 = (int32_t)pml_y0;
  g.pml_y1 = (int32_t)pml_y1;
  g.pml_x0 = (int32_t)pml_x0;
  g.pml_x1 = (int32_t)pml_x1;
  g.pml_y1h = (int32_t)(pml_y1 > pml_y0 ? pml_y1 - 1 : pml_y0);
  g.pml_x1h = (int32_t)(pml_x1 > pml_x0 ? pml_x1 - 1 : pml_x0);
  g.ca_b = ca_b;
  g.cb_b = cb_b;
  g.cq_b = cq_b;
  p.ca = ca;
  p.cb = cb;
  p.cq = cq;
  for (int k = 0; k < 12; ++k) p.prof[k] = prof[k];
  p.n_shots = (int32_t)n_shots;
  p.n_src = n_src;
  p.n_rec = n_rec;
  return p;
}

template <typename T>
int forward_entry(T const *ca, T const *cb, T const *cq, T const *f, T *ey, T *hx, T *hz,
                  T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x, T *r, T const *ay,
                  T const *by, T const *ayh, T const *byh, T const *ax, T const *bx,
                  T const *axh, T const *bxh, T const *ky, T const *kyh, T const *kx,
                  T const *kxh, int64_t const *sources_i, int64_t const *receivers_i, T rdy,
                  T rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
                  int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b, bool cq_b,
                  int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
                  int64_t pml_x1, int64_t device, void *stream_handle, void *ey_store,
                  void *curl_store, int fmt, size_t store_step_bytes) {
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};
  Problem<T> p = make_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx, n_src, n_rec,
                                 ca_b, cb_b, cq_b, pml_y0, pml_x0, pml_y1, pml_x1);
  return forward_impl<T>(p, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, f, r, sources_i,
                         receivers_i, nt, start_t, step_ratio, ey_store, curl_store, fmt,
                         store_step_bytes, (hipStream_t)stream_handle);
}

// Shared backward body.  When gca64/gcb64 are given (FP64 accumulators, +=)
// they receive the gradient and grad_ca/grad_cb are left untouched;
// otherwise FP64 accumulators are allocated here and folded into
// grad_ca/grad_cb (+=) at the end.
template <typename T>
int backward_entry(T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey,
                   T *lambda_hx, T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z,
                   T *m_lambda_hx_z, T *m_lambda_hz_x, void *ey_store_1, void *curl_store_1,
                   T *grad_f, T *grad_ca, T *grad_cb, double *gca64, double *gcb64,
                   T const *ay, T const *by, T const *ayh, T const *byh, T const *ax,
                   T const *bx, T const *axh, T const *bxh, T const *ky, T const *kyh,
                   T const *kx, T const *kxh, int64_t const *sources_i,
                   int64_t const *receivers_i, T rdy, T rdx, int64_t nt, int64_t n_shots,
                   int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio,
                   int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
                   bool ca_rg, bool cb_rg, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,
                   int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
                   int64_t device, void *stream_handle) {
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  hipStream_t stream = (hipStream_t)stream_handle;
  const bool want = ca_rg || cb_rg;
  if (want && storage_mode != kStorageDevice) {
    set_error(storage_mode == kStorageNone
                  ? "model gradients need stored snapshots (storage_mode 'none')"
                  : "host- and disk-staged snapshot storage is not built; use "
                    "storage_mode 'device'");
    return storage_mode == kStorageNone ? kInvalidArgument : kUnsupported;
  }
  const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};
  Problem<T> p = make_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx, n_src, n_rec,
                                 ca_b, cb_b, cq_b, pml_y0, pml_x0, pml_y1, pml_x1);
  const int64_t numel = ny * nx;
  const int64_t na = ca_b ? n_shots * numel : numel;
  const int64_t nb = cb_b ? n_shots * numel : numel;
  Scratch scr(stream);
  double *ga = gca64, *gb = gcb64;
  const bool own = gca64 == nullptr && gcb64 == nullptr;
  if (own) {
    if (ca_rg) {
      TIDE_CHECK(scr.alloc(&ga, (size_t)na));
      TIDE_HIP_TRY(hipMemsetAsync(ga, 0, (size_t)na * sizeof(double), stream));
    }
    if (cb_rg) {
      TIDE_CHECK(scr.alloc(&gb, (size_t)nb));
      TIDE_HIP_TRY(hipMemsetAsync(gb, 0, (size_t)nb * sizeof(double), stream));
    }
  }
  const int fmt = storage_format == kFormatBf16 ? 2 : 1;
  TIDE_CHECK(backward_impl<T>(p, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,
                              m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, ey_store_1,
                              curl_store_1, fmt, (size_t)shot_bytes_uncomp * (size_t)n_shots,
                              grad_f, ca_rg ? ga : nullptr, cb_rg ? gb : nullptr, sources_i,
                              receivers_i, nt, start_t, step_ratio, stream));
  if (own) {
    if (ca_rg) TIDE_CHECK(fold_into<T>(grad_ca, ga, na, stream));
    if (cb_rg) TIDE_CHECK(fold_into<T>(grad_cb, gb, nb, stream));
  }
  return kOk;
}

}  // namespace

#define TIDE_TM2D_ENTRY_POINTS(T, NAME)                                                    \
  extern "C" int maxwell_tm_2_##NAME##_forward_cuda(                                      \
      T const *ca, T const *cb, T const *cq, T const *f, T *ey, T *hx, T *hz, T *m_ey_x,  \
      T *m_ey_z, T *m_hx_z, T *m_hz_x, T const *debye_a, T const *debye_b,                \
      T const *debye_cp, T *polarization, T *ey_prev, T *r, int64_t n_poles, T const *ay, \
      T const *by, T const *ayh, T const *byh, T const *ax, T const *bx, T const *axh,    \
      T const *bxh, T const *ky, T const *kyh, T const *kx, T const *kxh,                 \
      int64_t const *sources_i, int64_t const *receivers_i, T rdy, T rdx, T dt,           \
      int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,  \
      int64_t step_ratio, bool has_dispersion, bool ca_b, bool cb_b, bool cq_b,           \
      int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,    \
      int64_t n_threads, int64_t device, void *compute_stream) {                          \
    (void)debye_a; (void)debye_b; (void)debye_cp; (void)polarization;                     \
    (void)ey_prev; (void)n_poles; (void)dt; (void)n_threads;                              \
    if (has_dispersion) {                                                                 \
      set_error("Debye dispersion is not part of this build");                            \
      return kUnsupported;                                                                \
    }                                                                                     \
    return forward_entry<T>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, \
                            ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh,         \
                            sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, \
                            n_rec, step_ratio, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, \
                            pml_y1, pml_x1, device, compute_stream, nullptr, nullptr, 0,  \
                            0);                                                           \
  }                                                                                       \
                                                                                          \
  extern "C" int maxwell_tm_2_##NAME##_forward_with_storage_cuda(                         \
      T const *ca, T const *cb, T const *cq, T const *f, T *ey, T *hx, T *hz, T *m_ey_x,  \
      T *m_ey_z, T *m_hx_z, T *m_hz_x, T *r, void *ey_store_1, void *ey_store_3,          \
      char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,            \
      char const *const *curl_filenames, T const *ay, T const *by, T const *ayh,          \
      T const *byh, T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky,    \
      T const *kyh, T const *kx, T const *kxh, int64_t const *sources_i,                  \
      int64_t const *receivers_i, T rdy, T rdx, T dt, int64_t nt, int64_t n_shots,        \
      int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio,           \
      int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,            \
      bool ca_rg, bool cb_rg, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,           \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t n_threads,  \
      int64_t device, void *compute_stream, void *storage_stream) {                       \
    (void)ey_store_3; (void)ey_filenames; (void)curl_store_3;                             \
    (void)curl_filenames; (void)dt; (void)n_threads; (void)storage_stream;                \
    if (storage_mode != kStorageDevice && storage_mode != kStorageNone) {                 \
      set_error("host- and disk-staged snapshot storage is not built; use "               \
                "storage_mode 'device' or 'none'");                                       \
      return kUnsupported;                                                                \
    }                                                                                     \
    const bool st = storage_mode == kStorageDevice;                                       \
    const int fmt = storage_format == kFormatBf16 ? 2 : 1;                                \
    return forward_entry<T>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, \
                            ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh,         \
                            sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, \
                            n_rec, step_ratio, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, \
                            pml_y1, pml_x1, device, compute_stream,                       \
                            st && ca_rg ? ey_store_1 : nullptr,                           \
                            st && cb_rg ? curl_store_1 : nullptr, fmt,                    \
                            (size_t)shot_bytes_uncomp * (size_t)n_shots);                 \
  }                                                                                       \
                                                                                          \
  extern "C" int maxwell_tm_2_##NAME##_backward_cuda(                                     \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey, T *lambda_hx, \
      T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z, T *m_lambda_hx_z,                 \
      T *m_lambda_hz_x, T *adjoint_memory_scratch, void *ey_store_1, void *ey_store_3,    \
      char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,            \
      char const *const *curl_filenames, T *grad_f, T *grad_ca, T *grad_cb,               \
      T *grad_ca_shot, T *grad_cb_shot, T const *ay, T const *by, T const *ayh,           \
      T const *byh, T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky,    \
      T const *kyh, T const *kx, T const *kxh, int64_t const *sources_i,                  \
      int64_t const *receivers_i, T rdy, T rdx, T dt, int64_t nt, int64_t n_shots,        \
      int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio,           \
      int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,            \
      bool ca_rg, bool cb_rg, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,           \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t n_threads,  \
      int64_t device, void *compute_stream, void *storage_stream) {                       \
    (void)adjoint_memory_scratch; (void)ey_store_3; (void)ey_filenames;                   \
    (void)curl_store_3; (void)curl_filenames; (void)grad_ca_shot;                         \
    (void)grad_cb_shot; (void)dt; (void)n_threads; (void)storage_stream;                  \
    return backward_entry<T>(ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz,         \
                             m_lambda_ey_x, m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x,  \
                             ey_store_1, curl_store_1, grad_f, grad_ca, grad_cb, nullptr, \
                             nullptr, ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx,    \
                             kxh, sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx,  \
                             n_src, n_rec, step_ratio, storage_mode, storage_format,      \
                             shot_bytes_uncomp, ca_rg, cb_rg, ca_b, cb_b, cq_b, start_t,  \
                             pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream);     \
  }

TIDE_TM2D_ENTRY_POINTS(float, float)
TIDE_TM2D_ENTRY_POINTS(double, double)

// FP64-accumulator backward used by the Python operator (and any caller that
// runs shot chunks and wants one FP64 sum across them); see tide_hip.h.
#define TIDE_TM2D_BACKWARD_F64(T, NAME)                                                    \
  extern "C" int tide_tm2d_##NAME##_backward_f64acc(                                      \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey, T *lambda_hx, \
      T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z, T *m_lambda_hx_z,                 \
      T *m_lambda_hz_x, void const *ey_store, void const *curl_store, T *grad_f,          \
      double *grad_ca, double *grad_cb, T const *ay, T const *by, T const *ayh,           \
      T const *byh, T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky,    \
      T const *kyh, T const *kx, T const *kxh, int64_t const *sources_i,                  \
      int64_t const *receivers_i, T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t ny,  \
      int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio,                       \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_b, bool cb_b, bool cq_b, \
      int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,    \
      int64_t device, void *compute_stream) {                                             \
    const bool want = grad_ca != nullptr || grad_cb != nullptr;                           \
    return backward_entry<T>(                                                             \
        ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,               \
        m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, (void *)ey_store, (void *)curl_store, \
        grad_f, nullptr, nullptr, grad_ca, grad_cb, ay, by, ayh, byh, ax, bx, axh, bxh,   \
        ky, kyh, kx, kxh, sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src,   \
        n_rec, step_ratio, want ? kStorageDevice : kStorageNone, storage_format,          \
        shot_bytes_uncomp, grad_ca != nullptr, grad_cb != nullptr, ca_b, cb_b, cq_b,      \
        start_t, pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream);                 \
  }

TIDE_TM2D_BACKWARD_F64(float, float)
TIDE_TM2D_BACKWARD_F64(double, double)

// Launch-shape tuning (rows per wavefront, shots per workgroup) for the
// forward and adjoint step kernels; see tide_hip.h.
extern "C" int tide_tm2d_set_tuning(int fwd_rows, int fwd_group, int adj_rows, int adj_group) {
  const bool ok = (fwd_rows == 2 || fwd_rows == 4 || fwd_rows == 8 || fwd_rows == 16) &&
                  (adj_rows == 2 || adj_rows == 4 || adj_rows == 8) && fwd_group >= 1 &&
                  fwd_group <= kMaxGroup && adj_group >= 1 && adj_group <= kMaxGroup;
  if (!ok) {
    set_error("tuning: rows must be fwd {2,4,8,16}, adj {2,4,8}; groups 1..8");
    return kInvalidArgument;
  }
  g_tuning = Tuning{fwd_rows, fwd_group, adj_rows, adj_group};
  return kOk;
}

extern "C" int tide_tm2d_get_tuning(int *out4) {
  if (out4 == nullptr) {
    set_error("tuning: null output");
    return kInvalidArgument;
  }
  out4[0] = g_tuning.fwd_rows;
  out4[1] = g_tuning.fwd_group;
  out4[2] = g_tuning.adj_rows;
  out4[3] = g_tuning.adj_group;
  return kOk;
}
