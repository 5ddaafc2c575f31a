// Two-step (temporally blocked) 2-D TM step kernels.
//
// A wavefront owns kOwn2 = 60 columns (lanes 2..61) of TY rows of one shot
// and advances them by two time steps in one launch: it loads the step-t
// state of its rows and columns plus a halo of two, computes step t+1 on the
// owned cells and a halo of one (recomputing what the neighbouring tiles
// own), and step t+2 on the owned cells.  HBM sees the field state once per
// two steps; the snapshots of both steps are written (forward) or read
// (adjoint) as before.  Every cell update is the single-step kernel's
// expression in the same operation order, so the results are bit-identical
// to two single steps.
//
// CPML memories ping-pong between an input and an output buffer (the halo
// cells of a tile read the memories its neighbours update).  Tiles whose
// rows and columns (halo included) touch neither a CPML strip nor the grid
// edge take a branch-free path.
#pragma once

#include "tm2d_device.h"

namespace tide_hip {
namespace tm2d {

constexpr int kOwn2 = 60;  // owned columns per wavefront (lanes 2..61)

// The interior tiles -- those whose loads and updates (halo of two
// included) stay inside the grid and outside every CPML strip, see
// make_split -- form the rectangle of tile columns [xa, xb) and rows
// [ya, yb) (cell rows, multiples of the interior tile height).  Interior and
// border tiles run as separate launches, so the interior kernel carries none
// of the CPML and edge state (fewer live scalar registers, no spills), and
// the border kernel may use lower tiles (more wavefronts for the short, heavy
// border work).
struct TileSplit {
  int32_t xa, xb, ya, yb;
};

// Tile sets of a two-step launch.
constexpr int kInnerTiles = 0;   // the interior rectangle (branch-free kernel)
constexpr int kBorderTiles = 1;  // every other tile (x and y checks)
constexpr int kAllTiles = 2;     // all tiles, the kind chosen per tile
constexpr int kSideTiles = 3;    // border tiles beside the interior rows (x checks only)
constexpr int kBandTiles = 4;    // border tiles above and below the interior rows
constexpr int kOrderedTiles = 5; // all tiles, border tiles first (one kernel)

template <int TY>
__device__ __forceinline__ bool in_split(const TileSplit &sp, int xt, int yt) {
  return xt >= sp.xa && xt < sp.xb && yt * TY >= sp.ya && yt * TY < sp.yb;
}

// Number of TY-row tiles of a tile set (TY divides the interior tile height).
template <int TY>
__host__ __device__ __forceinline__ int64_t split_tiles(const TileSplit &sp, int n_xt, int ny, int set) {
  const int64_t iw = sp.xb - sp.xa, ih = (sp.yb - sp.ya) / TY;
  const int64_t top = (int64_t)(sp.ya / TY) * n_xt, mid = ih * (n_xt - iw);
  const int64_t bottom = (int64_t)((ny - sp.yb + TY - 1) / TY) * n_xt;
  switch (set) {
    case kInnerTiles: return iw * ih;
    case kSideTiles: return mid;
    case kBandTiles: return top + bottom;
    default: return top + mid + bottom;  // kBorderTiles
  }
}

// Tile (of height TY) and shot group of this workgroup within tile set SET
// (kInnerTiles, kBorderTiles, kSideTiles or kBandTiles).
template <int SET, int TY>
__device__ __forceinline__ TileCoord split_tile(const TileSplit &sp, int n_xt, int ny, int n_groups) {
  const int64_t n_tiles = split_tiles<TY>(sp, n_xt, ny, SET);
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, n_tiles * n_groups);
  TileCoord c;
  c.grp = (int)(flat / n_tiles);
  int i = (int)(flat - (int64_t)c.grp * n_tiles);
  const int iw = sp.xb - sp.xa;
  if (SET == kInnerTiles) {
    c.yt = sp.ya / TY + i / iw;
    c.xt = sp.xa + i % iw;
    return c;
  }
  const int top = (sp.ya / TY) * n_xt, w = n_xt - iw, mid = ((sp.yb - sp.ya) / TY) * w;
  if (SET == kSideTiles) i += top;                       // skip the top band
  if (SET == kBandTiles && i >= top) i += mid;           // skip the side tiles
  if (i < top) {
    c.yt = i / n_xt;
    c.xt = i % n_xt;
  } else if (i < top + mid) {
    i -= top;
    c.yt = sp.ya / TY + i / w;
    const int col = i % w;
    c.xt = col < sp.xa ? col : sp.xb + (col - sp.xa);
  } else {
    i -= top + mid;
    c.yt = sp.yb / TY + i / n_xt;
    c.xt = i % n_xt;
  }
  return c;
}

// Edge handling of the tiles of a set: 0 interior, 1 x only, 3 x and y.
template <int SET>
constexpr int set_edge() {
  return SET == kInnerTiles ? 0 : (SET == kSideTiles ? 1 : 3);
}

// Tile b of the ordered tile set and the edge handling it needs: border
// tiles of TYB rows first (edge 3 in the bands above and below the interior
// rows, 1 beside them), then the interior tiles of TY rows (edge 0); each set
// repeated per shot group.  With `remap`, b is a launch's workgroup id and the
// interior tiles are grouped by XCD; without it b is an index that is
// already XCD-grouped.
template <int TY, int TYB>
__device__ __forceinline__ TileCoord ordered_tile(const TileSplit &sp, int n_xt, int ny, int n_groups,
                                                  int64_t b, bool remap, int *edge) {
  const int64_t nb = split_tiles<TYB>(sp, n_xt, ny, kBorderTiles);
  const int64_t ni = split_tiles<TY>(sp, n_xt, ny, kInnerTiles);
  // indices 0 .. nb*n_groups-1 (dispatched first, spread over the XCDs):
  // border tiles; the rest: the interior
  const int64_t nbg = nb * n_groups;
  TileCoord c;
  int64_t i;
  const bool border = b < nbg;
  if (border) {
    c.grp = (int)(b / nb);
    i = b - (int64_t)c.grp * nb;
  } else {
    const int64_t f = remap ? xcd_remap(b - nbg, ni * n_groups) : b - nbg;
    c.grp = (int)(f / ni);
    i = f - (int64_t)c.grp * ni;
  }
  const int iw = sp.xb - sp.xa;
  if (!border) {
    c.yt = sp.ya / TY + (int)(i / iw);
    c.xt = sp.xa + (int)(i % iw);
    *edge = 0;
    return c;
  }
  const int top = (sp.ya / TYB) * n_xt, w = n_xt - iw, mid = ((sp.yb - sp.ya) / TYB) * w;
  int k = (int)i;
  if (k < top) {
    c.yt = k / n_xt;
    c.xt = k % n_xt;
    *edge = 3;
  } else if (k < top + mid) {
    k -= top;
    c.yt = sp.ya / TYB + k / w;
    const int col = k % w;
    c.xt = col < sp.xa ? col : sp.xb + (col - sp.xa);
    *edge = 1;
  } else {
    k -= top + mid;
    c.yt = sp.yb / TYB + k / n_xt;
    c.xt = k % n_xt;
    *edge = 3;
  }
  return c;
}

// Number of tiles of the ordered set per shot group.
template <int TY, int TYB>
__host__ __device__ __forceinline__ int64_t ordered_tiles(const TileSplit &sp, int n_xt, int ny) {
  return split_tiles<TYB>(sp, n_xt, ny, kBorderTiles) + split_tiles<TY>(sp, n_xt, ny, kInnerTiles);
}

// This workgroup's tile for tile set SET; for the one-kernel sets also the
// edge handling the tile needs.  kOrderedTiles hands out the border tiles
// (their long per-wavefront latency) before the interior ones, so the
// interior work overlaps them.
// For kOrderedTiles the border tiles are TYB rows high (TYB divides TY).
template <int SET, int TY, int TYB = TY>
__device__ __forceinline__ TileCoord pick_tile(const TileSplit &sp, int n_xt, int n_yt, int ny,
                                               int n_groups, int *edge) {
  if constexpr (SET == kAllTiles) {
    const TileCoord tc = decode_tile(n_xt, n_yt, n_groups);
    *edge = in_split<TY>(sp, tc.xt, tc.yt) ? 0 : 3;
    return tc;
  } else if constexpr (SET == kOrderedTiles) {
    return ordered_tile<TY, TYB>(sp, n_xt, ny, n_groups, (int64_t)blockIdx.x, true, edge);
  } else {
    *edge = set_edge<SET>();
    return split_tile<SET, TY>(sp, n_xt, ny, n_groups);
  }
}

template <typename T> struct Fwd2Args {
  Geom<T> g;
  Points pts;  // built with build_point_table_halo(tile_w 60, lane offset 2)
  const T *ca, *cb, *cq;
  const T *ey_i, *hx_i, *hz_i;
  T *ey_o, *hx_o, *hz_o;
  const T *mex_i, *mez_i, *mhzx_i, *mhxz_i;  // m_ey_x m_ey_z m_hz_x m_hx_z, step t
  T *mex_o, *mez_o, *mhzx_o, *mhxz_o;        // step t+2
  const T *f0, *f1;                          // source terms of steps t, t+1
  T *r0, *r1;                                // receiver rows of steps t, t+1
  void *ey_store0, *curl_store0;             // snapshots of step t (null: none)
  void *ey_store1, *curl_store1;             // snapshots of step t+1
  int32_t n_shots, n_yt;
  TileSplit split;
};

// Point-table slots of the N rows y_first .. y_first+N-1 (-1: no points, or
// a row outside the grid), read at the start of a tile so that these scalar
// loads overlap the row loads instead of stalling each row's update.
template <int N, bool GY>
__device__ __forceinline__ void tile_slots(const Points &pts, int y_first, int ny, int (&slot)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int y = y_first + k;
    slot[k] = (GY && (y < 0 || y >= ny)) ? -1 : pts.row_slot[y];
  }
}

// Applies the point entries of (row slot `slot`, shot s, tile xt) to `ey`:
// sources of this lane (owner entries and halo copies) add f[q]; receivers
// (owner entries only, and only when `record`) store ey to r[q].
template <typename T>
__device__ __forceinline__ void fused_points(const Points &pts, int slot, int n_shots, int s, int xt,
                                             int lane, const T *f, T *r, bool record, T &ey) {
  if (slot < 0) return;
  const int64_t si = ((int64_t)slot * n_shots + s) * pts.n_xt + xt;
  const int e1 = pts.seg[si + 1];
  for (int e = pts.seg[si]; e < e1; ++e) {
    const int2 en = pts.ent[e];
    if ((en.x & 0xff) != lane) continue;
    const int kind = (en.x >> 8) & 1;
    if (kind == 0) {
      if (f != nullptr) ey += f[en.y];
    } else if (record && (en.x >> 9) == 0 && r != nullptr) {
      r[en.y] = ey;
    }
  }
}

// EDGE bit 0: the tile's columns touch the grid edge or an x CPML strip;
// bit 1: its rows touch the grid edge or a y strip (per-cell checks for
// that axis).  EDGE 0 is the branch-free interior path.
template <typename T, int TY, int FMT, int EDGE>
__device__ __forceinline__ void fwd2_tile(const Fwd2Args<T> &a, int s, int xt, int yt) {
  constexpr bool GX = (EDGE & 1) != 0, GY = (EDGE & 2) != 0, GEN = EDGE != 0;
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * kOwn2 - 2 + lane;
  const int y0 = yt * TY;
  const bool own = lane >= 2 && lane <= 61 && x < nx;
  const bool e_col = !GX || (x >= 1 && x < nx);       // Ey and Hx columns
  const bool hz_col = !GX || (x >= 1 && x < nx - 1);  // Hz columns
  const bool pml_xe = GX && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = GX && (x < g.pml_x0 || x >= g.pml_x1h);
  // x profiles of this lane's column, loaded once (they do not change with
  // the row; the clamped column keeps halo lanes in bounds)
  T xb = 0, xa = 0, xk = 1, xbh = 0, xah = 0, xkh = 1;
  if constexpr (GX) {
    const int xc = min(max(x, 0), nx - 1);
    xb = g.X(kB, xc), xa = g.X(kA, xc), xk = g.X(kK, xc);
    xbh = g.X(kBh, xc), xah = g.X(kAh, xc), xkh = g.X(kKh, xc);
  }

  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ey_i = a.ey_i + base;
  const T *__restrict__ hx_i = a.hx_i + base;
  const T *__restrict__ hz_i = a.hz_i + base;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, GEN> L(x, nx, ny);
  void *const es0 = snap_plane<T, FMT>(a.ey_store0, base);
  void *const cs0 = snap_plane<T, FMT>(a.curl_store0, base);
  void *const es1 = snap_plane<T, FMT>(a.ey_store1, base);
  void *const cs1 = snap_plane<T, FMT>(a.curl_store1, base);
  int slot[TY + 2];  // rows y0-1 .. y0+TY
  tile_slots<TY + 2, GY>(a.pts, y0 - 1, ny, slot);
  LaneRows<T> yp;  // y profiles of rows y0-2 .. y0+TY+1
  if constexpr (GY) yp.load(g, y0 - 2, lane);

  // Rows: Ey y0-2..y0+TY+1, Hx and cq y0-2..y0+TY, Hz, ca and cb y0-1..y0+TY.
  constexpr int NE = TY + 4, NX = TY + 3, NZ = TY + 2;
  T ey0[NE], hx0[NX], cqv[NX], hz0[NZ], cav[NZ], cbv[NZ];
#pragma unroll
  for (int k = 0; k < NE; ++k) ey0[k] = L(ey_i, y0 - 2 + k);
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    hx0[k] = L(hx_i, y0 - 2 + k);
    cqv[k] = L(cq, y0 - 2 + k);
  }
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    hz0[k] = L(hz_i, y0 - 1 + k);
    cav[k] = L(ca, y0 - 1 + k);
    cbv[k] = L(cb, y0 - 1 + k);
  }

  // CPML memories of step t (GEN only): m_ey_z at the Hx rows, m_ey_x at the
  // Hz rows, m_hz_x and m_hx_z at the Ey rows of step t+1.  The memory
  // arrays cover the whole grid, so every row is loaded (no per-lane or
  // per-row branch around the loads, which would wait for all loads in
  // flight) and the values outside the strips are replaced by 0.
  T mez[NX], mex[NZ], mhzx[NZ], mhxz[NZ];
  if constexpr (GY) {
    const T *__restrict__ mez_i = a.mez_i + base;
    const T *__restrict__ mhxz_i = a.mhxz_i + base;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int y = y0 - 2 + k;
      const T v = L(mez_i, y);
      mez[k] = keep_or_zero(v, y < g.pml_y0 || y >= g.pml_y1h);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int y = y0 - 1 + k;
      const T v = L(mhxz_i, y);
      mhxz[k] = keep_or_zero(v, y < g.pml_y0 || y >= g.pml_y1);
    }
  }
  if constexpr (GX) {
    const T *__restrict__ mex_i = a.mex_i + base;
    const T *__restrict__ mhzx_i = a.mhzx_i + base;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int y = y0 - 1 + k;
      const T v = L(mex_i, y), w = L(mhzx_i, y);
      mex[k] = keep_or_zero(v, pml_xh);
      mhzx[k] = keep_or_zero(w, pml_xe);
    }
  }

  // ------------------------------------------------------------- step t
  // Hx^{t+1/2} rows y0-2..y0+TY.
  T hx1[NX];
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int y = y0 - 2 + k;
    hx1[k] = hx0[k];
    const bool pyh = GY && (y < g.pml_y0 || y >= g.pml_y1h);
    T pa, pb, pk;  // read with every lane active
    yp.row3(kAh, kBh, kKh, y, pyh, pa, pb, pk);
    if (e_col && (!GY || (y >= 1 && y < ny - 1))) {
      T d = (ey0[k + 1] - ey0[k]) * g.rdy;
      if (pyh) {
        const T m = pb * mez[k] + pa * d;
        mez[k] = m;
        d = d / pk + m;
      }
      hx1[k] = hx0[k] - cqv[k] * d;
    }
  }
  // Hz^{t+1/2} and Ey^{t+1} rows y0-1..y0+TY.
  T hz1[NZ], ey1[NZ];
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    const int y = y0 - 1 + k;
    const bool row = !GY || (y >= 1 && y < ny);
    const T ey_c = ey0[k + 1];
    const T ey_r = shfl_dn1(ey_c);  // unused at lane 63
    hz1[k] = hz0[k];
    if (hz_col && row) {
      T d = (ey_r - ey_c) * g.rdx;
      if (pml_xh) {
        const T m = xbh * mex[k] + xah * d;
        mex[k] = m;
        d = d / xkh + m;
      }
      hz1[k] = hz0[k] + cqv[k + 1] * d;
    }
    const T hz_l = shfl_up1(hz1[k]);  // unused at lane 0
    T ey_n = ey_c;
    const bool pye = GY && (y < g.pml_y0 || y >= g.pml_y1);
    T pa, pb, pk;  // read with every lane active
    yp.row3(kA, kB, kK, y, pye, pa, pb, pk);
    if (e_col && row) {
      T dhz = (hz1[k] - hz_l) * g.rdx;
      T dhx = (hx1[k + 1] - hx1[k]) * g.rdy;
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
        if (own && k >= 1 && k <= TY) {
          const uint32_t i = (uint32_t)(y * nx + x);
          if (es0) put_store32<T, FMT>(es0, i, ey_c);
          if (cs0) put_store32<T, FMT>(cs0, i, curl);
        }
      }
      ey_n = cav[k] * ey_c + cbv[k] * curl;
    }
    fused_points<T>(a.pts, slot[k], a.n_shots, s, xt, lane, a.f0, a.r0, own && k >= 1 && k <= TY, ey_n);
    ey1[k] = ey_n;
  }

  // ----------------------------------------------------------- step t+1
  // Hx^{t+3/2} rows y0-1..y0+TY-1 (index k <-> row y0-1+k, from Hx^{t+1/2}
  // row index k+1).
  T hx2[TY + 1];
#pragma unroll
  for (int k = 0; k < TY + 1; ++k) {
    const int y = y0 - 1 + k;
    hx2[k] = hx1[k + 1];
    const bool pyh = GY && (y < g.pml_y0 || y >= g.pml_y1h);
    T pa, pb, pk;  // read with every lane active
    yp.row3(kAh, kBh, kKh, y, pyh, pa, pb, pk);
    if (e_col && (!GY || (y >= 1 && y < ny - 1))) {
      T d = (ey1[k + 1] - ey1[k]) * g.rdy;
      if (pyh) {
        const T m = pb * mez[k + 1] + pa * d;
        mez[k + 1] = m;
        d = d / pk + m;
      }
      hx2[k] = hx1[k + 1] - cqv[k + 1] * d;
    }
  }
  T *__restrict__ ey_o = a.ey_o + base;
  T *__restrict__ hx_o = a.hx_o + base;
  T *__restrict__ hz_o = a.hz_o + base;
#pragma unroll
  for (int k = 0; k < TY; ++k) {
    const int y = y0 + k;
    if (GY && y >= ny) break;  // uniform
    const bool row = !GY || y >= 1;
    const int j = y * nx + x;
    const T ey_c = ey1[k + 1];
    const T ey_r = shfl_dn1(ey_c);
    T hz_n = hz1[k + 1];
    if (hz_col && row) {
      T d = (ey_r - ey_c) * g.rdx;
      if (pml_xh) {
        const T m = xbh * mex[k + 1] + xah * d;
        mex[k + 1] = m;
        d = d / xkh + m;
      }
      hz_n = hz1[k + 1] + cqv[k + 2] * d;
    }
    const T hz_l = shfl_up1(hz_n);
    T ey_n = ey_c;
    const bool pye = GY && (y < g.pml_y0 || y >= g.pml_y1);
    T pa, pb, pk;  // read with every lane active
    yp.row3(kA, kB, kK, y, pye, pa, pb, pk);
    if (e_col && row) {
      T dhz = (hz_n - hz_l) * g.rdx;
      T dhx = (hx2[k + 1] - hx2[k]) * g.rdy;
      if (pml_xe) {
        const T m = xb * mhzx[k + 1] + xa * dhz;
        mhzx[k + 1] = m;
        dhz = dhz / xk + m;
      }
      if (pye) {
        const T m = pb * mhxz[k + 1] + pa * dhx;
        mhxz[k + 1] = m;
        dhx = dhx / pk + m;
      }
      const T curl = dhz - dhx;
      if constexpr (FMT != 0) {
        if (own) {
          if (es1) put_store32<T, FMT>(es1, (uint32_t)j, ey_c);
          if (cs1) put_store32<T, FMT>(cs1, (uint32_t)j, curl);
        }
      }
      ey_n = cav[k + 1] * ey_c + cbv[k + 1] * curl;
    }
    fused_points<T>(a.pts, slot[k + 1], a.n_shots, s, xt, lane, a.f1, a.r1, own, ey_n);
    if (own) {
      const uint32_t o = (uint32_t)j;
      *elem32(ey_o, o) = ey_n;
      *elem32(hx_o, o) = hx2[k + 1];
      *elem32(hz_o, o) = hz_n;
      if constexpr (GY) {
        if (y < g.pml_y0 || y >= g.pml_y1h) *elem32(a.mez_o + base, o) = mez[k + 2];
        if (y < g.pml_y0 || y >= g.pml_y1) *elem32(a.mhxz_o + base, o) = mhxz[k + 1];
      }
      if constexpr (GX) {
        if (pml_xh) *elem32(a.mex_o + base, o) = mex[k + 1];
        if (pml_xe) *elem32(a.mhzx_o + base, o) = mhzx[k + 1];
      }
    }
  }
}

// SET selects the tiles (kInnerTiles, kBorderTiles, kSideTiles, kBandTiles,
// kAllTiles, kOrderedTiles); TYB is the border tile height of kOrderedTiles.
template <typename T, int TY, int FMT, int SET, int TYB = TY>
__global__ __launch_bounds__(512) void fwd2_step(Fwd2Args<T> a) {
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  int edge = 3;
  const TileCoord tc = pick_tile<SET, TY, TYB>(a.split, a.pts.n_xt, a.n_yt, a.g.ny, n_groups, &edge);
  // threadIdx.y is one value per wavefront (64-wide rows): made explicit so
  // the shot's plane pointers stay scalar.
  const int s = tc.grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
  if constexpr (SET == kAllTiles) {
    if (edge == 0)
      fwd2_tile<T, TY, FMT, 0>(a, s, tc.xt, tc.yt);
    else
      fwd2_tile<T, TY, FMT, 3>(a, s, tc.xt, tc.yt);
  } else if constexpr (SET == kOrderedTiles) {
    if (edge == 0)
      fwd2_tile<T, TY, FMT, 0>(a, s, tc.xt, tc.yt);
    else if (edge == 1)
      fwd2_tile<T, TYB, FMT, 1>(a, s, tc.xt, tc.yt);
    else
      fwd2_tile<T, TYB, FMT, 3>(a, s, tc.xt, tc.yt);
  } else {
    fwd2_tile<T, TY, FMT, set_edge<SET>()>(a, s, tc.xt, tc.yt);
  }
}

// ----------------------------------------------------------------- adjoint
template <typename T> struct Adj2Args {
  Geom<T> g;
  Points pts;  // built with build_point_table_halo(tile_w 60, lane offset 2)
  const T *ca, *cb, *cq;
  const T *ley_i, *lhx_i, *lhz_i;
  T *ley_o, *lhx_o, *lhz_o;
  // adjoint CPML memories (mx: m_hz_x, my: m_hx_z, nx: m_ey_x, nz: m_ey_z)
  const T *mx_i, *my_i, *nx_i, *nz_i;
  T *mx_o, *my_o, *nx_o, *nz_o;
  // Iteration A is the transpose of forward step t, iteration B of step t-1.
  const void *ey_store0, *curl_store0;  // snapshots of step t (null: no gradient term)
  const void *ey_store1, *curl_store1;  // snapshots of step t-1
  double *gca, *gcb;
  int32_t per_shot_grad;
  double step_scale;
  const T *grad_r0, *grad_r1;  // receiver cotangents of steps t, t-1
  T *grad_f0, *grad_f1;        // source-term gradients of steps t, t-1
  int32_t n_shots, n_yt;
  int64_t group_stride;
  TileSplit split;
};

// Applies the point entries of (row slot `slot`, shot s, tile xt) to `ley`: receivers
// of this lane (owner entries and halo copies) inject grad_r[q]; sources
// (owner entries only, and only when `record`) read grad_f[q] = ley.
template <typename T>
__device__ __forceinline__ void fused_adj_points(const Points &pts, int slot, int n_shots, int s, int xt,
                                                 int lane, const T *gr, T *gf, bool record, T &ley) {
  if (slot < 0) return;
  const int64_t si = ((int64_t)slot * n_shots + s) * pts.n_xt + xt;
  const int e0 = pts.seg[si], e1 = pts.seg[si + 1];
  if (gr != nullptr) {
    for (int e = e0; e < e1; ++e) {
      const int2 en = pts.ent[e];
      if ((en.x & 0xff) == lane && ((en.x >> 8) & 1) == 1) ley += gr[en.y];
    }
  }
  if (record && gf != nullptr) {
    for (int e = e0; e < e1; ++e) {
      const int2 en = pts.ent[e];
      if ((en.x & 0xff) == lane && (en.x >> 8) == 0) gf[en.y] = ley;
    }
  }
}

// Two reverse iterations of TY rows of one shot.  When `reduce`, the
// per-shot FP64 gradient terms of iteration A are stored to red_ca / red_cb
// ([TY][64] LDS slots per coefficient) and those of iteration B added to
// them; otherwise both go straight to the per-shot gradient arrays, A then B.
template <typename T, int TY, int FMT, int EDGE>
__device__ __forceinline__ void adj2_tile(const Adj2Args<T> &a, int s, int xt, int yt,
                                          double *red_ca, double *red_cb, bool reduce) {
  constexpr bool GX = (EDGE & 1) != 0, GY = (EDGE & 2) != 0, GEN = EDGE != 0;
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * kOwn2 - 2 + lane;
  const int y0 = yt * TY;
  const bool own = lane >= 2 && lane <= 61 && x < nx;
  const bool e_col = !GX || (x >= 1 && x < nx);
  const bool pml_xe = GX && (x < g.pml_x0 || x >= g.pml_x1);
  const bool pml_xh = GX && (x < g.pml_x0 || x >= g.pml_x1h);
  T xb = 0, xa = 0, xk = 1, xbh = 0, xah = 0, xkh = 1;  // this column's x profiles
  if constexpr (GX) {
    const int xc = min(max(x, 0), nx - 1);
    xb = g.X(kB, xc), xa = g.X(kA, xc), xk = g.X(kK, xc);
    xbh = g.X(kBh, xc), xah = g.X(kAh, xc), xkh = g.X(kKh, xc);
  }
  const bool grad_a = FMT != 0 && (a.ey_store0 != nullptr || a.curl_store0 != nullptr);
  const bool grad_b = FMT != 0 && (a.ey_store1 != nullptr || a.curl_store1 != nullptr);

  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const RowLoader<T, GEN> L(x, nx, ny);
  const void *const es0 = snap_plane<T, FMT>(a.ey_store0, base);
  const void *const cs0 = snap_plane<T, FMT>(a.curl_store0, base);
  const void *const es1 = snap_plane<T, FMT>(a.ey_store1, base);
  const void *const cs1 = snap_plane<T, FMT>(a.curl_store1, base);
  int slot[TY + 2];  // rows y0-1 .. y0+TY
  tile_slots<TY + 2, GY>(a.pts, y0 - 1, ny, slot);
  LaneRows<T> yp;  // y profiles of rows y0-2 .. y0+TY+1
  if constexpr (GY) yp.load(g, y0 - 2, lane);

  // Rows: lambda_Ey and cb y0-2..y0+TY+1, lambda_Hx and cq y0-2..y0+TY,
  // lambda_Hz and ca y0-1..y0+TY.
  constexpr int NE = TY + 4, NX = TY + 3, NZ = TY + 2;
  T ley0[NE], cbv[NE], lhx0[NX], cqv[NX], lhz0[NZ], cav[NZ];
  {
    const T *__restrict__ ley_i = a.ley_i + base;
    const T *__restrict__ lhx_i = a.lhx_i + base;
    const T *__restrict__ lhz_i = a.lhz_i + base;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      ley0[k] = L(ley_i, y0 - 2 + k);
      cbv[k] = L(cb, y0 - 2 + k);
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      lhx0[k] = L(lhx_i, y0 - 2 + k);
      cqv[k] = L(cq, y0 - 2 + k);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      lhz0[k] = L(lhz_i, y0 - 1 + k);
      cav[k] = L(ca, y0 - 1 + k);
    }
  }
  // Snapshots of the owned rows (exact in fp32 for fp32 / BF16 payloads).
  // They are loaded without per-lane or per-buffer branches (an absent
  // buffer reads this shot's lambda_Ey plane, a lane outside the grid its
  // clamped column; keep_or_zero clears those values) so that the loads
  // issue with the row loads above and complete under one wait.
  float se0[TY], sc0[TY], se1[TY], sc1[TY];
#pragma unroll
  for (int k = 0; k < TY; ++k) {
    se0[k] = sc0[k] = se1[k] = sc1[k] = 0.0f;
    if constexpr (sizeof(T) == 4 && FMT != 0) {
      const int y = y0 + k;
      const void *const dflt = a.ley_i + base;
      const int xc = GX ? min(max(x, 0), nx - 1) : x;
      const int yc = GY ? min(y, ny - 1) : y;
      const uint32_t i = (uint32_t)(yc * nx + xc);
      const float v0 = (float)get_store32<T, FMT>(es0 ? es0 : dflt, i);
      const float v1 = (float)get_store32<T, FMT>(cs0 ? cs0 : dflt, i);
      const float v2 = (float)get_store32<T, FMT>(es1 ? es1 : dflt, i);
      const float v3 = (float)get_store32<T, FMT>(cs1 ? cs1 : dflt, i);
      const bool in = (!GX || (x >= 1 && x < nx)) && (!GY || y < ny);
      se0[k] = keep_or_zero(v0, grad_a && es0 && in);
      sc0[k] = keep_or_zero(v1, grad_a && cs0 && in);
      se1[k] = keep_or_zero(v2, grad_b && es1 && in);
      sc1[k] = keep_or_zero(v3, grad_b && cs1 && in);
    }
  }
  // CPML memories (GEN only): my and nz indexed from row y0-2, mx and nx
  // from row y0-1.  Loaded for every row and lane (the arrays cover the
  // grid; no branch around the loads), zero outside the strips.
  T my[NE], nz[NX], mx[NZ], nxm[NZ];
  if constexpr (GY) {
    const T *__restrict__ my_i = a.my_i + base;
    const T *__restrict__ nz_i = a.nz_i + base;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int y = y0 - 2 + k;
      const T v = L(my_i, y);
      my[k] = keep_or_zero(v, y < g.pml_y0 || y >= g.pml_y1);
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int y = y0 - 2 + k;
      const T v = L(nz_i, y);
      nz[k] = keep_or_zero(v, y < g.pml_y0 || y >= g.pml_y1h);
    }
  }
  if constexpr (GX) {
    const T *__restrict__ mx_i = a.mx_i + base;
    const T *__restrict__ nx_i = a.nx_i + base;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int y = y0 - 1 + k;
      const T v = L(mx_i, y), w = L(nx_i, y);
      mx[k] = keep_or_zero(v, pml_xe);
      nxm[k] = keep_or_zero(w, pml_xh);
    }
  }

  // The single-step kernel's terms (adj_tile), each also advancing its CPML
  // memory `m` where it is updated.
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
  auto grad_terms = [&](bool on, int y, T lm_t, float se, float sc, const void *es, const void *cs,
                        int k, bool add) {
    if (!on || !own || !in_e(y)) return;
    const double lm = (double)lm_t;
    double ve, vc;
    if constexpr (sizeof(T) == 4) {
      ve = (double)se;
      vc = (double)sc;
    } else {
      const uint32_t i = (uint32_t)(y * nx + x);
      ve = es ? get_store32<T, FMT>(es, i) : 0.0;
      vc = cs ? get_store32<T, FMT>(cs, i) : 0.0;
    }
    const double pa = lm * ve * a.step_scale;
    const double pb = lm * vc * a.step_scale;
    if (reduce) {
      double *const ra = red_ca + k * 64 + lane;
      double *const rb = red_cb + k * 64 + lane;
      *ra = add ? *ra + pa : pa;
      *rb = add ? *rb + pb : pb;
    } else {
      const int64_t gb = a.per_shot_grad ? base : 0;
      const uint32_t gi = (uint32_t)(y * nx + x);
      if (a.gca && es) *elem32(a.gca + gb, gi) += pa;
      if (a.gcb && cs) *elem32(a.gcb + gb, gi) += pb;
    }
  };

  // ---------------------------------------------------- iteration A (t)
  T nrA[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) nrA[k] = nraw_y(y0 - 2 + k, ley0[k], cbv[k], my[k]);
  T lhxA[NX], rzA[NX];
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int y = y0 - 2 + k;
    lhxA[k] = lhx0[k];
    if (in_e(y)) lhxA[k] = lhx0[k] + (nrA[k + 1] - nrA[k]) * g.rdy;
    rzA[k] = raw_z(y, lhxA[k], cqv[k], nz[k]);
  }
  T lhzA[NZ], leyA[NZ];
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    const int y = y0 - 1 + k;
    const T ley_c = ley0[k + 1];
    const T rx = raw_x(y, ley_c, cbv[k + 1], mx[k]);
    const T rx_r = shfl_dn1(rx);  // unused at lane 63
    lhzA[k] = lhz0[k];
    if (in_e(y)) lhzA[k] = lhz0[k] + (rx - rx_r) * g.rdx;
    const T rxh = raw_xh(y, lhzA[k], cqv[k + 1], nxm[k]);
    const T rxh_l = shfl_up1(rxh);  // unused at lane 0
    T ley_n = ley_c;
    if (in_e(y)) {
      const T t1 = cav[k] * ley_c;
      const T t2 = (rzA[k] - rzA[k + 1]) * g.rdy + (rxh_l - rxh) * g.rdx;
      ley_n = t1 + t2;
    }
    const bool owned_row = k >= 1 && k <= TY;
    fused_adj_points<T>(a.pts, slot[k], a.n_shots, s, xt, lane, a.grad_r0, a.grad_f0, own && owned_row, ley_n);
    if (owned_row)
      grad_terms(grad_a, y, ley_n, se0[k - 1], sc0[k - 1], es0, cs0, k - 1, false);
    leyA[k] = ley_n;
  }

  // -------------------------------------------------- iteration B (t-1)
  T nrB[TY + 2];
#pragma unroll
  for (int k = 0; k < TY + 2; ++k) nrB[k] = nraw_y(y0 - 1 + k, leyA[k], cbv[k + 1], my[k + 1]);
  T lhxB[TY + 1], rzB[TY + 1];
#pragma unroll
  for (int k = 0; k < TY + 1; ++k) {
    const int y = y0 - 1 + k;
    lhxB[k] = lhxA[k + 1];
    if (in_e(y)) lhxB[k] = lhxA[k + 1] + (nrB[k + 1] - nrB[k]) * g.rdy;
    rzB[k] = raw_z(y, lhxB[k], cqv[k + 1], nz[k + 1]);
  }
  T *__restrict__ ley_o = a.ley_o + base;
  T *__restrict__ lhx_o = a.lhx_o + base;
  T *__restrict__ lhz_o = a.lhz_o + base;
#pragma unroll
  for (int k = 0; k < TY; ++k) {
    const int y = y0 + k;
    if (GY && y >= ny) break;  // uniform
    const int j = y * nx + x;
    const T ley_c = leyA[k + 1];
    const T rx = raw_x(y, ley_c, cbv[k + 2], mx[k + 1]);
    const T rx_r = shfl_dn1(rx);
    T lhz_n = lhzA[k + 1];
    if (in_e(y)) lhz_n = lhzA[k + 1] + (rx - rx_r) * g.rdx;
    const T rxh = raw_xh(y, lhz_n, cqv[k + 2], nxm[k + 1]);
    const T rxh_l = shfl_up1(rxh);
    T ley_n = ley_c;
    if (own && in_e(y)) {
      const T t1 = cav[k + 1] * ley_c;
      const T t2 = (rzB[k] - rzB[k + 1]) * g.rdy + (rxh_l - rxh) * g.rdx;
      ley_n = t1 + t2;
    }
    fused_adj_points<T>(a.pts, slot[k + 1], a.n_shots, s, xt, lane, a.grad_r1, a.grad_f1, own, ley_n);
    grad_terms(grad_b, y, ley_n, se1[k], sc1[k], es1, cs1, k, true);
    if (own) {
      const uint32_t o = (uint32_t)j;
      *elem32(ley_o, o) = ley_n;
      *elem32(lhx_o, o) = lhxB[k + 1];
      *elem32(lhz_o, o) = lhz_n;
      if constexpr (GEN) {
        if (in_e(y)) {
          if constexpr (GX) {
            if (pml_xe) *elem32(a.mx_o + base, o) = mx[k + 1];
            if (x < nx - 1 && pml_xh) *elem32(a.nx_o + base, o) = nxm[k + 1];
          }
          if constexpr (GY) {
            if (y < g.pml_y0 || y >= g.pml_y1) *elem32(a.my_o + base, o) = my[k + 2];
            if (y < ny - 1 && (y < g.pml_y0 || y >= g.pml_y1h)) *elem32(a.nz_o + base, o) = nz[k + 2];
          }
        }
      }
    }
  }
}

// One two-step adjoint workgroup; `red` is the dynamic LDS block
// [2 coefficients][G][TY][64].  Each slot holds one shot's terms of both
// iterations; the shots' slots are summed in shot order and added to the
// cell once per launch (FP64 throughout, so the result differs from two
// single steps only by the rounding of the summation order).
// ALLOW: bit k set = edge variant k (0 interior, 1 x checks, 3 x and y
// checks: bit 3) may occur; `edge` picks one at run time.
template <typename T, int TY, int FMT, int ALLOW>
__device__ __forceinline__ void adj2_workgroup(const Adj2Args<T> &a, const TileCoord &tc, int edge,
                                               double *red) {
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  const int G = (int)blockDim.y;
  const int s = tc.grp * G + w;
  const bool any_grad = FMT != 0 && (a.gca != nullptr || a.gcb != nullptr);
  const bool reduce = any_grad && !a.per_shot_grad;
  const int64_t slab = (int64_t)G * TY * 64;  // one [G][TY][64] block
  double *red_ca = red + (int64_t)w * TY * 64;
  double *red_cb = red_ca + slab;
  if (reduce) {
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      red_ca[r * 64 + lane] = 0.0;
      red_cb[r * 64 + lane] = 0.0;
    }
  }
  // The group's accumulators of this wavefront's first reduction row (row
  // w) are read now, so the final add after the barrier waits on no load.
  // Only this workgroup writes them.
  const int x = tc.xt * kOwn2 - 2 + lane;
  const int y0 = tc.yt * TY;
  const bool red_lane = lane >= 2 && lane <= 61 && x >= 1 && x < a.g.nx;
  const bool step_a = a.ey_store0 != nullptr || a.curl_store0 != nullptr;
  const bool step_b = a.ey_store1 != nullptr || a.curl_store1 != nullptr;
  double pre[2] = {0.0, 0.0};
  if (reduce && red_lane && w < TY && y0 + w >= 1 && y0 + w < a.g.ny) {
    const uint32_t j = (uint32_t)((y0 + w) * a.g.nx + x);
    if (a.gca && (a.ey_store0 || a.ey_store1)) pre[0] = *elem32(a.gca + tc.grp * a.group_stride, j);
    if (a.gcb && (a.curl_store0 || a.curl_store1)) pre[1] = *elem32(a.gcb + tc.grp * a.group_stride, j);
  }
  if (s < a.n_shots) {  // inactive wavefronts still join the barrier below
    constexpr bool a0 = ALLOW & 1, a1 = ALLOW & 2, a3 = ALLOW & 8;
    if (a0 && (!(a1 || a3) || edge == 0)) {
      if constexpr (a0) adj2_tile<T, TY, FMT, 0>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    } else if (a1 && (!a3 || edge == 1)) {
      if constexpr (a1) adj2_tile<T, TY, FMT, 1>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    } else {
      if constexpr (a3) adj2_tile<T, TY, FMT, 3>(a, s, tc.xt, tc.yt, red_ca, red_cb, reduce);
    }
  }
  if (!reduce) return;  // uniform over the workgroup
  __syncthreads();
  if (!red_lane) return;
  for (int pass = 0; pass < 2; ++pass) {
    double *out = pass == 0 ? a.gca : a.gcb;
    const void *sa = pass == 0 ? a.ey_store0 : a.curl_store0;
    const void *sb = pass == 0 ? a.ey_store1 : a.curl_store1;
    if (out == nullptr || (!sa && !sb)) continue;
    out += tc.grp * a.group_stride;
    if (!((step_a && sa) || (step_b && sb))) continue;
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

template <typename T, int TY, int FMT, int SET, int TYB = TY>
__global__ __launch_bounds__(512) void adj2_step(Adj2Args<T> a) {
  extern __shared__ double red[];
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  int edge = 3;
  const TileCoord tc = pick_tile<SET, TY, TYB>(a.split, a.pts.n_xt, a.n_yt, a.g.ny, n_groups, &edge);
  if constexpr (SET == kAllTiles) {
    adj2_workgroup<T, TY, FMT, 1 | 8>(a, tc, edge, red);
  } else if constexpr (SET == kOrderedTiles && TYB == TY) {
    adj2_workgroup<T, TY, FMT, 1 | 2 | 8>(a, tc, edge, red);
  } else if constexpr (SET == kOrderedTiles) {
    if (edge == 0)  // uniform over the workgroup
      adj2_workgroup<T, TY, FMT, 1>(a, tc, edge, red);
    else
      adj2_workgroup<T, TYB, FMT, 2 | 8>(a, tc, edge, red);
  } else {
    constexpr int e = set_edge<SET>();
    adj2_workgroup<T, TY, FMT, e == 0 ? 1 : (e == 1 ? 2 : 8)>(a, tc, edge, red);
  }
}

// Host side of the split: the interior rectangle for TY-row tiles (rows
// returned in cells).  A tile
// is interior when its columns x0-2..x0+61 and rows y0-2..y0+TY+1 lie in
// [1, n-2] and outside the CPML strips (H-phase bounds, which contain the
// E-phase ones); the test separates into a column and a row condition.
// Empty when no tile qualifies.
template <typename T>
inline TileSplit make_split(const Geom<T> &g, int n_xt, int n_yt, int ty) {
  auto col_ok = [&](int xt) {
    const int x0 = xt * kOwn2 - 2;
    return x0 >= 1 && x0 + 63 <= g.nx - 2 && x0 >= g.pml_x0 && x0 + 63 < g.pml_x1h;
  };
  auto row_ok = [&](int yt) {
    const int y0 = yt * ty;
    return y0 - 2 >= 1 && y0 + ty + 1 <= g.ny - 2 && y0 - 2 >= g.pml_y0 && y0 + ty < g.pml_y1h;
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

}  // namespace tm2d
}  // namespace tide_hip
