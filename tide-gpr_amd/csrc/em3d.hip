// 3-D Maxwell FDTD propagator for MI355X (gfx950), stencil order 2.
//
// Two kernels per forward step (H update, then E update) and two per
// adjoint step (transpose of the E update, then of the H update).  Each
// kernel only writes the field family it updates and reads the other, so
// all of them work in place; the CPML adjoint memories, which neighbours
// read while the owner writes, ping-pong between two buffers.
//
// Layout: fields are [n_shots, nz, ny, nx] (x fastest).  A workgroup is one
// (62-column x tile, y row, z block) with one wavefront per shot; the 64
// lanes cover columns x0-1 .. x0+62 (lanes 1..62 own their cells), the
// x-neighbour comes from one cross-lane shuffle, the z-neighbour is carried
// in registers while the wavefront marches through its z block, and the
// y-neighbour row is read directly (it is the neighbouring workgroup's own
// row, so it comes from L2).  Coefficients are read by every wavefront of a
// workgroup from the same lines.  Sources, receivers, the adjoint injection
// and the FP64 coefficient gradient (summed over the workgroup's shots in
// LDS in a fixed order) are handled by the thread that owns the cell.
//
// Semantics follow the reference propagator (paths relative to src/tide/):
//   H / E updates     csrc/em3d/maxwell_3d_inst.cpp:541-670 (CPU loops),
//                     storing E and curl H before the update :1000-1061
//   point ops         csrc/em3d/maxwell_3d_inst.cpp:300-348
//   gradient          csrc/em3d/maxwell_3d_inst.cpp:221-298
//   exact adjoint     csrc/em3d/maxwell_3d_cuda_inst.cu:1420-2187
// with the same floating-point operation order (no FMA contraction).

#include "tide_common.h"
#include "tide_hip.h"

#include <algorithm>
#include <cstring>

namespace tide_hip {
namespace em3d {

constexpr int kOwn = 62;
constexpr int kMaxGroup = 8;
enum { kA = 0, kB = 1, kK = 2, kAh = 3, kBh = 4, kKh = 5 };
enum { EX = 0, EY, EZ, HX, HY, HZ,
       M_HZ_Y, M_HY_Z, M_HX_Z, M_HZ_X, M_HY_X, M_HX_Y,
       M_EY_Z, M_EZ_Y, M_EZ_X, M_EX_Z, M_EX_Y, M_EY_X, N_SLOTS };

// CPML profiles packed per axis as [a, b, k, ah, bh, kh] x n.
template <typename T> struct Geom3 {
  const T *pz, *py, *px;
  T rdz, rdy, rdx;
  int32_t nz, ny, nx, numel;
  int32_t z0, z1, z1h, y0, y1, y1h, x0, x1, x1h;
  int32_t ca_b, cb_b, cq_b;
  __device__ __forceinline__ T Z(int k, int z) const { return pz[k * nz + z]; }
  __device__ __forceinline__ T Y(int k, int y) const { return py[k * ny + y]; }
  __device__ __forceinline__ T X(int k, int x) const { return px[k * nx + x]; }
};

struct Points {
  const int32_t *row_slot;  // indexed by line z*ny + y
  const int32_t *seg;
  const int2 *ent;
  int32_t n_xt;
};

template <typename T> struct Args3 {
  Geom3<T> g;
  Points pts;
  const T *ca, *cb, *cq;
  T *f[N_SLOTS];        // forward state, or adjoint state (same slots)
  T *mn[N_SLOTS];       // adjoint: next CPML memories (memory slots only)
  void *st[6];          // E x/y/z, curl x/y/z snapshots of this step
  const T *src_amp;     // f[t] (forward) or grad_r[t] (adjoint)
  T *rec_out;           // r[t] (forward) or grad_f[t] (adjoint)
  double *gca, *gcb;
  double step_scale;
  int32_t per_shot_grad;
  int32_t n_shots, n_zb, src_comp, rec_comp;
  int64_t group_stride;  // shot group g adds into gca/gcb + g*group_stride
  int32_t shot_fast;     // tile order (decode3)
  // Born scattered E update (e3_step<.., BORN>): dca, dcb [nz, ny, nx] and
  // the background's pre-update E x/y/z and curl H x/y/z of this step,
  // indexed like the fields
  const T *dca, *dcb;
  const T *bg[6];
};

// Snapshots are written once and read once, much later: non-temporal stores
// keep them from evicting the fields.
template <typename T, int FMT>
__device__ __forceinline__ void put_store(void *base, int64_t i, T v) {
  if constexpr (FMT == 2) {
    __builtin_nontemporal_store(f32_to_bf16((float)v), (uint16_t *)base + i);
  } else {
    __builtin_nontemporal_store(v, (T *)base + i);
  }
}

// The adjoint reads each snapshot once: non-temporal loads.
template <typename T, int FMT>
__device__ __forceinline__ double get_store(const void *base, int64_t i) {
  if constexpr (FMT == 2) {
    return (double)bf16_to_f32(__builtin_nontemporal_load((const uint16_t *)base + i));
  } else {
    return (double)__builtin_nontemporal_load((const T *)base + i);
  }
}

struct Tile3 {
  int xt, y, zb, grp;
};

// Tile and shot group of this workgroup.  Shot-major order (shot_fast 0)
// puts all tiles of one shot group before the next group; shot-minor order
// (shot_fast 1) puts the shot groups of one tile next to each other, so they
// run on the same XCD at about the same time and share its L2 for the
// coefficient lines and the neighbouring rows.
__device__ __forceinline__ Tile3 decode3(int n_xt, int ny, int n_zb, int n_groups,
                                         int shot_fast) {
  const int64_t n_tiles = (int64_t)n_xt * ny * n_zb;
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, n_tiles * n_groups);
  Tile3 c;
  int64_t t;
  if (shot_fast) {
    t = flat / n_groups;
    c.grp = (int)(flat - t * n_groups);
  } else {
    c.grp = (int)(flat / n_tiles);
    t = flat - (int64_t)c.grp * n_tiles;
  }
  c.zb = (int)(t / ((int64_t)n_xt * ny));
  t -= (int64_t)c.zb * n_xt * ny;
  c.y = (int)(t / n_xt);
  c.xt = (int)(t - (int64_t)c.y * n_xt);
  return c;
}

// Reads p[z, y, x] of one shot volume (p wave-uniform); zero outside the
// grid.  Branch-free (clamped line and column, then a mask), so a wavefront
// issues all its loads before waiting on any; the line start is a uniform
// 64-bit pointer and the column a 32-bit lane offset.
template <typename T> struct Ld3 {
  int xc, nx, ny, nz;
  bool xv;
  __device__ __forceinline__ Ld3(int x, int nx_, int ny_, int nz_) : nx(nx_), ny(ny_), nz(nz_) {
    xv = x >= 0 && x < nx_;
    xc = min(max(x, 0), nx_ - 1);
  }
  __device__ __forceinline__ T operator()(const T *__restrict__ p, int z, int y) const {
    const int zc = min(max(z, 0), nz - 1), yc = min(max(y, 0), ny - 1);
    const T v = *elem32(p + (int64_t)(zc * ny + yc) * nx, (uint32_t)xc);
    return keep_or_zero(v, xv && z == zc && y == yc);
  }
  // Element x of line (z, y), both inside the grid, at the clamped column.
  __device__ __forceinline__ T at(const T *__restrict__ p, int z, int y) const {
    return *elem32(p + (int64_t)(z * ny + y) * nx, (uint32_t)xc);
  }
};

// CPML correction of a raw difference d with memory m (updated in place):
// m = b m + a d; d = d / k + m.
template <typename T>
__device__ __forceinline__ T cpml(T d, T *__restrict__ m, int j, T a, T b, T k, bool own) {
  const T mv = b * m[j] + a * d;
  if (own) m[j] = mv;
  return d / k + mv;
}

// ----------------------------------------------------------- forward: H
template <typename T, int TZ>
__global__ __launch_bounds__(512) void h3_step(Args3<T> a) {
  const Geom3<T> &g = a.g;
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const Tile3 tc = decode3(a.pts.n_xt, g.ny, a.n_zb, n_groups, a.shot_fast);
  // threadIdx.y is one value per wavefront (64-wide rows): made explicit so
  // the shot's volume pointers stay scalar.
  const int s = tc.grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (s >= a.n_shots) return;
  const int nz = g.nz, ny = g.ny, nx = g.nx, y = tc.y;
  const int lane = threadIdx.x;
  const int x = tc.xt * kOwn - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && x >= 1 && x < nx && y >= 1;
  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ex = a.f[EX] + base;
  const T *__restrict__ ey = a.f[EY] + base;
  const T *__restrict__ ez = a.f[EZ] + base;
  T *__restrict__ hx = a.f[HX] + base;
  T *__restrict__ hy = a.f[HY] + base;
  T *__restrict__ hz = a.f[HZ] + base;
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const Ld3<T> L(x, nx, ny, nz);
  const bool has_dy = y < ny - 1, has_dx = x < nx - 1;
  const bool pyh = y < g.y0 || y >= g.y1h;
  const bool pxh = x < g.x0 || x >= g.x1h;
  const int zs = tc.zb * TZ;

  // Inputs of plane z: E at z + 1 (z-difference) and at y + 1, and this
  // cell's coefficient and H.  Plane z + 1's inputs are loaded before plane
  // z's stores (software pipelining: the compiler cannot move the loads
  // above stores that might alias them).
  struct In {
    T exn, eyn, ezn, ex_yp, ez_yp, cqv, hx0, hy0, hz0;
  };
  auto load_in = [&](int z) {
    In v;
    const int zc = min(z, nz - 1);
    v.exn = L(ex, z + 1, y);
    v.eyn = L(ey, z + 1, y);
    v.ezn = L(ez, z + 1, y);
    v.ex_yp = L(ex, zc, y + 1);
    v.ez_yp = L(ez, zc, y + 1);
    v.cqv = L.at(cq, zc, y);
    v.hx0 = L.at(hx, zc, y);
    v.hy0 = L.at(hy, zc, y);
    v.hz0 = L.at(hz, zc, y);
    return v;
  };
  T exc = L(ex, zs, y), eyc = L(ey, zs, y), ezc = L(ez, zs, y);
  In next = load_in(zs);
#pragma unroll
  for (int r = 0; r < TZ; ++r) {
    const int z = zs + r;
    if (z >= nz) break;  // uniform
    const In in = next;
    if (r + 1 < TZ) next = load_in(z + 1);
    const T exn = in.exn, eyn = in.eyn, ezn = in.ezn, ex_yp = in.ex_yp, ez_yp = in.ez_yp;
    const T cqv = in.cqv, hx0 = in.hx0, hy0 = in.hy0, hz0 = in.hz0;
    const T ez_xp = shfl_dn1(ezc), ey_xp = shfl_dn1(eyc);
    if (own && z >= 1) {
      const int j = (z * ny + y) * nx + x;
      const bool has_dz = z < nz - 1;
      const bool pzh = z < g.z0 || z >= g.z1h;
      T dey_dz = (T)0, dez_dy = (T)0, dez_dx = (T)0, dex_dz = (T)0, dex_dy = (T)0, dey_dx = (T)0;
      if (has_dz) {
        dey_dz = (eyn - eyc) * g.rdz;
        dex_dz = (exn - exc) * g.rdz;
        if (pzh) {
          const T az = g.Z(kAh, z), bz = g.Z(kBh, z), kz = g.Z(kKh, z);
          dey_dz = cpml(dey_dz, a.f[M_EY_Z] + base, j, az, bz, kz, true);
          dex_dz = cpml(dex_dz, a.f[M_EX_Z] + base, j, az, bz, kz, true);
        }
      }
      if (has_dy) {
        dez_dy = (ez_yp - ezc) * g.rdy;
        dex_dy = (ex_yp - exc) * g.rdy;
        if (pyh) {
          const T ay = g.Y(kAh, y), by = g.Y(kBh, y), ky = g.Y(kKh, y);
          dez_dy = cpml(dez_dy, a.f[M_EZ_Y] + base, j, ay, by, ky, true);
          dex_dy = cpml(dex_dy, a.f[M_EX_Y] + base, j, ay, by, ky, true);
        }
      }
      if (has_dx) {
        dez_dx = (ez_xp - ezc) * g.rdx;
        dey_dx = (ey_xp - eyc) * g.rdx;
        if (pxh) {
          const T ax = g.X(kAh, x), bx = g.X(kBh, x), kx = g.X(kKh, x);
          dez_dx = cpml(dez_dx, a.f[M_EZ_X] + base, j, ax, bx, kx, true);
          dey_dx = cpml(dey_dx, a.f[M_EY_X] + base, j, ax, bx, kx, true);
        }
      }
      const int64_t line = (int64_t)(z * ny + y) * nx;
      *elem32(hx + line, (uint32_t)x) = hx0 - cqv * (dey_dz - dez_dy);
      *elem32(hy + line, (uint32_t)x) = hy0 - cqv * (dez_dx - dex_dz);
      *elem32(hz + line, (uint32_t)x) = hz0 - cqv * (dex_dy - dey_dx);
    }
    exc = exn;
    eyc = eyn;
    ezc = ezn;
  }
}

// ----------------------------------------------------------- forward: E
// BORN: the scattered field of the Born forward, updated as
// ca E + cb curl + dca E_bg + dcb curl_bg in that order (reference
// maxwell_3d_inst.cpp:1444-1456).
template <typename T, int TZ, int FMT, bool BORN = false>
__global__ __launch_bounds__(512) void e3_step(Args3<T> a) {
  const Geom3<T> &g = a.g;
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const Tile3 tc = decode3(a.pts.n_xt, g.ny, a.n_zb, n_groups, a.shot_fast);
  const int s = tc.grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  if (s >= a.n_shots) return;
  const int nz = g.nz, ny = g.ny, nx = g.nx, y = tc.y;
  const int lane = threadIdx.x;
  const int x = tc.xt * kOwn - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && x >= 1 && x < nx && y >= 1;
  const int64_t base = (int64_t)s * g.numel;
  T *__restrict__ ex = a.f[EX] + base;
  T *__restrict__ ey = a.f[EY] + base;
  T *__restrict__ ez = a.f[EZ] + base;
  const T *__restrict__ hx = a.f[HX] + base;
  const T *__restrict__ hy = a.f[HY] + base;
  const T *__restrict__ hz = a.f[HZ] + base;
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const Ld3<T> L(x, nx, ny, nz);
  const bool py = y < g.y0 || y >= g.y1;
  const bool px = x < g.x0 || x >= g.x1;
  const int zs = tc.zb * TZ;

  // Inputs of plane z: H at z and y - 1, and this cell's E and
  // coefficients; plane z + 1's are loaded before plane z's stores.
  struct In {
    T hxc, hyc, hzc, hz_ym, hx_ym, ex0, ey0, ez0, cav, cbv;
  };
  auto load_in = [&](int z) {
    In v;
    const int zc = min(z, nz - 1);
    v.hxc = L(hx, z, y);
    v.hyc = L(hy, z, y);
    v.hzc = L(hz, z, y);
    v.hz_ym = L(hz, z, y - 1);
    v.hx_ym = L(hx, z, y - 1);
    v.ex0 = L.at(ex, zc, y);
    v.ey0 = L.at(ey, zc, y);
    v.ez0 = L.at(ez, zc, y);
    v.cav = L.at(ca, zc, y);
    v.cbv = L.at(cb, zc, y);
    return v;
  };
  T hx_zm = L(hx, zs - 1, y), hy_zm = L(hy, zs - 1, y);
  In next = load_in(zs);
#pragma unroll
  for (int r = 0; r < TZ; ++r) {
    const int z = zs + r;
    if (z >= nz) break;
    const In in = next;
    if (r + 1 < TZ) next = load_in(z + 1);
    const T hxc = in.hxc, hyc = in.hyc, hzc = in.hzc, hz_ym = in.hz_ym, hx_ym = in.hx_ym;
    const T ex0 = in.ex0, ey0 = in.ey0, ez0 = in.ez0, cav = in.cav, cbv = in.cbv;
    const T hz_xm = shfl_up1(hzc), hy_xm = shfl_up1(hyc);
    if (own && z >= 1) {
      const int j = (z * ny + y) * nx + x;
      const bool pz = z < g.z0 || z >= g.z1;
      T dhy_dz = (hyc - hy_zm) * g.rdz;
      T dhz_dy = (hzc - hz_ym) * g.rdy;
      T dhz_dx = (hzc - hz_xm) * g.rdx;
      T dhx_dz = (hxc - hx_zm) * g.rdz;
      T dhx_dy = (hxc - hx_ym) * g.rdy;
      T dhy_dx = (hyc - hy_xm) * g.rdx;
      if (pz) {
        const T az = g.Z(kA, z), bz = g.Z(kB, z), kz = g.Z(kK, z);
        dhy_dz = cpml(dhy_dz, a.f[M_HY_Z] + base, j, az, bz, kz, true);
        dhx_dz = cpml(dhx_dz, a.f[M_HX_Z] + base, j, az, bz, kz, true);
      }
      if (py) {
        const T ay = g.Y(kA, y), by = g.Y(kB, y), ky = g.Y(kK, y);
        dhz_dy = cpml(dhz_dy, a.f[M_HZ_Y] + base, j, ay, by, ky, true);
        dhx_dy = cpml(dhx_dy, a.f[M_HX_Y] + base, j, ay, by, ky, true);
      }
      if (px) {
        const T ax = g.X(kA, x), bx = g.X(kB, x), kx = g.X(kK, x);
        dhz_dx = cpml(dhz_dx, a.f[M_HZ_X] + base, j, ax, bx, kx, true);
        dhy_dx = cpml(dhy_dx, a.f[M_HY_X] + base, j, ax, bx, kx, true);
      }
      const T cx = dhy_dz - dhz_dy;
      const T cy = dhz_dx - dhx_dz;
      const T cz = dhx_dy - dhy_dx;
      T e0 = ex0, e1 = ey0, e2 = ez0;
      if constexpr (FMT != 0) {
        const int64_t i = base + j;
        if (a.st[0]) {
          put_store<T, FMT>(a.st[0], i, e0);
          put_store<T, FMT>(a.st[1], i, e1);
          put_store<T, FMT>(a.st[2], i, e2);
        }
        if (a.st[3]) {
          put_store<T, FMT>(a.st[3], i, cx);
          put_store<T, FMT>(a.st[4], i, cy);
          put_store<T, FMT>(a.st[5], i, cz);
        }
      }
      if constexpr (BORN) {
        const int64_t i = base + j;
        const T dcav = a.dca[j], dcbv = a.dcb[j];
        e0 = cav * e0 + cbv * cx + dcav * a.bg[0][i] + dcbv * a.bg[3][i];
        e1 = cav * e1 + cbv * cy + dcav * a.bg[1][i] + dcbv * a.bg[4][i];
        e2 = cav * e2 + cbv * cz + dcav * a.bg[2][i] + dcbv * a.bg[5][i];
      } else {
        e0 = cav * e0 + cbv * cx;
        e1 = cav * e1 + cbv * cy;
        e2 = cav * e2 + cbv * cz;
      }
      // Sources of this shot at this cell, then receivers.
      const int slot = a.pts.row_slot[z * ny + y];
      if (slot >= 0) {
        const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + tc.xt;
        const int e1_ = a.pts.seg[si + 1];
        for (int e = a.pts.seg[si]; e < e1_; ++e) {
          const int2 en = a.pts.ent[e];
          if ((en.x & 0xff) != lane) continue;
          if ((en.x >> 8) == 0) {
            const T v = a.src_amp[en.y];
            if (a.src_comp == 0) e0 += v;
            else if (a.src_comp == 1) e1 += v;
            else e2 += v;
          } else {
            a.rec_out[en.y] = a.rec_comp == 0 ? e0 : (a.rec_comp == 1 ? e1 : e2);
          }
        }
      }
      const int64_t line = (int64_t)(z * ny + y) * nx;
      *elem32(ex + line, (uint32_t)x) = e0;
      *elem32(ey + line, (uint32_t)x) = e1;
      *elem32(ez + line, (uint32_t)x) = e2;
    }
    hx_zm = hxc;
    hy_zm = hyc;
  }
}

// ------------------------------------------------------------- adjoint
// Adjoint of a CPML-corrected derivative whose adjoint is q:
// strip: a (b M + q) + q / k, with the next memory b M + q; else q.
template <typename T>
__device__ __forceinline__ T adj_raw(T q, T m, T a, T b, T k, bool strip) {
  if (!strip) return q;
  const T mn = b * m + q;
  return a * mn + q / k;
}

// Transpose of the E update: lambda_H += D^T (raw E-curl adjoints); the
// E-phase adjoint memories advance into a.mn.  Every load is issued without
// a per-lane branch (clamped cell, masked value); the CPML memories are read
// only behind wave-uniform strip tests, so interior tiles never touch them.
template <typename T, int TZ>
__global__ __launch_bounds__(512) void adj_e3(Args3<T> a) {
  const Geom3<T> &g = a.g;
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const Tile3 tc = decode3(a.pts.n_xt, g.ny, a.n_zb, n_groups, a.shot_fast);
  const int s = tc.grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  if (s >= a.n_shots) return;
  const int nz = g.nz, ny = g.ny, nx = g.nx, y = tc.y;
  const int lane = threadIdx.x;
  const int x = tc.xt * kOwn - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && x >= 1 && x < nx && y >= 1;
  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ lex = a.f[EX] + base;
  const T *__restrict__ ley = a.f[EY] + base;
  const T *__restrict__ lez = a.f[EZ] + base;
  T *__restrict__ lhx = a.f[HX] + base;
  T *__restrict__ lhy = a.f[HY] + base;
  T *__restrict__ lhz = a.f[HZ] + base;
  const T *__restrict__ m_hy_z = a.f[M_HY_Z] + base;
  const T *__restrict__ m_hx_z = a.f[M_HX_Z] + base;
  const T *__restrict__ m_hz_y = a.f[M_HZ_Y] + base;
  const T *__restrict__ m_hx_y = a.f[M_HX_Y] + base;
  const T *__restrict__ m_hz_x = a.f[M_HZ_X] + base;
  const T *__restrict__ m_hy_x = a.f[M_HY_X] + base;
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const Ld3<T> L(x, nx, ny, nz);
  const bool xdom = x >= 1 && x < nx;
  const bool py = y < g.y0 || y >= g.y1, py1 = y + 1 < g.y0 || y + 1 >= g.y1;
  const bool px = x < g.x0 || x >= g.x1;
  const int x_first = tc.xt * kOwn - 1;
  const bool tile_px = x_first < g.x0 || x_first + 63 >= g.x1;  // some lane in an x strip
  const int zs = tc.zb * TZ;
  // y profiles at y and y + 1; x profiles at x (clamped for halo lanes)
  const int xc = min(max(x, 0), nx - 1);
  const T ay = g.Y(kA, y), by = g.Y(kB, y), ky = g.Y(kK, y);
  const int yp = min(y + 1, ny - 1);
  const T ay1 = g.Y(kA, yp), by1 = g.Y(kB, yp), ky1 = g.Y(kK, yp);
  const T ax = g.X(kA, xc), bx = g.X(kB, xc), kx = g.X(kK, xc);

  // z-derivative raws at plane zz: (dHy/dz from Ex, +) and (dHx/dz from Ey, -);
  // loads (z_load) kept apart from arithmetic (z_raws) so a later plane's
  // loads can be issued early.
  struct ZIn {
    T cbv, l0, l1, m_yz, m_xz;
  };
  auto z_load = [&](int zz) {
    ZIn v;
    v.cbv = L(cb, zz, y);
    v.l0 = L(lex, zz, y);
    v.l1 = L(ley, zz, y);
    v.m_yz = v.m_xz = (T)0;
    if (zz < g.z0 || zz >= g.z1) {  // uniform
      v.m_yz = L(m_hy_z, zz, y);
      v.m_xz = L(m_hx_z, zz, y);
    }
    return v;
  };
  auto z_raws = [&](int zz, const ZIn &v, T &r_yz, T &r_xz) {
    const bool in = zz >= 1 && zz < nz && y >= 1 && xdom;
    const T q_yz = v.cbv * v.l0;
    const T q_xz = -(v.cbv * v.l1);
    const bool pz = zz < g.z0 || zz >= g.z1;  // uniform
    const int zzc = min(max(zz, 0), nz - 1);
    const T az = g.Z(kA, zzc), bz = g.Z(kB, zzc), kz = g.Z(kK, zzc);
    r_yz = keep_or_zero(adj_raw(q_yz, v.m_yz, az, bz, kz, pz), in);
    r_xz = keep_or_zero(adj_raw(q_xz, v.m_xz, az, bz, kz, pz), in);
  };
  // this cell and its y + 1 neighbour (plane z), loaded before any branch
  struct PIn {
    T l2, cbp, l0p, l2p, h0, h1, h2, m_zx, m_yx, m_zy, m_xy, m_zy1, m_xy1;
  };
  auto p_load = [&](int z) {
    PIn v;
    const int zc = min(z, nz - 1);
    v.l2 = L(lez, zc, y);
    v.cbp = L(cb, zc, y + 1);
    v.l0p = L(lex, zc, y + 1);
    v.l2p = L(lez, zc, y + 1);
    v.h0 = L.at(lhx, zc, y);
    v.h1 = L.at(lhy, zc, y);
    v.h2 = L.at(lhz, zc, y);
    v.m_zx = v.m_yx = v.m_zy = v.m_xy = v.m_zy1 = v.m_xy1 = (T)0;
    if (tile_px) {
      v.m_zx = keep_or_zero(L(m_hz_x, zc, y), px);
      v.m_yx = keep_or_zero(L(m_hy_x, zc, y), px);
    }
    if (py) {
      v.m_zy = L(m_hz_y, zc, y);
      v.m_xy = L(m_hx_y, zc, y);
    }
    if (py1) {
      v.m_zy1 = L(m_hz_y, zc, y + 1);
      v.m_xy1 = L(m_hx_y, zc, y + 1);
    }
    return v;
  };

  // Plane z + 1's loads are issued before plane z's stores (software
  // pipelining: the compiler cannot move loads above stores that might
  // alias them).
  ZIn zc_in = z_load(zs), zn_in = z_load(zs + 1);
  PIn p_in = p_load(zs);
  T ryz_c, rxz_c;
  z_raws(zs, zc_in, ryz_c, rxz_c);
#pragma unroll
  for (int r = 0; r < TZ; ++r) {
    const int z = zs + r;
    if (z >= nz) break;
    const ZIn zcur = zc_in, znext = zn_in;
    const PIn pin = p_in;
    if (r + 1 < TZ) {
      zc_in = zn_in;
      zn_in = z_load(z + 2);
      p_in = p_load(z + 1);
    }
    T ryz_n, rxz_n;
    z_raws(z + 1, znext, ryz_n, rxz_n);
    const bool zdom = z >= 1;
    const int j = (z * ny + y) * nx + x;
    const bool pz = z < g.z0 || z >= g.z1;  // uniform
    const T cbv = zcur.cbv, l0 = zcur.l0, l1 = zcur.l1, l2 = pin.l2;
    const T cbp = pin.cbp, l0p = pin.l0p, l2p = pin.l2p;
    const T h0 = pin.h0, h1 = pin.h1, h2 = pin.h2;
    const T m_zx = pin.m_zx, m_yx = pin.m_yx, m_zy = pin.m_zy, m_xy = pin.m_xy;
    const T m_zy1 = pin.m_zy1, m_xy1 = pin.m_xy1, m_yz = zcur.m_yz, m_xz = zcur.m_xz;
    // x-derivative raws at (z, y, x) for every lane (the left lane needs ours)
    const bool inx = zdom && y >= 1 && xdom;
    const T r_zx = keep_or_zero(adj_raw(cbv * l1, m_zx, ax, bx, kx, px), inx);
    const T r_yx = keep_or_zero(adj_raw(-(cbv * l2), m_yx, ax, bx, kx, px), inx);
    const T r_zx_xp = shfl_dn1(r_zx);
    const T r_yx_xp = shfl_dn1(r_yx);
    if (own && zdom) {
      // y-derivative raws at y and y + 1: (dHz/dy from Ex, -) and (dHx/dy from Ez, +)
      const T q_zy = -(cbv * l0);
      const T q_xy = cbv * l2;
      const T r_zy = adj_raw(q_zy, m_zy, ay, by, ky, py);
      const T r_xy = adj_raw(q_xy, m_xy, ay, by, ky, py);
      T r_zy_yp = (T)0, r_xy_yp = (T)0;
      if (y + 1 < ny) {
        const T q_zy1 = -(cbp * l0p);
        const T q_xy1 = cbp * l2p;
        r_zy_yp = adj_raw(q_zy1, m_zy1, ay1, by1, ky1, py1);
        r_xy_yp = adj_raw(q_xy1, m_xy1, ay1, by1, ky1, py1);
      }
      // next E-phase memories at this cell
      if (pz) {
        const T bz = g.Z(kB, z);
        a.mn[M_HY_Z][base + j] = bz * m_yz + cbv * l0;
        a.mn[M_HX_Z][base + j] = bz * m_xz + -(cbv * l1);
      }
      if (py) {
        a.mn[M_HZ_Y][base + j] = by * m_zy + -(cbv * l0);
        a.mn[M_HX_Y][base + j] = by * m_xy + cbv * l2;
      }
      if (px) {
        a.mn[M_HZ_X][base + j] = bx * m_zx + cbv * l1;
        a.mn[M_HY_X][base + j] = bx * m_yx + -(cbv * l2);
      }
      T vx = h0;
      vx = vx + (rxz_c - rxz_n) * g.rdz;
      vx = vx + (r_xy - r_xy_yp) * g.rdy;
      T vy = h1;
      vy = vy + (ryz_c - ryz_n) * g.rdz;
      vy = vy + (r_yx - r_yx_xp) * g.rdx;
      T vz = h2;
      vz = vz + (r_zy - r_zy_yp) * g.rdy;
      vz = vz + (r_zx - r_zx_xp) * g.rdx;
      const int64_t line = (int64_t)(z * ny + y) * nx;
      *elem32(lhx + line, (uint32_t)x) = vx;
      *elem32(lhy + line, (uint32_t)x) = vy;
      *elem32(lhz + line, (uint32_t)x) = vz;
    }
    ryz_c = ryz_n;
    rxz_c = rxz_n;
  }
}

// Transpose of the H update: lambda_E = ca lambda_E + D^T (raw H-term
// adjoints); H-phase adjoint memories advance into a.mn; then receivers
// inject grad_r[t], sources read grad_f[t], and the coefficient gradient
// accumulates against the stored forward state.  Loads are issued as in
// adj_e3 (no per-lane branch; memories behind wave-uniform strip tests; the
// snapshots of the cell read before the owner's update).
template <typename T, int TZ, int FMT>
__global__ __launch_bounds__(512) void adj_h3(Args3<T> a) {
  extern __shared__ double red[];  // [2][G][TZ][64] when reducing
  const Geom3<T> &g = a.g;
  const int G = (int)blockDim.y;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  const int n_groups = (a.n_shots + G - 1) / G;
  const Tile3 tc = decode3(a.pts.n_xt, g.ny, a.n_zb, n_groups, a.shot_fast);
  const int s = tc.grp * G + w;
  const int nz = g.nz, ny = g.ny, nx = g.nx, y = tc.y;
  const int lane = threadIdx.x;
  const int x = tc.xt * kOwn - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && x >= 1 && x < nx && y >= 1;
  const bool want_grad = (a.gca != nullptr || a.gcb != nullptr) && FMT != 0;
  const bool reduce = want_grad && !a.per_shot_grad;
  const int zs = tc.zb * TZ;
  double *red_ca = red + (int64_t)w * TZ * 64;
  double *red_cb = red + ((int64_t)G + w) * TZ * 64;
  if (reduce) {
#pragma unroll
    for (int r = 0; r < TZ; ++r) {
      red_ca[r * 64 + lane] = 0.0;
      red_cb[r * 64 + lane] = 0.0;
    }
  }
  if (s < a.n_shots) {  // inactive wavefronts still join the barrier below
    const int64_t base = (int64_t)s * g.numel;
    T *__restrict__ lex = a.f[EX] + base;
    T *__restrict__ ley = a.f[EY] + base;
    T *__restrict__ lez = a.f[EZ] + base;
    const T *__restrict__ lhx = a.f[HX] + base;
    const T *__restrict__ lhy = a.f[HY] + base;
    const T *__restrict__ lhz = a.f[HZ] + base;
    const T *__restrict__ m_ey_z = a.f[M_EY_Z] + base;
    const T *__restrict__ m_ex_z = a.f[M_EX_Z] + base;
    const T *__restrict__ m_ez_y = a.f[M_EZ_Y] + base;
    const T *__restrict__ m_ex_y = a.f[M_EX_Y] + base;
    const T *__restrict__ m_ez_x = a.f[M_EZ_X] + base;
    const T *__restrict__ m_ey_x = a.f[M_EY_X] + base;
    const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
    const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
    const Ld3<T> L(x, nx, ny, nz);
    const bool xdom = x >= 1 && x < nx;
    const bool has_dx = x < nx - 1;
    const bool pyh = y < g.y0 || y >= g.y1h;
    const int ym = y - 1;
    const bool pyh_m = ym < g.y0 || ym >= g.y1h;
    const int xc = min(max(x, 0), nx - 1);
    const int ymc = max(ym, 0);
    const T ay = g.Y(kAh, y), by = g.Y(kBh, y), ky = g.Y(kKh, y);
    const T aym = g.Y(kAh, ymc), bym = g.Y(kBh, ymc), kym = g.Y(kKh, ymc);
    const bool pxh = x < g.x0 || x >= g.x1h;
    const int x_first = tc.xt * kOwn - 1;
    const bool tile_pxh = x_first < g.x0 || x_first + 63 >= g.x1h;  // some lane in an x strip
    const T ax = g.X(kAh, xc), bx = g.X(kBh, xc), kx = g.X(kKh, xc);

    // z-derivative raws at plane zz (cell (zz, y, x)):
    //   Hx term dEy/dz (first, -cq lHx) and Hy term dEx/dz (second, +cq lHy);
    // loads (z_load) kept apart from arithmetic (z_raws).
    struct ZIn {
      T cqv, h0, h1, m_yz, m_xz;
    };
    auto z_load = [&](int zz) {
      ZIn v;
      v.cqv = L(cq, zz, y);
      v.h0 = L(lhx, zz, y);
      v.h1 = L(lhy, zz, y);
      v.m_yz = v.m_xz = (T)0;
      if (zz < g.z0 || zz >= g.z1h) {  // uniform
        v.m_yz = L(m_ey_z, zz, y);
        v.m_xz = L(m_ex_z, zz, y);
      }
      return v;
    };
    auto z_raws = [&](int zz, const ZIn &v, T &r_yz, T &r_xz) {
      const bool in = zz >= 1 && zz < nz - 1 && y >= 1 && xdom;
      const T adj_yz = -v.cqv * v.h0;
      const T adj_xz = v.cqv * v.h1;
      const bool pz = zz < g.z0 || zz >= g.z1h;  // uniform
      const int zzc = min(max(zz, 0), nz - 1);
      const T az = g.Z(kAh, zzc), bz = g.Z(kBh, zzc), kz = g.Z(kKh, zzc);
      r_yz = keep_or_zero(adj_raw(adj_yz, v.m_yz, az, bz, kz, pz), in);
      r_xz = keep_or_zero(adj_raw(adj_xz, v.m_xz, az, bz, kz, pz), in);
    };
    // the rest of plane z: this cell, its y - 1 neighbour, memories and
    // snapshots, loaded before any branch
    struct PIn {
      T h2, cqm, h0m, h2m, cav, e0, e1, e2, m_zx, m_yx, m_zy, m_xy, m_zym, m_xym;
      float sa[3], sb[3];
    };
    auto p_load = [&](int z) {
      PIn v;
      const int zc = min(z, nz - 1);
      const int64_t line = (int64_t)(zc * ny + y) * nx;
      v.h2 = L(lhz, zc, y);
      v.cqm = L(cq, zc, ym);
      v.h0m = L(lhx, zc, ym);
      v.h2m = L(lhz, zc, ym);
      v.cav = L.at(ca, zc, y);
      v.e0 = L.at(lex, zc, y);
      v.e1 = L.at(ley, zc, y);
      v.e2 = L.at(lez, zc, y);
      v.m_zx = v.m_yx = v.m_zy = v.m_xy = v.m_zym = v.m_xym = (T)0;
      if (tile_pxh) {
        v.m_zx = keep_or_zero(L(m_ez_x, zc, y), pxh);
        v.m_yx = keep_or_zero(L(m_ey_x, zc, y), pxh);
      }
      if (pyh) {
        v.m_zy = L(m_ez_y, zc, y);
        v.m_xy = L(m_ex_y, zc, y);
      }
      if (pyh_m) {
        v.m_zym = L(m_ez_y, zc, ym);
        v.m_xym = L(m_ex_y, zc, ym);
      }
      // snapshots of this cell (forward state before step t), when wanted;
      // exact in float (fp32 or BF16 payloads of fp32 fields)
#pragma unroll
      for (int c = 0; c < 3; ++c) v.sa[c] = v.sb[c] = 0.0f;
      if constexpr (FMT != 0 && sizeof(T) == 4) {
        if (want_grad) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const void *const es = a.st[c], *const cs = a.st[3 + c];
            if (a.gca && es) v.sa[c] = (float)get_store32<T, FMT>(snap_plane<T, FMT>(es, base + line), (uint32_t)xc);
            if (a.gcb && cs) v.sb[c] = (float)get_store32<T, FMT>(snap_plane<T, FMT>(cs, base + line), (uint32_t)xc);
          }
        }
      }
      return v;
    };

    // Plane z + 1's loads are issued before plane z's stores.
    ZIn z_in = z_load(zs);
    PIn p_in = p_load(zs);
    T ryz_p, rxz_p;
    z_raws(zs - 1, z_load(zs - 1), ryz_p, rxz_p);
#pragma unroll
    for (int r = 0; r < TZ; ++r) {
      const int z = zs + r;
      if (z >= nz) break;
      const ZIn zcur = z_in;
      const PIn pin = p_in;
      if (r + 1 < TZ) {
        z_in = z_load(z + 1);
        p_in = p_load(z + 1);
      }
      const bool zdom = z >= 1;
      const int j = (z * ny + y) * nx + x;
      const int64_t line = (int64_t)(z * ny + y) * nx;
      T ryz_c, rxz_c;
      z_raws(z, zcur, ryz_c, rxz_c);
      const bool pzh = z < g.z0 || z >= g.z1h;  // uniform
      const T cqv = zcur.cqv, h0 = zcur.h0, h1 = zcur.h1, h2 = pin.h2;
      const T cqm = pin.cqm, h0m = pin.h0m, h2m = pin.h2m;
      const T cav = pin.cav, e0 = pin.e0, e1 = pin.e1, e2 = pin.e2;
      const T m_zx = pin.m_zx, m_yx = pin.m_yx, m_zy = pin.m_zy, m_xy = pin.m_xy;
      const T m_zym = pin.m_zym, m_xym = pin.m_xym, m_yz = zcur.m_yz, m_xz = zcur.m_xz;
      double sa[3], sb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sa[c] = (double)pin.sa[c];
        sb[c] = (double)pin.sb[c];
      }
      if constexpr (sizeof(T) == 8 && FMT != 0) {  // fp64 payloads: read at use
        if (want_grad) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const void *const es = a.st[c], *const cs = a.st[3 + c];
            if (a.gca && es) sa[c] = get_store32<T, FMT>(snap_plane<T, FMT>(es, base + line), (uint32_t)xc);
            if (a.gcb && cs) sb[c] = get_store32<T, FMT>(snap_plane<T, FMT>(cs, base + line), (uint32_t)xc);
          }
        }
      }
      // x-derivative raws at (z, y, x): Hy term dEz/dx (first, -cq lHy) and
      // Hz term dEy/dx (second, +cq lHz); the right lane needs ours
      const bool inx = zdom && y >= 1 && xdom && has_dx;
      const T r_zx = keep_or_zero(adj_raw(-cqv * h1, m_zx, ax, bx, kx, pxh), inx);
      const T r_yx = keep_or_zero(adj_raw(cqv * h2, m_yx, ax, bx, kx, pxh), inx);
      const T r_zx_xm = shfl_up1(r_zx);
      const T r_yx_xm = shfl_up1(r_yx);
      if (own && zdom) {
        // y-derivative raws at y and y - 1: Hx term dEz/dy (second, +cq lHx)
        // and Hz term dEx/dy (first, -cq lHz); defined for y < ny - 1
        T r_zy = (T)0, r_xy = (T)0;
        if (y < ny - 1) {
          const T adj_zy = cqv * h0;
          const T adj_xy = -cqv * h2;
          r_zy = adj_raw(adj_zy, m_zy, ay, by, ky, pyh);
          r_xy = adj_raw(adj_xy, m_xy, ay, by, ky, pyh);
        }
        T r_zy_ym = (T)0, r_xy_ym = (T)0;
        if (ym >= 1) {
          const T adj_zy1 = cqm * h0m;
          const T adj_xy1 = -cqm * h2m;
          r_zy_ym = adj_raw(adj_zy1, m_zym, aym, bym, kym, pyh_m);
          r_xy_ym = adj_raw(adj_xy1, m_xym, aym, bym, kym, pyh_m);
        }
        const T rzx_m = x - 1 >= 1 ? r_zx_xm : (T)0;
        const T ryx_m = x - 1 >= 1 ? r_yx_xm : (T)0;
        // next H-phase memories at this cell
        if (z < nz - 1 && pzh) {
          const T bz = g.Z(kBh, z);
          a.mn[M_EY_Z][base + j] = bz * m_yz + -cqv * h0;
          a.mn[M_EX_Z][base + j] = bz * m_xz + cqv * h1;
        }
        if (y < ny - 1 && pyh) {
          a.mn[M_EZ_Y][base + j] = by * m_zy + cqv * h0;
          a.mn[M_EX_Y][base + j] = by * m_xy + -cqv * h2;
        }
        if (has_dx && pxh) {
          a.mn[M_EZ_X][base + j] = bx * m_zx + -cqv * h1;
          a.mn[M_EY_X][base + j] = bx * m_yx + cqv * h2;
        }
        const T tx = (rxz_p - rxz_c) * g.rdz + (r_xy_ym - r_xy) * g.rdy;
        const T ty = (ryz_p - ryz_c) * g.rdz + (ryx_m - r_yx) * g.rdx;
        const T tz = (r_zy_ym - r_zy) * g.rdy + (rzx_m - r_zx) * g.rdx;
        T v0 = cav * e0 + tx;
        T v1 = cav * e1 + ty;
        T v2 = cav * e2 + tz;
        // receivers inject grad_r[t]; sources read grad_f[t]
        const int slot = a.pts.row_slot[z * ny + y];
        if (slot >= 0) {
          const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + tc.xt;
          const int p0 = a.pts.seg[si], p1 = a.pts.seg[si + 1];
          for (int e = p0; e < p1; ++e) {
            const int2 en = a.pts.ent[e];
            if ((en.x >> 8) != 1 || (en.x & 0xff) != lane) continue;
            const T v = a.src_amp[en.y];
            if (a.rec_comp == 0) v0 += v;
            else if (a.rec_comp == 1) v1 += v;
            else v2 += v;
          }
          if (a.rec_out != nullptr) {
            for (int e = p0; e < p1; ++e) {
              const int2 en = a.pts.ent[e];
              if ((en.x >> 8) != 0 || (en.x & 0xff) != lane) continue;
              a.rec_out[en.y] = a.src_comp == 0 ? v0 : (a.src_comp == 1 ? v1 : v2);
            }
          }
        }
        *elem32(lex + line, (uint32_t)x) = v0;
        *elem32(ley + line, (uint32_t)x) = v1;
        *elem32(lez + line, (uint32_t)x) = v2;
        if (want_grad) {
          const int64_t i = base + j;
          double pa = 0.0, pb = 0.0;
          if (a.gca) pa = ((double)v0 * sa[0] + (double)v1 * sa[1] + (double)v2 * sa[2]) * a.step_scale;
          if (a.gcb) pb = ((double)v0 * sb[0] + (double)v1 * sb[1] + (double)v2 * sb[2]) * a.step_scale;
          if (reduce) {
            red_ca[r * 64 + lane] = pa;
            red_cb[r * 64 + lane] = pb;
          } else {
            if (a.gca) a.gca[i] += pa;
            if (a.gcb) a.gcb[i] += pb;
          }
        }
      }
      ryz_p = ryz_c;
      rxz_p = rxz_c;
    }
  }
  if (!reduce) return;  // uniform over the workgroup
  __syncthreads();
  const int x_ = tc.xt * kOwn - 1 + lane;
  if (lane < 1 || lane > kOwn || x_ < 1 || x_ >= nx || y < 1) return;
  for (int pass = 0; pass < 2; ++pass) {
    double *out = pass == 0 ? a.gca : a.gcb;
    if (out == nullptr) continue;
    const double *slots = red + (int64_t)pass * G * TZ * 64;
    for (int r = w; r < TZ; r += G) {
      const int z = zs + r;
      if (z >= 1 && z < nz) {
        double sum = 0.0;
        for (int k = 0; k < G; ++k) sum += slots[(k * TZ + r) * 64 + lane];
        const int64_t j = ((int64_t)z * ny + y) * nx + x_ + tc.grp * a.group_stride;
        out[j] += sum;
      }
    }
  }
}

// out[i] += sum over g of acc[g*n + i], summed in group order.
__global__ void fold_groups3(double *__restrict__ out, const double *__restrict__ acc, int64_t n,
                             int groups) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sum = acc[i];
  for (int g = 1; g < groups; ++g) sum += acc[(int64_t)g * n + i];
  out[i] += sum;
}

}  // namespace em3d
}  // namespace tide_hip

#include "em3d_fused.h"

namespace tide_hip {
namespace em3d {

// ------------------------------------------------------------------ driver

struct Tuning3 {
  // adjoint 4 planes per wavefront: about 5% more gradient throughput than 8
  // on a 281^3 grid with 4 shots; the forward runs one shot per workgroup
  // (6% faster there than 8), the adjoint sums 8 shots' gradient terms per
  // workgroup
  int fwd_z = 8, adj_z = 4, group = 8, fwd_group = 1;
  // one-launch forward step (em3d_fused.h): rows per wavefront (2 or 4; 0
  // runs the two-kernel step) and planes per march (8 or 16)
  int fused_rows = 0, fused_z = 8;
  // workgroup order: shot groups of one tile adjacent (1) or all tiles of
  // one shot group first (0)
  int shot_fast = 0;
};
Tuning3 g_tuning3;

template <typename T> struct Problem3 {
  Geom3<T> g;
  const T *ca, *cb, *cq;
  const T *prof[18];  // az bz azh bzh ay by ayh byh ax bx axh bxh kz kzh ky kyh kx kxh
  int32_t n_shots;
  int64_t n_src, n_rec;
  int32_t src_comp, rec_comp;
};

#define TIDE_CHECK(expr)                    \
  do {                                      \
    const int _rc = (expr);                 \
    if (_rc != ::tide_hip::kOk) return _rc; \
  } while (0)

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

template <typename T>
int prepare(Problem3<T> &p, const int64_t *src, const int64_t *rec, Scratch &scr,
            DevicePointTable &dpt, hipStream_t stream) {
  Geom3<T> &g = p.g;
  if (p.n_shots <= 0 || g.nz < 3 || g.ny < 3 || g.nx < 3) {
    set_error("n_shots must be positive and the padded grid at least 3x3x3");
    return kInvalidArgument;
  }
  if ((int64_t)g.nz * g.ny * g.nx >= ((int64_t)1 << 31)) {
    set_error("a shot volume must hold fewer than 2^31 cells");
    return kInvalidArgument;
  }
  if (p.src_comp < 0 || p.src_comp > 2 || p.rec_comp < 0 || p.rec_comp > 2) {
    set_error("source/receiver component must be 0 (x), 1 (y) or 2 (z)");
    return kInvalidArgument;
  }
  // profiles: [a, b, k, ah, bh, kh] per axis
  const int n[3] = {g.nz, g.ny, g.nx};
  T *buf;
  TIDE_CHECK(scr.alloc(&buf, (size_t)6 * (n[0] + n[1] + n[2])));
  T *axis_buf[3] = {buf, buf + 6 * n[0], buf + 6 * (n[0] + n[1])};
  for (int ax = 0; ax < 3; ++ax) {
    const int src_k[6] = {4 * ax + 0, 4 * ax + 1, 12 + 2 * ax, 4 * ax + 2, 4 * ax + 3, 13 + 2 * ax};
    for (int k = 0; k < 6; ++k)
      TIDE_HIP_TRY(hipMemcpyAsync(axis_buf[ax] + (size_t)k * n[ax], p.prof[src_k[k]],
                                  (size_t)n[ax] * sizeof(T), hipMemcpyDefault, stream));
  }
  g.pz = axis_buf[0];
  g.py = axis_buf[1];
  g.px = axis_buf[2];
  std::vector<int64_t> hs, hr;
  TIDE_CHECK(fetch_indices(src, (int64_t)p.n_shots * p.n_src, hs, stream));
  TIDE_CHECK(fetch_indices(rec, (int64_t)p.n_shots * p.n_rec, hr, stream));
  PointTable table;
  if (!build_point_table(hs.data(), p.n_src, hr.data(), p.n_rec, p.n_shots,
                         (int64_t)g.nz * g.ny, g.nx, kOwn, 1, table))
    return kInvalidArgument;
  return dpt.upload(table, stream);
}

template <typename T>
void fill_common(Args3<T> &a, const Problem3<T> &p, const DevicePointTable &dpt, int tz) {
  a.g = p.g;
  a.pts = Points{dpt.row_slot, dpt.seg, dpt.ent, dpt.n_xt};
  a.ca = p.ca;
  a.cb = p.cb;
  a.cq = p.cq;
  a.n_shots = p.n_shots;
  a.n_zb = (p.g.nz + tz - 1) / tz;
  a.src_comp = p.src_comp;
  a.rec_comp = p.rec_comp;
  a.gca = a.gcb = nullptr;
  a.per_shot_grad = 0;
  a.group_stride = 0;
  a.shot_fast = g_tuning3.shot_fast;
  a.step_scale = 1.0;
  for (int k = 0; k < 6; ++k) a.st[k] = nullptr;
  for (int k = 0; k < N_SLOTS; ++k) a.mn[k] = nullptr;
  a.dca = a.dcb = nullptr;
  for (int k = 0; k < 6; ++k) a.bg[k] = nullptr;
}

template <typename T>
void launch_h3(int tz, dim3 grid, dim3 block, hipStream_t st, const Args3<T> &a) {
  switch (tz) {
    case 4: h3_step<T, 4><<<grid, block, 0, st>>>(a); break;
    case 16: h3_step<T, 16><<<grid, block, 0, st>>>(a); break;
    default: h3_step<T, 8><<<grid, block, 0, st>>>(a); break;
  }
}

template <typename T, int FMT, bool BORN = false>
void launch_e3(int tz, dim3 grid, dim3 block, hipStream_t st, const Args3<T> &a) {
  switch (tz) {
    case 4: e3_step<T, 4, FMT, BORN><<<grid, block, 0, st>>>(a); break;
    case 16: e3_step<T, 16, FMT, BORN><<<grid, block, 0, st>>>(a); break;
    default: e3_step<T, 8, FMT, BORN><<<grid, block, 0, st>>>(a); break;
  }
}

template <typename T>
void launch_adj_e3(int tz, dim3 grid, dim3 block, hipStream_t st, const Args3<T> &a) {
  switch (tz) {
    case 4: adj_e3<T, 4><<<grid, block, 0, st>>>(a); break;
    default: adj_e3<T, 8><<<grid, block, 0, st>>>(a); break;
  }
}

template <typename T, int FMT>
void launch_adj_h3(int tz, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                   const Args3<T> &a) {
  switch (tz) {
    case 4: adj_h3<T, 4, FMT><<<grid, block, lds, st>>>(a); break;
    default: adj_h3<T, 8, FMT><<<grid, block, lds, st>>>(a); break;
  }
}

template <typename T, int FMT>
void launch_fused(int ty, int tz, dim3 grid, dim3 block, hipStream_t st, const Fused3<T> &a) {
  if (ty >= 4) {
    if (tz >= 16) em3_fused<T, 4, 16, FMT><<<grid, block, 0, st>>>(a);
    else em3_fused<T, 4, 8, FMT><<<grid, block, 0, st>>>(a);
  } else {
    if (tz >= 16) em3_fused<T, 2, 16, FMT><<<grid, block, 0, st>>>(a);
    else em3_fused<T, 2, 8, FMT><<<grid, block, 0, st>>>(a);
  }
}

// The forward time loop on the one-launch step: the fields and the H-phase
// memories ping-pong between the caller's arrays and a second copy (both
// start equal, so cells a step never writes agree in both); the final state
// ends in the caller's arrays.
template <typename T>
int forward_fused(const Problem3<T> &p, T *const *fields, const T *f, T *r, int64_t nt,
                  int64_t start_t, int64_t step_ratio, void *const *stores, int fmt,
                  size_t store_step_bytes, int64_t store_t0, hipStream_t stream, Scratch &scr,
                  const DevicePointTable &dpt) {
  static constexpr int kHMem[6] = {M_EY_Z, M_EX_Z, M_EZ_Y, M_EX_Y, M_EZ_X, M_EY_X};
  static constexpr int kEMem[6] = {M_HY_Z, M_HX_Z, M_HZ_Y, M_HX_Y, M_HZ_X, M_HY_X};
  const int64_t total = (int64_t)p.n_shots * p.g.numel;
  const size_t bytes = (size_t)total * sizeof(T);
  T *fa[6], *fb[6], *ma[6], *mb[6];
  for (int k = 0; k < 6; ++k) {
    fa[k] = fields[k];
    ma[k] = fields[kHMem[k]];
    TIDE_CHECK(scr.alloc(&fb[k], (size_t)total));
    TIDE_CHECK(scr.alloc(&mb[k], (size_t)total));
    TIDE_HIP_TRY(hipMemcpyAsync(fb[k], fa[k], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(mb[k], ma[k], bytes, hipMemcpyDeviceToDevice, stream));
  }
  const int ty = g_tuning3.fused_rows >= 4 ? 4 : 2;
  const int tz = g_tuning3.fused_z >= 16 ? 16 : 8;
  Fused3<T> a;
  a.g = p.g;
  a.pts = Points{dpt.row_slot, dpt.seg, dpt.ent, dpt.n_xt};
  a.ca = p.ca;
  a.cb = p.cb;
  a.cq = p.cq;
  a.n_shots = p.n_shots;
  a.n_yt = (p.g.ny + ty - 1) / ty;
  a.n_zb = (p.g.nz + tz - 1) / tz;
  a.src_comp = p.src_comp;
  a.rec_comp = p.rec_comp;
  a.shot_fast = g_tuning3.shot_fast;
  for (int k = 0; k < 6; ++k) a.me[k] = fields[kEMem[k]];
  const int group = std::min<int>(p.n_shots, g_tuning3.fwd_group);
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * a.n_yt * a.n_zb * n_groups));
  bool storing = false;
  for (int k = 0; k < 6; ++k) storing = storing || (stores && stores[k]);
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    for (int k = 0; k < 6; ++k) {
      a.fi[k] = fa[k];
      a.fo[k] = fb[k];
      a.mhi[k] = ma[k];
      a.mho[k] = mb[k];
    }
    a.src_amp = p.n_src > 0 ? f + t * p.n_shots * p.n_src : nullptr;
    a.rec_out = p.n_rec > 0 ? r + t * p.n_shots * p.n_rec : nullptr;
    const bool st = storing && (t % step_ratio) == 0;
    for (int k = 0; k < 6; ++k)
      a.st[k] = st && stores[k]
                    ? (char *)stores[k] + (size_t)((t - store_t0) / step_ratio) * store_step_bytes
                    : nullptr;
    if (!st) {
      launch_fused<T, 0>(ty, tz, grid, block, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_fused<T, 2>(ty, tz, grid, block, stream, a);
    } else {
      launch_fused<T, 1>(ty, tz, grid, block, stream, a);
    }
    for (int k = 0; k < 6; ++k) {
      std::swap(fa[k], fb[k]);
      std::swap(ma[k], mb[k]);
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  for (int k = 0; k < 6; ++k) {
    if (fa[k] != fields[k])
      TIDE_HIP_TRY(hipMemcpyAsync(fields[k], fa[k], bytes, hipMemcpyDeviceToDevice, stream));
    if (ma[k] != fields[kHMem[k]])
      TIDE_HIP_TRY(hipMemcpyAsync(fields[kHMem[k]], ma[k], bytes, hipMemcpyDeviceToDevice, stream));
  }
  return kOk;
}

// Forward over [start_t, start_t + nt).  Snapshot slot of step t is
// (t - store_t0) / step_ratio (t a multiple of step_ratio).
template <typename T>
int forward_impl(Problem3<T> p, T *const *fields, const T *f, T *r, const int64_t *src,
                 const int64_t *rec, int64_t nt, int64_t start_t, int64_t step_ratio,
                 void *const *stores, int fmt, size_t store_step_bytes, int64_t store_t0,
                 hipStream_t stream) {
  if (nt <= 0) return kOk;
  if (step_ratio <= 0) {
    set_error("step_ratio must be positive");
    return kInvalidArgument;
  }
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  Scratch scr(stream);
  DevicePointTable dpt;
  TIDE_CHECK(prepare(p, src, rec, scr, dpt, stream));
  if (g_tuning3.fused_rows > 0) {
    TIDE_CHECK(forward_fused(p, fields, f, r, nt, start_t, step_ratio, stores, fmt,
                             store_step_bytes, store_t0, stream, scr, dpt));
    dpt.release(stream);
    return kOk;
  }
  const int tz = g_tuning3.fwd_z;
  const int group = std::min<int>(p.n_shots, g_tuning3.fwd_group);
  Args3<T> a;
  fill_common(a, p, dpt, tz);
  for (int k = 0; k < N_SLOTS; ++k) a.f[k] = fields[k];
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * p.g.ny * a.n_zb * n_groups));
  bool storing = false;
  for (int k = 0; k < 6; ++k) storing = storing || (stores && stores[k]);
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    a.src_amp = p.n_src > 0 ? f + t * p.n_shots * p.n_src : nullptr;
    a.rec_out = p.n_rec > 0 ? r + t * p.n_shots * p.n_rec : nullptr;
    const bool st = storing && (t % step_ratio) == 0;
    for (int k = 0; k < 6; ++k)
      a.st[k] = st && stores[k]
                    ? (char *)stores[k] + (size_t)((t - store_t0) / step_ratio) * store_step_bytes
                    : nullptr;
    launch_h3<T>(tz, grid, block, stream, a);
    if (!st) {
      launch_e3<T, 0>(tz, grid, block, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_e3<T, 2>(tz, grid, block, stream, a);
    } else {
      launch_e3<T, 1>(tz, grid, block, stream, a);
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  dpt.release(stream);
  return kOk;
}

// Born forward over [start_t, start_t + nt) (reference
// maxwell_3d_inst.cpp:1100-1527): each step the background H and E kernels
// (the E kernel keeping the background's pre-update E and curl H in a
// one-step buffer), then the scattered H kernel and the scattered E kernel
// with the Born terms; f0 / df enter the background / scattered field, and
// background_r / r record them.  The scattered field starts from `dfields`.
template <typename T>
int born_forward_impl(Problem3<T> p, const T *dca, const T *dcb, T *const *fields,
                      T *const *dfields, const T *f0, const T *df, T *r, T *bg_r,
                      const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                      hipStream_t stream) {
  if (nt <= 0) return kOk;
  if (dca == nullptr || dcb == nullptr) {
    set_error("Born forward: dca and dcb are required");
    return kInvalidArgument;
  }
  if (p.n_src > 0 && (f0 == nullptr || df == nullptr)) {
    set_error("Born forward: f0 and df are required when there are sources");
    return kInvalidArgument;
  }
  Scratch scr(stream);
  DevicePointTable dpt;
  TIDE_CHECK(prepare(p, src, rec, scr, dpt, stream));
  const int64_t total = (int64_t)p.n_shots * p.g.numel;
  const int64_t rec_total = (start_t + nt) * p.n_shots * p.n_rec;
  T *snap = nullptr;
  TIDE_CHECK(scr.alloc(&snap, (size_t)(6 * total)));
  // receivers nobody asked for still have somewhere to go
  if (p.n_rec > 0 && r == nullptr) TIDE_CHECK(scr.alloc(&r, (size_t)rec_total));
  if (p.n_rec > 0 && bg_r == nullptr) TIDE_CHECK(scr.alloc(&bg_r, (size_t)rec_total));
  const int tz = g_tuning3.fwd_z;
  const int group = std::min<int>(p.n_shots, g_tuning3.fwd_group);
  Args3<T> a, b;
  fill_common(a, p, dpt, tz);
  fill_common(b, p, dpt, tz);
  for (int k = 0; k < N_SLOTS; ++k) {
    a.f[k] = fields[k];
    b.f[k] = dfields[k];
  }
  for (int k = 0; k < 6; ++k) {
    a.st[k] = snap + k * total;
    b.bg[k] = snap + k * total;
  }
  b.dca = dca;
  b.dcb = dcb;
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * p.g.ny * a.n_zb * n_groups));
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    const int64_t so = t * p.n_shots * p.n_src, ro = t * p.n_shots * p.n_rec;
    a.src_amp = p.n_src > 0 ? f0 + so : nullptr;
    b.src_amp = p.n_src > 0 ? df + so : nullptr;
    a.rec_out = p.n_rec > 0 ? bg_r + ro : nullptr;
    b.rec_out = p.n_rec > 0 ? r + ro : nullptr;
    launch_h3<T>(tz, grid, block, stream, a);
    launch_e3<T, 1>(tz, grid, block, stream, a);
    launch_h3<T>(tz, grid, block, stream, b);
    launch_e3<T, 0, true>(tz, grid, block, stream, b);
  }
  TIDE_HIP_TRY(hipGetLastError());
  dpt.release(stream);
  return kOk;
}

// Adjoint over steps start_t-1 down to start_t-nt.  lam holds the 18
// adjoint arrays (advanced in place); gca/gcb are FP64 accumulators
// ([nz,ny,nx] summed over shots, or [n_shots,...] when per shot), added to.
template <typename T>
int backward_impl(Problem3<T> p, const T *grad_r, T *const *lam, const void *const *stores,
                  int fmt, size_t store_step_bytes, int64_t store_t0, T *grad_f, double *gca,
                  double *gcb, const int64_t *src, const int64_t *rec, int64_t nt,
                  int64_t start_t, int64_t step_ratio, hipStream_t stream) {
  if (nt <= 0) return kOk;
  if (step_ratio <= 0) {
    set_error("step_ratio must be positive");
    return kInvalidArgument;
  }
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  if ((gca && (!stores || !stores[0] || !stores[1] || !stores[2])) ||
      (gcb && (!stores || !stores[3] || !stores[4] || !stores[5]))) {
    set_error("coefficient gradients require the stored forward snapshots");
    return kInvalidArgument;
  }
  // One gradient layout serves both accumulators (per shot when a
  // coefficient is per shot); mixed batching would overrun the shared one.
  if (gca && gcb && p.g.ca_b != p.g.cb_b) {
    set_error("ca and cb gradients need the same batching (ca_batched == cb_batched)");
    return kInvalidArgument;
  }
  Scratch scr(stream);
  DevicePointTable dpt;
  TIDE_CHECK(prepare(p, src, rec, scr, dpt, stream));
  const int64_t total = (int64_t)p.n_shots * p.g.numel;
  const size_t bytes = (size_t)total * sizeof(T);
  T *cur[N_SLOTS], *nxt[N_SLOTS];
  for (int k = 0; k < N_SLOTS; ++k) {
    cur[k] = lam[k];
    nxt[k] = nullptr;
  }
  for (int k = M_HZ_Y; k < N_SLOTS; ++k) {
    TIDE_CHECK(scr.alloc(&nxt[k], (size_t)total));
    // Memories are only written inside their strips; both copies start
    // equal so untouched cells agree whichever buffer ends up current.
    TIDE_HIP_TRY(hipMemcpyAsync(nxt[k], cur[k], bytes, hipMemcpyDeviceToDevice, stream));
  }
  const int tz = g_tuning3.adj_z;
  const int group = std::min<int>(p.n_shots, g_tuning3.group);
  Args3<T> a;
  fill_common(a, p, dpt, tz);
  a.per_shot_grad = ((gca && p.g.ca_b) || (gcb && p.g.cb_b)) ? 1 : 0;
  a.step_scale = (double)step_ratio;
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  // several shot groups: per-group FP64 slices, folded in group order below
  double *gacc = nullptr, *gbcc = nullptr;
  const int64_t numel = p.g.numel;
  if (n_groups > 1 && !a.per_shot_grad && (gca || gcb)) {
    a.group_stride = numel;
    if (gca) {
      TIDE_CHECK(scr.alloc(&gacc, (size_t)(n_groups * numel)));
      TIDE_HIP_TRY(hipMemsetAsync(gacc, 0, (size_t)(n_groups * numel) * sizeof(double), stream));
    }
    if (gcb) {
      TIDE_CHECK(scr.alloc(&gbcc, (size_t)(n_groups * numel)));
      TIDE_HIP_TRY(hipMemsetAsync(gbcc, 0, (size_t)(n_groups * numel) * sizeof(double), stream));
    }
  }
  double *const ga_k = gacc ? gacc : gca, *const gb_k = gbcc ? gbcc : gcb;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * p.g.ny * a.n_zb * n_groups));
  const size_t lds = 2 * (size_t)group * tz * 64 * sizeof(double);
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
    for (int k = 0; k < N_SLOTS; ++k) {
      a.f[k] = cur[k];
      a.mn[k] = nxt[k];
    }
    launch_adj_e3<T>(tz, grid, block, stream, a);
    for (int k = M_HZ_Y; k <= M_HX_Y; ++k) std::swap(cur[k], nxt[k]);
    for (int k = 0; k < N_SLOTS; ++k) {
      a.f[k] = cur[k];
      a.mn[k] = nxt[k];
    }
    a.src_amp = p.n_rec > 0 ? grad_r + t * p.n_shots * p.n_rec : nullptr;
    a.rec_out = (p.n_src > 0 && grad_f) ? grad_f + t * p.n_shots * p.n_src : nullptr;
    const bool g_step = (gca || gcb) && (t % step_ratio) == 0;
    for (int k = 0; k < 6; ++k)
      a.st[k] = g_step && stores[k]
                    ? (void *)((const char *)stores[k] +
                               (size_t)((t - store_t0) / step_ratio) * store_step_bytes)
                    : nullptr;
    a.gca = g_step ? ga_k : nullptr;
    a.gcb = g_step ? gb_k : nullptr;
    if (!g_step) {
      launch_adj_h3<T, 0>(tz, grid, block, lds, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_adj_h3<T, 2>(tz, grid, block, lds, stream, a);
    } else {
      launch_adj_h3<T, 1>(tz, grid, block, lds, stream, a);
    }
    for (int k = M_EY_Z; k < N_SLOTS; ++k) std::swap(cur[k], nxt[k]);
  }
  TIDE_HIP_TRY(hipGetLastError());
  if (gacc || gbcc) {
    const unsigned blocks = (unsigned)((numel + 255) / 256);
    if (gacc) fold_groups3<<<blocks, 256, 0, stream>>>(gca, gacc, numel, (int)n_groups);
    if (gbcc) fold_groups3<<<blocks, 256, 0, stream>>>(gcb, gbcc, numel, (int)n_groups);
    TIDE_HIP_TRY(hipGetLastError());
  }
  for (int k = M_HZ_Y; k < N_SLOTS; ++k)
    if (cur[k] != lam[k])
      TIDE_HIP_TRY(hipMemcpyAsync(lam[k], cur[k], bytes, hipMemcpyDeviceToDevice, stream));
  dpt.release(stream);
  return kOk;
}

template <typename T>
__global__ void fold_grad3(T *__restrict__ out, const double *__restrict__ acc, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)((double)out[i] + acc[i]);
}

template <typename T>
Problem3<T> make_problem(const T *ca, const T *cb, const T *cq, const T *const *prof, T rdz,
                         T rdy, T rdx, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
                         int64_t n_src, int64_t n_rec, bool ca_b, bool cb_b, bool cq_b,
                         const int64_t *pml, int64_t src_comp, int64_t rec_comp) {
  Problem3<T> p;
  Geom3<T> &g = p.g;
  g.pz = g.py = g.px = nullptr;
  g.rdz = rdz;
  g.rdy = rdy;
  g.rdx = rdx;
  g.nz = (int32_t)nz;
  g.ny = (int32_t)ny;
  g.nx = (int32_t)nx;
  g.numel = (int32_t)(nz * ny * nx);
  auto h1 = [](int64_t p0, int64_t p1) { return (int32_t)(p1 > p0 ? p1 - 1 : p0); };
  g.z0 = (int32_t)pml[0];
  g.y0 = (int32_t)pml[1];
  g.x0 = (int32_t)pml[2];
  g.z1 = (int32_t)pml[3];
  g.y1 = (int32_t)pml[4];
  g.x1 = (int32_t)pml[5];
  g.z1h = h1(pml[0], pml[3]);
  g.y1h = h1(pml[1], pml[4]);
  g.x1h = h1(pml[2], pml[5]);
  g.ca_b = ca_b;
  g.cb_b = cb_b;
  g.cq_b = cq_b;
  p.ca = ca;
  p.cb = cb;
  p.cq = cq;
  for (int k = 0; k < 18; ++k) p.prof[k] = prof[k];
  p.n_shots = (int32_t)n_shots;
  p.n_src = n_src;
  p.n_rec = n_rec;
  p.src_comp = (int32_t)src_comp;
  p.rec_comp = (int32_t)rec_comp;
  return p;
}

}  // namespace em3d
}  // namespace tide_hip

// ===================================================================== C ABI

using namespace tide_hip;
using namespace tide_hip::em3d;

namespace {

inline int set_device(int64_t device) {
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  return kOk;
}

template <typename T>
int em3d_forward(const T *ca, const T *cb, const T *cq, const T *f, T *const *fields, T *r,
                 const T *const *prof, const int64_t *src, const int64_t *rec, T rdz, T rdy,
                 T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
                 int64_t n_src, int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b,
                 bool cq_b, int64_t start_t, const int64_t *pml, int64_t src_comp,
                 int64_t rec_comp, void *const *stores, int64_t storage_format,
                 int64_t store_step_bytes, int64_t store_t0, int64_t device, void *stream) {
  TIDE_CHECK(set_device(device));
  if (fields == nullptr || prof == nullptr || pml == nullptr) {
    set_error("fields, profiles and pml bounds are required");
    return kInvalidArgument;
  }
  Problem3<T> p = make_problem<T>(ca, cb, cq, prof, rdz, rdy, rdx, n_shots, nz, ny, nx, n_src,
                                  n_rec, ca_b, cb_b, cq_b, pml, src_comp, rec_comp);
  const int fmt = storage_format == 1 ? 2 : 1;
  return forward_impl<T>(p, fields, f, r, src, rec, nt, start_t, step_ratio, stores, fmt,
                         (size_t)store_step_bytes, store_t0, (hipStream_t)stream);
}

template <typename T>
int em3d_backward(const T *ca, const T *cb, const T *cq, const T *grad_r, T *const *lam,
                  const void *const *stores, int64_t storage_format, int64_t store_step_bytes,
                  int64_t store_t0, T *grad_f, double *gca, double *gcb, const T *const *prof,
                  const int64_t *src, const int64_t *rec, T rdz, T rdy, T rdx, int64_t nt,
                  int64_t n_shots, int64_t nz, int64_t ny, int64_t nx, int64_t n_src,
                  int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b, bool cq_b,
                  int64_t start_t, const int64_t *pml, int64_t src_comp, int64_t rec_comp,
                  int64_t device, void *stream) {
  TIDE_CHECK(set_device(device));
  if (lam == nullptr || prof == nullptr || pml == nullptr) {
    set_error("adjoint fields, profiles and pml bounds are required");
    return kInvalidArgument;
  }
  Problem3<T> p = make_problem<T>(ca, cb, cq, prof, rdz, rdy, rdx, n_shots, nz, ny, nx, n_src,
                                  n_rec, ca_b, cb_b, cq_b, pml, src_comp, rec_comp);
  const int fmt = storage_format == 1 ? 2 : 1;
  return backward_impl<T>(p, grad_r, lam, stores, fmt, (size_t)store_step_bytes, store_t0,
                          grad_f, gca, gcb, src, rec, nt, start_t, step_ratio,
                          (hipStream_t)stream);
}

}  // namespace

#define TIDE_EM3D_ENTRY_POINTS(T, NAME)                                                     \
  extern "C" int tide_em3d_##NAME##_born_forward(                                          \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *f0,      \
      T const *df, T *const *fields, T *const *dfields, T *r, T *background_r,             \
      T const *const *prof, int64_t const *sources_i, int64_t const *receivers_i, T rdz,   \
      T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,       \
      int64_t n_src, int64_t n_rec, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,      \
      int64_t const *pml, int64_t source_component, int64_t receiver_component,            \
      int64_t device, void *compute_stream) {                                              \
    TIDE_CHECK(set_device(device));                                                        \
    if (fields == nullptr || dfields == nullptr || prof == nullptr || pml == nullptr) {    \
      set_error("fields, scattered fields, profiles and pml bounds are required");         \
      return kInvalidArgument;                                                             \
    }                                                                                      \
    Problem3<T> p = make_problem<T>(ca, cb, cq, prof, rdz, rdy, rdx, n_shots, nz, ny, nx,  \
                                    n_src, n_rec, ca_b, cb_b, cq_b, pml, source_component, \
                                    receiver_component);                                   \
    return born_forward_impl<T>(p, dca, dcb, fields, dfields, f0, df, r, background_r,     \
                                sources_i, receivers_i, nt, start_t,                       \
                                (hipStream_t)compute_stream);                              \
  }                                                                                        \
  extern "C" int tide_em3d_##NAME##_forward(                                               \
      T const *ca, T const *cb, T const *cq, T const *f, T *const *fields, T *r,           \
      T const *const *prof, int64_t const *sources_i, int64_t const *receivers_i, T rdz,   \
      T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,       \
      int64_t n_src, int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b, bool cq_b,   \
      int64_t start_t, int64_t const *pml, int64_t source_component,                      \
      int64_t receiver_component, void *const *stores, int64_t storage_format,             \
      int64_t store_step_bytes, int64_t store_t0, int64_t device, void *compute_stream) {  \
    return em3d_forward<T>(ca, cb, cq, f, fields, r, prof, sources_i, receivers_i, rdz,    \
                           rdy, rdx, nt, n_shots, nz, ny, nx, n_src, n_rec, step_ratio,    \
                           ca_b, cb_b, cq_b, start_t, pml, source_component,               \
                           receiver_component, stores, storage_format, store_step_bytes,   \
                           store_t0, device, compute_stream);                              \
  }                                                                                        \
  extern "C" int tide_em3d_##NAME##_backward_f64acc(                                       \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *const *lam,               \
      void const *const *stores, int64_t storage_format, int64_t store_step_bytes,         \
      int64_t store_t0, T *grad_f, double *grad_ca, double *grad_cb, T const *const *prof, \
      int64_t const *sources_i, int64_t const *receivers_i, T rdz, T rdy, T rdx,           \
      int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx, int64_t n_src,      \
      int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b, bool cq_b, int64_t start_t, \
      int64_t const *pml, int64_t source_component, int64_t receiver_component,            \
      int64_t device, void *compute_stream) {                                              \
    return em3d_backward<T>(ca, cb, cq, grad_r, lam, stores, storage_format,               \
                            store_step_bytes, store_t0, grad_f, grad_ca, grad_cb, prof,    \
                            sources_i, receivers_i, rdz, rdy, rdx, nt, n_shots, nz, ny,    \
                            nx, n_src, n_rec, step_ratio, ca_b, cb_b, cq_b, start_t, pml,  \
                            source_component, receiver_component, device,                  \
                            compute_stream);                                               \
  }

TIDE_EM3D_ENTRY_POINTS(float, float)
TIDE_EM3D_ENTRY_POINTS(double, double)

// ------------------------------------------------ reference-order 3-D ABI
//
// maxwell_3d_2_{float,double}_{forward,forward_with_storage,backward}_cuda
// take the reference's arguments in the order of its ctypes specs
// (src/tide/backend_utils.py:446-555: _3D_FORWARD_SPEC,
// _3D_FORWARD_WITH_STORAGE_SPEC, _3D_BACKWARD_SPEC) and return a status
// instead of void.  They run the tide_em3d_* drivers above.  Snapshots are
// device resident (storage_mode 0): the snapshot of step t goes to slot
// t / step_ratio of each *_s1 buffer, as in the reference's device mode.
// CPML memories (and their adjoint counterparts) are in the reference's
// compact layout.  Debye dispersion, host/disk staging and the
// reference's scratch and per-shot gradient buffers are not used (the
// adjoint keeps its own ping-pong memories and FP64 accumulators).

namespace {

constexpr int64_t kRefStorageDevice = 0;
constexpr int64_t kRefStorageNone = 3;

// d(ca, cb)/d(eps_r, sigma) chain rule of one cell (reference
// maxwell_3d_cuda_inst.cu:2729-2777 and utils.py:185-217):
//   dca/deps = (1 - ca) cb / dt, dcb/deps = -cb^2 / dt (per unit of
//   absolute permittivity, times EP0 per unit of eps_r),
//   dca/dsigma = -(1 + ca) cb / 2, dcb/dsigma = -cb^2 / 2.
template <typename T>
__global__ void coeff_grad_to_material3(const T *__restrict__ ca, const T *__restrict__ cb,
                                        const T *__restrict__ gca, const T *__restrict__ gcb,
                                        T *__restrict__ geps, T *__restrict__ gsig, T dt,
                                        int64_t numel, int64_t total, int ca_b, int cb_b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t shot = ca_b ? i / numel : 0;
  const int64_t j = i - shot * numel;
  const T a = ca[ca_b ? i : j];
  const T b = cb[cb_b ? shot * numel + j : j];
  const T b2 = b * b;
  const T inv_dt = (T)1 / dt;
  const T g_a = gca ? gca[i] : (T)0;
  const T g_b = gcb ? gcb[i] : (T)0;
  const T da_de = ((T)1 - a) * b * inv_dt;
  const T db_de = -b2 * inv_dt;
  const T da_ds = -((T)0.5) * ((T)1 + a) * b;
  const T db_ds = -((T)0.5) * b2;
  if (geps) geps[i] = (g_a * da_de + g_b * db_de) * (T)8.8541878128e-12;
  if (gsig) gsig[i] = g_a * da_ds + g_b * db_ds;
}

// The reference keeps each 3-D CPML memory compact along the axis of its
// strips (src/tide/grid_utils.py:96-160, maxwell_3d_cuda_inst.cu:307-363):
// along that axis only the cells [0, lo) and [hs, n) are stored, hs =
// max(lo, hi - 1), as [n_shots, .., lo + n - hs, ..].  The drivers here keep
// full-size memories and touch only the strips, so the reference-order
// entry points copy the strips between the two layouts.
__host__ __device__ inline int memory_axis3(int slot) {
  // slots M_HZ_Y .. M_EY_X: y z z x x y z y x z y x
  constexpr int axis[12] = {1, 0, 0, 2, 2, 1, 0, 1, 2, 0, 1, 2};
  return axis[slot - M_HZ_Y];
}

template <typename T>
__global__ void strip_copy3(T *__restrict__ full, T *__restrict__ compact, int64_t n_shots,
                            int nz, int ny, int nx, int axis, int lo, int hs, int to_full) {
  const int n[3] = {nz, ny, nx};
  const int ext = lo + n[axis] - hs;
  int dims[3] = {nz, ny, nx};
  dims[axis] = ext;
  const int64_t total = n_shots * (int64_t)dims[0] * dims[1] * dims[2];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int64_t rest = i;
  int c[3];
  c[2] = (int)(rest % dims[2]); rest /= dims[2];
  c[1] = (int)(rest % dims[1]); rest /= dims[1];
  c[0] = (int)(rest % dims[0]); rest /= dims[0];
  const int64_t shot = rest;
  if (c[axis] >= lo) c[axis] += hs - lo;
  const int64_t j = ((shot * nz + c[0]) * ny + c[1]) * (int64_t)nx + c[2];
  if (to_full) full[j] = compact[i];
  else compact[i] = full[j];
}

// Full-size working copies of the 12 compact memories of `state`
// (slots M_HZ_Y..M_EY_X), filled from them; `back` copies them back.
template <typename T> struct FullMemories {
  T *full[N_SLOTS] = {};
  int64_t n_shots = 0;
  int n[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hs[3] = {0, 0, 0};
  int open(T *const *state, int64_t shots, int nz, int ny, int nx, const int64_t *pml,
           Scratch &scr, hipStream_t stream) {
    n_shots = shots;
    n[0] = nz; n[1] = ny; n[2] = nx;
    for (int a = 0; a < 3; ++a) {
      lo[a] = (int)pml[a];
      hs[a] = std::max<int>((int)pml[a], (int)pml[3 + a] - 1);
      if (lo[a] < 0 || hs[a] > n[a]) {
        set_error("pml bounds outside the padded grid");
        return kInvalidArgument;
      }
    }
    const int64_t numel = (int64_t)nz * ny * nx;
    for (int k = M_HZ_Y; k < N_SLOTS; ++k) {
      TIDE_CHECK(scr.alloc(&full[k], (size_t)(shots * numel)));
      if (state[k] != nullptr) TIDE_CHECK(copy(k, state[k], 1, stream));
    }
    return kOk;
  }
  int copy(int k, T *compact, int to_full, hipStream_t stream) {
    const int a = memory_axis3(k);
    int dims[3] = {n[0], n[1], n[2]};
    dims[a] = lo[a] + n[a] - hs[a];
    const int64_t total = n_shots * (int64_t)dims[0] * dims[1] * dims[2];
    if (total == 0) return kOk;
    strip_copy3<T><<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(
        full[k], compact, n_shots, n[0], n[1], n[2], a, lo[a], hs[a], to_full);
    TIDE_HIP_TRY(hipGetLastError());
    return kOk;
  }
  int back(T *const *state, hipStream_t stream) {
    for (int k = M_HZ_Y; k < N_SLOTS; ++k)
      if (state[k] != nullptr) TIDE_CHECK(copy(k, state[k], 0, stream));
    return kOk;
  }
};

template <typename T>
int ref3_forward(const T *ca, const T *cb, const T *cq, const T *f, T *const fields[18], T *r,
                 const T *const prof[18], const int64_t *src, const int64_t *rec, T rdz, T rdy,
                 T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
                 int64_t n_src, int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b,
                 bool cq_b, int64_t start_t, const int64_t pml[6], int64_t src_comp,
                 int64_t rec_comp, int64_t device, void *stream, void *const stores[6],
                 int64_t storage_format, int64_t shot_bytes) {
  TIDE_CHECK(set_device(device));
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  FullMemories<T> mem;
  TIDE_CHECK(mem.open(fields, n_shots, (int)nz, (int)ny, (int)nx, pml, scr, s));
  T *work[N_SLOTS];
  for (int k = 0; k < N_SLOTS; ++k) work[k] = k < M_HZ_Y ? fields[k] : mem.full[k];
  TIDE_CHECK(em3d_forward<T>(ca, cb, cq, f, work, r, prof, src, rec, rdz, rdy, rdx, nt, n_shots,
                             nz, ny, nx, n_src, n_rec, step_ratio, ca_b, cb_b, cq_b, start_t,
                             pml, src_comp, rec_comp, stores, storage_format,
                             shot_bytes * n_shots, 0, device, stream));
  return mem.back(fields, s);
}

template <typename T>
int ref3_born_forward(const T *ca, const T *cb, const T *cq, const T *dca, const T *dcb,
                      const T *f0, const T *df, T *const fields[18], T *const dfields[18], T *r,
                      T *bg_r, const T *const prof[18], const int64_t *src, const int64_t *rec,
                      T rdz, T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny,
                      int64_t nx, int64_t n_src, int64_t n_rec, bool ca_b, bool cb_b, bool cq_b,
                      int64_t start_t, const int64_t pml[6], int64_t src_comp, int64_t rec_comp,
                      int64_t device, void *stream) {
  TIDE_CHECK(set_device(device));
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  FullMemories<T> mem, dmem;
  TIDE_CHECK(mem.open(fields, n_shots, (int)nz, (int)ny, (int)nx, pml, scr, s));
  TIDE_CHECK(dmem.open(dfields, n_shots, (int)nz, (int)ny, (int)nx, pml, scr, s));
  T *work[N_SLOTS], *dwork[N_SLOTS];
  for (int k = 0; k < N_SLOTS; ++k) {
    work[k] = k < M_HZ_Y ? fields[k] : mem.full[k];
    dwork[k] = k < M_HZ_Y ? dfields[k] : dmem.full[k];
  }
  Problem3<T> p = make_problem<T>(ca, cb, cq, prof, rdz, rdy, rdx, n_shots, nz, ny, nx, n_src,
                                  n_rec, ca_b, cb_b, cq_b, pml, src_comp, rec_comp);
  TIDE_CHECK(born_forward_impl<T>(p, dca, dcb, work, dwork, f0, df, r, bg_r, src, rec, nt,
                                  start_t, s));
  TIDE_CHECK(mem.back(fields, s));
  return dmem.back(dfields, s);
}

template <typename T>
int ref3_backward(const T *ca, const T *cb, const T *cq, const T *grad_r, T *const lam[18],
                  const void *const stores[6], T *grad_f, T *grad_ca, T *grad_cb, T *grad_eps,
                  T *grad_sigma, bool zero_grad_on_entry, const T *const prof[18],
                  const int64_t *src, const int64_t *rec, T rdz, T rdy, T rdx, T dt, int64_t nt,
                  int64_t n_shots, int64_t nz, int64_t ny, int64_t nx, int64_t n_src,
                  int64_t n_rec, int64_t step_ratio, int64_t storage_mode,
                  int64_t storage_format, int64_t shot_bytes, bool ca_rg, bool cb_rg,
                  bool ca_b, bool cb_b, bool cq_b, int64_t start_t, const int64_t pml[6],
                  int64_t src_comp, int64_t rec_comp, int64_t device, void *stream_handle) {
  if (ca_rg && cb_rg && ca_b != cb_b) {  // checked before any device call
    set_error("ca and cb gradients need the same batching (ca_batched == cb_batched)");
    return kInvalidArgument;
  }
  const bool want = ca_rg || cb_rg;
  if (want && storage_mode != kRefStorageDevice) {
    set_error(storage_mode == kRefStorageNone
                  ? "model gradients need stored snapshots (storage_mode 'none')"
                  : "host- and disk-staged 3-D snapshots are not built; use storage_mode 'device'");
    return storage_mode == kRefStorageNone ? kInvalidArgument : kUnsupported;
  }
  TIDE_CHECK(set_device(device));
  hipStream_t stream = (hipStream_t)stream_handle;
  const int64_t numel = nz * ny * nx;
  const int64_t na = ca_b ? n_shots * numel : numel;
  const int64_t nb = cb_b ? n_shots * numel : numel;
  if (zero_grad_on_entry) {
    if (grad_f && nt > 0 && n_src > 0)
      TIDE_HIP_TRY(hipMemsetAsync(grad_f, 0, (size_t)(nt * n_shots * n_src) * sizeof(T), stream));
    if (ca_rg && grad_ca) TIDE_HIP_TRY(hipMemsetAsync(grad_ca, 0, (size_t)na * sizeof(T), stream));
    if (cb_rg && grad_cb) TIDE_HIP_TRY(hipMemsetAsync(grad_cb, 0, (size_t)nb * sizeof(T), stream));
    if (want && grad_eps) TIDE_HIP_TRY(hipMemsetAsync(grad_eps, 0, (size_t)na * sizeof(T), stream));
    if (want && grad_sigma)
      TIDE_HIP_TRY(hipMemsetAsync(grad_sigma, 0, (size_t)na * sizeof(T), stream));
  }
  if (want && ((ca_rg && !grad_ca) || (cb_rg && !grad_cb))) {
    set_error("grad_ca / grad_cb buffers are required for the requested gradients");
    return kInvalidArgument;
  }
  Scratch scr(stream);
  double *ga = nullptr, *gb = nullptr;
  if (ca_rg) {
    TIDE_CHECK(scr.alloc(&ga, (size_t)na));
    TIDE_HIP_TRY(hipMemsetAsync(ga, 0, (size_t)na * sizeof(double), stream));
  }
  if (cb_rg) {
    TIDE_CHECK(scr.alloc(&gb, (size_t)nb));
    TIDE_HIP_TRY(hipMemsetAsync(gb, 0, (size_t)nb * sizeof(double), stream));
  }
  FullMemories<T> mem;
  TIDE_CHECK(mem.open(lam, n_shots, (int)nz, (int)ny, (int)nx, pml, scr, stream));
  T *work[N_SLOTS];
  for (int k = 0; k < N_SLOTS; ++k) work[k] = k < M_HZ_Y ? lam[k] : mem.full[k];
  TIDE_CHECK(em3d_backward<T>(ca, cb, cq, grad_r, work, want ? stores : nullptr, storage_format,
                              shot_bytes * n_shots, 0, grad_f, ga, gb, prof, src, rec, rdz, rdy,
                              rdx, nt, n_shots, nz, ny, nx, n_src, n_rec, step_ratio, ca_b,
                              cb_b, cq_b, start_t, pml, src_comp, rec_comp, device,
                              stream_handle));
  TIDE_CHECK(mem.back(lam, stream));
  const unsigned blocks_a = (unsigned)((na + 255) / 256), blocks_b = (unsigned)((nb + 255) / 256);
  if (ca_rg) fold_grad3<T><<<blocks_a, 256, 0, stream>>>(grad_ca, ga, na);
  if (cb_rg) fold_grad3<T><<<blocks_b, 256, 0, stream>>>(grad_cb, gb, nb);
  if (want && (grad_eps || grad_sigma))
    coeff_grad_to_material3<T><<<blocks_a, 256, 0, stream>>>(
        ca, cb, ca_rg ? grad_ca : nullptr, cb_rg ? grad_cb : nullptr, grad_eps, grad_sigma, dt,
        numel, na, ca_b ? 1 : 0, cb_b ? 1 : 0);
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace

#define TIDE_REF3_FIELDS(T, P)                                                                  \
  T *P##ex, T *P##ey, T *P##ez, T *P##hx, T *P##hy, T *P##hz, T *P##m_hz_y, T *P##m_hy_z,   \
      T *P##m_hx_z, T *P##m_hz_x, T *P##m_hy_x, T *P##m_hx_y, T *P##m_ey_z, T *P##m_ez_y,   \
      T *P##m_ez_x, T *P##m_ex_z, T *P##m_ex_y, T *P##m_ey_x
#define TIDE_REF3_PROFILES(T)                                                                   \
  T const *az, T const *bz, T const *azh, T const *bzh, T const *ay, T const *by,            \
      T const *ayh, T const *byh, T const *ax, T const *bx, T const *axh, T const *bxh,      \
      T const *kz, T const *kzh, T const *ky, T const *kyh, T const *kx, T const *kxh,       \
      int64_t const *sources_i, int64_t const *receivers_i
#define TIDE_REF3_PROF_ARRAY(T)                                                                 \
  const T *prof[18] = {az, bz, azh, bzh, ay, by, ayh, byh, ax, bx, axh, bxh,                 \
                       kz, kzh, ky, kyh, kx, kxh}
#define TIDE_REF3_TAIL(T)                                                                       \
  T rdz, T rdy, T rdx, T dt, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx, \
      int64_t n_src, int64_t n_rec, int64_t step_ratio
#define TIDE_REF3_FLAGS                                                                      \
  bool ca_b, bool cb_b, bool cq_b, int64_t start_t, int64_t pml_z0, int64_t pml_y0,         \
      int64_t pml_x0, int64_t pml_z1, int64_t pml_y1, int64_t pml_x1,                        \
      int64_t source_component, int64_t receiver_component, int64_t n_threads,               \
      int64_t device, int64_t execution_backend
#define TIDE_REF3_STORES(P)                                                                  \
  void *P##ex_s1, void *P##ex_s3, char const *const *P##ex_fn, void *P##ey_s1,               \
      void *P##ey_s3, char const *const *P##ey_fn, void *P##ez_s1, void *P##ez_s3,           \
      char const *const *P##ez_fn, void *P##cx_s1, void *P##cx_s3,                           \
      char const *const *P##cx_fn, void *P##cy_s1, void *P##cy_s3,                           \
      char const *const *P##cy_fn, void *P##cz_s1, void *P##cz_s3, char const *const *P##cz_fn

#define TIDE_REF3_ENTRY_POINTS(T, NAME)                                                      \
  extern "C" int maxwell_3d_2_##NAME##_forward_cuda(                                        \
      T const *ca, T const *cb, T const *cq, T const *f, TIDE_REF3_FIELDS(T, ),                \
      T const *debye_a, T const *debye_b, T const *debye_cp, T *pol_ex, T *pol_ey,          \
      T *pol_ez, T *ex_prev, T *ey_prev, T *ez_prev, T *r, int64_t n_poles,                 \
      TIDE_REF3_PROFILES(T), TIDE_REF3_TAIL(T), bool has_dispersion, TIDE_REF3_FLAGS,             \
      void *compute_stream) {                                                               \
    (void)debye_a; (void)debye_b; (void)debye_cp; (void)pol_ex; (void)pol_ey;             \
    (void)pol_ez; (void)ex_prev; (void)ey_prev; (void)ez_prev; (void)n_poles; (void)dt;   \
    (void)n_threads; (void)execution_backend;                                               \
    if (has_dispersion) {                                                                   \
      set_error("Debye dispersion is not part of this build");                              \
      return kUnsupported;                                                                  \
    }                                                                                       \
    T *const fields[18] = {ex, ey, ez, hx, hy, hz, m_hz_y, m_hy_z, m_hx_z, m_hz_x,         \
                           m_hy_x, m_hx_y, m_ey_z, m_ez_y, m_ez_x, m_ex_z, m_ex_y, m_ey_x}; \
    TIDE_REF3_PROF_ARRAY(T);                                                                   \
    const int64_t pml[6] = {pml_z0, pml_y0, pml_x0, pml_z1, pml_y1, pml_x1};               \
    return ref3_forward<T>(ca, cb, cq, f, fields, r, prof, sources_i, receivers_i, rdz,     \
                           rdy, rdx, nt, n_shots, nz, ny, nx, n_src, n_rec, step_ratio,     \
                           ca_b, cb_b, cq_b, start_t, pml, source_component,                \
                           receiver_component, device, compute_stream, nullptr, 0, 0);      \
  }                                                                                         \
  extern "C" int maxwell_3d_2_##NAME##_born_forward_cuda(                                   \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *f0,       \
      T const *df, TIDE_REF3_FIELDS(T, ), TIDE_REF3_FIELDS(T, d), T *r, T *background_r,     \
      TIDE_REF3_PROFILES(T), TIDE_REF3_TAIL(T), TIDE_REF3_FLAGS, void *compute_stream) {     \
    (void)dt; (void)step_ratio; (void)n_threads; (void)execution_backend;                   \
    T *const fields[18] = {ex, ey, ez, hx, hy, hz, m_hz_y, m_hy_z, m_hx_z, m_hz_x,         \
                           m_hy_x, m_hx_y, m_ey_z, m_ez_y, m_ez_x, m_ex_z, m_ex_y, m_ey_x}; \
    T *const dfields[18] = {dex, dey, dez, dhx, dhy, dhz, dm_hz_y, dm_hy_z, dm_hx_z,        \
                            dm_hz_x, dm_hy_x, dm_hx_y, dm_ey_z, dm_ez_y, dm_ez_x, dm_ex_z,  \
                            dm_ex_y, dm_ey_x};                                              \
    TIDE_REF3_PROF_ARRAY(T);                                                                   \
    const int64_t pml[6] = {pml_z0, pml_y0, pml_x0, pml_z1, pml_y1, pml_x1};               \
    return ref3_born_forward<T>(ca, cb, cq, dca, dcb, f0, df, fields, dfields, r,           \
                                background_r, prof, sources_i, receivers_i, rdz, rdy, rdx,  \
                                nt, n_shots, nz, ny, nx, n_src, n_rec, ca_b, cb_b, cq_b,    \
                                start_t, pml, source_component, receiver_component, device, \
                                compute_stream);                                            \
  }                                                                                         \
  extern "C" int maxwell_3d_2_##NAME##_forward_with_storage_cuda(                           \
      T const *ca, T const *cb, T const *cq, T const *f, TIDE_REF3_FIELDS(T, ), T *r,          \
      TIDE_REF3_STORES(), TIDE_REF3_PROFILES(T), TIDE_REF3_TAIL(T), int64_t storage_mode,        \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_rg, bool cb_rg,           \
      TIDE_REF3_FLAGS, void *compute_stream, void *storage_stream) {                        \
    (void)ex_s3; (void)ey_s3; (void)ez_s3; (void)cx_s3; (void)cy_s3; (void)cz_s3;          \
    (void)ex_fn; (void)ey_fn; (void)ez_fn; (void)cx_fn; (void)cy_fn; (void)cz_fn;          \
    (void)dt; (void)n_threads; (void)execution_backend; (void)storage_stream;              \
    const bool want = ca_rg || cb_rg;                                                       \
    if (want && storage_mode != kRefStorageDevice && storage_mode != kRefStorageNone) {    \
      set_error("host- and disk-staged 3-D snapshots are not built; use storage_mode "      \
                "'device'");                                                                \
      return kUnsupported;                                                                  \
    }                                                                                       \
    T *const fields[18] = {ex, ey, ez, hx, hy, hz, m_hz_y, m_hy_z, m_hx_z, m_hz_x,         \
                           m_hy_x, m_hx_y, m_ey_z, m_ez_y, m_ez_x, m_ex_z, m_ex_y, m_ey_x}; \
    TIDE_REF3_PROF_ARRAY(T);                                                                   \
    const int64_t pml[6] = {pml_z0, pml_y0, pml_x0, pml_z1, pml_y1, pml_x1};               \
    const bool store = want && storage_mode == kRefStorageDevice;                          \
    void *const stores[6] = {store && ca_rg ? ex_s1 : nullptr, store && ca_rg ? ey_s1 : nullptr, \
                             store && ca_rg ? ez_s1 : nullptr, store && cb_rg ? cx_s1 : nullptr, \
                             store && cb_rg ? cy_s1 : nullptr, store && cb_rg ? cz_s1 : nullptr}; \
    return ref3_forward<T>(ca, cb, cq, f, fields, r, prof, sources_i, receivers_i, rdz,     \
                           rdy, rdx, nt, n_shots, nz, ny, nx, n_src, n_rec, step_ratio,     \
                           ca_b, cb_b, cq_b, start_t, pml, source_component,                \
                           receiver_component, device, compute_stream,                      \
                           store ? stores : nullptr, storage_format, shot_bytes_uncomp);    \
  }                                                                                         \
  extern "C" int maxwell_3d_2_##NAME##_backward_cuda(                                       \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ex, T *lambda_ey,   \
      T *lambda_ez, T *lambda_hx, T *lambda_hy, T *lambda_hz, T *m_lambda_ey_z,             \
      T *m_lambda_ez_y, T *m_lambda_ez_x, T *m_lambda_ex_z, T *m_lambda_ex_y,               \
      T *m_lambda_ey_x, T *m_lambda_hz_y, T *m_lambda_hy_z, T *m_lambda_hx_z,               \
      T *m_lambda_hz_x, T *m_lambda_hy_x, T *m_lambda_hx_y,                                 \
      T *adjoint_memory_scratch, TIDE_REF3_STORES(), T *grad_f, T *grad_ca, T *grad_cb,     \
      T *grad_eps, T *grad_sigma, T *grad_ca_shot, T *grad_cb_shot,                         \
      bool zero_grad_on_entry, TIDE_REF3_PROFILES(T), TIDE_REF3_TAIL(T), int64_t storage_mode,   \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_rg, bool cb_rg,           \
      TIDE_REF3_FLAGS, void *compute_stream, void *storage_stream) {                        \
    (void)adjoint_memory_scratch; (void)grad_ca_shot; (void)grad_cb_shot;                   \
    (void)ex_s3; (void)ey_s3; (void)ez_s3; (void)cx_s3; (void)cy_s3; (void)cz_s3;          \
    (void)ex_fn; (void)ey_fn; (void)ez_fn; (void)cx_fn; (void)cy_fn; (void)cz_fn;          \
    (void)n_threads; (void)execution_backend; (void)storage_stream;                         \
    /* slots in the forward state order: each adjoint memory sits in the */              \
    /* slot of the forward memory it is the transpose of                   */              \
    T *const lam[18] = {lambda_ex, lambda_ey, lambda_ez, lambda_hx, lambda_hy, lambda_hz,  \
                        m_lambda_hz_y, m_lambda_hy_z, m_lambda_hx_z, m_lambda_hz_x,         \
                        m_lambda_hy_x, m_lambda_hx_y, m_lambda_ey_z, m_lambda_ez_y,         \
                        m_lambda_ez_x, m_lambda_ex_z, m_lambda_ex_y, m_lambda_ey_x};        \
    TIDE_REF3_PROF_ARRAY(T);                                                                   \
    const int64_t pml[6] = {pml_z0, pml_y0, pml_x0, pml_z1, pml_y1, pml_x1};               \
    const void *const stores[6] = {ex_s1, ey_s1, ez_s1, cx_s1, cy_s1, cz_s1};              \
    return ref3_backward<T>(ca, cb, cq, grad_r, lam, stores, grad_f, grad_ca, grad_cb,      \
                            grad_eps, grad_sigma, zero_grad_on_entry, prof, sources_i,      \
                            receivers_i, rdz, rdy, rdx, dt, nt, n_shots, nz, ny, nx, n_src, \
                            n_rec, step_ratio, storage_mode, storage_format,                \
                            shot_bytes_uncomp, ca_rg, cb_rg, ca_b, cb_b, cq_b, start_t, pml,\
                            source_component, receiver_component, device, compute_stream);  \
  }

TIDE_REF3_ENTRY_POINTS(float, float)
TIDE_REF3_ENTRY_POINTS(double, double)

extern "C" int tide_em3d_set_tuning(int fwd_z, int adj_z, int group) {
  const bool ok = (fwd_z == 4 || fwd_z == 8 || fwd_z == 16) && (adj_z == 4 || adj_z == 8) &&
                  group >= 1 && group <= kMaxGroup;
  if (!ok) {
    set_error("tuning: fwd_z in {4,8,16}, adj_z in {4,8}, group 1..8");
    return kInvalidArgument;
  }
  g_tuning3.fwd_z = fwd_z;
  g_tuning3.adj_z = adj_z;
  g_tuning3.group = group;
  g_tuning3.fwd_group = group;
  return kOk;
}

extern "C" int tide_em3d_set_fused(int rows, int planes) {
  if (!(rows == 0 || rows == 2 || rows == 4) || !(planes == 8 || planes == 16)) {
    set_error("fused step: rows in {0 (off), 2, 4}, planes in {8, 16}");
    return kInvalidArgument;
  }
  g_tuning3.fused_rows = rows;
  g_tuning3.fused_z = planes;
  return kOk;
}

extern "C" int tide_em3d_set_tile_order(int shot_fast) {
  if (shot_fast != 0 && shot_fast != 1) {
    set_error("tile order: 0 (shot groups outermost) or 1 (shot groups of a tile adjacent)");
    return kInvalidArgument;
  }
  g_tuning3.shot_fast = shot_fast;
  return kOk;
}

extern "C" int tide_em3d_get_fused(int *out2) {
  if (out2 == nullptr) {
    set_error("get_fused: null output");
    return kInvalidArgument;
  }
  out2[0] = g_tuning3.fused_rows;
  out2[1] = g_tuning3.fused_z;
  return kOk;
}

extern "C" int tide_em3d_set_forward_group(int group) {
  if (group < 1 || group > kMaxGroup) {
    set_error("forward group: 1..8 shots per workgroup");
    return kInvalidArgument;
  }
  g_tuning3.fwd_group = group;
  return kOk;
}
