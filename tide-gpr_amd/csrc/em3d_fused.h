// One-launch 3-D forward step: the H update and the E update of a time step
// in one kernel (included by em3d.hip, after the two-kernel step).
//
// A wavefront owns 62 columns (lanes 1..62) of TY rows of one shot and
// marches through TZ planes.  At plane z it updates H on rows y0-1 .. y0+TY-1
// (the row above the tile and the left lane 0 are a recomputed halo, as is
// plane zs-1 before the first plane of the march), then E on its own rows
// from the new H of plane z, of the row above and of plane z-1 (carried in
// registers).  HBM therefore sees each field once per step instead of twice
// (the two-kernel step reads E in both kernels and H in both).
//
// The halo cells are owned by neighbouring workgroups, which may already
// have written their new values: the fields and the H-phase CPML memories
// (read at halo cells) therefore ping-pong between an input and an output
// buffer; the E-phase memories are read and written only by the owner of a
// cell and stay in place.  Every cell update is the two-kernel step's
// expression in the same operation order, so results are bit-identical.
#pragma once

namespace tide_hip {
namespace em3d {

// H-phase memories, in the order the fused kernel indexes them
enum { FH_EY_Z = 0, FH_EX_Z, FH_EZ_Y, FH_EX_Y, FH_EZ_X, FH_EY_X };
// E-phase memories
enum { FE_HY_Z = 0, FE_HX_Z, FE_HZ_Y, FE_HX_Y, FE_HZ_X, FE_HY_X };

template <typename T> struct Fused3 {
  Geom3<T> g;
  Points pts;
  const T *ca, *cb, *cq;
  const T *fi[6];   // Ex Ey Ez Hx Hy Hz at step t
  T *fo[6];         // the same at step t + 1
  const T *mhi[6];  // H-phase memories at step t (FH_* order)
  T *mho[6];        // ... at step t + 1
  T *me[6];         // E-phase memories (FE_* order), updated in place
  void *st[6];      // E x/y/z, curl x/y/z snapshots of this step
  const T *src_amp;
  T *rec_out;
  int32_t n_shots, n_yt, n_zb, src_comp, rec_comp;
  int32_t shot_fast;  // workgroup order, as decode3
};

// CPML correction reading the memory from `mi` and writing the new one to
// `mo` when `own`: m' = b m + a d; d' = d / k + m'.
template <typename T>
__device__ __forceinline__ T cpml_pp(T d, const T *__restrict__ mi, T *__restrict__ mo, int j,
                                     T a, T b, T k, bool own) {
  const T mv = b * mi[j] + a * d;
  if (own) mo[j] = mv;
  return d / k + mv;
}

template <typename T, int TY, int TZ, int FMT>
__global__ __launch_bounds__(256) void em3_fused(Fused3<T> a) {
  const Geom3<T> &g = a.g;
  const int G = (int)blockDim.y;
  const int n_groups = (a.n_shots + G - 1) / G;
  const int n_xt = a.pts.n_xt;
  const int64_t n_tiles = (int64_t)n_xt * a.n_yt * a.n_zb;
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, n_tiles * n_groups);
  int grp;
  int64_t rem;
  if (a.shot_fast) {
    rem = flat / n_groups;
    grp = (int)(flat - rem * n_groups);
  } else {
    grp = (int)(flat / n_tiles);
    rem = flat - (int64_t)grp * n_tiles;
  }
  const int zb = (int)(rem / ((int64_t)n_xt * a.n_yt));
  rem -= (int64_t)zb * n_xt * a.n_yt;
  const int yt = (int)(rem / n_xt);
  const int xt = (int)(rem - (int64_t)yt * n_xt);
  const int s = grp * G + __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one per wavefront
  if (s >= a.n_shots) return;

  const int nz = g.nz, ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int x = xt * kOwn - 1 + lane;
  const int y0 = yt * TY;
  const bool xin = x >= 1 && x < nx;              // columns H and E update
  const bool xown = xin && lane >= 1 && lane <= kOwn;  // columns this wavefront writes
  const int64_t base = (int64_t)s * g.numel;
  const T *__restrict__ ex = a.fi[0] + base;
  const T *__restrict__ ey = a.fi[1] + base;
  const T *__restrict__ ez = a.fi[2] + base;
  const T *__restrict__ hx = a.fi[3] + base;
  const T *__restrict__ hy = a.fi[4] + base;
  const T *__restrict__ hz = a.fi[5] + base;
  const T *__restrict__ cq = a.cq + (g.cq_b ? base : 0);
  const T *__restrict__ ca = a.ca + (g.ca_b ? base : 0);
  const T *__restrict__ cb = a.cb + (g.cb_b ? base : 0);
  const Ld3<T> L(x, nx, ny, nz);
  const bool has_dx = x < nx - 1;
  const bool pxh = x < g.x0 || x >= g.x1h;
  const bool px = x < g.x0 || x >= g.x1;
  const int xc = min(max(x, 0), nx - 1);
  const T axh = g.X(kAh, xc), bxh = g.X(kBh, xc), kxh = g.X(kKh, xc);
  const T axe = g.X(kA, xc), bxe = g.X(kB, xc), kxe = g.X(kK, xc);
  const int zs = zb * TZ;
  const int ze = min(zs + TZ, nz);

  // E of plane z on rows y0-1 .. y0+TY (index r+1 for row y0+r)
  T e0c[TY + 2], e1c[TY + 2], e2c[TY + 2];
#pragma unroll
  for (int i = 0; i < TY + 2; ++i) {
    e0c[i] = L(ex, zs - 1, y0 - 1 + i);
    e1c[i] = L(ey, zs - 1, y0 - 1 + i);
    e2c[i] = L(ez, zs - 1, y0 - 1 + i);
  }
  // new H of plane z - 1 on the tile's rows (Hx, Hy), for the E z-difference
  T hx_zm[TY], hy_zm[TY];
#pragma unroll
  for (int i = 0; i < TY; ++i) hx_zm[i] = hy_zm[i] = (T)0;

#pragma unroll
  for (int it = 0; it <= TZ; ++it) {
    const int z = zs - 1 + it;
    if (z >= ze) break;  // uniform
    // E of plane z + 1 (the H z-difference)
    T e0n[TY + 2], e1n[TY + 2], e2n[TY + 2];
#pragma unroll
    for (int i = 0; i < TY + 2; ++i) {
      e0n[i] = L(ex, z + 1, y0 - 1 + i);
      e1n[i] = L(ey, z + 1, y0 - 1 + i);
      e2n[i] = L(ez, z + 1, y0 - 1 + i);
    }
    const bool zin = z >= 1 && z < nz;
    const bool has_dz = z < nz - 1;
    const bool pzh = z < g.z0 || z >= g.z1h;  // uniform
    const int zc = min(max(z, 0), nz - 1);
    const T azh = g.Z(kAh, zc), bzh = g.Z(kBh, zc), kzh = g.Z(kKh, zc);
    const bool emit = it >= 1;  // plane of the march (not the halo plane)

    // ---- H on rows y0-1 .. y0+TY-1 (index i for row y0-1+i)
    T hxn[TY + 1], hyn[TY + 1], hzn[TY + 1];
#pragma unroll
    for (int i = 0; i < TY + 1; ++i) {
      const int y = y0 - 1 + i;
      const T hx0 = L(hx, z, y), hy0 = L(hy, z, y), hz0 = L(hz, z, y);
      const T cqv = L(cq, z, y);
      const T exc = e0c[i], eyc = e1c[i], ezc = e2c[i];
      const T ez_xp = shfl_dn1(ezc), ey_xp = shfl_dn1(eyc);
      const bool upd = xin && zin && y >= 1 && y < ny;
      const bool own = upd && emit && xown && i >= 1;
      const int j = (zc * ny + min(max(y, 0), ny - 1)) * nx + xc;
      T dey_dz = (T)0, dez_dy = (T)0, dez_dx = (T)0, dex_dz = (T)0, dex_dy = (T)0, dey_dx = (T)0;
      if (upd) {
        if (has_dz) {
          dey_dz = (e1n[i] - eyc) * g.rdz;
          dex_dz = (e0n[i] - exc) * g.rdz;
          if (pzh) {
            dey_dz = cpml_pp(dey_dz, a.mhi[FH_EY_Z] + base, a.mho[FH_EY_Z] + base, j, azh, bzh, kzh, own);
            dex_dz = cpml_pp(dex_dz, a.mhi[FH_EX_Z] + base, a.mho[FH_EX_Z] + base, j, azh, bzh, kzh, own);
          }
        }
        if (y < ny - 1) {
          dez_dy = (e2c[i + 1] - ezc) * g.rdy;
          dex_dy = (e0c[i + 1] - exc) * g.rdy;
          if (y < g.y0 || y >= g.y1h) {
            const T ay = g.Y(kAh, y), by = g.Y(kBh, y), ky = g.Y(kKh, y);
            dez_dy = cpml_pp(dez_dy, a.mhi[FH_EZ_Y] + base, a.mho[FH_EZ_Y] + base, j, ay, by, ky, own);
            dex_dy = cpml_pp(dex_dy, a.mhi[FH_EX_Y] + base, a.mho[FH_EX_Y] + base, j, ay, by, ky, own);
          }
        }
        if (has_dx) {
          dez_dx = (ez_xp - ezc) * g.rdx;
          dey_dx = (ey_xp - eyc) * g.rdx;
          if (pxh) {
            dez_dx = cpml_pp(dez_dx, a.mhi[FH_EZ_X] + base, a.mho[FH_EZ_X] + base, j, axh, bxh, kxh, own);
            dey_dx = cpml_pp(dey_dx, a.mhi[FH_EY_X] + base, a.mho[FH_EY_X] + base, j, axh, bxh, kxh, own);
          }
        }
      }
      hxn[i] = upd ? hx0 - cqv * (dey_dz - dez_dy) : hx0;
      hyn[i] = upd ? hy0 - cqv * (dez_dx - dex_dz) : hy0;
      hzn[i] = upd ? hz0 - cqv * (dex_dy - dey_dx) : hz0;
      if (own) {
        const int64_t line = (int64_t)(z * ny + y) * nx;
        *elem32(a.fo[3] + base + line, (uint32_t)x) = hxn[i];
        *elem32(a.fo[4] + base + line, (uint32_t)x) = hyn[i];
        *elem32(a.fo[5] + base + line, (uint32_t)x) = hzn[i];
      }
    }

    // ---- E on rows y0 .. y0+TY-1 (index r)
    if (emit) {
      const bool pz = z < g.z0 || z >= g.z1;  // uniform
      const T aze = g.Z(kA, zc), bze = g.Z(kB, zc), kze = g.Z(kK, zc);
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        const int y = y0 + r;
        const T hxc = hxn[r + 1], hyc = hyn[r + 1], hzc = hzn[r + 1];
        const T hz_ym = hzn[r], hx_ym = hxn[r];
        const T hz_xm = shfl_up1(hzc), hy_xm = shfl_up1(hyc);
        const T cav = L(ca, z, y), cbv = L(cb, z, y);
        const bool own = xown && zin && y >= 1 && y < ny;
        if (own) {
          const int j = (z * ny + y) * nx + x;
          T dhy_dz = (hyc - hy_zm[r]) * g.rdz;
          T dhz_dy = (hzc - hz_ym) * g.rdy;
          T dhz_dx = (hzc - hz_xm) * g.rdx;
          T dhx_dz = (hxc - hx_zm[r]) * g.rdz;
          T dhx_dy = (hxc - hx_ym) * g.rdy;
          T dhy_dx = (hyc - hy_xm) * g.rdx;
          if (pz) {
            dhy_dz = cpml(dhy_dz, a.me[FE_HY_Z] + base, j, aze, bze, kze, true);
            dhx_dz = cpml(dhx_dz, a.me[FE_HX_Z] + base, j, aze, bze, kze, true);
          }
          if (y < g.y0 || y >= g.y1) {
            const T ay = g.Y(kA, y), by = g.Y(kB, y), ky = g.Y(kK, y);
            dhz_dy = cpml(dhz_dy, a.me[FE_HZ_Y] + base, j, ay, by, ky, true);
            dhx_dy = cpml(dhx_dy, a.me[FE_HX_Y] + base, j, ay, by, ky, true);
          }
          if (px) {
            dhz_dx = cpml(dhz_dx, a.me[FE_HZ_X] + base, j, axe, bxe, kxe, true);
            dhy_dx = cpml(dhy_dx, a.me[FE_HY_X] + base, j, axe, bxe, kxe, true);
          }
          const T cx = dhy_dz - dhz_dy;
          const T cy = dhz_dx - dhx_dz;
          const T cz = dhx_dy - dhy_dx;
          T v0 = e0c[r + 1], v1 = e1c[r + 1], v2 = e2c[r + 1];
          if constexpr (FMT != 0) {
            const int64_t i = base + j;
            if (a.st[0]) {
              put_store<T, FMT>(a.st[0], i, v0);
              put_store<T, FMT>(a.st[1], i, v1);
              put_store<T, FMT>(a.st[2], i, v2);
            }
            if (a.st[3]) {
              put_store<T, FMT>(a.st[3], i, cx);
              put_store<T, FMT>(a.st[4], i, cy);
              put_store<T, FMT>(a.st[5], i, cz);
            }
          }
          v0 = cav * v0 + cbv * cx;
          v1 = cav * v1 + cbv * cy;
          v2 = cav * v2 + cbv * cz;
          const int slot = a.pts.row_slot[z * ny + y];
          if (slot >= 0) {
            const int64_t si = ((int64_t)slot * a.n_shots + s) * n_xt + xt;
            const int e1_ = a.pts.seg[si + 1];
            for (int e = a.pts.seg[si]; e < e1_; ++e) {
              const int2 en = a.pts.ent[e];
              if ((en.x & 0xff) != lane) continue;
              if ((en.x >> 8) == 0) {
                const T v = a.src_amp[en.y];
                if (a.src_comp == 0) v0 += v;
                else if (a.src_comp == 1) v1 += v;
                else v2 += v;
              } else {
                a.rec_out[en.y] = a.rec_comp == 0 ? v0 : (a.rec_comp == 1 ? v1 : v2);
              }
            }
          }
          const int64_t line = (int64_t)(z * ny + y) * nx;
          *elem32(a.fo[0] + base + line, (uint32_t)x) = v0;
          *elem32(a.fo[1] + base + line, (uint32_t)x) = v1;
          *elem32(a.fo[2] + base + line, (uint32_t)x) = v2;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      hx_zm[r] = hxn[r + 1];
      hy_zm[r] = hyn[r + 1];
    }
#pragma unroll
    for (int i = 0; i < TY + 2; ++i) {
      e0c[i] = e0n[i];
      e1c[i] = e1n[i];
      e2c[i] = e2n[i];
    }
  }
}

}  // namespace em3d
}  // namespace tide_hip
