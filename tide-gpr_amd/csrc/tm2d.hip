// 2-D TM Maxwell FDTD propagator for MI355X (gfx950).
//
// One fused kernel per forward time step and one per adjoint time step.
//
// Layout: every field is [n_shots, ny, nx] fp32 (or fp64), x fastest.  A
// workgroup is W wavefronts of 64 lanes; a wavefront owns 64 consecutive
// columns and marches down TY consecutive rows, so each cell is read once
// from HBM: the row below is carried in registers, the x-neighbours come
// through cross-lane shuffles, and only the first row of a march and the
// lanes at the wavefront edges recompute one neighbouring H (or adjoint raw)
// value.  Fields ping-pong between two buffers per step (the caller's arrays
// and a scratch copy), which removes every inter-workgroup read/write race
// of fusing the H and E half-steps.  Sources, receivers, the adjoint
// injection and the FP64 coefficient-gradient accumulation are applied by
// the thread that owns the cell, inside the same kernel.
//
// Semantics follow the reference propagator (paths relative to src/tide/):
//   H update          csrc/tm2d/maxwell_tm_core.cuh:448-502
//   E update/storage  csrc/tm2d/maxwell_tm_core.cuh:505-599
//   point ops         csrc/tm2d/maxwell.cu:89-161
//   exact adjoint     csrc/tm2d/maxwell_tm_cuda_inst.cu:922-1254
//   reverse loop      csrc/tm2d/maxwell_tm_cuda_inst.cu:2540-2731
// Stencil order 2 (the reference default) is the one built here.

#include "tide_common.h"
#include "tide_hip.h"

#include <algorithm>
#include <cstring>

namespace tide_hip {
namespace tm2d {

constexpr int kTileW = 64;

template <typename T> struct Geom {
  const T *ay, *by, *ayh, *byh, *ax, *bx, *axh, *bxh, *ky, *kyh, *kx, *kxh;
  T rdy, rdx;
  int32_t ny, nx;
  int64_t numel;
  int32_t pml_y0, pml_y1, pml_x0, pml_x1, pml_y1h, pml_x1h;
  int64_t ca_stride, cb_stride, cq_stride;  // 0 (shared) or numel (per shot)
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
  int32_t n_shots, group, n_yt;
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
  int64_t g_stride;  // 0: summed over shots, numel: per shot
  double step_scale;
  const T *grad_r_t;
  T *grad_f_t;
  int32_t n_shots, group, n_yt;
  int32_t use_atomics;
};

template <typename T>
__device__ __forceinline__ T ld(const T *p, int y, int x, int ny, int nx) {
  return (y >= 0 && y < ny && x >= 0 && x < nx) ? p[(int64_t)y * nx + x] : (T)0;
}

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

// ----------------------------------------------------------------- forward
// FMT: 0 = no snapshot, 1 = T payload, 2 = BF16 payload.
template <typename T, int TY, int W, int FMT>
__global__ __launch_bounds__(64 * W) void fwd_step(FwdArgs<T> a) {
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int n_groups = (a.n_shots + a.group - 1) / a.group;
  const TileCoord tc = decode_tile(a.pts.n_xt, a.n_yt, n_groups);
  const int x = tc.xt * kTileW + lane;
  const int y0 = (tc.yt * W + (int)threadIdx.y) * TY;
  if (y0 >= ny) return;  // uniform over the wavefront
  const bool xin = x < nx;
  // Update domains for stencil 2 (FD_PAD = 1).
  const bool e_col = x >= 1 && x < nx;       // Ey and Hx columns
  const bool hz_col = x >= 1 && x < nx - 1;  // Hz columns
  const bool pml_xe = x < g.pml_x0 || x >= g.pml_x1;
  const bool pml_xh = x < g.pml_x0 || x >= g.pml_x1h;
  const int s_begin = tc.grp * a.group;
  const int s_end = min(a.n_shots, s_begin + a.group);

  for (int s = s_begin; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.numel;
    const T *ey_i = a.ey_i + base;
    const T *hx_i = a.hx_i + base;
    const T *hz_i = a.hz_i + base;
    const T *mex_i = a.mex_i + base;
    const T *mez_i = a.mez_i + base;
    const T *ca = a.ca + s * g.ca_stride;
    const T *cb = a.cb + s * g.cb_stride;
    const T *cq = a.cq + s * g.cq_stride;

    T ey_c = ld(ey_i, y0, x, ny, nx);
    // Hx^{n+1/2}(y0-1, x): recomputed here, written by the tile above.
    T hx_up = (T)0;
    if (y0 >= 1) {
      const int y = y0 - 1;
      T v = ld(hx_i, y, x, ny, nx);
      if (e_col && y >= 1 && y < ny - 1) {
        const int64_t j = (int64_t)y * nx + x;
        T d = (ey_c - ey_i[j]) * g.rdy;
        if (y < g.pml_y0 || y >= g.pml_y1h) {
          const T m = g.byh[y] * mez_i[j] + g.ayh[y] * d;
          d = d / g.kyh[y] + m;
        }
        v = v - cq[j] * d;
      }
      hx_up = v;
    }

#pragma unroll
    for (int r = 0; r < TY; ++r) {
      const int y = y0 + r;
      if (y >= ny) break;  // uniform
      const int64_t j = (int64_t)y * nx + x;
      const int64_t i = base + j;
      const bool e_row = y >= 1;            // y < ny always holds here
      const bool hx_row = y >= 1 && y < ny - 1;
      const T ey_dn = ld(ey_i, y + 1, x, ny, nx);
      T ey_rt = __shfl_down(ey_c, 1, 64);
      if (lane == 63) ey_rt = ld(ey_i, y, x + 1, ny, nx);
      const T cqv = xin ? cq[j] : (T)0;

      // ---- H half-step for the owned cell.
      T hx_n = xin ? hx_i[j] : (T)0;
      if (e_col && hx_row) {
        T d = (ey_dn - ey_c) * g.rdy;
        if (y < g.pml_y0 || y >= g.pml_y1h) {
          const T m = g.byh[y] * mez_i[j] + g.ayh[y] * d;
          a.mez_o[i] = m;
          d = d / g.kyh[y] + m;
        }
        hx_n = hx_n - cqv * d;
      }
      T hz_n = xin ? hz_i[j] : (T)0;
      if (hz_col && e_row) {
        T d = (ey_rt - ey_c) * g.rdx;
        if (pml_xh) {
          const T m = g.bxh[x] * mex_i[j] + g.axh[x] * d;
          a.mex_o[i] = m;
          d = d / g.kxh[x] + m;
        }
        hz_n = hz_n + cqv * d;
      }
      // Hz^{n+1/2}(y, x-1): neighbour lane, or recomputed at lane 0.
      T hz_lf = __shfl_up(hz_n, 1, 64);
      if (lane == 0) {
        const int xl = x - 1;
        hz_lf = (T)0;
        if (xl >= 0) {
          const int64_t jl = j - 1;
          T v = hz_i[jl];
          if (xl >= 1 && xl < nx - 1 && e_row) {
            T d = (ey_c - ey_i[jl]) * g.rdx;
            if (xl < g.pml_x0 || xl >= g.pml_x1h) {
              const T m = g.bxh[xl] * mex_i[jl] + g.axh[xl] * d;
              d = d / g.kxh[xl] + m;
            }
            v = v + cq[jl] * d;
          }
          hz_lf = v;
        }
      }

      // ---- E step for the owned cell.
      T ey_n = ey_c;
      if (e_col && e_row) {
        T dhz = (hz_n - hz_lf) * g.rdx;
        T dhx = (hx_n - hx_up) * g.rdy;
        if (pml_xe) {
          const T m = g.bx[x] * a.mhzx[i] + g.ax[x] * dhz;
          a.mhzx[i] = m;
          dhz = dhz / g.kx[x] + m;
        }
        if (y < g.pml_y0 || y >= g.pml_y1) {
          const T m = g.by[y] * a.mhxz[i] + g.ay[y] * dhx;
          a.mhxz[i] = m;
          dhx = dhx / g.ky[y] + m;
        }
        const T curl = dhz - dhx;
        if constexpr (FMT != 0) {
          if (a.ey_store) put_store<T, FMT>(a.ey_store, i, ey_c);
          if (a.curl_store) put_store<T, FMT>(a.curl_store, i, curl);
        }
        ey_n = ca[j] * ey_c + cb[j] * curl;
      }

      // ---- sources, then receivers, of this shot in this row segment.
      const int slot = a.pts.row_slot[y];
      if (slot >= 0) {
        const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + tc.xt;
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

      if (xin) {
        a.hx_o[i] = hx_n;
        a.hz_o[i] = hz_n;
        a.ey_o[i] = ey_n;
      }
      hx_up = hx_n;
      ey_c = ey_dn;
    }
  }
}

// ----------------------------------------------------------------- adjoint
// One reverse iteration: transpose of the E step, transpose of the H
// half-step, receiver injection, source read-out, coefficient gradient.
template <typename T, int TY, int W, int FMT>
__global__ __launch_bounds__(64 * W) void adj_step(AdjArgs<T> a) {
  const Geom<T> &g = a.g;
  const int ny = g.ny, nx = g.nx;
  const int lane = threadIdx.x;
  const int n_groups = (a.n_shots + a.group - 1) / a.group;
  const TileCoord tc = decode_tile(a.pts.n_xt, a.n_yt, n_groups);
  const int x = tc.xt * kTileW + lane;
  const int y0 = (tc.yt * W + (int)threadIdx.y) * TY;
  if (y0 >= ny) return;
  const bool xin = x < nx;
  const bool e_col = x >= 1 && x < nx;
  const bool pml_xe = x < g.pml_x0 || x >= g.pml_x1;
  const bool pml_xh = x < g.pml_x0 || x >= g.pml_x1h;
  const int s_begin = tc.grp * a.group;
  const int s_end = min(a.n_shots, s_begin + a.group);
  const bool want_grad = (a.gca != nullptr || a.gcb != nullptr) && FMT != 0;
  const bool reg_sum = want_grad && a.g_stride == 0;

  double acc_ca[TY], acc_cb[TY];
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    acc_ca[r] = 0.0;
    acc_cb[r] = 0.0;
  }

  for (int s = s_begin; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.numel;
    const T *ley_i = a.ley_i + base;
    const T *lhx_i = a.lhx_i + base;
    const T *lhz_i = a.lhz_i + base;
    const T *mx_i = a.mx_i + base;
    const T *my_i = a.my_i + base;
    const T *nx_i = a.nx_i + base;
    const T *nz_i = a.nz_i + base;
    const T *ca = a.ca + s * g.ca_stride;
    const T *cb = a.cb + s * g.cb_stride;
    const T *cq = a.cq + s * g.cq_stride;

    // d/d(dHz/dx) adjoint at (y, xx) (reference FusedRawXAcc).
    auto raw_x = [&](int y, int xx) -> T {
      if (y < 1 || y >= ny || xx < 1 || xx >= nx) return (T)0;
      const int64_t jj = (int64_t)y * nx + xx;
      const T q = cb[jj] * ley_i[jj];
      if (xx >= g.pml_x0 && xx < g.pml_x1) return q;
      const T mn = g.bx[xx] * mx_i[jj] + q;
      return g.ax[xx] * mn + q / g.kx[xx];
    };
    // d/d(dHx/dz) adjoint at (y, x) (reference FusedRawYAcc).
    auto raw_y = [&](int y) -> T {
      if (y < 1 || y >= ny || !e_col) return (T)0;
      const int64_t jj = (int64_t)y * nx + x;
      const T q = cb[jj] * ley_i[jj];
      if (y >= g.pml_y0 && y < g.pml_y1) return -q;
      const T mn = g.by[y] * my_i[jj] - q;
      return g.ay[y] * mn - q / g.ky[y];
    };
    // d/d(dEy/dz at half rows) adjoint of the Hx update (FusedRawEyZAcc).
    auto raw_z = [&](int y, T lhx) -> T {
      if (y < 1 || y >= ny - 1 || !e_col) return (T)0;
      const int64_t jj = (int64_t)y * nx + x;
      const T adj = -cq[jj] * lhx;
      if (y >= g.pml_y0 && y < g.pml_y1h) return adj;
      const T mn = g.byh[y] * nz_i[jj] + adj;
      return g.ayh[y] * mn + adj / g.kyh[y];
    };
    // d/d(dEy/dx at half columns) adjoint of the Hz update (FusedRawEyXAcc).
    auto raw_xh = [&](int y, int xx, T lhz) -> T {
      if (y < 1 || y >= ny || xx < 1 || xx >= nx - 1) return (T)0;
      const int64_t jj = (int64_t)y * nx + xx;
      const T adj = cq[jj] * lhz;
      if (xx >= g.pml_x0 && xx < g.pml_x1h) return adj;
      const T mn = g.bxh[xx] * nx_i[jj] + adj;
      return g.axh[xx] * mn + adj / g.kxh[xx];
    };

    T ley_c = ld(ley_i, y0, x, ny, nx);
    T rawy_c = raw_y(y0);
    T rawz_up = (T)0;
    if (y0 >= 1) {
      const int y = y0 - 1;
      T lhx = ld(lhx_i, y, x, ny, nx);
      if (e_col && y >= 1) lhx = lhx + (raw_y(y) - rawy_c) * g.rdy;
      rawz_up = raw_z(y, lhx);
    }

#pragma unroll
    for (int r = 0; r < TY; ++r) {
      const int y = y0 + r;
      if (y >= ny) break;
      const int64_t j = (int64_t)y * nx + x;
      const int64_t i = base + j;
      const bool in_e = e_col && y >= 1;

      // ---- transpose of the E step.
      const T rawx_c = raw_x(y, x);
      T rawx_rt = __shfl_down(rawx_c, 1, 64);
      if (lane == 63) rawx_rt = raw_x(y, x + 1);
      T lhz_n = xin ? lhz_i[j] : (T)0;
      if (in_e) lhz_n = lhz_n + (rawx_c - rawx_rt) * g.rdx;
      const T ley_dn = ld(ley_i, y + 1, x, ny, nx);
      const T rawy_dn = raw_y(y + 1);
      T lhx_n = xin ? lhx_i[j] : (T)0;
      if (in_e) lhx_n = lhx_n + (rawy_c - rawy_dn) * g.rdy;
      if (in_e) {
        const T q = cb[j] * ley_c;
        if (pml_xe) a.mx_o[i] = g.bx[x] * mx_i[j] + q;
        if (y < g.pml_y0 || y >= g.pml_y1) a.my_o[i] = g.by[y] * my_i[j] - q;
      }

      // ---- transpose of the H half-step.
      const T rawz_c = raw_z(y, lhx_n);
      const T rawxh_c = raw_xh(y, x, lhz_n);
      T rawxh_lf = __shfl_up(rawxh_c, 1, 64);
      if (lane == 0) {
        const int xl = x - 1;
        rawxh_lf = (T)0;
        if (xl >= 0) {
          T lhz_l = lhz_i[j - 1];
          if (xl >= 1 && y >= 1) lhz_l = lhz_l + (raw_x(y, xl) - rawx_c) * g.rdx;
          rawxh_lf = raw_xh(y, xl, lhz_l);
        }
      }
      T ley_n = ley_c;
      if (in_e) {
        const T cqv = cq[j];
        const T adj_y = -cqv * lhx_n;
        const T adj_x = cqv * lhz_n;
        if (y < ny - 1 && (y < g.pml_y0 || y >= g.pml_y1h))
          a.nz_o[i] = g.byh[y] * nz_i[j] + adj_y;
        if (x < nx - 1 && pml_xh) a.nx_o[i] = g.bxh[x] * nx_i[j] + adj_x;
        const T t1 = ca[j] * ley_c;
        const T t2 = (rawz_up - rawz_c) * g.rdy + (rawxh_lf - rawxh_c) * g.rdx;
        ley_n = t1 + t2;
      }

      // ---- receivers inject grad_r[t]; sources read grad_f[t].
      const int slot = a.pts.row_slot[y];
      if (slot >= 0) {
        const int64_t si = ((int64_t)slot * a.n_shots + s) * a.pts.n_xt + tc.xt;
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
      if (want_grad && in_e) {
        const double lam = (double)ley_n;
        double pa = 0.0, pb = 0.0;
        if (a.gca) pa = lam * get_store<T, FMT>(a.ey_store, i) * a.step_scale;
        if (a.gcb) pb = lam * get_store<T, FMT>(a.curl_store, i) * a.step_scale;
        if (reg_sum) {
          acc_ca[r] += pa;
          acc_cb[r] += pb;
        } else {
          if (a.gca) a.gca[s * a.g_stride + j] += pa;
          if (a.gcb) a.gcb[s * a.g_stride + j] += pb;
        }
      }

      if (xin) {
        a.ley_o[i] = ley_n;
        a.lhx_o[i] = lhx_n;
        a.lhz_o[i] = lhz_n;
      }
      ley_c = ley_dn;
      rawy_c = rawy_dn;
      rawz_up = rawz_c;
    }
  }

  if (reg_sum && e_col) {
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      const int y = y0 + r;
      if (y < 1 || y >= ny) continue;
      const int64_t j = (int64_t)y * nx + x;
      if (a.use_atomics) {
        if (a.gca) atomicAdd(&a.gca[j], acc_ca[r]);
        if (a.gcb) atomicAdd(&a.gcb[j], acc_cb[r]);
      } else {
        if (a.gca) a.gca[j] += acc_ca[r];
        if (a.gcb) a.gcb[j] += acc_cb[r];
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
  int32_t n_shots;
  int64_t n_src, n_rec;
};

constexpr int kW = 4;
constexpr int kTY = 8;   // rows per wavefront, forward
constexpr int kTYA = 4;  // rows per wavefront, adjoint (all shots per tile)

template <typename T>
Geom<T> make_geom(const T *const prof[12], T rdy, T rdx, int64_t ny, int64_t nx,
                  int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
                  bool ca_b, bool cb_b, bool cq_b) {
  Geom<T> g;
  g.ay = prof[0]; g.by = prof[1]; g.ayh = prof[2]; g.byh = prof[3];
  g.ax = prof[4]; g.bx = prof[5]; g.axh = prof[6]; g.bxh = prof[7];
  g.ky = prof[8]; g.kyh = prof[9]; g.kx = prof[10]; g.kxh = prof[11];
  g.rdy = rdy;
  g.rdx = rdx;
  g.ny = (int32_t)ny;
  g.nx = (int32_t)nx;
  g.numel = ny * nx;
  g.pml_y0 = (int32_t)pml_y0;
  g.pml_y1 = (int32_t)pml_y1;
  g.pml_x0 = (int32_t)pml_x0;
  g.pml_x1 = (int32_t)pml_x1;
  g.pml_y1h = (int32_t)(pml_y1 > pml_y0 ? pml_y1 - 1 : pml_y0);
  g.pml_x1h = (int32_t)(pml_x1 > pml_x0 ? pml_x1 - 1 : pml_x0);
  g.ca_stride = ca_b ? g.numel : 0;
  g.cb_stride = cb_b ? g.numel : 0;
  g.cq_stride = cq_b ? g.numel : 0;
  return g;
}

inline int check_geometry(int64_t n_shots, int64_t ny, int64_t nx) {
  if (n_shots <= 0 || ny < 3 || nx < 3) {
    set_error("n_shots must be positive and the padded grid at least 3x3");
    return kInvalidArgument;
  }
  if (ny > (1 << 30) || nx > (1 << 30) || n_shots > (1 << 30)) {
    set_error("grid dimension too large");
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

#define TIDE_CHECK(expr)            \
  do {                              \
    const int _rc = (expr);         \
    if (_rc != ::tide_hip::kOk) return _rc; \
  } while (0)

template <typename T>
int forward_impl(const Problem<T> &p, T *ey, T *hx, T *hz, T *m_ey_x,
                 T *m_ey_z, T *m_hx_z, T *m_hz_x, const T *f, T *r,
                 const int64_t *src, const int64_t *rec, int64_t nt,
                 int64_t start_t, int64_t step_ratio, void *ey_store,
                 void *curl_store, int fmt, size_t store_step_bytes,
                 hipStream_t stream) {
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
  std::vector<int64_t> hsrc, hrec;
  TIDE_CHECK(fetch_indices(src, n_shots * p.n_src, hsrc, stream));
  TIDE_CHECK(fetch_indices(rec, n_shots * p.n_rec, hrec, stream));
  PointTable table;
  if (!build_point_table(hsrc.data(), p.n_src, hrec.data(), p.n_rec, n_shots,
                         p.g.ny, p.g.nx, kTileW, table))
    return kInvalidArgument;
  DevicePointTable dpt;
  TIDE_CHECK(dpt.upload(table, stream));

  Scratch scr(stream);
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
  a.group = 1;
  a.n_yt = (p.g.ny + kW * kTY - 1) / (kW * kTY);
  const int64_t n_groups = (p.n_shots + a.group - 1) / a.group;
  const dim3 block(64, kW);
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
      fwd_step<T, kTY, kW, 0><<<grid, block, 0, stream>>>(a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) fwd_step<T, kTY, kW, 2><<<grid, block, 0, stream>>>(a);
    } else {
      fwd_step<T, kTY, kW, 1><<<grid, block, 0, stream>>>(a);
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
int backward_impl(const Problem<T> &p, const T *grad_r, T *l_ey, T *l_hx,
                  T *l_hz, T *ml_ey_x, T *ml_ey_z, T *ml_hx_z, T *ml_hz_x,
                  const void *ey_store, const void *curl_store, int fmt,
                  size_t store_step_bytes, T *grad_f, double *gca,
                  double *gcb, const int64_t *src, const int64_t *rec,
                  int64_t nt, int64_t start_t, int64_t step_ratio,
                  hipStream_t stream) {
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
  std::vector<int64_t> hsrc, hrec;
  TIDE_CHECK(fetch_indices(src, n_shots * p.n_src, hsrc, stream));
  TIDE_CHECK(fetch_indices(rec, n_shots * p.n_rec, hrec, stream));
  PointTable table;
  if (!build_point_table(hsrc.data(), p.n_src, hrec.data(), p.n_rec, n_shots,
                         p.g.ny, p.g.nx, kTileW, table))
    return kInvalidArgument;
  DevicePointTable dpt;
  TIDE_CHECK(dpt.upload(table, stream));

  Scratch scr(stream);
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
  a.gca = gca;
  a.gcb = gcb;
  const bool per_shot = (gca && p.g.ca_stride != 0) || (gcb && p.g.cb_stride != 0);
  a.g_stride = per_shot ? p.g.numel : 0;
  a.step_scale = (double)step_ratio;
  a.n_shots = p.n_shots;
  a.group = p.n_shots;  // one workgroup sums every shot of its tile
  a.use_atomics = 0;
  a.n_yt = (p.g.ny + kW * kTYA - 1) / (kW * kTYA);
  const int64_t n_groups = (p.n_shots + a.group - 1) / a.group;
  const dim3 block(64, kW);
  const dim3 grid((unsigned)((int64_t)dpt.n_xt * a.n_yt * n_groups));

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
      adj_step<T, kTYA, kW, 0><<<grid, block, 0, stream>>>(a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) adj_step<T, kTYA, kW, 2><<<grid, block, 0, stream>>>(a);
    } else {
      adj_step<T, kTYA, kW, 1><<<grid, block, 0, stream>>>(a);
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
Problem<T> make_problem(const T *ca, const T *cb, const T *cq,
                        const T *const prof[12], T rdy, T rdx, int64_t n_shots,
                        int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,
                        bool ca_b, bool cb_b, bool cq_b, int64_t pml_y0,
                        int64_t pml_x0, int64_t pml_y1, int64_t pml_x1) {
  Problem<T> p;
  p.g = make_geom<T>(prof, rdy, rdx, ny, nx, pml_y0, pml_x0, pml_y1, pml_x1,
                     ca_b, cb_b, cq_b);
  p.ca = ca;
  p.cb = cb;
  p.cq = cq;
  p.n_shots = (int32_t)n_shots;
  p.n_src = n_src;
  p.n_rec = n_rec;
  return p;
}

template <typename T>
int forward_entry(T const *ca, T const *cb, T const *cq, T const *f, T *ey,
                  T *hx, T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x,
                  T *r, T const *ay, T const *by, T const *ayh, T const *byh,
                  T const *ax, T const *bx, T const *axh, T const *bxh,
                  T const *ky, T const *kyh, T const *kx, T const *kxh,
                  int64_t const *sources_i, int64_t const *receivers_i, T rdy,
                  T rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
                  int64_t n_src, int64_t n_rec, int64_t step_ratio,
                  bool ca_b, bool cb_b, bool cq_b, int64_t start_t,
                  int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
                  int64_t pml_x1, int64_t device, void *stream_handle,
                  void *ey_store, void *curl_store, int fmt,
                  size_t store_step_bytes) {
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};
  Problem<T> p = make_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx,
                                 n_src, n_rec, ca_b, cb_b, cq_b, pml_y0,
                                 pml_x0, pml_y1, pml_x1);
  return forward_impl<T>(p, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, f, r,
                         sources_i, receivers_i, nt, start_t, step_ratio,
                         ey_store, curl_store, fmt, store_step_bytes,
                         (hipStream_t)stream_handle);
}

}  // namespace

#define TIDE_TM2D_ENTRY_POINTS(T, NAME)                                         \
  extern "C" int maxwell_tm_2_##NAME##_forward_cuda(                           \
      T const *ca, T const *cb, T const *cq, T const *f, T *ey, T *hx, T *hz,  \
      T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x, T const *debye_a,            \
      T const *debye_b, T const *debye_cp, T *polarization, T *ey_prev, T *r,  \
      int64_t n_poles, T const *ay, T const *by, T const *ayh, T const *byh,   \
      T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky,       \
      T const *kyh, T const *kx, T const *kxh, int64_t const *sources_i,       \
      int64_t const *receivers_i, T rdy, T rdx, T dt, int64_t nt,              \
      int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,   \
      int64_t step_ratio, bool has_dispersion, bool ca_b, bool cb_b,           \
      bool cq_b, int64_t start_t, int64_t pml_y0, int64_t pml_x0,              \
      int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,       \
      void *compute_stream) {                                                  \
    (void)debye_a; (void)debye_b; (void)debye_cp; (void)polarization;          \
    (void)ey_prev; (void)n_poles; (void)dt; (void)n_threads;                   \
    if (has_dispersion) {                                                      \
      set_error("Debye dispersion is not part of this build");                 \
      return kUnsupported;                                                     \
    }                                                                          \
    return forward_entry<T>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, \
                            m_hz_x, r, ay, by, ayh, byh, ax, bx, axh, bxh, ky, \
                            kyh, kx, kxh, sources_i, receivers_i, rdy, rdx,    \
                            nt, n_shots, ny, nx, n_src, n_rec, step_ratio,     \
                            ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, pml_y1, \
                            pml_x1, device, compute_stream, nullptr, nullptr,  \
                            0, 0);                                             \
  }                                                                            \
                                                                               \
  extern "C" int maxwell_tm_2_##NAME##_forward_with_storage_cuda(              \
      T const *ca, T const *cb, T const *cq, T const *f, T *ey, T *hx, T *hz,  \
      T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x, T *r, void *ey_store_1,      \
      void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,   \
      void *curl_store_3, char const *const *curl_filenames, T const *ay,      \
      T const *by, T const *ayh, T const *byh, T const *ax, T const *bx,       \
      T const *axh, T const *bxh, T const *ky, T const *kyh, T const *kx,      \
      T const *kxh, int64_t const *sources_i, int64_t const *receivers_i,      \
      T rdy, T rdx, T dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, \
      int64_t n_src, int64_t n_rec, int64_t step_ratio, int64_t storage_mode,  \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_rg,           \
      bool cb_rg, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,            \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,          \
      int64_t n_threads, int64_t device, void *compute_stream,                 \
      void *storage_stream) {                                                  \
    (void)ey_store_3; (void)ey_filenames; (void)curl_store_3;                  \
    (void)curl_filenames; (void)dt; (void)n_threads; (void)storage_stream;     \
    if (storage_mode != kStorageDevice && storage_mode != kStorageNone) {      \
      set_error("host- and disk-staged snapshot storage is not built; use "    \
                "storage_mode 'device' or 'none'");                            \
      return kUnsupported;                                                     \
    }                                                                          \
    const bool st = storage_mode == kStorageDevice;                            \
    const int fmt = storage_format == kFormatBf16 ? 2 : 1;                     \
    return forward_entry<T>(                                                   \
        ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, ay, by,  \
        ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh, sources_i, receivers_i,  \
        rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec, step_ratio, ca_b, cb_b,   \
        cq_b, start_t, pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream, \
        st && ca_rg ? ey_store_1 : nullptr,                                    \
        st && cb_rg ? curl_store_1 : nullptr, fmt,                             \
        (size_t)shot_bytes_uncomp * (size_t)n_shots);                          \
  }                                                                            \
                                                                               \
  extern "C" int maxwell_tm_2_##NAME##_backward_cuda(                          \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey,    \
      T *lambda_hx, T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z,          \
      T *m_lambda_hx_z, T *m_lambda_hz_x, T *adjoint_memory_scratch,           \
      void *ey_store_1, void *ey_store_3, char const *const *ey_filenames,     \
      void *curl_store_1, void *curl_store_3,                                  \
      char const *const *curl_filenames, T *grad_f, T *grad_ca, T *grad_cb,    \
      T *grad_ca_shot, T *grad_cb_shot, T const *ay, T const *by,              \
      T const *ayh, T const *byh, T const *ax, T const *bx, T const *axh,      \
      T const *bxh, T const *ky, T const *kyh, T const *kx, T const *kxh,      \
      int64_t const *sources_i, int64_t const *receivers_i, T rdy, T rdx,      \
      T dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,               \
      int64_t n_src, int64_t n_rec, int64_t step_ratio, int64_t storage_mode,  \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_rg,           \
      bool cb_rg, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,            \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,          \
      int64_t n_threads, int64_t device, void *compute_stream,                 \
      void *storage_stream) {                                                  \
    (void)adjoint_memory_scratch; (void)ey_store_3; (void)ey_filenames;        \
    (void)curl_store_3; (void)curl_filenames; (void)grad_ca_shot;              \
    (void)grad_cb_shot; (void)dt; (void)n_threads; (void)storage_stream;       \
    return tide_tm2d_backward_impl_entry<T>(                                   \
        ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,    \
        m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, ey_store_1, curl_store_1, \
        grad_f, grad_ca, grad_cb, nullptr, nullptr, ay, by, ayh, byh, ax, bx,  \
        axh, bxh, ky, kyh, kx, kxh, sources_i, receivers_i, rdy, rdx, nt,      \
        n_shots, ny, nx, n_src, n_rec, step_ratio, storage_mode,               \
        storage_format, shot_bytes_uncomp, ca_rg, cb_rg, ca_b, cb_b, cq_b,     \
        start_t, pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream);      \
  }

namespace {

// Shared backward body.  When gca64/gcb64 are given (FP64 accumulators, +=)
// they receive the gradient and grad_ca/grad_cb are left untouched;
// otherwise FP64 accumulators are allocated here and folded into
// grad_ca/grad_cb (+=) at the end.
template <typename T>
int tide_tm2d_backward_impl_entry(
    T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey,
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
  Problem<T> p = make_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx,
                                 n_src, n_rec, ca_b, cb_b, cq_b, pml_y0,
                                 pml_x0, pml_y1, pml_x1);
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
  TIDE_CHECK(backward_impl<T>(
      p, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x, m_lambda_ey_z,
      m_lambda_hx_z, m_lambda_hz_x, ey_store_1, curl_store_1, fmt,
      (size_t)shot_bytes_uncomp * (size_t)n_shots, grad_f, ca_rg ? ga : nullptr,
      cb_rg ? gb : nullptr, sources_i, receivers_i, nt, start_t, step_ratio,
      stream));
  if (own) {
    if (ca_rg) TIDE_CHECK(fold_into<T>(grad_ca, ga, na, stream));
    if (cb_rg) TIDE_CHECK(fold_into<T>(grad_cb, gb, nb, stream));
  }
  return kOk;
}

}  // namespace

TIDE_TM2D_ENTRY_POINTS(float, float)
TIDE_TM2D_ENTRY_POINTS(double, double)

// FP64-accumulator backward used by the Python operator (and any caller that
// runs shot chunks and wants one FP64 sum across them).  Same arguments as
// maxwell_tm_2_<dtype>_backward_cuda except that grad_ca / grad_cb are FP64
// accumulators that are added to, and the unused per-shot workspaces and
// host-staging arguments are dropped.
#define TIDE_TM2D_BACKWARD_F64(T, NAME)                                         \
  extern "C" int tide_tm2d_##NAME##_backward_f64acc(                            \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey,     \
      T *lambda_hx, T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z,           \
      T *m_lambda_hx_z, T *m_lambda_hz_x, void const *ey_store,                 \
      void const *curl_store, T *grad_f, double *grad_ca, double *grad_cb,      \
      T const *ay, T const *by, T const *ayh, T const *byh, T const *ax,        \
      T const *bx, T const *axh, T const *bxh, T const *ky, T const *kyh,       \
      T const *kx, T const *kxh, int64_t const *sources_i,                      \
      int64_t const *receivers_i, T rdy, T rdx, int64_t nt, int64_t n_shots,    \
      int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,                     \
      int64_t step_ratio, int64_t storage_format, int64_t shot_bytes_uncomp,    \
      bool ca_b, bool cb_b, bool cq_b, int64_t start_t, int64_t pml_y0,         \
      int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t device,           \
      void *compute_stream) {                                                   \
    if (grad_ca == nullptr && grad_cb == nullptr && ey_store == nullptr &&     \
        curl_store == nullptr) {                                               \
      return tide_tm2d_backward_impl_entry<T>(                                 \
          ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,  \
          m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, nullptr, nullptr,       \
          grad_f, nullptr, nullptr, nullptr, nullptr, ay, by, ayh, byh, ax,    \
          bx, axh, bxh, ky, kyh, kx, kxh, sources_i, receivers_i, rdy, rdx,    \
          nt, n_shots, ny, nx, n_src, n_rec, step_ratio, kStorageNone,         \
          storage_format, shot_bytes_uncomp, false, false, ca_b, cb_b, cq_b,   \
          start_t, pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream);    \
    }                                                                          \
    return tide_tm2d_backward_impl_entry<T>(                                   \
        ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,    \
        m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, (void *)ey_store,         \
        (void *)curl_store, grad_f, nullptr, nullptr, grad_ca, grad_cb, ay,    \
        by, ayh,                                                               \
        byh, ax, bx, axh, bxh, ky, kyh, kx, kxh, sources_i, receivers_i, rdy,  \
        rdx, nt, n_shots, ny, nx, n_src, n_rec, step_ratio, kStorageDevice,    \
        storage_format, shot_bytes_uncomp, grad_ca != nullptr,                 \
        grad_cb != nullptr, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, pml_y1, \
        pml_x1, device, compute_stream);                                       \
  }

TIDE_TM2D_BACKWARD_F64(float, float)
TIDE_TM2D_BACKWARD_F64(double, double)
