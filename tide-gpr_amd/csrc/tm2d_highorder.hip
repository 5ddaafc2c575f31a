// Stencil orders 4, 6 and 8 of the 2-D TM propagator and its exact adjoint
// (reference entry points maxwell_tm_{4,6,8}_{float,double}_*_cuda,
// src/tide/csrc/tm2d/maxwell_tm_cuda_instantiations.inc; staggered
// differences src/tide/csrc/tm2d/maxwell_tm_core.cuh:61-365; update order
// maxwell_tm_core.cuh:448-599; exact adjoint and reverse loop
// maxwell_tm_cuda_inst.cu:922-1254, 2540-2731).
//
// These orders are the uncommon ones (the reference default and every
// benchmark workload is order 2, which has the fused, temporally blocked
// kernels of tm2d_fused.h), so they run one cell per lane: a workgroup is a
// 64 x 4 tile of one shot (64 consecutive cells of a row per wavefront,
// coalesced), and every time step is a short sequence of launches:
//   forward  H half-step, E step (+ snapshots), source injection, receivers;
//   adjoint  transpose of the E step, transpose of the H half-step,
//            receiver injection + source-gradient read, FP64 gradient.
// The H and E updates write only the cell they own and read only the field
// family the launch does not write, so they run in place; the adjoint's
// CPML memories are read at neighbouring cells while their owner writes the
// new value, so they ping-pong between the caller's arrays and a scratch
// copy.  Point operations run one shot per thread in point order, so
// duplicate points add in a fixed order.  Every expression keeps the
// reference's operation order (no FMA contraction), so results equal the
// reference CPU arithmetic.

#include "tide_common.h"
#include "tide_hip.h"
#include "tide_storage.h"
#include "tm2d_highorder.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace tide_hip {
namespace tm2d_ho {

// Coefficients of the staggered first difference of half-width P.  P = 1
// (order 2) multiplies the one difference by 1.0, which is exact, so these
// kernels also reproduce the order-2 arithmetic (the Born driver uses them
// for every order).
template <int P> struct Coef;
template <> struct Coef<1> {
  static constexpr double c[1] = {1.0};
};
template <> struct Coef<2> {
  static constexpr double c[2] = {9.0 / 8.0, -1.0 / 24.0};
};
template <> struct Coef<3> {
  static constexpr double c[3] = {75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0};
};
template <> struct Coef<4> {
  static constexpr double c[4] = {1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};
};

// The four difference forms, summed term by term from the nearest pair out
// and scaled by r last.  `at(o)` reads the operand at offset o along the axis.
template <typename T, int P, typename A>
__device__ __forceinline__ T d_int(A at, T r) {  // F(k) - F(k-1) form
  T s = (T)Coef<P>::c[0] * (at(0) - at(-1));
#pragma unroll
  for (int k = 1; k < P; ++k) s = s + (T)Coef<P>::c[k] * (at(k) - at(-1 - k));
  return s * r;
}
template <typename T, int P, typename A>
__device__ __forceinline__ T d_half(A at, T r) {  // F(k+1) - F(k) form
  T s = (T)Coef<P>::c[0] * (at(1) - at(0));
#pragma unroll
  for (int k = 1; k < P; ++k) s = s + (T)Coef<P>::c[k] * (at(1 + k) - at(-k));
  return s * r;
}
template <typename T, int P, typename A>
__device__ __forceinline__ T d_int_adj(A c, T r) {  // transpose of d_int
  T s = (T)Coef<P>::c[0] * (c(0) - c(1));
#pragma unroll
  for (int k = 1; k < P; ++k) s = s + (T)Coef<P>::c[k] * (c(-k) - c(1 + k));
  return s * r;
}
template <typename T, int P, typename A>
__device__ __forceinline__ T d_half_adj(A c, T r) {  // transpose of d_half
  T s = (T)Coef<P>::c[0] * (c(-1) - c(0));
#pragma unroll
  for (int k = 1; k < P; ++k) s = s + (T)Coef<P>::c[k] * (c(-1 - k) - c(k));
  return s * r;
}

template <typename T> struct Grid {
  const T *ca, *cb, *cq;
  // ay by ayh byh ax bx axh bxh ky kyh kx kxh
  const T *ay, *by, *ayh, *byh, *ax, *bx, *axh, *bxh, *ky, *kyh, *kx, *kxh;
  T rdy, rdx;
  int32_t ny, nx, numel, n_shots;
  int32_t pml_y0, pml_y1, pml_x0, pml_x1, y1h, x1h;
  int32_t ca_b, cb_b, cq_b;
};

// Thread -> (shot, y, x): 64 x 4 tiles, z = shot.
struct Cell {
  int s, y, x;
  uint32_t j;  // y * nx + x
};
template <typename T>
__device__ __forceinline__ Cell cell_of(const Grid<T> &g) {
  Cell c;
  c.x = (int)(blockIdx.x * 64 + threadIdx.x);
  c.y = (int)(blockIdx.y * 4 + threadIdx.y);
  c.s = (int)blockIdx.z;
  c.j = (uint32_t)(c.y * g.nx + c.x);
  return c;
}

template <typename T, int P>
__global__ __launch_bounds__(256) void h_step(Grid<T> g, const T *__restrict__ ey,
                                              T *__restrict__ hx, T *__restrict__ hz,
                                              T *__restrict__ m_ey_x, T *__restrict__ m_ey_z) {
  const Cell c = cell_of(g);
  const int ny = g.ny, nx = g.nx;
  if (c.y < P || c.y >= ny - P + 1 || c.x < P || c.x >= nx - P + 1) return;
  const size_t base = (size_t)c.s * g.numel;
  const T *e = ey + base + c.j;
  const T cq = g.cq[(g.cq_b ? base : 0) + c.j];
  const size_t i = base + c.j;
  if (c.y < ny - P) {
    T d = d_half<T, P>([&](int o) { return e[(ptrdiff_t)o * nx]; }, g.rdy);
    if (c.y < g.pml_y0 || c.y >= g.y1h) {
      const T m = g.byh[c.y] * m_ey_z[i] + g.ayh[c.y] * d;
      m_ey_z[i] = m;
      d = d / g.kyh[c.y] + m;
    }
    hx[i] = hx[i] - cq * d;
  }
  if (c.x < nx - P) {
    T d = d_half<T, P>([&](int o) { return e[o]; }, g.rdx);
    if (c.x < g.pml_x0 || c.x >= g.x1h) {
      const T m = g.bxh[c.x] * m_ey_x[i] + g.axh[c.x] * d;
      m_ey_x[i] = m;
      d = d / g.kxh[c.x] + m;
    }
    hz[i] = hz[i] + cq * d;
  }
}

template <typename T>
__device__ __forceinline__ void put_snap(void *plane, size_t i, T v, int fmt) {
  if (fmt == 2)
    ((uint16_t *)plane)[i] = f32_to_bf16((float)v);
  else
    ((T *)plane)[i] = v;
}

template <typename T>
__device__ __forceinline__ double get_snap(const void *plane, size_t i, int fmt) {
  if (fmt == 2) return (double)bf16_to_f32(((const uint16_t *)plane)[i]);
  return (double)((const T *)plane)[i];
}

// Born scattering source of one step: the tangent E update adds
// dca * Ey^n + dcb * curl H^n of the background (its pre-update Ey and curl
// at this step, [n_shots, ny, nx]); dca / dcb per shot when batched.
template <typename T> struct BornSrc {
  const T *dca, *dcb;  // null: no source (plain forward)
  const T *bg_ey, *bg_curl;
};

template <typename T, int P>
__global__ __launch_bounds__(256) void e_step(Grid<T> g, T *__restrict__ ey,
                                              const T *__restrict__ hx, const T *__restrict__ hz,
                                              T *__restrict__ m_hx_z, T *__restrict__ m_hz_x,
                                              void *ey_store, void *curl_store, int fmt,
                                              BornSrc<T> born = BornSrc<T>{nullptr, nullptr,
                                                                           nullptr, nullptr}) {
  const Cell c = cell_of(g);
  const int ny = g.ny, nx = g.nx;
  if (c.y < P || c.y >= ny - P + 1 || c.x < P || c.x >= nx - P + 1) return;
  const size_t base = (size_t)c.s * g.numel, i = base + c.j;
  const T ca = g.ca[(g.ca_b ? base : 0) + c.j];
  const T cb = g.cb[(g.cb_b ? base : 0) + c.j];
  T dhz = d_int<T, P>([&](int o) { return hz[i + o]; }, g.rdx);
  T dhx = d_int<T, P>([&](int o) { return hx[i + (ptrdiff_t)o * nx]; }, g.rdy);
  if (c.x < g.pml_x0 || c.x >= g.pml_x1) {
    const T m = g.bx[c.x] * m_hz_x[i] + g.ax[c.x] * dhz;
    m_hz_x[i] = m;
    dhz = dhz / g.kx[c.x] + m;
  }
  if (c.y < g.pml_y0 || c.y >= g.pml_y1) {
    const T m = g.by[c.y] * m_hx_z[i] + g.ay[c.y] * dhx;
    m_hx_z[i] = m;
    dhx = dhx / g.ky[c.y] + m;
  }
  const T curl = dhz - dhx;
  const T e0 = ey[i];
  if (ey_store != nullptr) put_snap<T>(ey_store, i, e0, fmt);
  if (curl_store != nullptr) put_snap<T>(curl_store, i, curl, fmt);
  T e1 = ca * e0 + cb * curl;
  if (born.dca != nullptr) {
    const T da = born.dca[(g.ca_b ? base : 0) + c.j];
    const T db = born.dcb[(g.cb_b ? base : 0) + c.j];
    e1 = e1 + (da * born.bg_ey[i] + db * born.bg_curl[i]);
  }
  ey[i] = e1;
}

// Debye E step: Ey^n into ey_prev, the plain update with the Debye
// coefficients, then + sum_p cp_p P_p (poles in order from zero); cp
// [n_poles, ny, nx], pol [n_shots, n_poles, ny, nx] (reference
// src/tide/csrc/tm2d/maxwell_tm_inst.cpp:214-272).
template <typename T, int P>
__global__ __launch_bounds__(256) void e_step_debye(Grid<T> g, T *__restrict__ ey,
                                                    const T *__restrict__ hx,
                                                    const T *__restrict__ hz,
                                                    T *__restrict__ m_hx_z, T *__restrict__ m_hz_x,
                                                    T *__restrict__ ey_prev,
                                                    const T *__restrict__ cp,
                                                    const T *__restrict__ pol, int n_poles) {
  const Cell c = cell_of(g);
  const int ny = g.ny, nx = g.nx;
  if (c.y < P || c.y >= ny - P + 1 || c.x < P || c.x >= nx - P + 1) return;
  const size_t base = (size_t)c.s * g.numel, i = base + c.j;
  const T ca = g.ca[c.j];
  const T cb = g.cb[c.j];
  T dhz = d_int<T, P>([&](int o) { return hz[i + o]; }, g.rdx);
  T dhx = d_int<T, P>([&](int o) { return hx[i + (ptrdiff_t)o * nx]; }, g.rdy);
  if (c.x < g.pml_x0 || c.x >= g.pml_x1) {
    const T m = g.bx[c.x] * m_hz_x[i] + g.ax[c.x] * dhz;
    m_hz_x[i] = m;
    dhz = dhz / g.kx[c.x] + m;
  }
  if (c.y < g.pml_y0 || c.y >= g.pml_y1) {
    const T m = g.by[c.y] * m_hx_z[i] + g.ay[c.y] * dhx;
    m_hx_z[i] = m;
    dhx = dhx / g.ky[c.y] + m;
  }
  const T curl = dhz - dhx;
  const T e0 = ey[i];
  ey_prev[i] = e0;
  T e1 = ca * e0 + cb * curl;
  T term = (T)0;
  for (int p = 0; p < n_poles; ++p)
    term = term + cp[(size_t)p * g.numel + c.j] * pol[((size_t)c.s * n_poles + p) * g.numel + c.j];
  ey[i] = e1 + term;
}

// Debye polarization of every cell after the sources:
// P_p = a_p P_p + b_p (Ey^{n+1} + Ey^n) (maxwell_tm_inst.cpp:275-300).
template <typename T>
__global__ void pol_update(const T *__restrict__ ey, const T *__restrict__ ey_prev,
                           const T *__restrict__ da, const T *__restrict__ db,
                           T *__restrict__ pol, int64_t numel, int n_shots, int n_poles) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = (int)blockIdx.y;
  if (j >= numel || s >= n_shots) return;
  const int64_t i = (int64_t)s * numel + j;
  const T sum = ey[i] + ey_prev[i];
  for (int p = 0; p < n_poles; ++p) {
    const int64_t q = ((int64_t)s * n_poles + p) * numel + j;
    pol[q] = da[(int64_t)p * numel + j] * pol[q] + db[(int64_t)p * numel + j] * sum;
  }
}

// One thread per shot, points in order (duplicates add in a fixed order).
template <typename T>
__global__ void add_points(T *field, int numel, int n_shots, const int64_t *idx, int n,
                           const T *vals) {
  const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= n_shots || vals == nullptr) return;
  for (int k = 0; k < n; ++k) {
    const int64_t q = idx[(size_t)s * n + k];
    if (q >= 0) field[(size_t)s * numel + q] += vals[(size_t)s * n + k];
  }
}

template <typename T>
__global__ void read_points(const T *field, int numel, int n_shots, const int64_t *idx, int n,
                            T *out) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n_shots * n || out == nullptr) return;
  const int64_t q = idx[k];
  if (q >= 0) out[k] = field[(size_t)(k / n) * numel + q];
}

// Transpose of the E step: lambda_Hz, lambda_Hx gather the x / y
// derivative adjoints of the neighbours' cb * lambda_Ey (through their CPML
// memories), and the E-phase adjoint memories advance into the `_n` copies.
// With `dcb` and `l2` (the Born adjoint's background pass) the adjoint of
// curl H is cb lambda_Ey + dcb lambda2: the scattered E update adds dcb curl H
// of the background, so the scattered adjoint lambda2 reaches it too.
template <typename T, int P>
__global__ __launch_bounds__(256) void e_adjoint(Grid<T> g, const T *__restrict__ l_ey,
                                                 T *__restrict__ l_hx, T *__restrict__ l_hz,
                                                 const T *__restrict__ m_hx_z,
                                                 const T *__restrict__ m_hz_x,
                                                 T *__restrict__ m_hx_z_n, T *__restrict__ m_hz_x_n,
                                                 const T *__restrict__ dcb,
                                                 const T *__restrict__ l2) {
  const Cell c = cell_of(g);
  const int ny = g.ny, nx = g.nx;
  if (c.y < P || c.y >= ny - P + 1 || c.x < P || c.x >= nx - P + 1) return;
  const size_t base = (size_t)c.s * g.numel;
  const T *cbp = g.cb + (g.cb_b ? base : 0);
  auto in_e = [&](int y, int x) { return y >= P && y < ny - P + 1 && x >= P && x < nx - P + 1; };
  auto q_at = [&](int y, int x) {
    const size_t j = (size_t)y * nx + x;
    const T q = cbp[j] * l_ey[base + j];
    return l2 != nullptr ? q + dcb[j] * l2[base + j] : q;
  };
  auto raw_x = [&](int y, int x) -> T {
    if (!in_e(y, x)) return (T)0;
    const T q = q_at(y, x);
    if (x >= g.pml_x0 && x < g.pml_x1) return q;
    const T mn = g.bx[x] * m_hz_x[base + (size_t)y * nx + x] + q;
    return g.ax[x] * mn + q / g.kx[x];
  };
  auto raw_y = [&](int y, int x) -> T {
    if (!in_e(y, x)) return (T)0;
    const T q = q_at(y, x);
    if (y >= g.pml_y0 && y < g.pml_y1) return -q;
    const T mn = g.by[y] * m_hx_z[base + (size_t)y * nx + x] - q;
    return g.ay[y] * mn - q / g.ky[y];
  };
  const size_t i = base + c.j;
  const T q = q_at(c.y, c.x);
  const bool px = c.x < g.pml_x0 || c.x >= g.pml_x1;
  const bool py = c.y < g.pml_y0 || c.y >= g.pml_y1;
  m_hz_x_n[i] = px ? g.bx[c.x] * m_hz_x[i] + q : (T)0;
  m_hx_z_n[i] = py ? g.by[c.y] * m_hx_z[i] - q : (T)0;
  l_hz[i] = l_hz[i] + d_int_adj<T, P>([&](int o) { return raw_x(c.y, c.x + o); }, g.rdx);
  l_hx[i] = l_hx[i] + d_int_adj<T, P>([&](int o) { return raw_y(c.y + o, c.x); }, g.rdy);
}

// Transpose of the H half-step: lambda_Ey = ca lambda_Ey + the half-point
// derivative adjoints of the neighbours' cq lambda_H (through their CPML
// memories); the H-phase adjoint memories advance into the `_n` copies.
// With `dca` and `l2`: ca lambda_Ey + dca lambda2 (the scattered E update
// adds dca Ey^n of the background) before the derivative terms.
template <typename T, int P>
__global__ __launch_bounds__(256) void h_adjoint(Grid<T> g, T *__restrict__ l_ey,
                                                 const T *__restrict__ l_hx,
                                                 const T *__restrict__ l_hz,
                                                 const T *__restrict__ m_ey_x,
                                                 const T *__restrict__ m_ey_z,
                                                 T *__restrict__ m_ey_x_n, T *__restrict__ m_ey_z_n,
                                                 const T *__restrict__ dca,
                                                 const T *__restrict__ l2) {
  const Cell c = cell_of(g);
  const int ny = g.ny, nx = g.nx;
  if (c.y < P || c.y >= ny - P + 1 || c.x < P || c.x >= nx - P + 1) return;
  const size_t base = (size_t)c.s * g.numel;
  const T *cqp = g.cq + (g.cq_b ? base : 0);
  auto raw_z = [&](int y, int x) -> T {  // Hx domain
    if (y < P || y >= ny - P || x < P || x >= nx - P + 1) return (T)0;
    const size_t j = (size_t)y * nx + x;
    const T adj = -cqp[j] * l_hx[base + j];
    if (y >= g.pml_y0 && y < g.y1h) return adj;
    const T mn = g.byh[y] * m_ey_z[base + j] + adj;
    return g.ayh[y] * mn + adj / g.kyh[y];
  };
  auto raw_xh = [&](int y, int x) -> T {  // Hz domain
    if (y < P || y >= ny - P + 1 || x < P || x >= nx - P) return (T)0;
    const size_t j = (size_t)y * nx + x;
    const T adj = cqp[j] * l_hz[base + j];
    if (x >= g.pml_x0 && x < g.x1h) return adj;
    const T mn = g.bxh[x] * m_ey_x[base + j] + adj;
    return g.axh[x] * mn + adj / g.kxh[x];
  };
  const size_t i = base + c.j;
  const T cq = cqp[c.j];
  const bool py = c.y < g.pml_y0 || c.y >= g.y1h;
  const bool px = c.x < g.pml_x0 || c.x >= g.x1h;
  const T adj_y = -cq * l_hx[i];
  const T adj_x = cq * l_hz[i];
  m_ey_z_n[i] = (c.y < ny - P && py) ? g.byh[c.y] * m_ey_z[i] + adj_y : (T)0;
  m_ey_x_n[i] = (c.x < nx - P && px) ? g.bxh[c.x] * m_ey_x[i] + adj_x : (T)0;
  const T ca = g.ca[(g.ca_b ? base : 0) + c.j];
  const T l = l2 != nullptr ? ca * l_ey[i] + dca[c.j] * l2[i] : ca * l_ey[i];
  l_ey[i] = l + (d_half_adj<T, P>([&](int o) { return raw_z(c.y + o, c.x); }, g.rdy) +
                 d_half_adj<T, P>([&](int o) { return raw_xh(c.y, c.x + o); }, g.rdx));
}

// gca += step_scale * sum_s lambda_Ey * Ey^n, gcb likewise with curl H^n,
// FP64, shots added in order (per shot when the coefficient is per shot).
template <typename T>
__global__ __launch_bounds__(256) void coeff_grad(Grid<T> g, int P, const T *__restrict__ l_ey,
                                                  const void *ey_snap, const void *curl_snap,
                                                  int fmt, double *gca, double *gcb,
                                                  int per_shot, double scale) {
  const int x = (int)(blockIdx.x * 64 + threadIdx.x);
  const int y = (int)(blockIdx.y * 4 + threadIdx.y);
  if (y < P || y >= g.ny - P + 1 || x < P || x >= g.nx - P + 1) return;
  const size_t j = (size_t)y * g.nx + x;
  double sa = 0.0, sb = 0.0;
  for (int s = 0; s < g.n_shots; ++s) {
    const size_t i = (size_t)s * g.numel + j;
    const double lam = (double)l_ey[i];
    const double pa = ey_snap ? lam * get_snap<T>(ey_snap, i, fmt) * scale : 0.0;
    const double pb = curl_snap ? lam * get_snap<T>(curl_snap, i, fmt) * scale : 0.0;
    if (per_shot) {
      if (gca && ey_snap) gca[i] += pa;
      if (gcb && curl_snap) gcb[i] += pb;
    } else {
      sa += pa;
      sb += pb;
    }
  }
  if (!per_shot) {
    if (gca && ey_snap) gca[j] += sa;
    if (gcb && curl_snap) gcb[j] += sb;
  }
}

// ------------------------------------------------------------------ driver

template <typename T>
Grid<T> make_grid(const Problem<T> &p) {
  Grid<T> g;
  g.ca = p.ca;
  g.cb = p.cb;
  g.cq = p.cq;
  const T *const *q = p.prof;
  g.ay = q[0]; g.by = q[1]; g.ayh = q[2]; g.byh = q[3];
  g.ax = q[4]; g.bx = q[5]; g.axh = q[6]; g.bxh = q[7];
  g.ky = q[8]; g.kyh = q[9]; g.kx = q[10]; g.kxh = q[11];
  g.rdy = p.rdy;
  g.rdx = p.rdx;
  g.ny = (int32_t)p.ny;
  g.nx = (int32_t)p.nx;
  g.numel = (int32_t)(p.ny * p.nx);
  g.n_shots = (int32_t)p.n_shots;
  g.pml_y0 = (int32_t)p.pml_y0;
  g.pml_x0 = (int32_t)p.pml_x0;
  g.pml_y1 = (int32_t)p.pml_y1;
  g.pml_x1 = (int32_t)p.pml_x1;
  g.y1h = (int32_t)(p.pml_y1 > p.pml_y0 ? p.pml_y1 - 1 : p.pml_y0);
  g.x1h = (int32_t)(p.pml_x1 > p.pml_x0 ? p.pml_x1 - 1 : p.pml_x0);
  g.ca_b = p.ca_b;
  g.cb_b = p.cb_b;
  g.cq_b = p.cq_b;
  return g;
}

inline int check(const int64_t n_shots, int64_t ny, int64_t nx, int P) {
  if (n_shots <= 0 || ny < 2 * P + 1 || nx < 2 * P + 1) {
    set_error("n_shots must be positive and the padded grid wider than the stencil");
    return kInvalidArgument;
  }
  if (ny * nx >= ((int64_t)1 << 31) / 2 || n_shots > 65535) {
    set_error("a shot plane must hold fewer than 2^30 cells and a call at most 65535 shots");
    return kInvalidArgument;
  }
  return kOk;
}

// Points arrive as host or device int64 arrays; they are checked on the
// host (an index outside the shot plane is an error, a negative one is
// skipped) and the kernels read a device copy.
struct DevIdx {
  int64_t *p = nullptr;
  hipStream_t s = nullptr;
  ~DevIdx() {
    if (p) (void)hipFreeAsync(p, s);
  }
  int take(const int64_t *src, int64_t n, int64_t numel, hipStream_t stream) {
    s = stream;
    if (n <= 0 || src == nullptr) return kOk;
    std::vector<int64_t> h((size_t)n);
    TIDE_HIP_TRY(hipMemcpyAsync(h.data(), src, (size_t)n * sizeof(int64_t), hipMemcpyDefault,
                                stream));
    TIDE_HIP_TRY(hipStreamSynchronize(stream));
    for (int64_t v : h)
      if (v >= numel) {
        set_error("source or receiver index " + std::to_string(v) +
                  " is outside the padded shot plane of " + std::to_string(numel) + " cells");
        return kInvalidArgument;
      }
    TIDE_HIP_TRY(hipMallocAsync((void **)&p, (size_t)n * sizeof(int64_t), stream));
    TIDE_HIP_TRY(hipMemcpyAsync(p, h.data(), (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice,
                                stream));
    // the host copy is pageable and freed when this call returns
    TIDE_HIP_TRY(hipStreamSynchronize(stream));
    return kOk;
  }
};

template <typename T, int P>
int forward_p(const Problem<T> &p, T *const fld[7], const T *f, T *r, const int64_t *src,
              const int64_t *rec, int64_t nt, int64_t start_t, int64_t step_ratio,
              void *ey_store, void *curl_store, int fmt, size_t step_bytes, hipStream_t stream,
              SnapshotStage *ey_stage, SnapshotStage *curl_stage) {
  TIDE_CHECK_HO(check(p.n_shots, p.ny, p.nx, P));
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  if (nt <= 0) return kOk;
  const Grid<T> g = make_grid(p);
  DevIdx ds, dr;
  TIDE_CHECK_HO(ds.take(src, p.n_shots * p.n_src, p.ny * p.nx, stream));
  TIDE_CHECK_HO(dr.take(rec, p.n_shots * p.n_rec, p.ny * p.nx, stream));
  const dim3 block(64, 4);
  const dim3 grid((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), (unsigned)p.n_shots);
  const unsigned pts_blocks_s = (unsigned)((p.n_shots + 63) / 64);
  const unsigned pts_blocks_r = (unsigned)((p.n_shots * p.n_rec + 255) / 256);
  T *ey = fld[0], *hx = fld[1], *hz = fld[2], *mex = fld[3], *mez = fld[4], *mhxz = fld[5],
    *mhzx = fld[6];
  SnapshotStage *stages[2] = {ey_stage, curl_stage};
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    const bool st = (ey_store || curl_store) && (t % step_ratio) == 0;
    const int64_t idx = t / step_ratio;
    void *es = nullptr, *cs = nullptr;
    if (st) {
      es = ey_store ? (ey_stage && ey_stage->active() ? ey_stage->slot(idx)
                                                      : (char *)ey_store + (size_t)idx * step_bytes)
                    : nullptr;
      cs = curl_store ? (curl_stage && curl_stage->active()
                             ? curl_stage->slot(idx)
                             : (char *)curl_store + (size_t)idx * step_bytes)
                      : nullptr;
      for (int k = 0; k < 2; ++k)
        if (stages[k] && stages[k]->active() && (k == 0 ? es : cs))
          TIDE_CHECK_HO(stages[k]->before_write(idx));
    }
    h_step<T, P><<<grid, block, 0, stream>>>(g, ey, hx, hz, mex, mez);
    e_step<T, P><<<grid, block, 0, stream>>>(g, ey, hx, hz, mhxz, mhzx, es, cs, fmt);
    if (st)
      for (int k = 0; k < 2; ++k)
        if (stages[k] && stages[k]->active() && (k == 0 ? es : cs))
          TIDE_CHECK_HO(stages[k]->after_write(idx));
    if (p.n_src > 0 && f != nullptr)
      add_points<T><<<pts_blocks_s, 64, 0, stream>>>(ey, g.numel, g.n_shots, ds.p, (int)p.n_src,
                                                     f + t * p.n_shots * p.n_src);
    if (p.n_rec > 0 && r != nullptr)
      read_points<T><<<pts_blocks_r, 256, 0, stream>>>(ey, g.numel, g.n_shots, dr.p, (int)p.n_rec,
                                                       r + t * p.n_shots * p.n_rec);
  }
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

// Debye-dispersive forward (forward inference, unbatched coefficients; the
// reference's native Debye path: src/tide/csrc/tm2d/maxwell_tm_inst.cpp:
// 780-805, CUDA src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:382-457).  Per
// step: H half-step, Debye E step, sources, polarization, receivers.
template <typename T, int P>
int forward_debye_p(const Problem<T> &p, T *const fld[7], const T *f, T *r, const int64_t *src,
                    const int64_t *rec, int64_t nt, int64_t start_t, const T *da, const T *db,
                    const T *cp, T *pol, T *ey_prev, int64_t n_poles, hipStream_t stream) {
  TIDE_CHECK_HO(check(p.n_shots, p.ny, p.nx, P));
  if (p.ca_b || p.cb_b || p.cq_b) {
    set_error("Debye dispersion: per-shot coefficients are not supported");
    return kUnsupported;
  }
  if (n_poles < 1 || n_poles > 64 || !da || !db || !cp || !pol || !ey_prev) {
    set_error("Debye dispersion needs 1..64 poles and the a, b, cp, polarization and ey_prev "
              "arrays");
    return kInvalidArgument;
  }
  if (nt <= 0) return kOk;
  const Grid<T> g = make_grid(p);
  DevIdx ds, dr;
  TIDE_CHECK_HO(ds.take(src, p.n_shots * p.n_src, p.ny * p.nx, stream));
  TIDE_CHECK_HO(dr.take(rec, p.n_shots * p.n_rec, p.ny * p.nx, stream));
  const dim3 block(64, 4);
  const dim3 grid((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), (unsigned)p.n_shots);
  const dim3 pgrid((unsigned)((p.ny * p.nx + 255) / 256), (unsigned)p.n_shots);
  const unsigned pts_blocks_s = (unsigned)((p.n_shots + 63) / 64);
  const unsigned pts_blocks_r = (unsigned)((p.n_shots * p.n_rec + 255) / 256);
  T *ey = fld[0], *hx = fld[1], *hz = fld[2], *mex = fld[3], *mez = fld[4], *mhxz = fld[5],
    *mhzx = fld[6];
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    h_step<T, P><<<grid, block, 0, stream>>>(g, ey, hx, hz, mex, mez);
    e_step_debye<T, P><<<grid, block, 0, stream>>>(g, ey, hx, hz, mhxz, mhzx, ey_prev, cp, pol,
                                                   (int)n_poles);
    if (p.n_src > 0 && f != nullptr)
      add_points<T><<<pts_blocks_s, 64, 0, stream>>>(ey, g.numel, g.n_shots, ds.p, (int)p.n_src,
                                                     f + t * p.n_shots * p.n_src);
    pol_update<T><<<pgrid, 256, 0, stream>>>(ey, ey_prev, da, db, pol, (int64_t)g.numel,
                                             g.n_shots, (int)n_poles);
    if (p.n_rec > 0 && r != nullptr)
      read_points<T><<<pts_blocks_r, 256, 0, stream>>>(ey, g.numel, g.n_shots, dr.p, (int)p.n_rec,
                                                       r + t * p.n_shots * p.n_rec);
  }
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

template <typename T, int P>
int backward_p(const Problem<T> &p, T *const lam[7], const T *grad_r, const void *ey_store,
               const void *curl_store, int fmt, size_t step_bytes, T *grad_f, double *gca,
               double *gcb, const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
               int64_t step_ratio, hipStream_t stream, SnapshotStage *ey_stage,
               SnapshotStage *curl_stage) {
  TIDE_CHECK_HO(check(p.n_shots, p.ny, p.nx, P));
  if (fmt == 2 && sizeof(T) != 4) {
    set_error("BF16 snapshots require fp32 fields");
    return kUnsupported;
  }
  if ((gca && !ey_store) || (gcb && !curl_store)) {
    set_error("coefficient gradients require the stored forward snapshots");
    return kInvalidArgument;
  }
  if (nt <= 0) return kOk;
  const Grid<T> g = make_grid(p);
  DevIdx ds, dr;
  TIDE_CHECK_HO(ds.take(src, p.n_shots * p.n_src, p.ny * p.nx, stream));
  TIDE_CHECK_HO(dr.take(rec, p.n_shots * p.n_rec, p.ny * p.nx, stream));
  const size_t total = (size_t)p.n_shots * p.ny * p.nx;
  T *scratch = nullptr;
  TIDE_HIP_TRY(hipMallocAsync((void **)&scratch, 4 * total * sizeof(T), stream));
  struct Free {
    T *q;
    hipStream_t s;
    ~Free() { (void)hipFreeAsync(q, s); }
  } free_scratch{scratch, stream};
  TIDE_HIP_TRY(hipMemsetAsync(scratch, 0, 4 * total * sizeof(T), stream));
  // lam: l_ey, l_hx, l_hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x
  T *l_ey = lam[0], *l_hx = lam[1], *l_hz = lam[2];
  T *cur[4] = {lam[3], lam[4], lam[5], lam[6]};  // m_ey_x m_ey_z m_hx_z m_hz_x
  T *nxt[4] = {scratch, scratch + total, scratch + 2 * total, scratch + 3 * total};
  const dim3 block(64, 4);
  const dim3 grid((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), (unsigned)p.n_shots);
  const dim3 grid2((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), 1);
  const unsigned pts_blocks_s = (unsigned)((p.n_shots * p.n_src + 255) / 256);
  const unsigned pts_blocks_r = (unsigned)((p.n_shots + 63) / 64);
  const int per_shot = ((gca && p.ca_b) || (gcb && p.cb_b)) ? 1 : 0;
  SnapshotStage *stages[2] = {gca ? ey_stage : nullptr, gcb ? curl_stage : nullptr};
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
    e_adjoint<T, P><<<grid, block, 0, stream>>>(g, l_ey, l_hx, l_hz, cur[2], cur[3], nxt[2],
                                                nxt[3], nullptr, nullptr);
    std::swap(cur[2], nxt[2]);
    std::swap(cur[3], nxt[3]);
    h_adjoint<T, P><<<grid, block, 0, stream>>>(g, l_ey, l_hx, l_hz, cur[0], cur[1], nxt[0],
                                                nxt[1], nullptr, nullptr);
    std::swap(cur[0], nxt[0]);
    std::swap(cur[1], nxt[1]);
    if (p.n_rec > 0 && grad_r != nullptr)
      add_points<T><<<pts_blocks_r, 64, 0, stream>>>(l_ey, g.numel, g.n_shots, dr.p, (int)p.n_rec,
                                                     grad_r + t * p.n_shots * p.n_rec);
    if (p.n_src > 0 && grad_f != nullptr)
      read_points<T><<<pts_blocks_s, 256, 0, stream>>>(l_ey, g.numel, g.n_shots, ds.p,
                                                       (int)p.n_src, grad_f + t * p.n_shots * p.n_src);
    if ((gca || gcb) && (t % step_ratio) == 0) {
      const int64_t idx = t / step_ratio;
      const void *es = nullptr, *cs = nullptr;
      if (gca)
        es = ey_stage && ey_stage->active() ? ey_stage->slot(idx)
                                            : (const char *)ey_store + (size_t)idx * step_bytes;
      if (gcb)
        cs = curl_stage && curl_stage->active()
                 ? curl_stage->slot(idx)
                 : (const char *)curl_store + (size_t)idx * step_bytes;
      for (int k = 0; k < 2; ++k)
        if (stages[k] && stages[k]->active()) TIDE_CHECK_HO(stages[k]->before_read(idx));
      coeff_grad<T><<<grid2, block, 0, stream>>>(g, P, l_ey, es, cs, fmt, gca, gcb, per_shot,
                                                 (double)step_ratio);
      for (int k = 0; k < 2; ++k)
        if (stages[k] && stages[k]->active()) {
          TIDE_CHECK_HO(stages[k]->after_read(idx));
          // the next stored step of the reverse loop
          const int64_t nxt_t = ((t - 1) / step_ratio) * step_ratio;
          if (t - 1 >= start_t - nt && nxt_t >= start_t - nt)
            TIDE_CHECK_HO(stages[k]->prefetch(nxt_t / step_ratio));
        }
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  const size_t bytes = total * sizeof(T);
  for (int k = 0; k < 4; ++k)
    if (cur[k] != lam[3 + k])
      TIDE_HIP_TRY(hipMemcpyAsync(lam[3 + k], cur[k], bytes, hipMemcpyDeviceToDevice, stream));
  return kOk;
}

// Born (tangent-linear) forward: the background and the scattered field
// advance together, one step at a time.  Per step: background H and E
// (its pre-update Ey and curl H kept in two planes), scattered H and E with
// the scattering source dca Ey^n + dcb curl H^n, then the sources (f0 into
// the background, df into the scattered field) and both receiver sets.
// With `hist` (four [nt, shots, ny, nx] arrays: background Ey^n and
// curl^n, scattered Ey^n and curl^n; slot t for step t) the planes of every
// step are kept for the Born adjoint.  Reference: the tangent recurrence of
// src/tide/maxwell/tm2d_born_python.py and the native born_forward
// (src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:533-742, 1611-1800).
template <typename T, int P>
int born_p(const Problem<T> &p, T *const bg[7], T *const sc[7], const T *dca, const T *dcb,
           const T *f0, const T *df, T *r, T *bg_r, const int64_t *src, const int64_t *rec,
           int64_t nt, int64_t start_t, hipStream_t stream, T *const *hist) {
  TIDE_CHECK_HO(check(p.n_shots, p.ny, p.nx, P));
  if (nt <= 0) return kOk;
  if ((dca == nullptr) != (dcb == nullptr)) {
    set_error("born_forward: dca and dcb are both given or both null");
    return kInvalidArgument;
  }
  const Grid<T> g = make_grid(p);
  DevIdx ds, dr;
  TIDE_CHECK_HO(ds.take(src, p.n_shots * p.n_src, p.ny * p.nx, stream));
  TIDE_CHECK_HO(dr.take(rec, p.n_shots * p.n_rec, p.ny * p.nx, stream));
  const size_t total = (size_t)p.n_shots * p.ny * p.nx;
  T *planes = nullptr;
  TIDE_HIP_TRY(hipMallocAsync((void **)&planes, 2 * total * sizeof(T), stream));
  struct Free {
    T *q;
    hipStream_t s;
    ~Free() { (void)hipFreeAsync(q, s); }
  } free_planes{planes, stream};
  const BornSrc<T> none{nullptr, nullptr, nullptr, nullptr};
  const dim3 block(64, 4);
  const dim3 grid((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), (unsigned)p.n_shots);
  const unsigned blocks_s = (unsigned)((p.n_shots + 63) / 64);
  const unsigned blocks_r = (unsigned)((p.n_shots * p.n_rec + 255) / 256);
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    T *be = planes, *bc = planes + total, *se = nullptr, *scu = nullptr;
    if (hist != nullptr) {
      be = hist[0] + (size_t)t * total;
      bc = hist[1] + (size_t)t * total;
      se = hist[2] + (size_t)t * total;
      scu = hist[3] + (size_t)t * total;
    }
    const BornSrc<T> born{dca, dcb, be, bc};
    h_step<T, P><<<grid, block, 0, stream>>>(g, bg[0], bg[1], bg[2], bg[3], bg[4]);
    e_step<T, P><<<grid, block, 0, stream>>>(g, bg[0], bg[1], bg[2], bg[5], bg[6], be, bc, 1,
                                             none);
    h_step<T, P><<<grid, block, 0, stream>>>(g, sc[0], sc[1], sc[2], sc[3], sc[4]);
    e_step<T, P><<<grid, block, 0, stream>>>(g, sc[0], sc[1], sc[2], sc[5], sc[6], se, scu, 1,
                                             dca != nullptr ? born : none);
    if (p.n_src > 0) {
      if (f0 != nullptr)
        add_points<T><<<blocks_s, 64, 0, stream>>>(bg[0], g.numel, g.n_shots, ds.p, (int)p.n_src,
                                                   f0 + t * p.n_shots * p.n_src);
      if (df != nullptr)
        add_points<T><<<blocks_s, 64, 0, stream>>>(sc[0], g.numel, g.n_shots, ds.p, (int)p.n_src,
                                                   df + t * p.n_shots * p.n_src);
    }
    if (p.n_rec > 0) {
      if (bg_r != nullptr)
        read_points<T><<<blocks_r, 256, 0, stream>>>(bg[0], g.numel, g.n_shots, dr.p,
                                                     (int)p.n_rec, bg_r + t * p.n_shots * p.n_rec);
      if (r != nullptr)
        read_points<T><<<blocks_r, 256, 0, stream>>>(sc[0], g.numel, g.n_shots, dr.p,
                                                     (int)p.n_rec, r + t * p.n_shots * p.n_rec);
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

// Adjoint of born_p (unbatched coefficients): gradients of <r, w> +
// <bg_r, w_bg> with respect to ca, cb, dca, dcb (FP64, added to; summed over
// shots), f0 and df (per step, written).  lam (scattered) and eta
// (background) adjoint states: Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x,
// advanced in place; `mem` is 8 [shots, ny, nx] arrays of scratch for the
// adjoint memories' second copies (null: allocated here).  Reverse step t
// transposes step t + 1 -- the background first, coupled to the scattered
// adjoint of the same step (the scattered E update reads the background's
// Ey^n and curl^n), then the scattered field's plain adjoint step -- then
// takes step t's receivers, source gradients and FP64 terms:
//   ca  += eta Ey^n + lam dEy^n     cb  += eta curl^n + lam dcurl^n
//   dca += lam Ey^n                 dcb += lam curl^n
// Reference counterpart: maxwell_tm_{S}_{T}_born_backward_bggrad_cuda
// (src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1611-2404).
template <typename T, int P>
int born_backward_p(const Problem<T> &p, T *const lam[7], T *const eta[7], const T *dca,
                    const T *dcb, const T *w, const T *w_bg, const T *const hist[4], T *g_df,
                    T *g_f0, double *g_ca, double *g_cb, double *g_dca, double *g_dcb,
                    const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                    hipStream_t stream, T *mem) {
  TIDE_CHECK_HO(check(p.n_shots, p.ny, p.nx, P));
  if (p.ca_b || p.cb_b || p.cq_b) {
    set_error("born_backward: per-shot coefficients are not supported");
    return kUnsupported;
  }
  if (dca == nullptr || dcb == nullptr || hist == nullptr || hist[0] == nullptr ||
      hist[1] == nullptr || hist[2] == nullptr || hist[3] == nullptr) {
    set_error("born_backward: dca, dcb and the four snapshot histories are required");
    return kInvalidArgument;
  }
  if (nt <= 0) return kOk;
  if (start_t - nt < 0) {
    set_error("born_backward: the reverse loop would run past step 0");
    return kInvalidArgument;
  }
  const Grid<T> g = make_grid(p);
  DevIdx ds, dr;
  TIDE_CHECK_HO(ds.take(src, p.n_shots * p.n_src, p.ny * p.nx, stream));
  TIDE_CHECK_HO(dr.take(rec, p.n_shots * p.n_rec, p.ny * p.nx, stream));
  const size_t total = (size_t)p.n_shots * p.ny * p.nx;
  T *own = nullptr;
  if (mem == nullptr) {
    TIDE_HIP_TRY(hipMallocAsync((void **)&own, 8 * total * sizeof(T), stream));
    mem = own;
  }
  struct Free {
    T *q;
    hipStream_t s;
    ~Free() {
      if (q) (void)hipFreeAsync(q, s);
    }
  } free_mem{own, stream};
  T *const *sys[2] = {eta, lam};  // 0: background (coupled), 1: scattered
  T *cur[2][4], *nxt[2][4];       // m_ey_x m_ey_z m_hx_z m_hz_x
  for (int k = 0; k < 2; ++k)
    for (int m = 0; m < 4; ++m) {
      cur[k][m] = sys[k][3 + m];
      nxt[k][m] = mem + (size_t)(4 * k + m) * total;
    }
  const dim3 block(64, 4);
  const dim3 grid((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), (unsigned)p.n_shots);
  const dim3 grid2((unsigned)((p.nx + 63) / 64), (unsigned)((p.ny + 3) / 4), 1);
  const unsigned blocks_s = (unsigned)((p.n_shots * p.n_src + 255) / 256);
  const unsigned blocks_r = (unsigned)((p.n_shots + 63) / 64);
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
    for (int k = 0; k < 2; ++k) {
      const T *l2 = k == 0 ? lam[0] : nullptr;
      e_adjoint<T, P><<<grid, block, 0, stream>>>(g, sys[k][0], sys[k][1], sys[k][2], cur[k][2],
                                                  cur[k][3], nxt[k][2], nxt[k][3], dcb, l2);
      std::swap(cur[k][2], nxt[k][2]);
      std::swap(cur[k][3], nxt[k][3]);
      h_adjoint<T, P><<<grid, block, 0, stream>>>(g, sys[k][0], sys[k][1], sys[k][2], cur[k][0],
                                                  cur[k][1], nxt[k][0], nxt[k][1], dca, l2);
      std::swap(cur[k][0], nxt[k][0]);
      std::swap(cur[k][1], nxt[k][1]);
    }
    if (p.n_rec > 0) {
      if (w != nullptr)
        add_points<T><<<blocks_r, 64, 0, stream>>>(lam[0], g.numel, g.n_shots, dr.p,
                                                   (int)p.n_rec, w + t * p.n_shots * p.n_rec);
      if (w_bg != nullptr)
        add_points<T><<<blocks_r, 64, 0, stream>>>(eta[0], g.numel, g.n_shots, dr.p,
                                                   (int)p.n_rec, w_bg + t * p.n_shots * p.n_rec);
    }
    if (p.n_src > 0) {
      if (g_df != nullptr)
        read_points<T><<<blocks_s, 256, 0, stream>>>(lam[0], g.numel, g.n_shots, ds.p,
                                                     (int)p.n_src, g_df + t * p.n_shots * p.n_src);
      if (g_f0 != nullptr)
        read_points<T><<<blocks_s, 256, 0, stream>>>(eta[0], g.numel, g.n_shots, ds.p,
                                                     (int)p.n_src, g_f0 + t * p.n_shots * p.n_src);
    }
    const T *be = hist[0] + (size_t)t * total, *bc = hist[1] + (size_t)t * total;
    const T *se = hist[2] + (size_t)t * total, *scu = hist[3] + (size_t)t * total;
    if (g_ca != nullptr || g_cb != nullptr) {
      coeff_grad<T><<<grid2, block, 0, stream>>>(g, P, eta[0], g_ca ? be : nullptr,
                                                 g_cb ? bc : nullptr, 1, g_ca, g_cb, 0, 1.0);
      coeff_grad<T><<<grid2, block, 0, stream>>>(g, P, lam[0], g_ca ? se : nullptr,
                                                 g_cb ? scu : nullptr, 1, g_ca, g_cb, 0, 1.0);
    }
    if (g_dca != nullptr || g_dcb != nullptr)
      coeff_grad<T><<<grid2, block, 0, stream>>>(g, P, lam[0], g_dca ? be : nullptr,
                                                 g_dcb ? bc : nullptr, 1, g_dca, g_dcb, 0, 1.0);
  }
  TIDE_HIP_TRY(hipGetLastError());
  const size_t bytes = total * sizeof(T);
  for (int k = 0; k < 2; ++k)
    for (int m = 0; m < 4; ++m)
      if (cur[k][m] != sys[k][3 + m])
        TIDE_HIP_TRY(hipMemcpyAsync(sys[k][3 + m], cur[k][m], bytes, hipMemcpyDeviceToDevice,
                                    stream));
  return kOk;
}

template <typename T>
int forward_debye(int stencil, const Problem<T> &p, T *const fld[7], const T *f, T *r,
                  const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                  const T *da, const T *db, const T *cp, T *pol, T *ey_prev, int64_t n_poles,
                  hipStream_t stream) {
#define TIDE_DEBYE_P(K) \
  forward_debye_p<T, K>(p, fld, f, r, src, rec, nt, start_t, da, db, cp, pol, ey_prev, n_poles, stream)
  switch (stencil) {
    case 2: return TIDE_DEBYE_P(1);
    case 4: return TIDE_DEBYE_P(2);
    case 6: return TIDE_DEBYE_P(3);
    case 8: return TIDE_DEBYE_P(4);
    default:
      set_error("stencil must be 2, 4, 6 or 8");
      return kInvalidArgument;
  }
#undef TIDE_DEBYE_P
}

template int forward_debye<float>(int, const Problem<float> &, float *const[7], const float *,
                                  float *, const int64_t *, const int64_t *, int64_t, int64_t,
                                  const float *, const float *, const float *, float *, float *,
                                  int64_t, hipStream_t);
template int forward_debye<double>(int, const Problem<double> &, double *const[7],
                                   const double *, double *, const int64_t *, const int64_t *,
                                   int64_t, int64_t, const double *, const double *,
                                   const double *, double *, double *, int64_t, hipStream_t);

template <typename T>
int born_forward(int stencil, const Problem<T> &p, T *const bg[7], T *const sc[7],
                 const T *dca, const T *dcb, const T *f0, const T *df, T *r, T *bg_r,
                 const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                 hipStream_t stream, T *const *hist) {
#define TIDE_BORN_P(K) \
  born_p<T, K>(p, bg, sc, dca, dcb, f0, df, r, bg_r, src, rec, nt, start_t, stream, hist)
  switch (stencil) {
    case 2: return TIDE_BORN_P(1);
    case 4: return TIDE_BORN_P(2);
    case 6: return TIDE_BORN_P(3);
    case 8: return TIDE_BORN_P(4);
    default:
      set_error("stencil must be 2, 4, 6 or 8");
      return kInvalidArgument;
  }
#undef TIDE_BORN_P
}

template <typename T>
int born_backward(int stencil, const Problem<T> &p, T *const lam[7], T *const eta[7],
                  const T *dca, const T *dcb, const T *w, const T *w_bg, const T *const hist[4],
                  T *g_df, T *g_f0, double *g_ca, double *g_cb, double *g_dca, double *g_dcb,
                  const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                  hipStream_t stream, T *mem) {
#define TIDE_BORN_B(K)                                                                     \
  born_backward_p<T, K>(p, lam, eta, dca, dcb, w, w_bg, hist, g_df, g_f0, g_ca, g_cb, g_dca, \
                        g_dcb, src, rec, nt, start_t, stream, mem)
  switch (stencil) {
    case 2: return TIDE_BORN_B(1);
    case 4: return TIDE_BORN_B(2);
    case 6: return TIDE_BORN_B(3);
    case 8: return TIDE_BORN_B(4);
    default:
      set_error("stencil must be 2, 4, 6 or 8");
      return kInvalidArgument;
  }
#undef TIDE_BORN_B
}

template int born_forward<float>(int, const Problem<float> &, float *const[7], float *const[7],
                                 const float *, const float *, const float *, const float *,
                                 float *, float *, const int64_t *, const int64_t *, int64_t,
                                 int64_t, hipStream_t, float *const *);
template int born_forward<double>(int, const Problem<double> &, double *const[7],
                                  double *const[7], const double *, const double *,
                                  const double *, const double *, double *, double *,
                                  const int64_t *, const int64_t *, int64_t, int64_t, hipStream_t,
                                  double *const *);
template int born_backward<float>(int, const Problem<float> &, float *const[7], float *const[7],
                                  const float *, const float *, const float *, const float *,
                                  const float *const[4], float *, float *, double *, double *,
                                  double *, double *, const int64_t *, const int64_t *, int64_t,
                                  int64_t, hipStream_t, float *);
template int born_backward<double>(int, const Problem<double> &, double *const[7],
                                   double *const[7], const double *, const double *,
                                   const double *, const double *, const double *const[4],
                                   double *, double *, double *, double *, double *, double *,
                                   const int64_t *, const int64_t *, int64_t, int64_t,
                                   hipStream_t, double *);

template <typename T>
int forward(int stencil, const Problem<T> &p, T *const fld[7], const T *f, T *r,
            const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
            int64_t step_ratio, void *ey_store, void *curl_store, int fmt, size_t step_bytes,
            hipStream_t stream, SnapshotStage *ey_stage, SnapshotStage *curl_stage) {
  switch (stencil) {
    case 4: return forward_p<T, 2>(p, fld, f, r, src, rec, nt, start_t, step_ratio, ey_store,
                                   curl_store, fmt, step_bytes, stream, ey_stage, curl_stage);
    case 6: return forward_p<T, 3>(p, fld, f, r, src, rec, nt, start_t, step_ratio, ey_store,
                                   curl_store, fmt, step_bytes, stream, ey_stage, curl_stage);
    case 8: return forward_p<T, 4>(p, fld, f, r, src, rec, nt, start_t, step_ratio, ey_store,
                                   curl_store, fmt, step_bytes, stream, ey_stage, curl_stage);
    default:
      set_error("stencil must be 2, 4, 6 or 8");
      return kInvalidArgument;
  }
}

template <typename T>
int backward(int stencil, const Problem<T> &p, T *const lam[7], const T *grad_r,
             const void *ey_store, const void *curl_store, int fmt, size_t step_bytes, T *grad_f,
             double *gca, double *gcb, const int64_t *src, const int64_t *rec, int64_t nt,
             int64_t start_t, int64_t step_ratio, hipStream_t stream, SnapshotStage *ey_stage,
             SnapshotStage *curl_stage) {
  switch (stencil) {
    case 4: return backward_p<T, 2>(p, lam, grad_r, ey_store, curl_store, fmt, step_bytes, grad_f,
                                    gca, gcb, src, rec, nt, start_t, step_ratio, stream, ey_stage,
                                    curl_stage);
    case 6: return backward_p<T, 3>(p, lam, grad_r, ey_store, curl_store, fmt, step_bytes, grad_f,
                                    gca, gcb, src, rec, nt, start_t, step_ratio, stream, ey_stage,
                                    curl_stage);
    case 8: return backward_p<T, 4>(p, lam, grad_r, ey_store, curl_store, fmt, step_bytes, grad_f,
                                    gca, gcb, src, rec, nt, start_t, step_ratio, stream, ey_stage,
                                    curl_stage);
    default:
      set_error("stencil must be 2, 4, 6 or 8");
      return kInvalidArgument;
  }
}

template int forward<float>(int, const Problem<float> &, float *const[7], const float *, float *,
                            const int64_t *, const int64_t *, int64_t, int64_t, int64_t, void *,
                            void *, int, size_t, hipStream_t, SnapshotStage *, SnapshotStage *);
template int forward<double>(int, const Problem<double> &, double *const[7], const double *,
                             double *, const int64_t *, const int64_t *, int64_t, int64_t,
                             int64_t, void *, void *, int, size_t, hipStream_t, SnapshotStage *,
                             SnapshotStage *);
template int backward<float>(int, const Problem<float> &, float *const[7], const float *,
                             const void *, const void *, int, size_t, float *, double *, double *,
                             const int64_t *, const int64_t *, int64_t, int64_t, int64_t,
                             hipStream_t, SnapshotStage *, SnapshotStage *);
template int backward<double>(int, const Problem<double> &, double *const[7], const double *,
                              const void *, const void *, int, size_t, double *, double *,
                              double *, const int64_t *, const int64_t *, int64_t, int64_t,
                              int64_t, hipStream_t, SnapshotStage *, SnapshotStage *);

}  // namespace tm2d_ho
}  // namespace tide_hip
