// TEST INFRASTRUCTURE ONLY -- never linked into, loaded by, or called from
// the product path.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use this file's library, and only as the checker.
//
// CPU restatement of TIDE's 2-D TM staggered-grid Maxwell propagator:
//   * forward step (H half-step then E step, CPML memories in the strips,
//     point sources, receiver gather), optionally storing the pre-update Ey
//     and the CPML curl of H for the gradient (fp32 or BF16 round-to-nearest-
//     even payload);
//   * the exact discrete adjoint of that step (gather form, ping-pong CPML
//     adjoint memories, the index-shifted reverse loop) accumulating
//     dJ/dca and dJ/dcb in FP64;
//   * the tangent-linear (Born) forward in the coefficients, used to pin the
//     adjoint with the dot-product identity <J dm, w> = <dm, J^T w>, and its
//     own adjoint (second-order products);
//   * the Debye-dispersive forward (forward inference).
//
// Followed reference code (paths relative to the reference's src/tide/):
//   stencils        csrc/tm2d/maxwell_tm_core.cuh:61-365 (DiffForward/Adjoint)
//   H update        csrc/tm2d/maxwell_tm_core.cuh:448-502
//   E update/store  csrc/tm2d/maxwell_tm_core.cuh:505-599
//   BF16 codec      csrc/tm2d/maxwell.cpp:63-83
//   forward loop    csrc/tm2d/maxwell_tm_inst.cpp:32-131, 727-820
//   point ops       csrc/tm2d/maxwell.cpp:85-160
//   Debye forward   csrc/tm2d/maxwell_tm_inst.cpp:214-300, 780-805
//   exact adjoint   csrc/tm2d/maxwell_tm_cuda_inst.cu:922-1254
//   reverse loop    csrc/tm2d/maxwell_tm_cuda_inst.cu:2540-2731
//
// Deliberate deviation: the coefficient gradients are accumulated in FP64
// (products of the stored fp32 values are exact in FP64), not in per-shot
// fp32 buffers.  Build with -ffp-contract=off so every a*b+c rounds twice,
// as the reference CPU build does on x86-64.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- stencils

template <int S> struct Coef;
template <> struct Coef<2> { static constexpr int P = 1; };
template <> struct Coef<4> { static constexpr int P = 2; };
template <> struct Coef<6> { static constexpr int P = 3; };
template <> struct Coef<8> { static constexpr int P = 4; };

// Integer-point derivative: F(k) - F(k-1) pattern (diff_x1 / diff_y1).
// `at(o)` returns the field at offset o along the axis.
template <typename T, int S, typename A>
inline T d_int(A at, T r) {
  if constexpr (S == 2) {
    return (at(0) - at(-1)) * r;
  } else if constexpr (S == 4) {
    return ((T)(9.0 / 8.0) * (at(0) - at(-1)) +
            (T)(-1.0 / 24.0) * (at(1) - at(-2))) *
           r;
  } else if constexpr (S == 6) {
    return ((T)(75.0 / 64.0) * (at(0) - at(-1)) +
            (T)(-25.0 / 384.0) * (at(1) - at(-2)) +
            (T)(3.0 / 640.0) * (at(2) - at(-3))) *
           r;
  } else {
    return ((T)(1225.0 / 1024.0) * (at(0) - at(-1)) +
            (T)(-245.0 / 3072.0) * (at(1) - at(-2)) +
            (T)(49.0 / 5120.0) * (at(2) - at(-3)) +
            (T)(-5.0 / 7168.0) * (at(3) - at(-4))) *
           r;
  }
}

// Half-point derivative: F(k+1) - F(k) pattern (diff_xh1 / diff_yh1).
template <typename T, int S, typename A>
inline T d_half(A at, T r) {
  if constexpr (S == 2) {
    return (at(1) - at(0)) * r;
  } else if constexpr (S == 4) {
    return ((T)(9.0 / 8.0) * (at(1) - at(0)) +
            (T)(-1.0 / 24.0) * (at(2) - at(-1))) *
           r;
  } else if constexpr (S == 6) {
    return ((T)(75.0 / 64.0) * (at(1) - at(0)) +
            (T)(-25.0 / 384.0) * (at(2) - at(-1)) +
            (T)(3.0 / 640.0) * (at(3) - at(-2))) *
           r;
  } else {
    return ((T)(1225.0 / 1024.0) * (at(1) - at(0)) +
            (T)(-245.0 / 3072.0) * (at(2) - at(-1)) +
            (T)(49.0 / 5120.0) * (at(3) - at(-2)) +
            (T)(-5.0 / 7168.0) * (at(4) - at(-3))) *
           r;
  }
}

// Transpose of d_int: (c(k) - c(k+1)) pattern (DiffAdjoint::diff_x1_adj).
template <typename T, int S, typename A>
inline T d_int_adj(A c, T r) {
  if constexpr (S == 2) {
    return (c(0) - c(1)) * r;
  } else if constexpr (S == 4) {
    return ((T)(9.0 / 8.0) * (c(0) - c(1)) +
            (T)(-1.0 / 24.0) * (c(-1) - c(2))) *
           r;
  } else if constexpr (S == 6) {
    return ((T)(75.0 / 64.0) * (c(0) - c(1)) +
            (T)(-25.0 / 384.0) * (c(-1) - c(2)) +
            (T)(3.0 / 640.0) * (c(-2) - c(3))) *
           r;
  } else {
    return ((T)(1225.0 / 1024.0) * (c(0) - c(1)) +
            (T)(-245.0 / 3072.0) * (c(-1) - c(2)) +
            (T)(49.0 / 5120.0) * (c(-2) - c(3)) +
            (T)(-5.0 / 7168.0) * (c(-3) - c(4))) *
           r;
  }
}

// Transpose of d_half: (c(k-1) - c(k)) pattern (DiffAdjoint::diff_xh1_adj).
template <typename T, int S, typename A>
inline T d_half_adj(A c, T r) {
  if constexpr (S == 2) {
    return (c(-1) - c(0)) * r;
  } else if constexpr (S == 4) {
    return ((T)(9.0 / 8.0) * (c(-1) - c(0)) +
            (T)(-1.0 / 24.0) * (c(-2) - c(1))) *
           r;
  } else if constexpr (S == 6) {
    return ((T)(75.0 / 64.0) * (c(-1) - c(0)) +
            (T)(-25.0 / 384.0) * (c(-2) - c(1)) +
            (T)(3.0 / 640.0) * (c(-3) - c(2))) *
           r;
  } else {
    return ((T)(1225.0 / 1024.0) * (c(-1) - c(0)) +
            (T)(-245.0 / 3072.0) * (c(-2) - c(1)) +
            (T)(49.0 / 5120.0) * (c(-3) - c(2)) +
            (T)(-5.0 / 7168.0) * (c(-4) - c(3))) *
           r;
  }
}

// ------------------------------------------------------------- BF16 codec
// Round to nearest even on the upper 16 bits (maxwell.cpp:63-83).
inline uint16_t f32_to_bf16(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}
inline float bf16_to_f32(uint16_t h) {
  uint32_t u = ((uint32_t)h) << 16;
  float v;
  std::memcpy(&v, &u, 4);
  return v;
}

// ------------------------------------------------------------------ params

template <typename T> struct Grid {
  const T *ay, *by, *ayh, *byh, *ax, *bx, *axh, *bxh;
  const T *ky, *kyh, *kx, *kxh;
  T rdy, rdx;
  int64_t n_shots, ny, nx;
  int64_t pml_y0, pml_y1, pml_x0, pml_x1;
  bool ca_b, cb_b, cq_b;
  int64_t numel() const { return ny * nx; }
};

// ------------------------------------------------------------- forward step

// H half-step for one shot (maxwell_tm_core.cuh:448-502).
template <typename T, int S>
void h_step(const Grid<T> &g, const T *cq, const T *ey, T *hx, T *hz,
            T *m_ey_x, T *m_ey_z, int64_t shot) {
  constexpr int P = Coef<S>::P;
  const int64_t ny = g.ny, nx = g.nx, base = shot * g.numel();
  const int64_t y1h = g.pml_y1 > g.pml_y0 ? g.pml_y1 - 1 : g.pml_y0;
  const int64_t x1h = g.pml_x1 > g.pml_x0 ? g.pml_x1 - 1 : g.pml_x0;
  for (int64_t y = P; y < ny - P + 1; ++y) {
    for (int64_t x = P; x < nx - P + 1; ++x) {
      const int64_t j = y * nx + x, i = base + j;
      const T cqv = g.cq_b ? cq[i] : cq[j];
      if (y < ny - P) {
        T d = d_half<T, S>([&](int o) { return ey[i + o * nx]; }, g.rdy);
        if (y < g.pml_y0 || y >= y1h) {
          m_ey_z[i] = g.byh[y] * m_ey_z[i] + g.ayh[y] * d;
          d = d / g.kyh[y] + m_ey_z[i];
        }
        hx[i] -= cqv * d;
      }
      if (x < nx - P) {
        T d = d_half<T, S>([&](int o) { return ey[i + o]; }, g.rdx);
        if (x < g.pml_x0 || x >= x1h) {
          m_ey_x[i] = g.bxh[x] * m_ey_x[i] + g.axh[x] * d;
          d = d / g.kxh[x] + m_ey_x[i];
        }
        hz[i] += cqv * d;
      }
    }
  }
}

// E step for one shot, optionally storing Ey^n and curl(H) first
// (maxwell_tm_core.cuh:505-599).  `fmt` 0 = T payload, 1 = BF16 payload.
template <typename T, int S>
void e_step(const Grid<T> &g, const T *ca, const T *cb, const T *hx,
            const T *hz, T *ey, T *m_hx_z, T *m_hz_x, void *ey_store,
            void *curl_store, int fmt, int64_t shot) {
  constexpr int P = Coef<S>::P;
  const int64_t ny = g.ny, nx = g.nx, base = shot * g.numel();
  for (int64_t y = P; y < ny - P + 1; ++y) {
    for (int64_t x = P; x < nx - P + 1; ++x) {
      const int64_t j = y * nx + x, i = base + j;
      const T cav = g.ca_b ? ca[i] : ca[j];
      const T cbv = g.cb_b ? cb[i] : cb[j];
      T dhz = d_int<T, S>([&](int o) { return hz[i + o]; }, g.rdx);
      T dhx = d_int<T, S>([&](int o) { return hx[i + o * nx]; }, g.rdy);
      if (x < g.pml_x0 || x >= g.pml_x1) {
        m_hz_x[i] = g.bx[x] * m_hz_x[i] + g.ax[x] * dhz;
        dhz = dhz / g.kx[x] + m_hz_x[i];
      }
      if (y < g.pml_y0 || y >= g.pml_y1) {
        m_hx_z[i] = g.by[y] * m_hx_z[i] + g.ay[y] * dhx;
        dhx = dhx / g.ky[y] + m_hx_z[i];
      }
      const T curl = dhz - dhx;
      if (ey_store != nullptr) {
        if (fmt == 1)
          ((uint16_t *)ey_store)[i] = f32_to_bf16((float)ey[i]);
        else
          ((T *)ey_store)[i] = ey[i];
      }
      if (curl_store != nullptr) {
        if (fmt == 1)
          ((uint16_t *)curl_store)[i] = f32_to_bf16((float)curl);
        else
          ((T *)curl_store)[i] = curl;
      }
      ey[i] = cav * ey[i] + cbv * curl;
    }
  }
}

template <typename T>
inline T load_store(const void *p, int fmt, int64_t i) {
  if (fmt == 1) return (T)bf16_to_f32(((const uint16_t *)p)[i]);
  return ((const T *)p)[i];
}

template <typename T, int S>
void forward(const T *ca, const T *cb, const T *cq, const T *f, T *ey, T *hx,
             T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x, T *r,
             const Grid<T> &g, const int64_t *src, const int64_t *rec,
             int64_t n_src, int64_t n_rec, int64_t nt, int64_t start_t,
             int64_t step_ratio, void *ey_store, void *curl_store, int fmt,
             int64_t store_bytes_per_step) {
  const int64_t numel = g.numel();
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    const bool store = (ey_store != nullptr || curl_store != nullptr) &&
                       (t % step_ratio) == 0;
    const int64_t idx = t / step_ratio;
    char *es = store && ey_store
                   ? (char *)ey_store + idx * store_bytes_per_step
                   : nullptr;
    char *cs = store && curl_store
                   ? (char *)curl_store + idx * store_bytes_per_step
                   : nullptr;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < g.n_shots; ++s) {
      h_step<T, S>(g, cq, ey, hx, hz, m_ey_x, m_ey_z, s);
      e_step<T, S>(g, ca, cb, hx, hz, ey, m_hx_z, m_hz_x, es, cs, fmt, s);
      for (int64_t k = 0; k < n_src; ++k) {
        const int64_t q = s * n_src + k;
        if (src[q] >= 0) ey[s * numel + src[q]] += f[t * g.n_shots * n_src + q];
      }
      for (int64_t k = 0; k < n_rec; ++k) {
        const int64_t q = s * n_rec + k;
        if (rec[q] >= 0) r[t * g.n_shots * n_rec + q] = ey[s * numel + rec[q]];
      }
    }
  }
}

// ----------------------------------------------------------- Debye forward
// Debye-dispersive forward (forward inference; reference
// csrc/tm2d/maxwell_tm_inst.cpp:214-300 and the time loop :780-805).  The E
// step keeps Ey^n in ey_prev, applies the plain update with the Debye
// coefficients, then adds sum_p cp_p P_p (poles in order, from zero); after
// the sources every cell's polarization advances
//   P_p = a_p P_p + b_p (Ey^{n+1} + Ey^n).
// cp, a, b: [n_poles, ny, nx]; pol: [shots, n_poles, ny, nx].
template <typename T, int S>
void e_step_debye(const Grid<T> &g, const T *ca, const T *cb, const T *hx, const T *hz, T *ey,
                  T *m_hx_z, T *m_hz_x, T *ey_prev, const T *cp, const T *pol,
                  int64_t n_poles, int64_t shot) {
  constexpr int P = Coef<S>::P;
  const int64_t ny = g.ny, nx = g.nx, numel = g.numel(), base = shot * numel;
  for (int64_t y = P; y < ny - P + 1; ++y) {
    for (int64_t x = P; x < nx - P + 1; ++x) {
      const int64_t j = y * nx + x, i = base + j;
      ey_prev[i] = ey[i];
      const T cav = g.ca_b ? ca[i] : ca[j];
      const T cbv = g.cb_b ? cb[i] : cb[j];
      T dhz = d_int<T, S>([&](int o) { return hz[i + o]; }, g.rdx);
      T dhx = d_int<T, S>([&](int o) { return hx[i + o * nx]; }, g.rdy);
      if (x < g.pml_x0 || x >= g.pml_x1) {
        m_hz_x[i] = g.bx[x] * m_hz_x[i] + g.ax[x] * dhz;
        dhz = dhz / g.kx[x] + m_hz_x[i];
      }
      if (y < g.pml_y0 || y >= g.pml_y1) {
        m_hx_z[i] = g.by[y] * m_hx_z[i] + g.ay[y] * dhx;
        dhx = dhx / g.ky[y] + m_hx_z[i];
      }
      const T curl = dhz - dhx;
      ey[i] = cav * ey[i] + cbv * curl;
      T pol_term = (T)0;
      for (int64_t p = 0; p < n_poles; ++p)
        pol_term += cp[p * numel + j] * pol[(shot * n_poles + p) * numel + j];
      ey[i] += pol_term;
    }
  }
}

template <typename T, int S>
void forward_debye(const T *ca, const T *cb, const T *cq, const T *f, T *ey, T *hx, T *hz,
                   T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x, const T *da, const T *db,
                   const T *cp, T *pol, T *ey_prev, int64_t n_poles, T *r, const Grid<T> &g,
                   const int64_t *src, const int64_t *rec, int64_t n_src, int64_t n_rec,
                   int64_t nt, int64_t start_t) {
  const int64_t numel = g.numel();
  for (int64_t t = start_t; t < start_t + nt; ++t) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < g.n_shots; ++s) {
      h_step<T, S>(g, cq, ey, hx, hz, m_ey_x, m_ey_z, s);
      e_step_debye<T, S>(g, ca, cb, hx, hz, ey, m_hx_z, m_hz_x, ey_prev, cp, pol, n_poles, s);
      for (int64_t k = 0; k < n_src; ++k) {
        const int64_t q = s * n_src + k;
        if (src[q] >= 0) ey[s * numel + src[q]] += f[t * g.n_shots * n_src + q];
      }
      for (int64_t p = 0; p < n_poles; ++p)
        for (int64_t j = 0; j < numel; ++j) {
          const int64_t i = s * numel + j, pi = (s * n_poles + p) * numel + j;
          pol[pi] = da[p * numel + j] * pol[pi] + db[p * numel + j] * (ey[i] + ey_prev[i]);
        }
      for (int64_t k = 0; k < n_rec; ++k) {
        const int64_t q = s * n_rec + k;
        if (rec[q] >= 0) r[t * g.n_shots * n_rec + q] = ey[s * numel + rec[q]];
      }
    }
  }
}

// ------------------------------------------------------------ exact adjoint
//
// Forward step n maps (Ey, Hx, Hz, psi) -> (Ey', Hx', Hz', psi').  Its
// transpose is applied E-phase first, then H-phase.  For a CPML memory
// psi' = b psi + a d the adjoint variable M carries the total adjoint of
// psi' and obeys M_n = b M_{n+1} + q_n; the derivative adjoint is then
// a M_n + q_n / k (strip) or q_n (interior).  M_{n+1} is read by the
// neighbours while M_n is written, hence the separate `next` arrays.

// With `dcb` and `l2` (the Born adjoint's background pass) the adjoint of
// curl H is cb lambda + dcb lambda2: the scattered E update adds dcb curl H of
// the background, so the scattered adjoint lambda2 reaches it too.
template <typename T, int S>
void e_adjoint(const Grid<T> &g, const T *cb, const T *l_ey, T *l_hx, T *l_hz,
               const T *m_hx_z, const T *m_hz_x, T *m_hx_z_n, T *m_hz_x_n,
               int64_t shot, const T *dcb = nullptr, const T *l2 = nullptr) {
  constexpr int P = Coef<S>::P;
  const int64_t ny = g.ny, nx = g.nx, base = shot * g.numel();
  auto in_e = [&](int64_t y, int64_t x) {
    return y >= P && y < ny - P + 1 && x >= P && x < nx - P + 1;
  };
  auto q_at = [&](int64_t y, int64_t x) {
    const int64_t j = y * nx + x;
    const T q = (g.cb_b ? cb[base + j] : cb[j]) * l_ey[base + j];
    return l2 ? q + dcb[j] * l2[base + j] : q;
  };
  // raw x-derivative adjoint at (y, x) (FusedRawXAcc).
  auto raw_x = [&](int64_t y, int64_t x) -> T {
    if (!in_e(y, x)) return (T)0;
    const T q = q_at(y, x);
    if (x >= g.pml_x0 && x < g.pml_x1) return q;
    const T mn = g.bx[x] * m_hz_x[base + y * nx + x] + q;
    return g.ax[x] * mn + q / g.kx[x];
  };
  auto raw_y = [&](int64_t y, int64_t x) -> T {
    if (!in_e(y, x)) return (T)0;
    const T q = q_at(y, x);
    if (y >= g.pml_y0 && y < g.pml_y1) return -q;
    const T mn = g.by[y] * m_hx_z[base + y * nx + x] - q;
    return g.ay[y] * mn - q / g.ky[y];
  };
  for (int64_t y = P; y < ny - P + 1; ++y) {
    for (int64_t x = P; x < nx - P + 1; ++x) {
      const int64_t i = base + y * nx + x;
      const T q = q_at(y, x);
      const bool px = x < g.pml_x0 || x >= g.pml_x1;
      const bool py = y < g.pml_y0 || y >= g.pml_y1;
      m_hz_x_n[i] = px ? g.bx[x] * m_hz_x[i] + q : (T)0;
      m_hx_z_n[i] = py ? g.by[y] * m_hx_z[i] - q : (T)0;
      l_hz[i] += d_int_adj<T, S>([&](int o) { return raw_x(y, x + o); }, g.rdx);
      l_hx[i] += d_int_adj<T, S>([&](int o) { return raw_y(y + o, x); }, g.rdy);
    }
  }
}

// With `dca` and `l2`: lambda_Ey = ca lambda_Ey + dca lambda2 (the scattered
// E update adds dca Ey^n of the background) before the derivative terms.
template <typename T, int S>
void h_adjoint(const Grid<T> &g, const T *ca, const T *cq, const T *l_hx,
               const T *l_hz, T *l_ey, const T *m_ey_x, const T *m_ey_z,
               T *m_ey_x_n, T *m_ey_z_n, int64_t shot, const T *dca = nullptr,
               const T *l2 = nullptr) {
  constexpr int P = Coef<S>::P;
  const int64_t ny = g.ny, nx = g.nx, base = shot * g.numel();
  const int64_t y1h = g.pml_y1 > g.pml_y0 ? g.pml_y1 - 1 : g.pml_y0;
  const int64_t x1h = g.pml_x1 > g.pml_x0 ? g.pml_x1 - 1 : g.pml_x0;
  auto cq_at = [&](int64_t j) { return g.cq_b ? cq[base + j] : cq[j]; };
  // FusedRawEyZAcc: Hx domain is y in [P, ny-P), x in [P, nx-P+1).
  auto raw_z = [&](int64_t y, int64_t x) -> T {
    if (y < P || y >= ny - P || x < P || x >= nx - P + 1) return (T)0;
    const int64_t j = y * nx + x;
    const T adj = -cq_at(j) * l_hx[base + j];
    if (y >= g.pml_y0 && y < y1h) return adj;
    const T mn = g.byh[y] * m_ey_z[base + j] + adj;
    return g.ayh[y] * mn + adj / g.kyh[y];
  };
  // FusedRawEyXAcc: Hz domain is y in [P, ny-P+1), x in [P, nx-P).
  auto raw_xh = [&](int64_t y, int64_t x) -> T {
    if (y < P || y >= ny - P + 1 || x < P || x >= nx - P) return (T)0;
    const int64_t j = y * nx + x;
    const T adj = cq_at(j) * l_hz[base + j];
    if (x >= g.pml_x0 && x < x1h) return adj;
    const T mn = g.bxh[x] * m_ey_x[base + j] + adj;
    return g.axh[x] * mn + adj / g.kxh[x];
  };
  for (int64_t y = P; y < ny - P + 1; ++y) {
    for (int64_t x = P; x < nx - P + 1; ++x) {
      const int64_t j = y * nx + x, i = base + j;
      const T cqv = cq_at(j);
      const bool py = y < g.pml_y0 || y >= y1h;
      const bool px = x < g.pml_x0 || x >= x1h;
      const T adj_y = -cqv * l_hx[i];
      const T adj_x = cqv * l_hz[i];
      m_ey_z_n[i] = (y < ny - P && py) ? g.byh[y] * m_ey_z[i] + adj_y : (T)0;
      m_ey_x_n[i] = (x < nx - P && px) ? g.bxh[x] * m_ey_x[i] + adj_x : (T)0;
      const T cav = g.ca_b ? ca[i] : ca[j];
      l_ey[i] = l2 ? cav * l_ey[i] + dca[j] * l2[i] : cav * l_ey[i];
      l_ey[i] += d_half_adj<T, S>([&](int o) { return raw_z(y + o, x); }, g.rdy) +
                 d_half_adj<T, S>([&](int o) { return raw_xh(y, x + o); }, g.rdx);
    }
  }
}

template <typename T, int S>
void backward(const T *ca, const T *cb, const T *cq, const T *grad_r, T *l_ey,
              T *l_hx, T *l_hz, T *ml_ey_x, T *ml_ey_z, T *ml_hx_z,
              T *ml_hz_x, const void *ey_store, const void *curl_store,
              int fmt, int64_t store_bytes_per_step, T *grad_f,
              double *grad_ca, double *grad_cb, const Grid<T> &g,
              const int64_t *src, const int64_t *rec, int64_t n_src,
              int64_t n_rec, int64_t nt, int64_t start_t, int64_t step_ratio,
              bool ca_rg, bool cb_rg) {
  constexpr int P = Coef<S>::P;
  const int64_t numel = g.numel(), total = g.n_shots * numel;
  std::vector<T> nxt(4 * total, (T)0);
  T *cur_ey_x = ml_ey_x, *cur_ey_z = ml_ey_z, *cur_hx_z = ml_hx_z,
    *cur_hz_x = ml_hz_x;
  T *n_ey_x = nxt.data(), *n_ey_z = nxt.data() + total,
    *n_hx_z = nxt.data() + 2 * total, *n_hz_x = nxt.data() + 3 * total;
  // Per-cell, per-shot FP64 accumulators; summed over shots at the end
  // unless the coefficients are per-shot (batched).
  std::vector<double> gca_s(ca_rg ? total : 0, 0.0), gcb_s(cb_rg ? total : 0, 0.0);
  const double scale = (double)step_ratio;
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < g.n_shots; ++s)
      e_adjoint<T, S>(g, cb, l_ey, l_hx, l_hz, cur_hx_z, cur_hz_x, n_hx_z,
                      n_hz_x, s);
    std::swap(cur_hx_z, n_hx_z);
    std::swap(cur_hz_x, n_hz_x);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < g.n_shots; ++s)
      h_adjoint<T, S>(g, ca, cq, l_hx, l_hz, l_ey, cur_ey_x, cur_ey_z, n_ey_x,
                      n_ey_z, s);
    std::swap(cur_ey_x, n_ey_x);
    std::swap(cur_ey_z, n_ey_z);
    for (int64_t s = 0; s < g.n_shots; ++s) {
      for (int64_t k = 0; k < n_rec; ++k) {
        const int64_t q = s * n_rec + k;
        if (rec[q] >= 0) l_ey[s * numel + rec[q]] += grad_r[t * g.n_shots * n_rec + q];
      }
      for (int64_t k = 0; k < n_src; ++k) {
        const int64_t q = s * n_src + k;
        if (grad_f != nullptr && src[q] >= 0)
          grad_f[t * g.n_shots * n_src + q] = l_ey[s * numel + src[q]];
      }
    }
    if ((t % step_ratio) == 0 && (ca_rg || cb_rg)) {
      const int64_t idx = t / step_ratio;
      const char *es = ey_store ? (const char *)ey_store + idx * store_bytes_per_step : nullptr;
      const char *cs = curl_store ? (const char *)curl_store + idx * store_bytes_per_step : nullptr;
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < g.n_shots; ++s) {
        for (int64_t y = P; y < g.ny - P + 1; ++y) {
          for (int64_t x = P; x < g.nx - P + 1; ++x) {
            const int64_t i = s * numel + y * g.nx + x;
            const double lam = (double)l_ey[i];
            if (ca_rg) gca_s[i] += lam * (double)load_store<T>(es, fmt, i) * scale;
            if (cb_rg) gcb_s[i] += lam * (double)load_store<T>(cs, fmt, i) * scale;
          }
        }
      }
    }
  }
  // Leave the final adjoint memories in the caller's arrays.
  if (cur_ey_x != ml_ey_x) std::memcpy(ml_ey_x, cur_ey_x, total * sizeof(T));
  if (cur_ey_z != ml_ey_z) std::memcpy(ml_ey_z, cur_ey_z, total * sizeof(T));
  if (cur_hx_z != ml_hx_z) std::memcpy(ml_hx_z, cur_hx_z, total * sizeof(T));
  if (cur_hz_x != ml_hz_x) std::memcpy(ml_hz_x, cur_hz_x, total * sizeof(T));
  auto reduce = [&](std::vector<double> &acc, double *out, bool batched) {
    if (batched) {
      for (int64_t i = 0; i < total; ++i) out[i] += acc[i];
    } else {
      for (int64_t s = 0; s < g.n_shots; ++s)
        for (int64_t j = 0; j < numel; ++j) out[j] += acc[s * numel + j];
    }
  };
  if (ca_rg) reduce(gca_s, grad_ca, g.ca_b);
  if (cb_rg) reduce(gcb_s, grad_cb, g.cb_b);
}

// --------------------------------------------------- tangent-linear forward
// Runs the background and the coefficient/source tangent together:
//   dEy' = ca dEy + cb dcurl + dca Ey^n + dcb curl^n;  dEy[src] += df.
// dca/dcb are [ny, nx] (unbatched) perturbations.  With the four history
// arrays ([nt, shots, ny, nx] each, null for none) it keeps, per step, the
// background's Ey^n and curl^n and the tangent's dEy^n and dcurl^n (the
// Born adjoint's snapshots).
template <typename T, int S>
void tangent(const T *ca, const T *cb, const T *cq, const T *f, const T *dca,
             const T *dcb, const T *df, T *ey, T *hx, T *hz, T *m_ey_x,
             T *m_ey_z, T *m_hx_z, T *m_hz_x, T *dey, T *dhx, T *dhz,
             T *dm_ey_x, T *dm_ey_z, T *dm_hx_z, T *dm_hz_x, T *r, T *dr,
             const Grid<T> &g, const int64_t *src, const int64_t *rec,
             int64_t n_src, int64_t n_rec, int64_t nt, T *bg_e_hist = nullptr,
             T *bg_c_hist = nullptr, T *sc_e_hist = nullptr, T *sc_c_hist = nullptr) {
  constexpr int P = Coef<S>::P;
  const int64_t numel = g.numel(), ny = g.ny, nx = g.nx, total = g.n_shots * numel;
  std::vector<T> ey_n(total), curl_n(total);
  for (int64_t t = 0; t < nt; ++t) {
    T *be = bg_e_hist ? bg_e_hist + t * total : ey_n.data();
    T *bc = bg_c_hist ? bg_c_hist + t * total : curl_n.data();
    T *se = sc_e_hist ? sc_e_hist + t * total : nullptr;
    T *scc = sc_c_hist ? sc_c_hist + t * total : nullptr;
    for (int64_t s = 0; s < g.n_shots; ++s) {
      const int64_t base = s * numel;
      // Background step, capturing Ey^n and curl^n.
      h_step<T, S>(g, cq, ey, hx, hz, m_ey_x, m_ey_z, s);
      e_step<T, S>(g, ca, cb, hx, hz, ey, m_hx_z, m_hz_x, be, bc, 0, s);
      // Tangent: the same linear step on the d-fields...
      h_step<T, S>(g, cq, dey, dhx, dhz, dm_ey_x, dm_ey_z, s);
      e_step<T, S>(g, ca, cb, dhx, dhz, dey, dm_hx_z, dm_hz_x, se, scc, 0, s);
      // ...plus the coefficient terms on the E-update domain.
      for (int64_t y = P; y < ny - P + 1; ++y)
        for (int64_t x = P; x < nx - P + 1; ++x) {
          const int64_t j = y * nx + x;
          dey[base + j] += dca[j] * be[base + j] + dcb[j] * bc[base + j];
        }
      for (int64_t k = 0; k < n_src; ++k) {
        const int64_t q = s * n_src + k;
        if (src[q] >= 0) {
          ey[base + src[q]] += f[t * g.n_shots * n_src + q];
          dey[base + src[q]] += df[t * g.n_shots * n_src + q];
        }
      }
      for (int64_t k = 0; k < n_rec; ++k) {
        const int64_t q = s * n_rec + k;
        if (rec[q] >= 0) {
          r[t * g.n_shots * n_rec + q] = ey[base + rec[q]];
          dr[t * g.n_shots * n_rec + q] = dey[base + rec[q]];
        }
      }
    }
  }
}

// ------------------------------------------------- adjoint of the tangent
// Transpose of `tangent` with respect to every input it is linear or
// differentiable in: given cotangents w (on dr) and w_bg (on r), returns
// d<dr, w> + <r, w_bg> with respect to ca, cb (FP64, through both the
// background and the tangent recurrences), dca, dcb (FP64), f and df (per
// step).  Per reverse step, the background adjoint eta runs first, coupled to
// the tangent adjoint lambda of the same step (the tangent E update reads the
// background's Ey^n and curl^n), then lambda runs the plain adjoint step;
// receivers, sources and the FP64 terms follow as in `backward`:
//   dJ/dca  += eta Ey^n + lambda dEy^n     dJ/dcb  += eta curl^n + lambda dcurl^n
//   dJ/ddca += lambda Ey^n                 dJ/ddcb += lambda curl^n
// lam / eta: 7 arrays each (Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x), in/out.
template <typename T, int S>
void tangent_adjoint(const T *ca, const T *cb, const T *cq, const T *dca, const T *dcb,
                     const T *w, const T *w_bg, const T *bg_e, const T *bg_c, const T *sc_e,
                     const T *sc_c, T *const *lam, T *const *eta, T *g_df, T *g_f0,
                     double *g_ca, double *g_cb, double *g_dca, double *g_dcb,
                     const Grid<T> &g, const int64_t *src, const int64_t *rec,
                     int64_t n_src, int64_t n_rec, int64_t nt) {
  constexpr int P = Coef<S>::P;
  const int64_t numel = g.numel(), total = g.n_shots * numel;
  std::vector<T> nxt(8 * total, (T)0);
  // memories: [system][ey_x, ey_z, hx_z, hz_x]
  T *cur[2][4], *nx_[2][4];
  for (int k = 0; k < 4; ++k) {
    cur[0][k] = eta[3 + k];
    cur[1][k] = lam[3 + k];
    nx_[0][k] = nxt.data() + (size_t)k * total;
    nx_[1][k] = nxt.data() + (size_t)(4 + k) * total;
  }
  std::vector<double> a_ca(total, 0.0), a_cb(total, 0.0), a_dca(total, 0.0), a_dcb(total, 0.0);
  T *const sys_f[2][3] = {{eta[0], eta[1], eta[2]}, {lam[0], lam[1], lam[2]}};
  for (int64_t t = nt - 1; t >= 0; --t) {
    for (int k = 0; k < 2; ++k) {  // 0: background (coupled), 1: tangent
      const T *l2 = k == 0 ? lam[0] : nullptr;
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < g.n_shots; ++s)
        e_adjoint<T, S>(g, cb, sys_f[k][0], sys_f[k][1], sys_f[k][2], cur[k][2], cur[k][3],
                        nx_[k][2], nx_[k][3], s, dcb, l2);
      std::swap(cur[k][2], nx_[k][2]);
      std::swap(cur[k][3], nx_[k][3]);
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < g.n_shots; ++s)
        h_adjoint<T, S>(g, ca, cq, sys_f[k][1], sys_f[k][2], sys_f[k][0], cur[k][0],
                        cur[k][1], nx_[k][0], nx_[k][1], s, dca, l2);
      std::swap(cur[k][0], nx_[k][0]);
      std::swap(cur[k][1], nx_[k][1]);
    }
    T *l_ey = lam[0], *e_ey = eta[0];
    for (int64_t s = 0; s < g.n_shots; ++s) {
      for (int64_t k = 0; k < n_rec; ++k) {
        const int64_t q = s * n_rec + k;
        if (rec[q] < 0) continue;
        if (w) l_ey[s * numel + rec[q]] += w[t * g.n_shots * n_rec + q];
        if (w_bg) e_ey[s * numel + rec[q]] += w_bg[t * g.n_shots * n_rec + q];
      }
      for (int64_t k = 0; k < n_src; ++k) {
        const int64_t q = s * n_src + k;
        if (src[q] < 0) continue;
        if (g_df) g_df[t * g.n_shots * n_src + q] = l_ey[s * numel + src[q]];
        if (g_f0) g_f0[t * g.n_shots * n_src + q] = e_ey[s * numel + src[q]];
      }
    }
    const int64_t o = t * total;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < g.n_shots; ++s)
      for (int64_t y = P; y < g.ny - P + 1; ++y)
        for (int64_t x = P; x < g.nx - P + 1; ++x) {
          const int64_t i = s * numel + y * g.nx + x;
          const double le = (double)l_ey[i], ee = (double)e_ey[i];
          a_ca[i] += ee * (double)bg_e[o + i];
          a_ca[i] += le * (double)sc_e[o + i];
          a_cb[i] += ee * (double)bg_c[o + i];
          a_cb[i] += le * (double)sc_c[o + i];
          a_dca[i] += le * (double)bg_e[o + i];
          a_dcb[i] += le * (double)bg_c[o + i];
        }
  }
  for (int k = 0; k < 4; ++k) {
    if (cur[0][k] != eta[3 + k]) std::memcpy(eta[3 + k], cur[0][k], total * sizeof(T));
    if (cur[1][k] != lam[3 + k]) std::memcpy(lam[3 + k], cur[1][k], total * sizeof(T));
  }
  for (int64_t s = 0; s < g.n_shots; ++s)
    for (int64_t j = 0; j < numel; ++j) {
      const int64_t i = s * numel + j;
      if (g_ca) g_ca[j] += a_ca[i];
      if (g_cb) g_cb[j] += a_cb[i];
      if (g_dca) g_dca[j] += a_dca[i];
      if (g_dcb) g_dcb[j] += a_dcb[i];
    }
}

template <typename T>
Grid<T> make_grid(const T *const *prof, T rdy, T rdx, int64_t n_shots,
                  int64_t ny, int64_t nx, const int64_t *pml, const bool *batched) {
  Grid<T> g;
  g.ay = prof[0]; g.by = prof[1]; g.ayh = prof[2]; g.byh = prof[3];
  g.ax = prof[4]; g.bx = prof[5]; g.axh = prof[6]; g.bxh = prof[7];
  g.ky = prof[8]; g.kyh = prof[9]; g.kx = prof[10]; g.kxh = prof[11];
  g.rdy = rdy; g.rdx = rdx;
  g.n_shots = n_shots; g.ny = ny; g.nx = nx;
  g.pml_y0 = pml[0]; g.pml_x0 = pml[1]; g.pml_y1 = pml[2]; g.pml_x1 = pml[3];
  g.ca_b = batched[0]; g.cb_b = batched[1]; g.cq_b = batched[2];
  return g;
}

template <typename T>
int dispatch_forward(int stencil, const T *ca, const T *cb, const T *cq,
                     const T *f, T *const *fields, T *r, const T *const *prof,
                     const int64_t *src, const int64_t *rec, T rdy, T rdx,
                     int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
                     int64_t n_src, int64_t n_rec, int64_t step_ratio,
                     int64_t start_t, const int64_t *pml, const bool *batched,
                     void *ey_store, void *curl_store, int fmt,
                     int64_t store_bytes_per_step) {
  Grid<T> g = make_grid<T>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_FWD(S)                                                     \
  forward<T, S>(ca, cb, cq, f, fields[0], fields[1], fields[2], fields[3],     \
                fields[4], fields[5], fields[6], r, g, src, rec, n_src, n_rec, \
                nt, start_t, step_ratio, ey_store, curl_store, fmt,            \
                store_bytes_per_step)
  switch (stencil) {
    case 2: TIDE_ORACLE_FWD(2); return 0;
    case 4: TIDE_ORACLE_FWD(4); return 0;
    case 6: TIDE_ORACLE_FWD(6); return 0;
    case 8: TIDE_ORACLE_FWD(8); return 0;
  }
#undef TIDE_ORACLE_FWD
  return 1;
}

template <typename T>
int dispatch_backward(int stencil, const T *ca, const T *cb, const T *cq,
                      const T *grad_r, T *const *lam, const void *ey_store,
                      const void *curl_store, int fmt,
                      int64_t store_bytes_per_step, T *grad_f, double *grad_ca,
                      double *grad_cb, const T *const *prof,
                      const int64_t *src, const int64_t *rec, T rdy, T rdx,
                      int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
                      int64_t n_src, int64_t n_rec, int64_t step_ratio,
                      int64_t start_t, const int64_t *pml, const bool *batched,
                      bool ca_rg, bool cb_rg) {
  Grid<T> g = make_grid<T>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_BWD(S)                                                     \
  backward<T, S>(ca, cb, cq, grad_r, lam[0], lam[1], lam[2], lam[3], lam[4],   \
                 lam[5], lam[6], ey_store, curl_store, fmt,                    \
                 store_bytes_per_step, grad_f, grad_ca, grad_cb, g, src, rec,  \
                 n_src, n_rec, nt, start_t, step_ratio, ca_rg, cb_rg)
  switch (stencil) {
    case 2: TIDE_ORACLE_BWD(2); return 0;
    case 4: TIDE_ORACLE_BWD(4); return 0;
    case 6: TIDE_ORACLE_BWD(6); return 0;
    case 8: TIDE_ORACLE_BWD(8); return 0;
  }
#undef TIDE_ORACLE_BWD
  return 1;
}

// Debye-dispersive forward (unbatched coefficients): debye = {a, b, cp}
// ([n_poles, ny, nx] each), pol [shots, n_poles, ny, nx] and ey_prev
// [shots, ny, nx] in/out.
template <typename T>
int dispatch_debye(int stencil, const T *ca, const T *cb, const T *cq, const T *f,
                   T *const *fields, const T *const *debye, T *pol, T *ey_prev, int64_t n_poles,
                   T *r, const T *const *prof, const int64_t *src, const int64_t *rec, T rdy,
                   T rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
                   int64_t n_rec, int64_t start_t, const int64_t *pml) {
  const bool batched[3] = {false, false, false};
  Grid<T> g = make_grid<T>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_DEB(S)                                                                \
  forward_debye<T, S>(ca, cb, cq, f, fields[0], fields[1], fields[2], fields[3], fields[4], \
                      fields[5], fields[6], debye[0], debye[1], debye[2], pol, ey_prev,    \
                      n_poles, r, g, src, rec, n_src, n_rec, nt, start_t)
  switch (stencil) {
    case 2: TIDE_ORACLE_DEB(2); return 0;
    case 4: TIDE_ORACLE_DEB(4); return 0;
    case 6: TIDE_ORACLE_DEB(6); return 0;
    case 8: TIDE_ORACLE_DEB(8); return 0;
  }
#undef TIDE_ORACLE_DEB
  return 1;
}

}  // namespace

// ------------------------------------------------------------------- C ABI
// fields: Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x (in/out, [shots,ny,nx])
// prof:   ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
// pml:    pml_y0, pml_x0, pml_y1, pml_x1
// batched: ca, cb, cq per-shot flags
// lam:    lambda Ey, Hx, Hz, m_lambda_Ey_x, m_lambda_Ey_z, m_lambda_Hx_z,
//         m_lambda_Hz_x (in/out)
extern "C" {

int oracle_tm2d_forward_f32(int stencil, const float *ca, const float *cb,
                            const float *cq, const float *f,
                            float *const *fields, float *r,
                            const float *const *prof, const int64_t *src,
                            const int64_t *rec, float rdy, float rdx,
                            int64_t nt, int64_t n_shots, int64_t ny,
                            int64_t nx, int64_t n_src, int64_t n_rec,
                            int64_t step_ratio, int64_t start_t,
                            const int64_t *pml, const bool *batched,
                            void *ey_store, void *curl_store, int fmt,
                            int64_t store_bytes_per_step) {
  return dispatch_forward<float>(stencil, ca, cb, cq, f, fields, r, prof, src,
                                 rec, rdy, rdx, nt, n_shots, ny, nx, n_src,
                                 n_rec, step_ratio, start_t, pml, batched,
                                 ey_store, curl_store, fmt,
                                 store_bytes_per_step);
}

int oracle_tm2d_forward_f64(int stencil, const double *ca, const double *cb,
                            const double *cq, const double *f,
                            double *const *fields, double *r,
                            const double *const *prof, const int64_t *src,
                            const int64_t *rec, double rdy, double rdx,
                            int64_t nt, int64_t n_shots, int64_t ny,
                            int64_t nx, int64_t n_src, int64_t n_rec,
                            int64_t step_ratio, int64_t start_t,
                            const int64_t *pml, const bool *batched,
                            void *ey_store, void *curl_store, int fmt,
                            int64_t store_bytes_per_step) {
  if (fmt == 1) return 2;  // BF16 payloads only for fp32 fields
  return dispatch_forward<double>(stencil, ca, cb, cq, f, fields, r, prof, src,
                                  rec, rdy, rdx, nt, n_shots, ny, nx, n_src,
                                  n_rec, step_ratio, start_t, pml, batched,
                                  ey_store, curl_store, fmt,
                                  store_bytes_per_step);
}

int oracle_tm2d_backward_f32(int stencil, const float *ca, const float *cb,
                             const float *cq, const float *grad_r,
                             float *const *lam, const void *ey_store,
                             const void *curl_store, int fmt,
                             int64_t store_bytes_per_step, float *grad_f,
                             double *grad_ca, double *grad_cb,
                             const float *const *prof, const int64_t *src,
                             const int64_t *rec, float rdy, float rdx,
                             int64_t nt, int64_t n_shots, int64_t ny,
                             int64_t nx, int64_t n_src, int64_t n_rec,
                             int64_t step_ratio, int64_t start_t,
                             const int64_t *pml, const bool *batched,
                             bool ca_rg, bool cb_rg) {
  return dispatch_backward<float>(stencil, ca, cb, cq, grad_r, lam, ey_store,
                                  curl_store, fmt, store_bytes_per_step,
                                  grad_f, grad_ca, grad_cb, prof, src, rec,
                                  rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,
                                  step_ratio, start_t, pml, batched, ca_rg,
                                  cb_rg);
}

int oracle_tm2d_backward_f64(int stencil, const double *ca, const double *cb,
                             const double *cq, const double *grad_r,
                             double *const *lam, const void *ey_store,
                             const void *curl_store, int fmt,
                             int64_t store_bytes_per_step, double *grad_f,
                             double *grad_ca, double *grad_cb,
                             const double *const *prof, const int64_t *src,
                             const int64_t *rec, double rdy, double rdx,
                             int64_t nt, int64_t n_shots, int64_t ny,
                             int64_t nx, int64_t n_src, int64_t n_rec,
                             int64_t step_ratio, int64_t start_t,
                             const int64_t *pml, const bool *batched,
                             bool ca_rg, bool cb_rg) {
  if (fmt == 1) return 2;
  return dispatch_backward<double>(stencil, ca, cb, cq, grad_r, lam, ey_store,
                                   curl_store, fmt, store_bytes_per_step,
                                   grad_f, grad_ca, grad_cb, prof, src, rec,
                                   rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,
                                   step_ratio, start_t, pml, batched, ca_rg,
                                   cb_rg);
}

int oracle_tm2d_forward_debye_f32(int stencil, const float *ca, const float *cb,
                                  const float *cq, const float *f, float *const *fields,
                                  const float *const *debye, float *pol, float *ey_prev,
                                  int64_t n_poles, float *r, const float *const *prof,
                                  const int64_t *src, const int64_t *rec, float rdy, float rdx,
                                  int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
                                  int64_t n_src, int64_t n_rec, int64_t start_t,
                                  const int64_t *pml) {
  return dispatch_debye<float>(stencil, ca, cb, cq, f, fields, debye, pol, ey_prev, n_poles, r,
                               prof, src, rec, rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,
                               start_t, pml);
}

int oracle_tm2d_forward_debye_f64(int stencil, const double *ca, const double *cb,
                                  const double *cq, const double *f, double *const *fields,
                                  const double *const *debye, double *pol, double *ey_prev,
                                  int64_t n_poles, double *r, const double *const *prof,
                                  const int64_t *src, const int64_t *rec, double rdy,
                                  double rdx, int64_t nt, int64_t n_shots, int64_t ny,
                                  int64_t nx, int64_t n_src, int64_t n_rec, int64_t start_t,
                                  const int64_t *pml) {
  return dispatch_debye<double>(stencil, ca, cb, cq, f, fields, debye, pol, ey_prev, n_poles,
                                r, prof, src, rec, rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,
                                start_t, pml);
}

// Tangent-linear forward in FP64 (unbatched coefficients).  `fields` and
// `dfields` are the 7 background and 7 tangent state arrays (in/out).
int oracle_tm2d_tangent_f64(int stencil, const double *ca, const double *cb,
                            const double *cq, const double *f,
                            const double *dca, const double *dcb,
                            const double *df, double *const *fields,
                            double *const *dfields, double *r, double *dr,
                            const double *const *prof, const int64_t *src,
                            const int64_t *rec, double rdy, double rdx,
                            int64_t nt, int64_t n_shots, int64_t ny,
                            int64_t nx, int64_t n_src, int64_t n_rec,
                            const int64_t *pml) {
  const bool batched[3] = {false, false, false};
  Grid<double> g = make_grid<double>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_TAN(S)                                                      \
  tangent<double, S>(ca, cb, cq, f, dca, dcb, df, fields[0], fields[1],         \
                     fields[2], fields[3], fields[4], fields[5], fields[6],     \
                     dfields[0], dfields[1], dfields[2], dfields[3],            \
                     dfields[4], dfields[5], dfields[6], r, dr, g, src, rec,    \
                     n_src, n_rec, nt)
  switch (stencil) {
    case 2: TIDE_ORACLE_TAN(2); return 0;
    case 4: TIDE_ORACLE_TAN(4); return 0;
    case 6: TIDE_ORACLE_TAN(6); return 0;
    case 8: TIDE_ORACLE_TAN(8); return 0;
  }
#undef TIDE_ORACLE_TAN
  return 1;
}

// Tangent-linear forward that also keeps the four snapshot histories of the
// tangent adjoint (bg_e, bg_c, sc_e, sc_c: [nt, shots, ny, nx] each).
int oracle_tm2d_tangent_store_f64(int stencil, const double *ca, const double *cb,
                                  const double *cq, const double *f, const double *dca,
                                  const double *dcb, const double *df,
                                  double *const *fields, double *const *dfields, double *r,
                                  double *dr, const double *const *prof, const int64_t *src,
                                  const int64_t *rec, double rdy, double rdx, int64_t nt,
                                  int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
                                  int64_t n_rec, const int64_t *pml, double *const *hist) {
  const bool batched[3] = {false, false, false};
  Grid<double> g = make_grid<double>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_TANS(S)                                                     \
  tangent<double, S>(ca, cb, cq, f, dca, dcb, df, fields[0], fields[1],         \
                     fields[2], fields[3], fields[4], fields[5], fields[6],     \
                     dfields[0], dfields[1], dfields[2], dfields[3],            \
                     dfields[4], dfields[5], dfields[6], r, dr, g, src, rec,    \
                     n_src, n_rec, nt, hist[0], hist[1], hist[2], hist[3])
  switch (stencil) {
    case 2: TIDE_ORACLE_TANS(2); return 0;
    case 4: TIDE_ORACLE_TANS(4); return 0;
    case 6: TIDE_ORACLE_TANS(6); return 0;
    case 8: TIDE_ORACLE_TANS(8); return 0;
  }
#undef TIDE_ORACLE_TANS
  return 1;
}

// Adjoint of the tangent-linear forward (FP64, unbatched); see
// tangent_adjoint.  hist as written by oracle_tm2d_tangent_store_f64; grads:
// g_df, g_f0 ([nt, shots, n_src]), g_ca, g_cb, g_dca, g_dcb ([ny, nx] FP64,
// added to); any may be null.
int oracle_tm2d_tangent_adjoint_f64(int stencil, const double *ca, const double *cb,
                                    const double *cq, const double *dca, const double *dcb,
                                    const double *w, const double *w_bg,
                                    const double *const *hist, double *const *lam,
                                    double *const *eta, double *g_df, double *g_f0,
                                    double *g_ca, double *g_cb, double *g_dca, double *g_dcb,
                                    const double *const *prof, const int64_t *src,
                                    const int64_t *rec, double rdy, double rdx, int64_t nt,
                                    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
                                    int64_t n_rec, const int64_t *pml) {
  const bool batched[3] = {false, false, false};
  Grid<double> g = make_grid<double>(prof, rdy, rdx, n_shots, ny, nx, pml, batched);
#define TIDE_ORACLE_TADJ(S)                                                           \
  tangent_adjoint<double, S>(ca, cb, cq, dca, dcb, w, w_bg, hist[0], hist[1], hist[2], \
                             hist[3], lam, eta, g_df, g_f0, g_ca, g_cb, g_dca, g_dcb, g, \
                             src, rec, n_src, n_rec, nt)
  switch (stencil) {
    case 2: TIDE_ORACLE_TADJ(2); return 0;
    case 4: TIDE_ORACLE_TADJ(4); return 0;
    case 6: TIDE_ORACLE_TADJ(6); return 0;
    case 8: TIDE_ORACLE_TADJ(8); return 0;
  }
#undef TIDE_ORACLE_TADJ
  return 1;
}

uint16_t oracle_f32_to_bf16(float v) { return f32_to_bf16(v); }
float oracle_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }

}  // extern "C"
