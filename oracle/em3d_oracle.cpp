// TEST INFRASTRUCTURE ONLY -- never linked into, loaded by, or called from
// the product path.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use this file's library, and only as the checker.
//
// CPU restatement of TIDE's 3-D Maxwell staggered-grid propagator, stencil
// order 2 (FD_PAD = 1):
//   * forward step: H update from half-grid E differences (six CPML
//     memories on the half profiles), then E update from integer-grid H
//     differences (six CPML memories), optionally storing the pre-update
//     (Ex, Ey, Ez) and the CPML curl of H for the gradient (T or BF16
//     payload); point sources add into, and receivers read, one chosen E
//     component;
//   * the exact discrete adjoint of that step in gather form (E phase then
//     H phase, ping-pong CPML adjoint memories, index-shifted reverse loop),
//     accumulating dJ/dca and dJ/dcb in FP64;
//   * the tangent-linear forward in (ca, cb, f), used only to pin the
//     adjoint with the dot-product identity <J dm, w> = <dm, J^T w>;
//   * the Born forward (background and scattered field in one time loop),
//     in the reference's operation order: csrc/em3d/maxwell_3d_inst.cpp:
//     1100-1527 (born_forward), pinned bit for bit to the reference's own
//     CPU build.
//
// Followed reference code (paths relative to the reference's src/tide/):
//   forward loops   csrc/em3d/maxwell_3d_inst.cpp:446-761 (forward),
//                   :763-1098 (forward_with_storage: store before update)
//   point ops       csrc/em3d/maxwell_3d_inst.cpp:300-348
//   coefficient
//   gradient        csrc/em3d/maxwell_3d_inst.cpp:221-298 (lambda . stored,
//                   scaled by step_ratio, E-update domain)
//   differences     csrc/common/staggered_grid.h:24-29 (DIFFZ1 .. DIFFXH1)
// The adjoint is the transpose of the forward recurrence above, derived the
// same way as the 2-D one (the reference's CPU 3-D backward re-runs the
// forward kernels and is not an exact transpose with CPML or heterogeneous
// coefficients).  Build with -ffp-contract=off.

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint16_t bf16_enc(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float bf16_dec(uint16_t h) {
  const uint32_t u = ((uint32_t)h) << 16;
  float v;
  std::memcpy(&v, &u, 4);
  return v;
}

// Field slots: E 0..2, H 3..5, then the CPML memories in the reference's
// argument order.
enum Slot {
  EX = 0, EY, EZ, HX, HY, HZ,
  M_HZ_Y, M_HY_Z, M_HX_Z, M_HZ_X, M_HY_X, M_HX_Y,  // E phase (H derivatives)
  M_EY_Z, M_EZ_Y, M_EZ_X, M_EX_Z, M_EX_Y, M_EY_X,  // H phase (E derivatives)
  N_SLOTS
};

// A derivative term: field `f` differenced along axis `ax` (0 z, 1 y, 2 x),
// CPML memory `m`, sign `sg` inside the bracket.
struct Term {
  int f, ax, m, sg;
};
// E update: Ex += cb (dHy/dz - dHz/dy), Ey: (dHz/dx - dHx/dz),
//           Ez: (dHx/dy - dHy/dx)       (integer-grid differences)
constexpr Term kE[3][2] = {
    {{HY, 0, M_HY_Z, +1}, {HZ, 1, M_HZ_Y, -1}},
    {{HZ, 2, M_HZ_X, +1}, {HX, 0, M_HX_Z, -1}},
    {{HX, 1, M_HX_Y, +1}, {HY, 2, M_HY_X, -1}},
};
// H update: Hx -= cq (dEy/dz - dEz/dy), Hy: (dEz/dx - dEx/dz),
//           Hz: (dEx/dy - dEy/dx)       (half-grid differences)
constexpr Term kH[3][2] = {
    {{EY, 0, M_EY_Z, +1}, {EZ, 1, M_EZ_Y, -1}},
    {{EZ, 2, M_EZ_X, +1}, {EX, 0, M_EX_Z, -1}},
    {{EX, 1, M_EX_Y, +1}, {EY, 2, M_EY_X, -1}},
};

template <typename T> struct Grid3 {
  // per axis (0 z, 1 y, 2 x): integer-point a, b, k and half-point ah, bh, kh
  const T *a[3], *b[3], *k[3], *ah[3], *bh[3], *kh[3];
  T rd[3];
  int64_t n_shots, n[3];
  int64_t p0[3], p1[3], p1h[3];
  bool ca_b, cb_b, cq_b;
  int64_t numel() const { return n[0] * n[1] * n[2]; }
  int64_t stride(int ax) const { return ax == 0 ? n[1] * n[2] : (ax == 1 ? n[2] : 1); }
  bool pml(int ax, int64_t c) const { return c < p0[ax] || c >= p1[ax]; }
  bool pmlh(int ax, int64_t c) const { return c < p0[ax] || c >= p1h[ax]; }
  bool dom(int64_t z, int64_t y, int64_t x) const {
    return z >= 1 && z < n[0] && y >= 1 && y < n[1] && x >= 1 && x < n[2];
  }
};

template <typename T>
inline T coef(const T *c, bool batched, int64_t i, int64_t j) {
  return batched ? c[i] : c[j];
}

// ------------------------------------------------------------- forward step

template <typename T>
void h_step(const Grid3<T> &g, const T *cq, T *const *fd, int64_t s) {
  const int64_t base = s * g.numel();
#pragma omp parallel for schedule(static)
  for (int64_t z = 1; z < g.n[0]; ++z)
    for (int64_t y = 1; y < g.n[1]; ++y)
      for (int64_t x = 1; x < g.n[2]; ++x) {
        const int64_t c[3] = {z, y, x};
        const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
        const T cqv = coef(cq, g.cq_b, i, j);
        T d[3][2];
        for (int h = 0; h < 3; ++h)
          for (int t = 0; t < 2; ++t) {
            const Term &tm = kH[h][t];
            const int ax = tm.ax;
            const int64_t p = c[ax];
            d[h][t] = (T)0;
            if (p < g.n[ax] - 1) {
              const T *e = fd[tm.f];
              T v = (e[i + g.stride(ax)] - e[i]) * g.rd[ax];
              if (g.pmlh(ax, p)) {
                T *m = fd[tm.m];
                m[i] = g.bh[ax][p] * m[i] + g.ah[ax][p] * v;
                v = v / g.kh[ax][p] + m[i];
              }
              d[h][t] = v;
            }
          }
        for (int h = 0; h < 3; ++h) fd[HX + h][i] -= cqv * (d[h][0] - d[h][1]);
      }
}

// Stores (when non-null): st[0..2] pre-update Ex/Ey/Ez, st[3..5] curl x/y/z.
// With `bg` (the Born scattered field): bg[0..2] the background's
// pre-update E, bg[3..5] its curl H of this step, indexed like the fields,
// and the update E = ca E + cb curl + dca E_bg + dcb curl_bg in that order
// (reference maxwell_3d_inst.cpp:1444-1456, dca/dcb [nz, ny, nx]).
template <typename T>
void e_step(const Grid3<T> &g, const T *ca, const T *cb, T *const *fd,
            void *const *st, int fmt, int64_t s, const T *dca = nullptr,
            const T *dcb = nullptr, const T *const *bg = nullptr) {
  const int64_t base = s * g.numel();
#pragma omp parallel for schedule(static)
  for (int64_t z = 1; z < g.n[0]; ++z)
    for (int64_t y = 1; y < g.n[1]; ++y)
      for (int64_t x = 1; x < g.n[2]; ++x) {
        const int64_t c[3] = {z, y, x};
        const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
        const T cav = coef(ca, g.ca_b, i, j), cbv = coef(cb, g.cb_b, i, j);
        T curl[3];
        for (int e = 0; e < 3; ++e) {
          T d[2];
          for (int t = 0; t < 2; ++t) {
            const Term &tm = kE[e][t];
            const int ax = tm.ax;
            const int64_t p = c[ax];
            const T *h = fd[tm.f];
            T v = (h[i] - h[i - g.stride(ax)]) * g.rd[ax];
            if (g.pml(ax, p)) {
              T *m = fd[tm.m];
              m[i] = g.b[ax][p] * m[i] + g.a[ax][p] * v;
              v = v / g.k[ax][p] + m[i];
            }
            d[t] = v;
          }
          curl[e] = d[0] - d[1];
        }
        for (int e = 0; e < 3; ++e) {
          if (st && st[e]) {
            if (fmt == 1) ((uint16_t *)st[e])[i] = bf16_enc((float)fd[EX + e][i]);
            else ((T *)st[e])[i] = fd[EX + e][i];
          }
          if (st && st[3 + e]) {
            if (fmt == 1) ((uint16_t *)st[3 + e])[i] = bf16_enc((float)curl[e]);
            else ((T *)st[3 + e])[i] = curl[e];
          }
        }
        if (bg) {
          for (int e = 0; e < 3; ++e)
            fd[EX + e][i] = cav * fd[EX + e][i] + cbv * curl[e] + dca[j] * bg[e][i] +
                            dcb[j] * bg[3 + e][i];
        } else {
          for (int e = 0; e < 3; ++e) fd[EX + e][i] = cav * fd[EX + e][i] + cbv * curl[e];
        }
      }
}

template <typename T>
inline double load_store(const void *p, int fmt, int64_t i) {
  if (fmt == 1) return (double)bf16_dec(((const uint16_t *)p)[i]);
  return (double)((const T *)p)[i];
}

template <typename T>
void forward(const Grid3<T> &g, const T *ca, const T *cb, const T *cq, const T *f,
             T *const *fd, T *r, const int64_t *src, const int64_t *rec, int64_t n_src,
             int64_t n_rec, int src_comp, int rec_comp, int64_t nt, int64_t start_t,
             int64_t step_ratio, void *const *stores, int fmt, int64_t store_step_bytes) {
  const int64_t numel = g.numel();
  bool any_store = false;
  for (int k = 0; k < 6; ++k) any_store = any_store || (stores && stores[k]);
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    void *st[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (any_store && (t % step_ratio) == 0)
      for (int k = 0; k < 6; ++k)
        st[k] = stores[k] ? (char *)stores[k] + (t / step_ratio) * store_step_bytes : nullptr;
    for (int64_t s = 0; s < g.n_shots; ++s) {
      h_step<T>(g, cq, fd, s);
      e_step<T>(g, ca, cb, fd, st, fmt, s);
    }
    T *es = fd[EX + src_comp], *er = fd[EX + rec_comp];
    for (int64_t s = 0; s < g.n_shots; ++s) {
      for (int64_t q = s * n_src; q < (s + 1) * n_src; ++q)
        if (src[q] >= 0) es[s * numel + src[q]] += f[t * g.n_shots * n_src + q];
      for (int64_t q = s * n_rec; q < (s + 1) * n_rec; ++q)
        r[t * g.n_shots * n_rec + q] = rec[q] >= 0 ? er[s * numel + rec[q]] : (T)0;
    }
  }
}

// ------------------------------------------------------------ exact adjoint
//
// For a CPML memory psi' = b psi + a d the adjoint variable M obeys
// M_n = b M_{n+1} + q_n (q = adjoint of the CPML-corrected derivative); the
// adjoint of the raw derivative is a M_n + q_n / k inside a strip and q_n
// outside.  Neighbours read M_{n+1} while M_n is written, so the memories
// ping-pong between `lm` (read) and `nm` (written).

template <typename T>
void e_adjoint(const Grid3<T> &g, const T *cb, T *const *lm, T *const *nm, int64_t s) {
  const int64_t base = s * g.numel();
  // adjoint of the CPML-corrected derivative of term (e, t) at cell i
  auto q_at = [&](int e, int t, int64_t i, int64_t j) -> T {
    const T q = coef(cb, g.cb_b, i, j) * lm[EX + e][i];
    return kE[e][t].sg > 0 ? q : -q;
  };
  auto raw = [&](int e, int t, int64_t z, int64_t y, int64_t x) -> T {
    if (!g.dom(z, y, x)) return (T)0;
    const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
    const int64_t c[3] = {z, y, x};
    const Term &tm = kE[e][t];
    const int64_t p = c[tm.ax];
    const T q = q_at(e, t, i, j);
    if (!g.pml(tm.ax, p)) return q;
    const T mn = g.b[tm.ax][p] * lm[tm.m][i] + q;
    return g.a[tm.ax][p] * mn + q / g.k[tm.ax][p];
  };
  // (e, t) pairs feeding each H component, in accumulation order.
  // Hx: (Ey, dHx/dz), (Ez, dHx/dy); Hy: (Ex, dHy/dz), (Ez, dHy/dx);
  // Hz: (Ex, dHz/dy), (Ey, dHz/dx)
  constexpr int kFeed[3][2][2] = {{{1, 1}, {2, 0}}, {{0, 0}, {2, 1}}, {{0, 1}, {1, 0}}};
#pragma omp parallel for schedule(static)
  for (int64_t z = 1; z < g.n[0]; ++z)
    for (int64_t y = 1; y < g.n[1]; ++y)
      for (int64_t x = 1; x < g.n[2]; ++x) {
        const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
        const int64_t c[3] = {z, y, x};
        for (int e = 0; e < 3; ++e)
          for (int t = 0; t < 2; ++t) {
            const Term &tm = kE[e][t];
            const int64_t p = c[tm.ax];
            nm[tm.m][i] = g.pml(tm.ax, p) ? g.b[tm.ax][p] * lm[tm.m][i] + q_at(e, t, i, j) : (T)0;
          }
        for (int h = 0; h < 3; ++h) {
          T acc = lm[HX + h][i];
          for (int u = 0; u < 2; ++u) {
            const int e = kFeed[h][u][0], t = kFeed[h][u][1];
            const int ax = kE[e][t].ax;
            const int64_t dz = ax == 0, dy = ax == 1, dx = ax == 2;
            acc = acc + (raw(e, t, z, y, x) - raw(e, t, z + dz, y + dy, x + dx)) * g.rd[ax];
          }
          nm[HX + h][i] = acc;
        }
      }
}

template <typename T>
void h_adjoint(const Grid3<T> &g, const T *ca, const T *cq, T *const *lm, T *const *nm,
               int64_t s) {
  const int64_t base = s * g.numel();
  // adjoint of the CPML-corrected E derivative of H-term (h, t) at cell i:
  // H -= cq (first - second)  =>  first: -cq lambda_H, second: +cq lambda_H
  auto adj_at = [&](int h, int t, int64_t i, int64_t j) -> T {
    const T cqv = coef(cq, g.cq_b, i, j);
    return kH[h][t].sg > 0 ? -cqv * lm[HX + h][i] : cqv * lm[HX + h][i];
  };
  auto raw = [&](int h, int t, int64_t z, int64_t y, int64_t x) -> T {
    if (!g.dom(z, y, x)) return (T)0;
    const int64_t c[3] = {z, y, x};
    const Term &tm = kH[h][t];
    const int64_t p = c[tm.ax];
    if (p >= g.n[tm.ax] - 1) return (T)0;
    const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
    const T adj = adj_at(h, t, i, j);
    if (!g.pmlh(tm.ax, p)) return adj;
    const T mn = g.bh[tm.ax][p] * lm[tm.m][i] + adj;
    return g.ah[tm.ax][p] * mn + adj / g.kh[tm.ax][p];
  };
  // (h, t) pairs feeding each E component, in accumulation order.
  // Ex: (Hy, dEx/dz), (Hz, dEx/dy); Ey: (Hx, dEy/dz), (Hz, dEy/dx);
  // Ez: (Hx, dEz/dy), (Hy, dEz/dx)
  constexpr int kFeed[3][2][2] = {{{1, 1}, {2, 0}}, {{0, 0}, {2, 1}}, {{0, 1}, {1, 0}}};
#pragma omp parallel for schedule(static)
  for (int64_t z = 1; z < g.n[0]; ++z)
    for (int64_t y = 1; y < g.n[1]; ++y)
      for (int64_t x = 1; x < g.n[2]; ++x) {
        const int64_t j = (z * g.n[1] + y) * g.n[2] + x, i = base + j;
        const int64_t c[3] = {z, y, x};
        for (int h = 0; h < 3; ++h)
          for (int t = 0; t < 2; ++t) {
            const Term &tm = kH[h][t];
            const int64_t p = c[tm.ax];
            nm[tm.m][i] = (p < g.n[tm.ax] - 1 && g.pmlh(tm.ax, p))
                              ? g.bh[tm.ax][p] * lm[tm.m][i] + adj_at(h, t, i, j)
                              : (T)0;
          }
        const T cav = coef(ca, g.ca_b, i, j);
        for (int e = 0; e < 3; ++e) {
          T sum = (T)0;
          for (int u = 0; u < 2; ++u) {
            const int h = kFeed[e][u][0], t = kFeed[e][u][1];
            const int ax = kH[h][t].ax;
            const int64_t dz = ax == 0, dy = ax == 1, dx = ax == 2;
            const T term = (raw(h, t, z - dz, y - dy, x - dx) - raw(h, t, z, y, x)) * g.rd[ax];
            sum = u == 0 ? term : sum + term;
          }
          nm[EX + e][i] = cav * lm[EX + e][i] + sum;
        }
      }
}

// `lam` holds the caller's 18 adjoint arrays (same slots as the forward
// state); it is advanced in place.  Coefficient gradients are FP64
// ([nz,ny,nx] summed over shots, or per shot when batched), added to.
template <typename T>
void backward(const Grid3<T> &g, const T *ca, const T *cb, const T *cq, const T *grad_r,
              T *const *lam, const void *const *stores, int fmt, int64_t store_step_bytes,
              T *grad_f, double *gca, double *gcb, const int64_t *src, const int64_t *rec,
              int64_t n_src, int64_t n_rec, int src_comp, int rec_comp, int64_t nt,
              int64_t start_t, int64_t step_ratio) {
  const int64_t numel = g.numel(), total = g.n_shots * numel;
  std::vector<T> buf((size_t)N_SLOTS * total);
  T *cur[N_SLOTS], *nxt[N_SLOTS];
  for (int k = 0; k < N_SLOTS; ++k) {
    cur[k] = lam[k];
    nxt[k] = buf.data() + (size_t)k * total;
  }
  std::vector<double> ga(gca ? total : 0, 0.0), gb(gcb ? total : 0, 0.0);
  const double scale = (double)step_ratio;
  for (int64_t t = start_t - 1; t >= start_t - nt; --t) {
    // E phase writes lambda_H and the E-phase memories of the update domain
    // into `nxt`; H phase then writes lambda_E and the H-phase memories.
    // Everything else is carried over.
    for (int k = 0; k < N_SLOTS; ++k) std::memcpy(nxt[k], cur[k], total * sizeof(T));
    for (int64_t s = 0; s < g.n_shots; ++s) e_adjoint<T>(g, cb, cur, nxt, s);
    for (int k = 0; k < N_SLOTS; ++k) std::swap(cur[k], nxt[k]);
    for (int k = 0; k < N_SLOTS; ++k) std::memcpy(nxt[k], cur[k], total * sizeof(T));
    for (int64_t s = 0; s < g.n_shots; ++s) h_adjoint<T>(g, ca, cq, cur, nxt, s);
    for (int k = 0; k < N_SLOTS; ++k) std::swap(cur[k], nxt[k]);
    T *lr = cur[EX + rec_comp], *ls = cur[EX + src_comp];
    for (int64_t s = 0; s < g.n_shots; ++s) {
      for (int64_t q = s * n_rec; q < (s + 1) * n_rec; ++q)
        if (rec[q] >= 0) lr[s * numel + rec[q]] += grad_r[t * g.n_shots * n_rec + q];
      for (int64_t q = s * n_src; q < (s + 1) * n_src; ++q)
        if (grad_f && src[q] >= 0) grad_f[t * g.n_shots * n_src + q] = ls[s * numel + src[q]];
    }
    if ((gca || gcb) && (t % step_ratio) == 0) {
      const int64_t off = (t / step_ratio) * store_step_bytes;
      const char *st[6];
      for (int k = 0; k < 6; ++k) st[k] = stores[k] ? (const char *)stores[k] + off : nullptr;
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < g.n_shots; ++s)
        for (int64_t z = 1; z < g.n[0]; ++z)
          for (int64_t y = 1; y < g.n[1]; ++y)
            for (int64_t x = 1; x < g.n[2]; ++x) {
              const int64_t i = s * numel + (z * g.n[1] + y) * g.n[2] + x;
              const double lx = cur[EX][i], ly = cur[EY][i], lz = cur[EZ][i];
              if (gca)
                ga[i] += (lx * load_store<T>(st[0], fmt, i) + ly * load_store<T>(st[1], fmt, i) +
                          lz * load_store<T>(st[2], fmt, i)) * scale;
              if (gcb)
                gb[i] += (lx * load_store<T>(st[3], fmt, i) + ly * load_store<T>(st[4], fmt, i) +
                          lz * load_store<T>(st[5], fmt, i)) * scale;
            }
    }
  }
  for (int k = 0; k < N_SLOTS; ++k)
    if (cur[k] != lam[k]) std::memcpy(lam[k], cur[k], total * sizeof(T));
  auto reduce = [&](const std::vector<double> &acc, double *out, bool batched) {
    if (batched) {
      for (int64_t i = 0; i < total; ++i) out[i] += acc[i];
    } else {
      for (int64_t s = 0; s < g.n_shots; ++s)
        for (int64_t j = 0; j < numel; ++j) out[j] += acc[s * numel + j];
    }
  };
  if (gca) reduce(ga, gca, g.ca_b);
  if (gcb) reduce(gb, gcb, g.cb_b);
}

// -------------------------------------------------- tangent-linear forward
// The step is linear in the state for fixed coefficients, so the tangent
// state follows the same step, plus dca E^n + dcb curl^n on the E-update
// domain and df at the sources.  dca/dcb are [nz, ny, nx].
template <typename T>
void tangent(const Grid3<T> &g, const T *ca, const T *cb, const T *cq, const T *f,
             const T *dca, const T *dcb, const T *df, T *const *fd, T *const *dfd, T *r,
             T *dr, const int64_t *src, const int64_t *rec, int64_t n_src, int64_t n_rec,
             int src_comp, int rec_comp, int64_t nt) {
  const int64_t numel = g.numel();
  std::vector<T> snap((size_t)6 * numel);
  for (int64_t t = 0; t < nt; ++t) {
    for (int64_t s = 0; s < g.n_shots; ++s) {
      const int64_t base = s * numel;
      void *st[6];
      for (int k = 0; k < 6; ++k) st[k] = snap.data() + (size_t)k * numel - base;
      h_step<T>(g, cq, fd, s);
      e_step<T>(g, ca, cb, fd, st, 0, s);
      h_step<T>(g, cq, dfd, s);
      e_step<T>(g, ca, cb, dfd, nullptr, 0, s);
      for (int64_t z = 1; z < g.n[0]; ++z)
        for (int64_t y = 1; y < g.n[1]; ++y)
          for (int64_t x = 1; x < g.n[2]; ++x) {
            const int64_t j = (z * g.n[1] + y) * g.n[2] + x;
            for (int e = 0; e < 3; ++e)
              dfd[EX + e][base + j] +=
                  dca[j] * snap[(size_t)e * numel + j] + dcb[j] * snap[(size_t)(3 + e) * numel + j];
          }
      for (int64_t q = s * n_src; q < (s + 1) * n_src; ++q)
        if (src[q] >= 0) {
          fd[EX + src_comp][base + src[q]] += f[t * g.n_shots * n_src + q];
          dfd[EX + src_comp][base + src[q]] += df[t * g.n_shots * n_src + q];
        }
      for (int64_t q = s * n_rec; q < (s + 1) * n_rec; ++q) {
        r[t * g.n_shots * n_rec + q] = rec[q] >= 0 ? fd[EX + rec_comp][base + rec[q]] : (T)0;
        dr[t * g.n_shots * n_rec + q] = rec[q] >= 0 ? dfd[EX + rec_comp][base + rec[q]] : (T)0;
      }
    }
  }
}

// ---------------------------------------------------- Born forward (3-D)
// Reference maxwell_3d_inst.cpp:1100-1527 (born_forward): per step the
// background and scattered H updates, then per cell the scattered E update
// ca dE + cb dcurl + dca E^n + dcb curl^n (E^n, curl^n the background's
// pre-update E and curl H of this step) and the background E update; then
// f0 into the background and df into the scattered field at the sources;
// r records the scattered field, background_r the background.
template <typename T>
void born_forward(const Grid3<T> &g, const T *ca, const T *cb, const T *cq, const T *dca,
                  const T *dcb, const T *f0, const T *df, T *const *fd, T *const *dfd, T *r,
                  T *bg_r, const int64_t *src, const int64_t *rec, int64_t n_src, int64_t n_rec,
                  int src_comp, int rec_comp, int64_t nt, int64_t start_t) {
  const int64_t numel = g.numel(), total = g.n_shots * numel;
  std::vector<T> snap((size_t)6 * total);
  void *st[6];
  const T *bg[6];
  for (int k = 0; k < 6; ++k) {
    st[k] = snap.data() + (size_t)k * total;
    bg[k] = snap.data() + (size_t)k * total;
  }
  for (int64_t t = start_t; t < start_t + nt; ++t) {
    for (int64_t s = 0; s < g.n_shots; ++s) {
      h_step<T>(g, cq, fd, s);
      h_step<T>(g, cq, dfd, s);
      e_step<T>(g, ca, cb, fd, st, 0, s);
      e_step<T>(g, ca, cb, dfd, nullptr, 0, s, dca, dcb, bg);
    }
    for (int64_t s = 0; s < g.n_shots; ++s) {
      const int64_t base = s * numel;
      for (int64_t q = s * n_src; q < (s + 1) * n_src; ++q)
        if (src[q] >= 0) {
          if (f0) fd[EX + src_comp][base + src[q]] += f0[t * g.n_shots * n_src + q];
          if (df) dfd[EX + src_comp][base + src[q]] += df[t * g.n_shots * n_src + q];
        }
      for (int64_t q = s * n_rec; q < (s + 1) * n_rec; ++q) {
        if (r) r[t * g.n_shots * n_rec + q] = rec[q] >= 0 ? dfd[EX + rec_comp][base + rec[q]] : (T)0;
        if (bg_r)
          bg_r[t * g.n_shots * n_rec + q] = rec[q] >= 0 ? fd[EX + rec_comp][base + rec[q]] : (T)0;
      }
    }
  }
}

// prof: az bz azh bzh ay by ayh byh ax bx axh bxh kz kzh ky kyh kx kxh
// (the reference's argument order); pml: z0 y0 x0 z1 y1 x1.
template <typename T>
Grid3<T> make_grid(const T *const *prof, T rdz, T rdy, T rdx, int64_t n_shots, int64_t nz,
                   int64_t ny, int64_t nx, const int64_t *pml, const bool *batched) {
  Grid3<T> g;
  for (int ax = 0; ax < 3; ++ax) {
    g.a[ax] = prof[4 * ax + 0];
    g.b[ax] = prof[4 * ax + 1];
    g.ah[ax] = prof[4 * ax + 2];
    g.bh[ax] = prof[4 * ax + 3];
    g.k[ax] = prof[12 + 2 * ax];
    g.kh[ax] = prof[13 + 2 * ax];
    g.p0[ax] = pml[ax];
    g.p1[ax] = pml[3 + ax];
    g.p1h[ax] = pml[3 + ax] > pml[ax] ? pml[3 + ax] - 1 : pml[ax];
  }
  g.rd[0] = rdz;
  g.rd[1] = rdy;
  g.rd[2] = rdx;
  g.n_shots = n_shots;
  g.n[0] = nz;
  g.n[1] = ny;
  g.n[2] = nx;
  g.ca_b = batched[0];
  g.cb_b = batched[1];
  g.cq_b = batched[2];
  return g;
}

}  // namespace

extern "C" {

#define EM3D_FORWARD(NAME, T)                                                               \
  int NAME(const T *ca, const T *cb, const T *cq, const T *f, T *const *fields, T *r,      \
           const T *const *prof, const int64_t *src, const int64_t *rec, T rdz, T rdy,      \
           T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,          \
           int64_t n_src, int64_t n_rec, int64_t step_ratio, int64_t start_t,               \
           const int64_t *pml, const bool *batched, int src_comp, int rec_comp,             \
           void *const *stores, int fmt, int64_t store_step_bytes) {                        \
    if (step_ratio <= 0 || src_comp < 0 || src_comp > 2 || rec_comp < 0 || rec_comp > 2)   \
      return 1;                                                                             \
    const Grid3<T> g = make_grid<T>(prof, rdz, rdy, rdx, n_shots, nz, ny, nx, pml, batched); \
    forward<T>(g, ca, cb, cq, f, fields, r, src, rec, n_src, n_rec, src_comp, rec_comp, nt, \
               start_t, step_ratio, stores, fmt, store_step_bytes);                         \
    return 0;                                                                               \
  }
EM3D_FORWARD(oracle_em3d_forward_f32, float)
EM3D_FORWARD(oracle_em3d_forward_f64, double)

#define EM3D_BACKWARD(NAME, T)                                                              \
  int NAME(const T *ca, const T *cb, const T *cq, const T *grad_r, T *const *lam,          \
           void *const *stores, int fmt, int64_t store_step_bytes, T *grad_f, double *gca, \
           double *gcb, const T *const *prof, const int64_t *src, const int64_t *rec,       \
           T rdz, T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny,        \
           int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio, int64_t start_t,   \
           const int64_t *pml, const bool *batched, int src_comp, int rec_comp) {           \
    if (step_ratio <= 0 || src_comp < 0 || src_comp > 2 || rec_comp < 0 || rec_comp > 2)   \
      return 1;                                                                             \
    const Grid3<T> g = make_grid<T>(prof, rdz, rdy, rdx, n_shots, nz, ny, nx, pml, batched); \
    backward<T>(g, ca, cb, cq, grad_r, lam, (const void *const *)stores, fmt,               \
                store_step_bytes, grad_f, gca, gcb, src, rec, n_src, n_rec, src_comp,       \
                rec_comp, nt, start_t, step_ratio);                                         \
    return 0;                                                                               \
  }
EM3D_BACKWARD(oracle_em3d_backward_f32, float)
EM3D_BACKWARD(oracle_em3d_backward_f64, double)

#define EM3D_BORN(NAME, T)                                                                  \
  int NAME(const T *ca, const T *cb, const T *cq, const T *dca, const T *dcb, const T *f0,  \
           const T *df, T *const *fields, T *const *dfields, T *r, T *bg_r,                 \
           const T *const *prof, const int64_t *src, const int64_t *rec, T rdz, T rdy,      \
           T rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,          \
           int64_t n_src, int64_t n_rec, int64_t start_t, const int64_t *pml,               \
           const bool *batched, int src_comp, int rec_comp) {                               \
    if (src_comp < 0 || src_comp > 2 || rec_comp < 0 || rec_comp > 2) return 1;            \
    const Grid3<T> g = make_grid<T>(prof, rdz, rdy, rdx, n_shots, nz, ny, nx, pml, batched); \
    born_forward<T>(g, ca, cb, cq, dca, dcb, f0, df, fields, dfields, r, bg_r, src, rec,    \
                    n_src, n_rec, src_comp, rec_comp, nt, start_t);                         \
    return 0;                                                                               \
  }
EM3D_BORN(oracle_em3d_born_forward_f32, float)
EM3D_BORN(oracle_em3d_born_forward_f64, double)

int oracle_em3d_tangent_f64(const double *ca, const double *cb, const double *cq,
                            const double *f, const double *dca, const double *dcb,
                            const double *df, double *const *fields, double *const *dfields,
                            double *r, double *dr, const double *const *prof,
                            const int64_t *src, const int64_t *rec, double rdz, double rdy,
                            double rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny,
                            int64_t nx, int64_t n_src, int64_t n_rec, const int64_t *pml,
                            int src_comp, int rec_comp) {
  const bool batched[3] = {false, false, false};
  const Grid3<double> g =
      make_grid<double>(prof, rdz, rdy, rdx, n_shots, nz, ny, nx, pml, batched);
  tangent<double>(g, ca, cb, cq, f, dca, dcb, df, fields, dfields, r, dr, src, rec, n_src,
                  n_rec, src_comp, rec_comp, nt);
  return 0;
}

}  // extern "C"
