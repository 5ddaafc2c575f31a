// Stencil orders 4, 6 and 8 of the 2-D TM propagator (tm2d_highorder.hip):
// one-cell-per-lane forward and exact-adjoint drivers behind the same C ABI
// as order 2.
#pragma once

#include "tide_common.h"
#include "tide_storage.h"

#include <cstdint>

namespace tide_hip {
namespace tm2d_ho {

template <typename T> struct Problem {
  const T *ca, *cb, *cq;
  const T *prof[12];  // ay by ayh byh ax bx axh bxh ky kyh kx kxh
  T rdy, rdx;
  int64_t n_shots, ny, nx, n_src, n_rec;
  int64_t pml_y0, pml_x0, pml_y1, pml_x1;
  bool ca_b, cb_b, cq_b;
};

#define TIDE_CHECK_HO(expr)                 \
  do {                                      \
    const int _rc = (expr);                 \
    if (_rc != ::tide_hip::kOk) return _rc; \
  } while (0)

// fld: Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x ([n_shots, ny, nx]);
// snapshots as in the order-2 driver (history buffers, or the ring slots
// of active stages); fmt 1 full precision, 2 BF16.
template <typename T>
int forward(int stencil, const Problem<T> &p, T *const fld[7], const T *f, T *r,
            const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
            int64_t step_ratio, void *ey_store, void *curl_store, int fmt, size_t step_bytes,
            hipStream_t stream, SnapshotStage *ey_stage, SnapshotStage *curl_stage);

// lam: lambda_Ey, lambda_Hx, lambda_Hz, m_lambda_Ey_x, m_lambda_Ey_z,
// m_lambda_Hx_z, m_lambda_Hz_x; gca / gcb FP64 accumulators (added to, per
// shot when the coefficient is per shot), null when not wanted.
template <typename T>
int backward(int stencil, const Problem<T> &p, T *const lam[7], const T *grad_r,
             const void *ey_store, const void *curl_store, int fmt, size_t step_bytes, T *grad_f,
             double *gca, double *gcb, const int64_t *src, const int64_t *rec, int64_t nt,
             int64_t start_t, int64_t step_ratio, hipStream_t stream, SnapshotStage *ey_stage,
             SnapshotStage *curl_stage);

// Debye-dispersive forward (any order, unbatched coefficients): da, db, cp
// [n_poles, ny, nx], pol [n_shots, n_poles, ny, nx] and ey_prev [n_shots,
// ny, nx] (in/out); ca / cb are the Debye coefficients.
template <typename T>
int forward_debye(int stencil, const Problem<T> &p, T *const fld[7], const T *f, T *r,
                  const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                  const T *da, const T *db, const T *cp, T *pol, T *ey_prev, int64_t n_poles,
                  hipStream_t stream);

// Born forward (any order): background fields bg and scattered fields sc
// (each Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x) advance together;
// dca / dcb the coefficient perturbation (both null: none), f0 / df the
// background and scattered source terms, r / bg_r the scattered and
// background receivers; hist (null, or four [nt, shots, ny, nx] arrays)
// keeps the background Ey^n, curl^n and scattered Ey^n, curl^n of step t in
// slot t for born_backward.
template <typename T>
int born_forward(int stencil, const Problem<T> &p, T *const bg[7], T *const sc[7],
                 const T *dca, const T *dcb, const T *f0, const T *df, T *r, T *bg_r,
                 const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                 hipStream_t stream, T *const *hist);

// Adjoint of born_forward over steps start_t-1 down to start_t-nt
// (unbatched coefficients): lam / eta the scattered / background adjoint
// states (7 arrays each, in place), w / w_bg the cotangents of r / bg_r,
// g_df / g_f0 per-step source gradients, g_ca, g_cb, g_dca, g_dcb FP64
// accumulators (added to; any may be null); mem: null or 8 [shots, ny, nx]
// scratch arrays.
template <typename T>
int born_backward(int stencil, const Problem<T> &p, T *const lam[7], T *const eta[7],
                  const T *dca, const T *dcb, const T *w, const T *w_bg, const T *const hist[4],
                  T *g_df, T *g_f0, double *g_ca, double *g_cb, double *g_dca, double *g_dcb,
                  const int64_t *src, const int64_t *rec, int64_t nt, int64_t start_t,
                  hipStream_t stream, T *mem);

}  // namespace tm2d_ho
}  // namespace tide_hip
