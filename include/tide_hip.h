/*
 * C ABI of the MI355X (gfx950) TIDE propagator library, libtide_hip.so.
 *
 * Conventions shared by every entry point:
 *  - all array arguments are device pointers to contiguous, caller-owned
 *    buffers that stay valid until the enqueued work completes; the library
 *    keeps no pointer after returning.  sources_i / receivers_i may be host
 *    or device pointers.
 *  - fields are [n_shots, ny, nx] with x fastest (padded grid); ca/cb/cq are
 *    [ny, nx], or [n_shots, ny, nx] when the matching *_batched flag is set.
 *  - work is enqueued on `compute_stream` (a hipStream_t; NULL = the null
 *    stream) and the call returns without waiting for it.
 *  - the return value is 0 on success or a nonzero status; the message of
 *    the last failure on the calling thread is returned by tide_last_error().
 *    (The reference functions return void and exit the process on error.)
 *  - storage_mode: 0 device (the whole history in *_store_1), 1 cpu (a
 *    three-step device ring in *_store_1 staged through the pinned host
 *    history *_store_3), 2 disk (the ring staged through a three-step
 *    pinned host ring *_store_3 to the file filenames[0]), 3 none;
 *    storage_format: 0 full precision, 1 BF16 (round-to-nearest-even).
 *  - stencil orders 2, 4, 6 and 8 are built; {dtype} is float or double.
 */
#ifndef TIDE_HIP_H
#define TIDE_HIP_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" if none). */
const char *tide_last_error(void);

/*
 * Forward propagation without snapshots.
 * Replaces maxwell_tm_2_{dtype}_forward_cuda: reference ABI
 * src/tide/backend_utils.py:256-269 (_TM_FORWARD_SPEC), CUDA implementation
 * src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1281-1368.  The Debye
 * arguments are accepted and must be unused (has_dispersion false).
 */
int maxwell_tm_2_float_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *polarization, float *ey_prev, float *r,
    int64_t n_poles, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_2_double_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double const *debye_a,
    double const *debye_b, double const *debye_cp, double *polarization,
    double *ey_prev, double *r, int64_t n_poles, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

/*
 * Forward propagation storing, every step_ratio-th step t, the pre-update Ey
 * and the CPML curl of H at ey_store_1 / curl_store_1 +
 * (t / step_ratio) * shot_bytes_uncomp * n_shots.
 * Replaces maxwell_tm_2_{dtype}_forward_with_storage_cuda: reference ABI
 * src/tide/backend_utils.py:271-287, CUDA implementation
 * src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1371-1609 (host staging
 * :1407-1571; storage helpers src/tide/csrc/storage/storage_utils.cu).
 */
int maxwell_tm_2_float_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_2_double_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

/*
 * Exact adjoint over nt steps ending at start_t: injects grad_r, writes
 * grad_f, and adds dJ/dca, dJ/dcb (accumulated in FP64, summed over shots
 * unless batched) into grad_ca / grad_cb.  grad_*_shot and
 * adjoint_memory_scratch are not needed (pass NULL).  Host-staged modes
 * read the snapshots back through the same buffers the forward used.
 * Replaces maxwell_tm_2_{dtype}_backward_cuda: reference ABI
 * src/tide/backend_utils.py:289-308, CUDA implementation
 * src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:2405-2783.
 */
int maxwell_tm_2_float_backward_cuda(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ey, float *lambda_hx, float *lambda_hz, float *m_lambda_ey_x,
    float *m_lambda_ey_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    float *adjoint_memory_scratch, void *ey_store_1, void *ey_store_3,
    char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,
    char const *const *curl_filenames, float *grad_f, float *grad_ca,
    float *grad_cb, float *grad_ca_shot, float *grad_cb_shot, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_2_double_backward_cuda(
    double const *ca, double const *cb, double const *cq, double const *grad_r,
    double *lambda_ey, double *lambda_hx, double *lambda_hz,
    double *m_lambda_ey_x, double *m_lambda_ey_z, double *m_lambda_hx_z,
    double *m_lambda_hz_x, double *adjoint_memory_scratch, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double *grad_f,
    double *grad_ca, double *grad_cb, double *grad_ca_shot,
    double *grad_cb_shot, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdy, double rdx, double dt, int64_t nt,
    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

/*
 * Stencil orders 4, 6 and 8: the same arguments and meaning as the order-2
 * entry points above (fd_pad = [S/2, S/2-1, S/2, S/2-1] and the profiles of
 * that order).  Replace maxwell_tm_{4,6,8}_{dtype}_*_cuda: reference
 * instantiations src/tide/csrc/tm2d/maxwell_tm_cuda_instantiations.inc and
 * names src/tide/backend_utils.py:771-781; staggered differences
 * src/tide/csrc/tm2d/maxwell_tm_core.cuh:61-365.
 */
int maxwell_tm_4_float_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *polarization, float *ey_prev, float *r,
    int64_t n_poles, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_4_float_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_4_float_backward_cuda(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ey, float *lambda_hx, float *lambda_hz, float *m_lambda_ey_x,
    float *m_lambda_ey_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    float *adjoint_memory_scratch, void *ey_store_1, void *ey_store_3,
    char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,
    char const *const *curl_filenames, float *grad_f, float *grad_ca,
    float *grad_cb, float *grad_ca_shot, float *grad_cb_shot, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_4_double_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double const *debye_a,
    double const *debye_b, double const *debye_cp, double *polarization,
    double *ey_prev, double *r, int64_t n_poles, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_4_double_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_4_double_backward_cuda(
    double const *ca, double const *cb, double const *cq, double const *grad_r,
    double *lambda_ey, double *lambda_hx, double *lambda_hz,
    double *m_lambda_ey_x, double *m_lambda_ey_z, double *m_lambda_hx_z,
    double *m_lambda_hz_x, double *adjoint_memory_scratch, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double *grad_f,
    double *grad_ca, double *grad_cb, double *grad_ca_shot,
    double *grad_cb_shot, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdy, double rdx, double dt, int64_t nt,
    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_float_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *polarization, float *ey_prev, float *r,
    int64_t n_poles, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_6_float_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_float_backward_cuda(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ey, float *lambda_hx, float *lambda_hz, float *m_lambda_ey_x,
    float *m_lambda_ey_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    float *adjoint_memory_scratch, void *ey_store_1, void *ey_store_3,
    char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,
    char const *const *curl_filenames, float *grad_f, float *grad_ca,
    float *grad_cb, float *grad_ca_shot, float *grad_cb_shot, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_double_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double const *debye_a,
    double const *debye_b, double const *debye_cp, double *polarization,
    double *ey_prev, double *r, int64_t n_poles, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_6_double_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_double_backward_cuda(
    double const *ca, double const *cb, double const *cq, double const *grad_r,
    double *lambda_ey, double *lambda_hx, double *lambda_hz,
    double *m_lambda_ey_x, double *m_lambda_ey_z, double *m_lambda_hx_z,
    double *m_lambda_hz_x, double *adjoint_memory_scratch, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double *grad_f,
    double *grad_ca, double *grad_cb, double *grad_ca_shot,
    double *grad_cb_shot, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdy, double rdx, double dt, int64_t nt,
    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_float_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *polarization, float *ey_prev, float *r,
    int64_t n_poles, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_8_float_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_float_backward_cuda(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ey, float *lambda_hx, float *lambda_hz, float *m_lambda_ey_x,
    float *m_lambda_ey_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    float *adjoint_memory_scratch, void *ey_store_1, void *ey_store_3,
    char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,
    char const *const *curl_filenames, float *grad_f, float *grad_ca,
    float *grad_cb, float *grad_ca_shot, float *grad_cb_shot, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_double_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double const *debye_a,
    double const *debye_b, double const *debye_cp, double *polarization,
    double *ey_prev, double *r, int64_t n_poles, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

int maxwell_tm_8_double_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *r, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax,
    double const *bx, double const *axh, double const *bxh, double const *ky,
    double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_double_backward_cuda(
    double const *ca, double const *cb, double const *cq, double const *grad_r,
    double *lambda_ey, double *lambda_hx, double *lambda_hz,
    double *m_lambda_ey_x, double *m_lambda_ey_z, double *m_lambda_hx_z,
    double *m_lambda_hz_x, double *adjoint_memory_scratch, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double *grad_f,
    double *grad_ca, double *grad_cb, double *grad_ca_shot,
    double *grad_cb_shot, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdy, double rdx, double dt, int64_t nt,
    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

/*
 * Born (tangent-linear) forward: the background fields (ey .. m_hz_x) and
 * the scattered fields (dey .. dm_hz_x) advance together over nt steps from
 * start_t; the scattered E update adds dca * Ey^n + dcb * curl H^n of the
 * background, df is injected into the scattered field and f0 into the
 * background.  r receives the scattered receivers, background_r the
 * background's.  dca / dcb NULL: no coefficient perturbation; step_ratio
 * must be 1.  Replaces maxwell_tm_{2,4,6,8}_{dtype}_born_forward_cuda:
 * reference ABI src/tide/backend_utils.py:336-349 (_TM_BORN_FORWARD_SPEC),
 * CUDA implementation src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1611-1800.
 */
int maxwell_tm_2_float_born_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x,
    float *m_ey_z, float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz,
    float *dm_ey_x, float *dm_ey_z, float *dm_hx_z, float *dm_hz_x, float *r,
    float *background_r, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_2_double_born_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x,
    double *m_ey_z, double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz,
    double *dm_ey_x, double *dm_ey_z, double *dm_hx_z, double *dm_hz_x, double *r,
    double *background_r, double const *ay, double const *by, double const *ayh,
    double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_4_float_born_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x,
    float *m_ey_z, float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz,
    float *dm_ey_x, float *dm_ey_z, float *dm_hx_z, float *dm_hz_x, float *r,
    float *background_r, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_4_double_born_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x,
    double *m_ey_z, double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz,
    double *dm_ey_x, double *dm_ey_z, double *dm_hx_z, double *dm_hz_x, double *r,
    double *background_r, double const *ay, double const *by, double const *ayh,
    double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_6_float_born_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x,
    float *m_ey_z, float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz,
    float *dm_ey_x, float *dm_ey_z, float *dm_hx_z, float *dm_hz_x, float *r,
    float *background_r, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_6_double_born_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x,
    double *m_ey_z, double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz,
    double *dm_ey_x, double *dm_ey_z, double *dm_hx_z, double *dm_hz_x, double *r,
    double *background_r, double const *ay, double const *by, double const *ayh,
    double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_8_float_born_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x,
    float *m_ey_z, float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz,
    float *dm_ey_x, float *dm_ey_z, float *dm_hx_z, float *dm_hz_x, float *r,
    float *background_r, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

int maxwell_tm_8_double_born_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x,
    double *m_ey_z, double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz,
    double *dm_ey_x, double *dm_ey_z, double *dm_hx_z, double *dm_hz_x, double *r,
    double *background_r, double const *ay, double const *by, double const *ayh,
    double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream);

/*
 * Born forward that keeps the Born adjoint's snapshots: ey_store_1 /
 * curl_store_1 receive the background's pre-update Ey^n and curl H^n,
 * dey_store / dcurl_store the scattered field's, one [n_shots, ny, nx] slot
 * per step (slot t for step t), on the device in full precision
 * (storage_mode 0, storage_format 0, shot_bytes_uncomp = ny * nx * sizeof
 * of the field type); all four stores or none.  Otherwise as
 * maxwell_tm_*_born_forward_cuda.  Replaces
 * maxwell_tm_{2,4,6,8}_{dtype}_born_forward_with_storage_cuda: reference ABI
 * src/tide/backend_utils.py:351-370 (_TM_BORN_FORWARD_WITH_STORAGE_SPEC).
 */
int maxwell_tm_2_float_born_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz, float *dm_ey_x, float *dm_ey_z,
    float *dm_hx_z, float *dm_hz_x, float *r, float *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_2_double_born_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz, double *dm_ey_x, double *dm_ey_z,
    double *dm_hx_z, double *dm_hz_x, double *r, double *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_4_float_born_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz, float *dm_ey_x, float *dm_ey_z,
    float *dm_hx_z, float *dm_hz_x, float *r, float *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_4_double_born_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz, double *dm_ey_x, double *dm_ey_z,
    double *dm_hx_z, double *dm_hz_x, double *r, double *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_float_born_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz, float *dm_ey_x, float *dm_ey_z,
    float *dm_hx_z, float *dm_hz_x, float *r, float *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_6_double_born_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz, double *dm_ey_x, double *dm_ey_z,
    double *dm_hx_z, double *dm_hz_x, double *r, double *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_float_born_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float *dey, float *dhx, float *dhz, float *dm_ey_x, float *dm_ey_z,
    float *dm_hx_z, float *dm_hz_x, float *r, float *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

int maxwell_tm_8_double_born_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double *ey, double *hx, double *hz, double *m_ey_x, double *m_ey_z,
    double *m_hx_z, double *m_hz_x, double *dey, double *dhx, double *dhz, double *dm_ey_x, double *dm_ey_z,
    double *dm_hx_z, double *dm_hz_x, double *r, double *background_r, void *ey_store_1,
    void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,
    void *curl_filenames, void *dey_store, void *dcurl_store, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, double dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot, int64_t step_ratio,
    int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,
    bool ca_requires_grad, bool cb_requires_grad, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t n_threads, int64_t device, void *compute_stream,
    void *storage_stream);

/*
 * Adjoint of the Born forward (second-order products): gradients of
 * <r, grad_r> + <background_r, grad_background_r> with respect to ca, cb
 * (through the background and the scattered recurrences), dca, dcb (FP64
 * accumulation, added to the outputs; grad_ca / grad_cb only when
 * ca/cb_requires_grad), f0 and df (grad_f0 / grad_df written per step).
 * ey, hx, hz + m_lambda_* hold the scattered adjoint state and dey, dhx, dhz
 * + m_eta_* the background's (in/out, the caller zeroes them for a fresh
 * run); adjoint_memory_scratch (8 [n_shots, ny, nx] arrays) or NULL.  The
 * stores are those of born_forward_with_storage; unbatched coefficients,
 * device storage, step_ratio 1, no background-adjoint capture or reuse;
 * lambda_store, eta_source_old, work_eta_*, grad_*_shot are not read.
 * Replaces maxwell_tm_{2,4,6,8}_{dtype}_born_backward_bggrad_cuda: reference
 * ABI src/tide/backend_utils.py:407-444 (_TM_BORN_BACKWARD_BGGRAD_SPEC),
 * CUDA implementation src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1611-2404.
 */
int maxwell_tm_2_float_born_backward_bggrad_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float const *grad_r, float const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, float *ey, float *hx, float *hz, float *dey,
    float *dhx, float *dhz, float *grad_f0, float *grad_df, float *grad_ca, float *grad_cb,
    float *grad_dca, float *grad_dcb, float *m_lambda_ey_x, float *m_lambda_ey_z,
    float *m_lambda_hx_z, float *m_lambda_hz_x, float *m_eta_ey_x, float *m_eta_ey_z,
    float *m_eta_hx_z, float *m_eta_hz_x, float *eta_source_old, float *work_eta_x,
    float *work_eta_z, float *grad_ca_shot, float *grad_cb_shot, float *grad_dca_shot,
    float *grad_dcb_shot, float *adjoint_memory_scratch, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_2_double_born_backward_bggrad_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double const *grad_r, double const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, double *ey, double *hx, double *hz, double *dey,
    double *dhx, double *dhz, double *grad_f0, double *grad_df, double *grad_ca, double *grad_cb,
    double *grad_dca, double *grad_dcb, double *m_lambda_ey_x, double *m_lambda_ey_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, double *m_eta_ey_x, double *m_eta_ey_z,
    double *m_eta_hx_z, double *m_eta_hz_x, double *eta_source_old, double *work_eta_x,
    double *work_eta_z, double *grad_ca_shot, double *grad_cb_shot, double *grad_dca_shot,
    double *grad_dcb_shot, double *adjoint_memory_scratch, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_4_float_born_backward_bggrad_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float const *grad_r, float const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, float *ey, float *hx, float *hz, float *dey,
    float *dhx, float *dhz, float *grad_f0, float *grad_df, float *grad_ca, float *grad_cb,
    float *grad_dca, float *grad_dcb, float *m_lambda_ey_x, float *m_lambda_ey_z,
    float *m_lambda_hx_z, float *m_lambda_hz_x, float *m_eta_ey_x, float *m_eta_ey_z,
    float *m_eta_hx_z, float *m_eta_hz_x, float *eta_source_old, float *work_eta_x,
    float *work_eta_z, float *grad_ca_shot, float *grad_cb_shot, float *grad_dca_shot,
    float *grad_dcb_shot, float *adjoint_memory_scratch, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_4_double_born_backward_bggrad_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double const *grad_r, double const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, double *ey, double *hx, double *hz, double *dey,
    double *dhx, double *dhz, double *grad_f0, double *grad_df, double *grad_ca, double *grad_cb,
    double *grad_dca, double *grad_dcb, double *m_lambda_ey_x, double *m_lambda_ey_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, double *m_eta_ey_x, double *m_eta_ey_z,
    double *m_eta_hx_z, double *m_eta_hz_x, double *eta_source_old, double *work_eta_x,
    double *work_eta_z, double *grad_ca_shot, double *grad_cb_shot, double *grad_dca_shot,
    double *grad_dcb_shot, double *adjoint_memory_scratch, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_6_float_born_backward_bggrad_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float const *grad_r, float const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, float *ey, float *hx, float *hz, float *dey,
    float *dhx, float *dhz, float *grad_f0, float *grad_df, float *grad_ca, float *grad_cb,
    float *grad_dca, float *grad_dcb, float *m_lambda_ey_x, float *m_lambda_ey_z,
    float *m_lambda_hx_z, float *m_lambda_hz_x, float *m_eta_ey_x, float *m_eta_ey_z,
    float *m_eta_hx_z, float *m_eta_hz_x, float *eta_source_old, float *work_eta_x,
    float *work_eta_z, float *grad_ca_shot, float *grad_cb_shot, float *grad_dca_shot,
    float *grad_dcb_shot, float *adjoint_memory_scratch, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_6_double_born_backward_bggrad_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double const *grad_r, double const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, double *ey, double *hx, double *hz, double *dey,
    double *dhx, double *dhz, double *grad_f0, double *grad_df, double *grad_ca, double *grad_cb,
    double *grad_dca, double *grad_dcb, double *m_lambda_ey_x, double *m_lambda_ey_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, double *m_eta_ey_x, double *m_eta_ey_z,
    double *m_eta_hx_z, double *m_eta_hz_x, double *eta_source_old, double *work_eta_x,
    double *work_eta_z, double *grad_ca_shot, double *grad_cb_shot, double *grad_dca_shot,
    double *grad_dcb_shot, double *adjoint_memory_scratch, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_8_float_born_backward_bggrad_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *f0, float const *df, float const *grad_r, float const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, float *ey, float *hx, float *hz, float *dey,
    float *dhx, float *dhz, float *grad_f0, float *grad_df, float *grad_ca, float *grad_cb,
    float *grad_dca, float *grad_dcb, float *m_lambda_ey_x, float *m_lambda_ey_z,
    float *m_lambda_hx_z, float *m_lambda_hz_x, float *m_eta_ey_x, float *m_eta_ey_z,
    float *m_eta_hx_z, float *m_eta_hz_x, float *eta_source_old, float *work_eta_x,
    float *work_eta_z, float *grad_ca_shot, float *grad_cb_shot, float *grad_dca_shot,
    float *grad_dcb_shot, float *adjoint_memory_scratch, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

int maxwell_tm_8_double_born_backward_bggrad_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *f0, double const *df, double const *grad_r, double const *grad_background_r,
    void *ey_store_1, void *ey_store_3, void *ey_filenames, void *curl_store_1,
    void *curl_store_3, void *curl_filenames, void *dey_store, void *dcurl_store,
    void *lambda_store, int64_t background_n_shots, double *ey, double *hx, double *hz, double *dey,
    double *dhx, double *dhz, double *grad_f0, double *grad_df, double *grad_ca, double *grad_cb,
    double *grad_dca, double *grad_dcb, double *m_lambda_ey_x, double *m_lambda_ey_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, double *m_eta_ey_x, double *m_eta_ey_z,
    double *m_eta_hx_z, double *m_eta_hz_x, double *eta_source_old, double *work_eta_x,
    double *work_eta_z, double *grad_ca_shot, double *grad_cb_shot, double *grad_dca_shot,
    double *grad_dcb_shot, double *adjoint_memory_scratch, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx, double const *axh,
    double const *bxh, double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool capture_background_adjoint,
    bool reuse_background_adjoint, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,
    int64_t n_threads, int64_t device, void *compute_stream, void *storage_stream);

/*
 * This library's form of the Born adjoint (used by its own autograd
 * function): the four snapshot histories as written by
 * born_forward_with_storage, the scattered (l_*) and background (e_*)
 * adjoint states, cotangents w (scattered receivers) and w_bg (background
 * receivers; either may be NULL), per-step source gradients grad_df /
 * grad_f0 (NULL: not wanted), FP64 accumulators gca, gcb, gdca, gdcb (added
 * to; NULL: not wanted) over steps start_t-1 .. start_t-nt.  No reference
 * counterpart (the reference folds into field-typed outputs).
 */
int tide_tm2d_2_float_born_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *w, float const *w_bg, float const *bg_ey_hist, float const *bg_curl_hist,
    float const *sc_ey_hist, float const *sc_curl_hist, float *l_ey, float *l_hx, float *l_hz,
    float *ml_ey_x, float *ml_ey_z, float *ml_hx_z, float *ml_hz_x, float *e_ey, float *e_hx,
    float *e_hz, float *me_ey_x, float *me_ey_z, float *me_hx_z, float *me_hz_x, float *grad_df,
    float *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_2_double_born_backward_f64acc(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *w, double const *w_bg, double const *bg_ey_hist, double const *bg_curl_hist,
    double const *sc_ey_hist, double const *sc_curl_hist, double *l_ey, double *l_hx, double *l_hz,
    double *ml_ey_x, double *ml_ey_z, double *ml_hx_z, double *ml_hz_x, double *e_ey, double *e_hx,
    double *e_hz, double *me_ey_x, double *me_ey_z, double *me_hx_z, double *me_hz_x, double *grad_df,
    double *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_4_float_born_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *w, float const *w_bg, float const *bg_ey_hist, float const *bg_curl_hist,
    float const *sc_ey_hist, float const *sc_curl_hist, float *l_ey, float *l_hx, float *l_hz,
    float *ml_ey_x, float *ml_ey_z, float *ml_hx_z, float *ml_hz_x, float *e_ey, float *e_hx,
    float *e_hz, float *me_ey_x, float *me_ey_z, float *me_hx_z, float *me_hz_x, float *grad_df,
    float *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_4_double_born_backward_f64acc(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *w, double const *w_bg, double const *bg_ey_hist, double const *bg_curl_hist,
    double const *sc_ey_hist, double const *sc_curl_hist, double *l_ey, double *l_hx, double *l_hz,
    double *ml_ey_x, double *ml_ey_z, double *ml_hx_z, double *ml_hz_x, double *e_ey, double *e_hx,
    double *e_hz, double *me_ey_x, double *me_ey_z, double *me_hx_z, double *me_hz_x, double *grad_df,
    double *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_6_float_born_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *w, float const *w_bg, float const *bg_ey_hist, float const *bg_curl_hist,
    float const *sc_ey_hist, float const *sc_curl_hist, float *l_ey, float *l_hx, float *l_hz,
    float *ml_ey_x, float *ml_ey_z, float *ml_hx_z, float *ml_hz_x, float *e_ey, float *e_hx,
    float *e_hz, float *me_ey_x, float *me_ey_z, float *me_hx_z, float *me_hz_x, float *grad_df,
    float *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_6_double_born_backward_f64acc(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *w, double const *w_bg, double const *bg_ey_hist, double const *bg_curl_hist,
    double const *sc_ey_hist, double const *sc_curl_hist, double *l_ey, double *l_hx, double *l_hz,
    double *ml_ey_x, double *ml_ey_z, double *ml_hx_z, double *ml_hz_x, double *e_ey, double *e_hx,
    double *e_hz, double *me_ey_x, double *me_ey_z, double *me_hx_z, double *me_hz_x, double *grad_df,
    double *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_8_float_born_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *dca, float const *dcb,
    float const *w, float const *w_bg, float const *bg_ey_hist, float const *bg_curl_hist,
    float const *sc_ey_hist, float const *sc_curl_hist, float *l_ey, float *l_hx, float *l_hz,
    float *ml_ey_x, float *ml_ey_z, float *ml_hx_z, float *ml_hz_x, float *e_ey, float *e_hx,
    float *e_hz, float *me_ey_x, float *me_ey_z, float *me_hx_z, float *me_hz_x, float *grad_df,
    float *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);

int tide_tm2d_8_double_born_backward_f64acc(
    double const *ca, double const *cb, double const *cq, double const *dca, double const *dcb,
    double const *w, double const *w_bg, double const *bg_ey_hist, double const *bg_curl_hist,
    double const *sc_ey_hist, double const *sc_curl_hist, double *l_ey, double *l_hx, double *l_hz,
    double *ml_ey_x, double *ml_ey_z, double *ml_hx_z, double *ml_hz_x, double *e_ey, double *e_hx,
    double *e_hz, double *me_ey_x, double *me_ey_z, double *me_hx_z, double *me_hz_x, double *grad_df,
    double *grad_f0, double *gca, double *gcb, double *gdca, double *gdcb, double const *ay,
    double const *by, double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh, double const *kx,
    double const *kxh, int64_t const *sources_i, int64_t const *receivers_i, double rdy,
    double rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,
    int64_t pml_x1, int64_t device, void *compute_stream);


/*
 * Allocation-free variant of maxwell_tm_2_float_forward_cuda: the same
 * arguments and results, with the call's scratch (ping-pong copies of the
 * fields and CPML memories, packed profiles, point tables) in a fixed
 * 64 MiB device array of the library instead of stream-ordered
 * allocations; no hipMalloc / hipFree or allocating library routine is
 * called.  A call whose scratch exceeds the array (7 x n_shots x ny x nx
 * fp32 values plus the tables) returns status 1 before any device work.
 * Each device holds one array for the life of the process: calls that
 * overlap on one device (other streams or threads) race on it.  No
 * reference counterpart; for callers that cannot allocate (code under
 * graph capture, tight serving loops, harnesses that forbid it).
 */
int tide_tm2d_float_forward_noalloc(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ey, float *hx, float *hz, float *m_ey_x, float *m_ey_z,
    float *m_hx_z, float *m_hz_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *polarization, float *ey_prev, float *r,
    int64_t n_poles, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,
    void *compute_stream);

/*
 * Same adjoint as maxwell_tm_2_{dtype}_backward_cuda, with grad_ca / grad_cb
 * as FP64 accumulators that are added to (so a caller running shot chunks
 * keeps one FP64 sum across them).  The six storage arguments and
 * storage_mode have the reference backward's meaning: mode 0 reads the
 * device history in *_store_1; modes 1 (cpu) and 2 (disk) read a
 * three-step device ring in *_store_1, filled from the pinned host history
 * (cpu) or from the file filenames[0] through a three-step pinned host ring
 * in *_store_3 (disk), on storage_stream.  stencil: 2, 4, 6 or 8.  NULL
 * grad_ca / grad_cb: that
 * coefficient's gradient is not wanted (its storage is not read).  No
 * reference counterpart; the reference keeps fp32 per-shot accumulators
 * (src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:834-871).
 */
int tide_tm2d_float_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ey, float *lambda_hx, float *lambda_hz, float *m_lambda_ey_x,
    float *m_lambda_ey_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    void *ey_store_1, void *ey_store_3, char const *const *ey_filenames,
    void *curl_store_1, void *curl_store_3, char const *const *curl_filenames,
    float *grad_f, double *grad_ca, double *grad_cb, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_y1, int64_t pml_x1, int64_t stencil, int64_t device,
    void *compute_stream, void *storage_stream);

int tide_tm2d_double_backward_f64acc(
    double const *ca, double const *cb, double const *cq,
    double const *grad_r, double *lambda_ey, double *lambda_hx,
    double *lambda_hz, double *m_lambda_ey_x, double *m_lambda_ey_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, void *ey_store_1,
    void *ey_store_3, char const *const *ey_filenames, void *curl_store_1,
    void *curl_store_3, char const *const *curl_filenames, double *grad_f,
    double *grad_ca, double *grad_cb, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdy, double rdx, int64_t nt,
    int64_t n_shots, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_batched,
    bool cb_batched, bool cq_batched, int64_t start_t, int64_t pml_y0,
    int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t stencil,
    int64_t device, void *compute_stream, void *storage_stream);

/*
 * Launch shape of the TM2D step kernels: rows marched per wavefront
 * (forward 2/4/8/16, adjoint 2/4/8; halved for fp64) and shots per workgroup
 * (1..8), process-wide.  Set by the autotuner.  No reference counterpart
 * (the reference fixes a 32x8 block at build time,
 * src/tide/csrc/CMakeLists.txt:24-25).
 */
int tide_tm2d_set_tuning(int fwd_rows, int fwd_group, int adj_rows, int adj_group);
int tide_tm2d_get_tuning(int *out4);

/*
 * Kernel family of the single-problem TM2D step, per pass (0 or 1): 0 puts
 * one shot per wavefront and sums a tile's FP64 gradient terms across the
 * workgroup in LDS; 1 gives each wavefront one tile and walks every shot of
 * the call, summing gradient terms in registers (one read-modify-write per
 * cell and step).  With 1 the group of tide_tm2d_set_tuning is the number
 * of tiles per workgroup (clamped to 4).  Results are identical in both
 * families up to the order of the FP64 shot sum.  No reference counterpart
 * (the reference launches one thread per cell and shot,
 * src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:263-270).
 */
int tide_tm2d_set_shot_loop(int fwd_loop, int adj_loop);
int tide_tm2d_get_shot_loop(int *out2);

/*
 * Time steps per launch of the single-problem TM2D path, per pass (1 to
 * 4).  With K >= 2, K steps run in one launch of the K-step kernels
 * (64 - 2K owned columns per wavefront, a recomputed halo of K), which read
 * and write the wavefield state once per K steps; results are bitwise
 * those of single steps.  A remainder runs as two-step and single-step
 * launches, and host-staged storage (modes 1, 2) caps K at 2.  No
 * reference counterpart (one launch per kernel and step,
 * src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:1371-1609).
 */
int tide_tm2d_set_time_block(int fwd_steps, int adj_steps);
int tide_tm2d_get_time_block(int *out2);

/*
 * Two-step launches per pass: 1 runs the interior tiles (no CPML strip, no
 * grid edge within the halo) and the border tiles as two kernels, each with
 * its own register budget; 0 runs one kernel over all tiles.  Results are
 * identical.  No reference counterpart.
 */
int tide_tm2d_set_tile_split(int fwd_split, int adj_split);
int tide_tm2d_get_tile_split(int *out2);
/* Height of the two-step border tiles per pass: 1 or 2 rows, or 0 for the
 * interior tile height (defaults: forward 0, adjoint 2). */
int tide_tm2d_set_border_rows(int fwd_rows, int adj_rows);
int tide_tm2d_get_border_rows(int *out2);

/*
 * Kernel timing without a profiler: when enabled, every single-problem
 * TM2D time loop is bracketed by HIP events on its stream (the launches
 * only).  take_loop_timing waits for the recorded loops of a pass (0
 * forward, 1 adjoint), returns their summed milliseconds, launch count and
 * loop count, and clears them.  No reference counterpart.
 */
int tide_tm2d_set_loop_timing(int enable);
int tide_tm2d_take_loop_timing(int pass, double *ms, int64_t *launches, int64_t *loops);

/* ------------------------------------------------------- batched 2-D plans
 * Many independent fp32 TM2D problems (different grid sizes, shot counts
 * and nt) advanced together: one launch per time step covers every problem
 * still running, and the forward and adjoint time loops are captured once
 * into HIP graphs and replayed by each run.  All pointers are device
 * pointers owned by the caller and must stay valid for the plan's life
 * (sources_i / receivers_i and the profiles are read once, at creation);
 * the kernels read ca/cb/cq/f/grad_r through them at every run, so a
 * caller updates models or sources in place between runs.  The forward
 * starts from `fields` (the caller's initial state, normally zero); the
 * backward zeroes `lambda`, grad_ca/grad_cb (FP64, [ny, nx], summed over
 * shots) and grad_f itself.  No reference counterpart (the reference runs
 * one problem per call, maxwell_tm_cuda_inst.cu:1371-1609, 2405-2734).
 */
typedef struct tide_tm2d_problem {
  float const *ca, *cb, *cq;          /* [ny, nx] */
  float const *f;                     /* [nt, n_shots, n_src] */
  int64_t const *sources_i;           /* [n_shots, n_src] (host or device) */
  int64_t const *receivers_i;         /* [n_shots, n_rec] */
  float const *prof[12];              /* ay by ayh byh ax bx axh bxh ky kyh kx kxh */
  float rdy, rdx;
  int64_t nt, n_shots, ny, nx, n_src, n_rec;
  int64_t interval;                   /* snapshot / gradient sampling (step_ratio) */
  int64_t pml[4];                     /* pml_y0, pml_x0, pml_y1, pml_x1 */
  float *fields[7];                   /* ey hx hz m_ey_x m_ey_z m_hx_z m_hz_x */
  float *r;                           /* [nt, n_shots, n_rec] */
  void *ey_store, *curl_store;        /* [ceil(nt/interval), n_shots, ny, nx] or NULL */
  float const *grad_r;                /* [nt, n_shots, n_rec] */
  float *lambda[7];                   /* adjoint state, same order as fields; NULL = forward only */
  double *grad_ca, *grad_cb;          /* [ny, nx] FP64, or NULL */
  float *grad_f;                      /* [nt, n_shots, n_src] or NULL */
} tide_tm2d_problem;

/* storage_format 0 = fp32 snapshots, 1 = BF16; use_graph 0 launches each
 * run eagerly (same kernels).  Launch shape from tide_tm2d_set_tuning. */
int tide_tm2d_plan_create(const tide_tm2d_problem *problems, int64_t n,
                          int64_t storage_format, int64_t use_graph, int64_t device,
                          void **plan);
int tide_tm2d_plan_forward(void *plan, void *stream);
int tide_tm2d_plan_backward(void *plan, void *stream);
int tide_tm2d_plan_destroy(void *plan);

/* ------------------------------------------------- 2-D serving engine
 * Continuous batching of independent fp32 TM2D solves that arrive and
 * leave at any time.  A job is one problem (or one shot chunk of a request)
 * whose device arrays the caller allocates and keeps alive until the job is
 * removed; the engine itself allocates per job only its packed CPML
 * profiles and source/receiver table (stream-ordered, released at removal).
 * sources_i / receivers_i must be host arrays here.  All calls of one
 * engine use one stream.
 * Every tick advances every running job by up to `max_steps` steps of its
 * current phase with ONE launch per step: each workgroup reads its job's
 * descriptor and runs that job's forward step or adjoint iteration, so
 * forward-only jobs, forward passes of gradient jobs and adjoints share the
 * same launches.  Job phases: 0 forward running, 1 forward done, 2 adjoint
 * running, 3 adjoint done.  A job with lambda[0] == NULL is forward-only.
 * Results are the kernels of the single-problem path with the same shot
 * grouping: receivers and gradients do not depend on what else runs.
 * No reference counterpart (the reference solves one request per call,
 * src/tide/maxwell/tm2d.py:258-745); the calls are not thread-safe.
 */
typedef struct tide_tm2d_job {
  tide_tm2d_problem p;       /* fields[0..4] / lambda: parity-0 buffers, zero at start */
  float *fields_b[7];        /* parity-1 copies of fields[7] (same order), zero at start */
  float *lambda_b[7];        /* parity-1 adjoint state, same order as lambda, zero at start */
  double *grad_ca_groups;    /* [ceil(n_shots/group), ny, nx] zeroed FP64, when n_shots > group */
  double *grad_cb_groups;
  int64_t storage_format;    /* 0 fp32 snapshots, 1 BF16 */
} tide_tm2d_job;

int tide_tm2d_engine_create(int64_t device, int64_t group, int64_t max_jobs, void **engine);
/* shots per workgroup (sizes grad_*_groups) */
int tide_tm2d_engine_group(void *engine, int64_t *group);
/* Most steps a small job advances per tick, as a multiple of the largest
 * running job's steps (1..64, default 4); jobs are boosted in inverse
 * proportion to their cells x shots per step. */
int tide_tm2d_engine_set_boost(void *engine, int64_t boost);
int tide_tm2d_engine_add(void *engine, const tide_tm2d_job *job, int64_t *job_id);
/* Step multiplier of one running job (1..64, default 1): its allotment per
 * tick is multiplied by it, a priority (the server gives its oldest
 * request's jobs more steps per tick than newer ones). */
int tide_tm2d_engine_set_weight(void *engine, int64_t job, int64_t weight);
/* Enqueues up to max_steps steps of every running job on `stream`;
 * *n_running = jobs still running after it. */
int tide_tm2d_engine_tick(void *engine, int64_t max_steps, void *stream, int64_t *n_running);
int tide_tm2d_engine_query(void *engine, int64_t job_id, int64_t *phase, int64_t *step);
/* Ids (up to cap) of the jobs whose phase ended in the ticks since the last
 * call, in the order they ended; *n = how many were written. */
int tide_tm2d_engine_finished(void *engine, int64_t *ids, int64_t cap, int64_t *n);
/* phase 1 -> 2; grad_r must hold the receiver cotangents by then (stream order). */
int tide_tm2d_engine_begin_adjoint(void *engine, int64_t job_id);
/* Forgets the job; after a finished adjoint enqueues the fold of the group
 * slices into grad_ca / grad_cb on `stream`. */
int tide_tm2d_engine_remove(void *engine, int64_t job_id, void *stream);
int tide_tm2d_engine_destroy(void *engine);

/* ------------------------------------------------------------------- 3-D
 * Fields are [n_shots, nz, ny, nx] (x fastest).  `fields` / `lambda` point
 * to 18 device arrays in the reference's state order:
 *   ex, ey, ez, hx, hy, hz,
 *   m_hz_y, m_hy_z, m_hx_z, m_hz_x, m_hy_x, m_hx_y,   (E-update memories)
 *   m_ey_z, m_ez_y, m_ez_x, m_ex_z, m_ex_y, m_ey_x    (H-update memories)
 * (the adjoint uses the same slots for the lambda counterparts).
 * `prof` points to 18 profile arrays in the reference's order
 *   az, bz, azh, bzh, ay, by, ayh, byh, ax, bx, axh, bxh,
 *   kz, kzh, ky, kyh, kx, kxh;
 * `pml` to int64 {z0, y0, x0, z1, y1, x1}; components are 0 x, 1 y, 2 z.
 * `stores` points to 6 snapshot buffers (Ex, Ey, Ez, curl x, y, z; NULL
 * entries are not written); the snapshot of step t (t a multiple of
 * step_ratio) sits at (t - store_t0) / step_ratio * store_step_bytes, so a
 * caller can recompute one checkpoint segment into a segment-sized buffer.
 * storage_format 0 = T payload, 1 = BF16 (fp32 fields only).
 */

/*
 * 3-D Born forward with full-size CPML memories and pointer arrays (the
 * layout of tide_em3d_*_forward): fields / dfields are the 18 background /
 * scattered arrays, advanced in place; otherwise as
 * maxwell_3d_2_*_born_forward_cuda.  No reference counterpart (the host
 * side of Maxwell3D.linearize().jvp calls it).
 */
int tide_em3d_float_born_forward(
    float const *ca, float const *cb, float const *cq, float const *dca,
    float const *dcb, float const *f0, float const *df, float *const *fields,
    float *const *dfields, float *r, float *background_r,
    float const *const *prof, int64_t const *sources_i,
    int64_t const *receivers_i, float rdz, float rdy, float rdx, int64_t nt,
    int64_t n_shots, int64_t nz, int64_t ny, int64_t nx, int64_t n_src,
    int64_t n_rec, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,
    int64_t const *pml, int64_t source_component, int64_t receiver_component,
    int64_t device, void *compute_stream);
int tide_em3d_double_born_forward(
    double const *ca, double const *cb, double const *cq, double const *dca,
    double const *dcb, double const *f0, double const *df,
    double *const *fields, double *const *dfields, double *r,
    double *background_r, double const *const *prof,
    int64_t const *sources_i, int64_t const *receivers_i, double rdz,
    double rdy, double rdx, int64_t nt, int64_t n_shots, int64_t nz,
    int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec, bool ca_b,
    bool cb_b, bool cq_b, int64_t start_t, int64_t const *pml,
    int64_t source_component, int64_t receiver_component, int64_t device,
    void *compute_stream);

/*
 * Forward propagation over [start_t, start_t + nt), in place.
 * Replaces maxwell_3d_2_{dtype}_forward_cuda and
 * maxwell_3d_2_{dtype}_forward_with_storage_cuda: reference ABI
 * src/tide/backend_utils.py:487-521 (_3D_FORWARD_SPEC,
 * _3D_FORWARD_WITH_STORAGE_SPEC), CUDA implementation
 * src/tide/csrc/em3d/maxwell_3d_cuda_inst.cu:3012-3765; pointer arrays
 * replace the 18 + 18 individual pointers and store_t0 is added.
 */
int tide_em3d_float_forward(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *const *fields, float *r, float const *const *prof,
    int64_t const *sources_i, int64_t const *receivers_i, float rdz, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t const *pml, int64_t source_component,
    int64_t receiver_component, void *const *stores, int64_t storage_format,
    int64_t store_step_bytes, int64_t store_t0, int64_t device,
    void *compute_stream);

int tide_em3d_double_forward(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *const *fields, double *r, double const *const *prof,
    int64_t const *sources_i, int64_t const *receivers_i, double rdz,
    double rdy, double rdx, int64_t nt, int64_t n_shots, int64_t nz,
    int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, bool ca_batched,
    bool cb_batched, bool cq_batched, int64_t start_t, int64_t const *pml,
    int64_t source_component, int64_t receiver_component,
    void *const *stores, int64_t storage_format, int64_t store_step_bytes,
    int64_t store_t0, int64_t device, void *compute_stream);

/*
 * Exact adjoint over steps start_t - 1 down to start_t - nt (receivers
 * inject grad_r[t] into lambda of the receiver component, sources read
 * grad_f[t]), accumulating dJ/dca and dJ/dcb into FP64 arrays (added to;
 * [nz, ny, nx] summed over shots, or per shot when batched; NULL when not
 * wanted).  Replaces maxwell_3d_2_{dtype}_backward_cuda: reference ABI
 * src/tide/backend_utils.py:523-555 (_3D_BACKWARD_SPEC), CUDA adjoint
 * src/tide/csrc/em3d/maxwell_3d_cuda_inst.cu:1420-2187, 2300-2450.
 */
int tide_em3d_float_backward_f64acc(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *const *lambda, void const *const *stores, int64_t storage_format,
    int64_t store_step_bytes, int64_t store_t0, float *grad_f,
    double *grad_ca, double *grad_cb, float const *const *prof,
    int64_t const *sources_i, int64_t const *receivers_i, float rdz, float rdy,
    float rdx, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t const *pml, int64_t source_component,
    int64_t receiver_component, int64_t device, void *compute_stream);

int tide_em3d_double_backward_f64acc(
    double const *ca, double const *cb, double const *cq,
    double const *grad_r, double *const *lambda, void const *const *stores,
    int64_t storage_format, int64_t store_step_bytes, int64_t store_t0,
    double *grad_f, double *grad_ca, double *grad_cb,
    double const *const *prof, int64_t const *sources_i,
    int64_t const *receivers_i, double rdz, double rdy, double rdx,
    int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t const *pml, int64_t source_component,
    int64_t receiver_component, int64_t device, void *compute_stream);

/*
 * The reference's own 3-D entry points, argument for argument
 * (src/tide/backend_utils.py:446-555: _3D_FORWARD_SPEC,
 * _3D_FORWARD_WITH_STORAGE_SPEC, _3D_BACKWARD_SPEC; symbol names :771-781;
 * CUDA implementations src/tide/csrc/em3d/maxwell_3d_cuda_inst.cu:3012-3289
 * (forward), :3290-3765 (forward_with_storage), :4542-5260 (backward)), so
 * the reference's maxwell3d_cuda.py binds them unchanged.  They run
 * tide_em3d_* and return a status.  Snapshots are device resident: step t
 * (t a multiple of step_ratio) goes to slot t / step_ratio of each *_s1
 * buffer (shot_bytes_uncomp * n_shots bytes per slot); *_s3 and *_fn
 * (host/disk staging) must be unused (storage_mode 0 or 3).  The backward
 * adds dJ/dca, dJ/dcb to grad_ca / grad_cb (accumulated in FP64), writes
 * grad_eps / grad_sigma (when non-NULL) from the totals with the
 * reference's closed-form chain rule, and clears grad_f, grad_ca, grad_cb,
 * grad_eps, grad_sigma first when zero_grad_on_entry is set.
 * adjoint_memory_scratch, grad_ca_shot, grad_cb_shot, n_threads and
 * execution_backend are unused; m_lambda_X_Y is the adjoint memory of the
 * forward memory m_X_Y.  CPML memories, forward and adjoint, are in the
 * reference's compact layout: [n_shots, ..] with only the strips [0, lo)
 * and [max(lo, hi - 1), n) of the memory's own axis stored.  Mixed ca_batched / cb_batched with both
 * gradients requested returns status 1.
 */
int maxwell_3d_2_float_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ex, float *ey, float *ez, float *hx, float *hy, float *hz,
    float *m_hz_y, float *m_hy_z, float *m_hx_z, float *m_hz_x, float *m_hy_x,
    float *m_hx_y, float *m_ey_z, float *m_ez_y, float *m_ez_x, float *m_ex_z,
    float *m_ex_y, float *m_ey_x, float const *debye_a, float const *debye_b,
    float const *debye_cp, float *pol_ex, float *pol_ey, float *pol_ez,
    float *ex_prev, float *ey_prev, float *ez_prev, float *r, int64_t n_poles,
    float const *az, float const *bz, float const *azh, float const *bzh,
    float const *ay, float const *by, float const *ayh, float const *byh,
    float const *ax, float const *bx, float const *axh, float const *bxh,
    float const *kz, float const *kzh, float const *ky, float const *kyh,
    float const *kx, float const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, float rdz, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_z0, int64_t pml_y0,
    int64_t pml_x0, int64_t pml_z1, int64_t pml_y1, int64_t pml_x1,
    int64_t source_component, int64_t receiver_component, int64_t n_threads,
    int64_t device, int64_t execution_backend, void *compute_stream);

int maxwell_3d_2_double_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ex, double *ey, double *ez, double *hx, double *hy, double *hz,
    double *m_hz_y, double *m_hy_z, double *m_hx_z, double *m_hz_x,
    double *m_hy_x, double *m_hx_y, double *m_ey_z, double *m_ez_y,
    double *m_ez_x, double *m_ex_z, double *m_ex_y, double *m_ey_x,
    double const *debye_a, double const *debye_b, double const *debye_cp,
    double *pol_ex, double *pol_ey, double *pol_ez, double *ex_prev,
    double *ey_prev, double *ez_prev, double *r, int64_t n_poles,
    double const *az, double const *bz, double const *azh, double const *bzh,
    double const *ay, double const *by, double const *ayh, double const *byh,
    double const *ax, double const *bx, double const *axh, double const *bxh,
    double const *kz, double const *kzh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdz, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool has_dispersion, bool ca_batched, bool cb_batched,
    bool cq_batched, int64_t start_t, int64_t pml_z0, int64_t pml_y0,
    int64_t pml_x0, int64_t pml_z1, int64_t pml_y1, int64_t pml_x1,
    int64_t source_component, int64_t receiver_component, int64_t n_threads,
    int64_t device, int64_t execution_backend, void *compute_stream);

/*
 * Born forward, 3-D (reference src/tide/backend_utils.py:556-577,
 * _3D_BORN_FORWARD_SPEC; CPU semantics src/tide/csrc/em3d/
 * maxwell_3d_inst.cpp:1100-1527): the background (ex .. m_ey_x) and the
 * Born-scattered field (dex .. dm_ey_x) advance together, the scattered E
 * update adding dca E^n + dcb curl H^n of the background (dca, dcb
 * [nz, ny, nx]); f0 is injected into the background and df into the
 * scattered field, background_r records the background and r the scattered
 * field.  CPML memories in the reference's compact layout, as above; dt,
 * step_ratio, n_threads and execution_backend are unused (the reference's
 * CPU version ignores step_ratio too).  Returns a status.
 */
int maxwell_3d_2_float_born_forward_cuda(
    float const *ca, float const *cb, float const *cq, float const *dca,
    float const *dcb, float const *f0, float const *df, float *ex, float *ey,
    float *ez, float *hx, float *hy, float *hz, float *m_hz_y, float *m_hy_z,
    float *m_hx_z, float *m_hz_x, float *m_hy_x, float *m_hx_y, float *m_ey_z,
    float *m_ez_y, float *m_ez_x, float *m_ex_z, float *m_ex_y, float *m_ey_x,
    float *dex, float *dey, float *dez, float *dhx, float *dhy, float *dhz,
    float *dm_hz_y, float *dm_hy_z, float *dm_hx_z, float *dm_hz_x,
    float *dm_hy_x, float *dm_hx_y, float *dm_ey_z, float *dm_ez_y,
    float *dm_ez_x, float *dm_ex_z, float *dm_ex_y, float *dm_ey_x, float *r,
    float *background_r, float const *az, float const *bz, float const *azh,
    float const *bzh, float const *ay, float const *by, float const *ayh,
    float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *kz, float const *kzh, float const *ky,
    float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdz, float rdy,
    float rdx, float dt, int64_t nt, int64_t n_shots, int64_t nz, int64_t ny,
    int64_t nx, int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_z0, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_z1, int64_t pml_y1, int64_t pml_x1, int64_t source_component,
    int64_t receiver_component, int64_t n_threads, int64_t device,
    int64_t execution_backend, void *compute_stream);

int maxwell_3d_2_double_born_forward_cuda(
    double const *ca, double const *cb, double const *cq, double const *dca,
    double const *dcb, double const *f0, double const *df, double *ex,
    double *ey, double *ez, double *hx, double *hy, double *hz, double *m_hz_y,
    double *m_hy_z, double *m_hx_z, double *m_hz_x, double *m_hy_x,
    double *m_hx_y, double *m_ey_z, double *m_ez_y, double *m_ez_x,
    double *m_ex_z, double *m_ex_y, double *m_ey_x, double *dex, double *dey,
    double *dez, double *dhx, double *dhy, double *dhz, double *dm_hz_y,
    double *dm_hy_z, double *dm_hx_z, double *dm_hz_x, double *dm_hy_x,
    double *dm_hx_y, double *dm_ey_z, double *dm_ez_y, double *dm_ez_x,
    double *dm_ex_z, double *dm_ex_y, double *dm_ey_x, double *r,
    double *background_r, double const *az, double const *bz,
    double const *azh, double const *bzh, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *kz, double const *kzh,
    double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdz,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots, int64_t nz,
    int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, bool ca_batched,
    bool cb_batched, bool cq_batched, int64_t start_t, int64_t pml_z0,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_z1, int64_t pml_y1,
    int64_t pml_x1, int64_t source_component, int64_t receiver_component,
    int64_t n_threads, int64_t device, int64_t execution_backend,
    void *compute_stream);

int maxwell_3d_2_float_forward_with_storage_cuda(
    float const *ca, float const *cb, float const *cq, float const *f,
    float *ex, float *ey, float *ez, float *hx, float *hy, float *hz,
    float *m_hz_y, float *m_hy_z, float *m_hx_z, float *m_hz_x, float *m_hy_x,
    float *m_hx_y, float *m_ey_z, float *m_ez_y, float *m_ez_x, float *m_ex_z,
    float *m_ex_y, float *m_ey_x, float *r, void *ex_s1, void *ex_s3,
    char const *const *ex_fn, void *ey_s1, void *ey_s3,
    char const *const *ey_fn, void *ez_s1, void *ez_s3,
    char const *const *ez_fn, void *cx_s1, void *cx_s3,
    char const *const *cx_fn, void *cy_s1, void *cy_s3,
    char const *const *cy_fn, void *cz_s1, void *cz_s3,
    char const *const *cz_fn, float const *az, float const *bz,
    float const *azh, float const *bzh, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx,
    float const *axh, float const *bxh, float const *kz, float const *kzh,
    float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdz,
    float rdy, float rdx, float dt, int64_t nt, int64_t n_shots, int64_t nz,
    int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_z0, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_z1, int64_t pml_y1, int64_t pml_x1, int64_t source_component,
    int64_t receiver_component, int64_t n_threads, int64_t device,
    int64_t execution_backend, void *compute_stream, void *storage_stream);

int maxwell_3d_2_double_forward_with_storage_cuda(
    double const *ca, double const *cb, double const *cq, double const *f,
    double *ex, double *ey, double *ez, double *hx, double *hy, double *hz,
    double *m_hz_y, double *m_hy_z, double *m_hx_z, double *m_hz_x,
    double *m_hy_x, double *m_hx_y, double *m_ey_z, double *m_ez_y,
    double *m_ez_x, double *m_ex_z, double *m_ex_y, double *m_ey_x, double *r,
    void *ex_s1, void *ex_s3, char const *const *ex_fn, void *ey_s1,
    void *ey_s3, char const *const *ey_fn, void *ez_s1, void *ez_s3,
    char const *const *ez_fn, void *cx_s1, void *cx_s3,
    char const *const *cx_fn, void *cy_s1, void *cy_s3,
    char const *const *cy_fn, void *cz_s1, void *cz_s3,
    char const *const *cz_fn, double const *az, double const *bz,
    double const *azh, double const *bzh, double const *ay, double const *by,
    double const *ayh, double const *byh, double const *ax, double const *bx,
    double const *axh, double const *bxh, double const *kz, double const *kzh,
    double const *ky, double const *kyh, double const *kx, double const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, double rdz,
    double rdy, double rdx, double dt, int64_t nt, int64_t n_shots,
    int64_t nz, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_z0, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_z1, int64_t pml_y1, int64_t pml_x1, int64_t source_component,
    int64_t receiver_component, int64_t n_threads, int64_t device,
    int64_t execution_backend, void *compute_stream, void *storage_stream);

int maxwell_3d_2_float_backward_cuda(
    float const *ca, float const *cb, float const *cq, float const *grad_r,
    float *lambda_ex, float *lambda_ey, float *lambda_ez, float *lambda_hx,
    float *lambda_hy, float *lambda_hz, float *m_lambda_ey_z,
    float *m_lambda_ez_y, float *m_lambda_ez_x, float *m_lambda_ex_z,
    float *m_lambda_ex_y, float *m_lambda_ey_x, float *m_lambda_hz_y,
    float *m_lambda_hy_z, float *m_lambda_hx_z, float *m_lambda_hz_x,
    float *m_lambda_hy_x, float *m_lambda_hx_y, float *adjoint_memory_scratch,
    void *ex_s1, void *ex_s3, char const *const *ex_fn, void *ey_s1,
    void *ey_s3, char const *const *ey_fn, void *ez_s1, void *ez_s3,
    char const *const *ez_fn, void *cx_s1, void *cx_s3,
    char const *const *cx_fn, void *cy_s1, void *cy_s3,
    char const *const *cy_fn, void *cz_s1, void *cz_s3,
    char const *const *cz_fn, float *grad_f, float *grad_ca, float *grad_cb,
    float *grad_eps, float *grad_sigma, float *grad_ca_shot,
    float *grad_cb_shot, bool zero_grad_on_entry, float const *az,
    float const *bz, float const *azh, float const *bzh, float const *ay,
    float const *by, float const *ayh, float const *byh, float const *ax,
    float const *bx, float const *axh, float const *bxh, float const *kz,
    float const *kzh, float const *ky, float const *kyh, float const *kx,
    float const *kxh, int64_t const *sources_i, int64_t const *receivers_i,
    float rdz, float rdy, float rdx, float dt, int64_t nt, int64_t n_shots,
    int64_t nz, int64_t ny, int64_t nx, int64_t n_sources_per_shot,
    int64_t n_receivers_per_shot, int64_t step_ratio, int64_t storage_mode,
    int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_requires_grad,
    bool cb_requires_grad, bool ca_batched, bool cb_batched, bool cq_batched,
    int64_t start_t, int64_t pml_z0, int64_t pml_y0, int64_t pml_x0,
    int64_t pml_z1, int64_t pml_y1, int64_t pml_x1, int64_t source_component,
    int64_t receiver_component, int64_t n_threads, int64_t device,
    int64_t execution_backend, void *compute_stream, void *storage_stream);

int maxwell_3d_2_double_backward_cuda(
    double const *ca, double const *cb, double const *cq,
    double const *grad_r, double *lambda_ex, double *lambda_ey,
    double *lambda_ez, double *lambda_hx, double *lambda_hy,
    double *lambda_hz, double *m_lambda_ey_z, double *m_lambda_ez_y,
    double *m_lambda_ez_x, double *m_lambda_ex_z, double *m_lambda_ex_y,
    double *m_lambda_ey_x, double *m_lambda_hz_y, double *m_lambda_hy_z,
    double *m_lambda_hx_z, double *m_lambda_hz_x, double *m_lambda_hy_x,
    double *m_lambda_hx_y, double *adjoint_memory_scratch, void *ex_s1,
    void *ex_s3, char const *const *ex_fn, void *ey_s1, void *ey_s3,
    char const *const *ey_fn, void *ez_s1, void *ez_s3,
    char const *const *ez_fn, void *cx_s1, void *cx_s3,
    char const *const *cx_fn, void *cy_s1, void *cy_s3,
    char const *const *cy_fn, void *cz_s1, void *cz_s3,
    char const *const *cz_fn, double *grad_f, double *grad_ca,
    double *grad_cb, double *grad_eps, double *grad_sigma,
    double *grad_ca_shot, double *grad_cb_shot, bool zero_grad_on_entry,
    double const *az, double const *bz, double const *azh, double const *bzh,
    double const *ay, double const *by, double const *ayh, double const *byh,
    double const *ax, double const *bx, double const *axh, double const *bxh,
    double const *kz, double const *kzh, double const *ky, double const *kyh,
    double const *kx, double const *kxh, int64_t const *sources_i,
    int64_t const *receivers_i, double rdz, double rdy, double rdx, double dt,
    int64_t nt, int64_t n_shots, int64_t nz, int64_t ny, int64_t nx,
    int64_t n_sources_per_shot, int64_t n_receivers_per_shot,
    int64_t step_ratio, int64_t storage_mode, int64_t storage_format,
    int64_t shot_bytes_uncomp, bool ca_requires_grad, bool cb_requires_grad,
    bool ca_batched, bool cb_batched, bool cq_batched, int64_t start_t,
    int64_t pml_z0, int64_t pml_y0, int64_t pml_x0, int64_t pml_z1,
    int64_t pml_y1, int64_t pml_x1, int64_t source_component,
    int64_t receiver_component, int64_t n_threads, int64_t device,
    int64_t execution_backend, void *compute_stream, void *storage_stream);

/*
 * Launch shape of the 3-D step kernels: z planes marched per wavefront
 * (forward 4/8/16, adjoint 4/8) and shots per workgroup (1..8, both
 * passes).  Defaults: 8 / 4 planes, 1 shot per forward workgroup, 8 per
 * adjoint workgroup.  No reference counterpart.
 */
int tide_em3d_set_tuning(int fwd_z, int adj_z, int group);

/*
 * Shots per workgroup of the 3-D forward kernels alone (1..8), after
 * tide_em3d_set_tuning.  No reference counterpart.
 */
int tide_em3d_set_forward_group(int group);

/*
 * Register budget of the fp32 K-step (3- or 4-step) 2-D adjoint: 4 (the
 * default) holds the kernel to four waves per SIMD (128 VGPRs, spilling the
 * rest; workgroups of up to 256 threads), 0 leaves it to the compiler.
 * fp64 kernels always use the compiler's count.  Results do not depend on
 * it.  No reference counterpart.
 */
int tide_tm2d_set_adjoint_waves(int waves);
int tide_tm2d_get_adjoint_waves(int *out);

/*
 * One-launch 3-D forward step (the H and E updates of a time step in one
 * kernel, fields and H-phase CPML memories ping-ponged through a second
 * copy): rows per wavefront 2 or 4 (0: the two-kernel step) and z planes
 * marched per wavefront 8 or 16.  Results are bit-identical to the
 * two-kernel step.  tide_em3d_get_fused writes {rows, planes}.  No
 * reference counterpart.
 */
int tide_em3d_set_fused(int rows, int planes);
int tide_em3d_get_fused(int *out2);

/*
 * Workgroup order of the 3-D kernels: 0 runs all tiles of one shot group
 * before the next group, 1 runs the shot groups of one tile next to each
 * other (same XCD, shared L2 for the coefficient lines).  Results do not
 * depend on it.  No reference counterpart.
 */
int tide_em3d_set_tile_order(int shot_fast);

#ifdef __cplusplus
}
#endif

#endif /* TIDE_HIP_H */
