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

#include "tm2d_device.h"
#include "tm2d_fused.h"
#include "tm2d_launch.h"
#include "tm2d_launch_k.h"
#include "tide_hip.h"
#include "tide_storage.h"
#include "tm2d_highorder.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace tide_hip {
namespace tm2d {

// ------------------------------------------------------------------ driver

template <typename T> struct Problem {
  Geom<T> g;
  const T *ca, *cb, *cq;
  const T *prof[12];  // ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
  int32_t n_shots;
  int64_t n_src, n_rec;
};

Tuning g_tuning;

// Optional HIP-event timing of the time loops (launches only, no set-up,
// fold or copies): a caller turns it on, runs, then collects the summed
// milliseconds and launch counts per pass.
struct LoopTimes {
  bool enabled = false;
  struct Rec {
    hipEvent_t start, stop;
    int64_t launches;
  };
  std::vector<Rec> recs[2];  // 0 forward, 1 adjoint
};
LoopTimes g_loop_times;

struct LoopTimer {
  int pass;
  hipStream_t stream;
  hipEvent_t start = nullptr, stop = nullptr;
  LoopTimer(int pass_, hipStream_t s) : pass(pass_), stream(s) {
    if (!g_loop_times.enabled) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
      (void)hipGetLastError();
      start = stop = nullptr;
      return;
    }
    (void)hipEventRecord(start, stream);
  }
  void finish(int64_t launches) {
    if (start == nullptr) return;
    (void)hipEventRecord(stop, stream);
    g_loop_times.recs[pass].push_back({start, stop, launches});
    start = stop = nullptr;
  }
};

constexpr int kMaxLoopTiles = 4;  // wavefronts per workgroup of the shot-looping kernels

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

// Scratch device memory of one call: stream-ordered allocations freed when
// the call returns, or, with an arena, consecutive 256-byte aligned pieces
// of caller-provided device memory (nothing allocated or freed).
struct Scratch {
  std::vector<void *> bufs;
  hipStream_t stream;
  char *arena = nullptr;
  size_t cap = 0, used = 0;
  explicit Scratch(hipStream_t s, char *arena_ = nullptr, size_t cap_ = 0)
      : stream(s), arena(arena_), cap(cap_) {}
  ~Scratch() {
    for (void *p : bufs) (void)hipFreeAsync(p, stream);
  }
  template <typename U> int alloc(U **out, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(U);
    if (arena != nullptr) {
      const size_t at = (used + 255) & ~(size_t)255;
      if (at + bytes > cap) {
        set_error("the call needs more scratch memory than the fixed scratch array holds (" +
                  std::to_string(cap >> 20) + " MiB)");
        return kInvalidArgument;
      }
      *out = (U *)(arena + at);
      used = at + bytes;
      return kOk;
    }
    void *p = nullptr;
    TIDE_HIP_TRY(hipMallocAsync(&p, bytes, stream));
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

// halo2: the table of the two-step kernels (60-column tiles, lane offset 2,
// halo copies) in `dpt2`.
// With a Scratch arena the tables are placed in it instead of allocated.
inline int put_table(const PointTable &t, DevicePointTable &d, Scratch *arena_scr,
                     hipStream_t stream) {
  if (arena_scr == nullptr || arena_scr->arena == nullptr) return d.upload(t, stream);
  int32_t *rs, *sg;
  int2 *en;
  TIDE_CHECK(arena_scr->alloc(&rs, DevicePointTable::row_slot_count(t)));
  TIDE_CHECK(arena_scr->alloc(&sg, DevicePointTable::seg_count(t)));
  TIDE_CHECK(arena_scr->alloc(&en, DevicePointTable::ent_count(t)));
  return d.place(t, rs, sg, en, stream);
}

// The table of the K-step kernels (64-2K column tiles, lane offset K, halo
// copies) in `dptk`.
template <typename T>
int upload_points_k(const Problem<T> &p, const int64_t *src, const int64_t *rec, int k,
                    DevicePointTable &dptk, hipStream_t stream, Scratch *arena_scr = nullptr) {
  std::vector<int64_t> hsrc, hrec;
  TIDE_CHECK(fetch_indices(src, (int64_t)p.n_shots * p.n_src, hsrc, stream));
  TIDE_CHECK(fetch_indices(rec, (int64_t)p.n_shots * p.n_rec, hrec, stream));
  PointTable table;
  if (!build_point_table_halo(hsrc.data(), p.n_src, hrec.data(), p.n_rec, p.n_shots, p.g.ny,
                              p.g.nx, 64 - 2 * k, k, table))
    return kInvalidArgument;
  return put_table(table, dptk, arena_scr, stream);
}

template <typename T>
int upload_points(const Problem<T> &p, const int64_t *src, const int64_t *rec,
                  DevicePointTable &dpt, hipStream_t stream, DevicePointTable *dpt2 = nullptr,
                  Scratch *arena_scr = nullptr) {
  std::vector<int64_t> hsrc, hrec;
  TIDE_CHECK(fetch_indices(src, (int64_t)p.n_shots * p.n_src, hsrc, stream));
  TIDE_CHECK(fetch_indices(rec, (int64_t)p.n_shots * p.n_rec, hrec, stream));
  PointTable table;
  if (!build_point_table(hsrc.data(), p.n_src, hrec.data(), p.n_rec, p.n_shots, p.g.ny,
                         p.g.nx, kOwn, 1, table))
    return kInvalidArgument;
  TIDE_CHECK(put_table(table, dpt, arena_scr, stream));
  if (dpt2 == nullptr) return kOk;
  PointTable table2;
  if (!build_point_table_halo(hsrc.data(), p.n_src, hrec.data(), p.n_rec, p.n_shots, p.g.ny,
                              p.g.nx, kOwn2, 2, table2))
    return kInvalidArgument;
  return put_table(table2, *dpt2, arena_scr, stream);
}

template <typename T>
int forward_impl(Problem<T> p, T *ey, T *hx, T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z,
                 T *m_hz_x, const T *f, T *r, const int64_t *src, const int64_t *rec,
                 int64_t nt, int64_t start_t, int64_t step_ratio, void *ey_store,
                 void *curl_store, int fmt, size_t store_step_bytes, hipStream_t stream,
                 char *arena = nullptr, size_t arena_cap = 0,
                 SnapshotStage *ey_stage = nullptr, SnapshotStage *curl_stage = nullptr) {
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
  const bool staged = (ey_stage && ey_stage->active()) || (curl_stage && curl_stage->active());
  // K-step launches (K = 3, 4) need K ring slots at once, more than a
  // host-staged ring holds: staged calls use two-step launches.
  const int kblk = (g_tuning.fwd_steps >= 3 && nt >= g_tuning.fwd_steps && !staged)
                       ? g_tuning.fwd_steps : 0;
  const bool fused = g_tuning.fwd_steps >= 2 && nt >= 2;
  Scratch scr(stream, arena, arena_cap);
  DevicePointTable dpt, dpt2, dptk;
  TIDE_CHECK(upload_points(p, src, rec, dpt, stream, fused ? &dpt2 : nullptr, &scr));
  if (kblk) TIDE_CHECK(upload_points_k(p, src, rec, kblk, dptk, stream, &scr));
  TIDE_CHECK(pack_profiles(p, scr, stream));
  T *ey2, *hx2, *hz2, *mex2, *mez2, *mhzx2 = nullptr, *mhxz2 = nullptr;
  TIDE_CHECK(scr.alloc(&ey2, (size_t)total));
  TIDE_CHECK(scr.alloc(&hx2, (size_t)total));
  TIDE_CHECK(scr.alloc(&hz2, (size_t)total));
  TIDE_CHECK(scr.alloc(&mex2, (size_t)total));
  TIDE_CHECK(scr.alloc(&mez2, (size_t)total));
  const size_t bytes = (size_t)total * sizeof(T);
  // CPML memories are only written inside the strips; both copies start
  // equal so the untouched cells agree whichever buffer ends up current.
  TIDE_HIP_TRY(hipMemcpyAsync(mex2, m_ey_x, bytes, hipMemcpyDeviceToDevice, stream));
  TIDE_HIP_TRY(hipMemcpyAsync(mez2, m_ey_z, bytes, hipMemcpyDeviceToDevice, stream));
  if (fused) {  // the two-step kernels also ping-pong the E-phase memories
    TIDE_CHECK(scr.alloc(&mhzx2, (size_t)total));
    TIDE_CHECK(scr.alloc(&mhxz2, (size_t)total));
    TIDE_HIP_TRY(hipMemcpyAsync(mhzx2, m_hz_x, bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(mhxz2, m_hx_z, bytes, hipMemcpyDeviceToDevice, stream));
  }

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
  const bool loop = g_tuning.fwd_loop != 0 && !fused;
  const int group = loop ? std::min<int>(g_tuning.fwd_group, kMaxLoopTiles)
                         : std::min<int>(p.n_shots, g_tuning.fwd_group);
  a.n_yt = (p.g.ny + rows - 1) / rows;
  const int64_t n_tiles = (int64_t)dpt.n_xt * a.n_yt;
  const int64_t n_groups = (p.n_shots + group - 1) / group;
  const dim3 block(64, (unsigned)group);
  const dim3 grid((unsigned)(loop ? (n_tiles + group - 1) / group : n_tiles * n_groups));

  T *cur[5] = {ey, hx, hz, m_ey_x, m_ey_z};
  T *nxt[5] = {ey2, hx2, hz2, mex2, mez2};
  T *cur_e[2] = {m_hz_x, m_hx_z};  // E-phase memories (swapped by two-step launches)
  T *nxt_e[2] = {mhzx2, mhxz2};
  const bool storing = ey_store != nullptr || curl_store != nullptr;
  // Device address of the snapshot of step t: the history buffer, or the
  // ring slot of a host-staged stream.
  auto store_at = [&](int64_t t, void *buf, SnapshotStage *stage = nullptr) -> void * {
    if (buf == nullptr || !storing || (t % step_ratio) != 0) return nullptr;
    if (stage != nullptr && stage->active()) return stage->slot(t / step_ratio);
    return (char *)buf + (size_t)(t / step_ratio) * store_step_bytes;
  };
  // Host-staged streams: a ring slot is written only after its previous
  // content has been copied out, and copied out after the launch.
  SnapshotStage *stages[2] = {ey_stage, curl_stage};
  void *bufs[2] = {ey_store, curl_store};
  auto stage_steps = [&](int64_t t0, int n, bool after) -> int {
    for (int k = 0; k < 2; ++k) {
      if (stages[k] == nullptr || !stages[k]->active() || bufs[k] == nullptr) continue;
      for (int64_t tt = t0; tt < t0 + n; ++tt) {
        if ((tt % step_ratio) != 0) continue;
        const int64_t idx = tt / step_ratio;
        const int rc = after ? stages[k]->after_write(idx) : stages[k]->before_write(idx);
        if (rc != kOk) return rc;
      }
    }
    return kOk;
  };
  Fwd2Args<T> a2;
  const bool split2 = g_tuning.fwd_split != 0;
  int64_t groups2 = 1;
  const dim3 block2(64, (unsigned)std::min<int>(p.n_shots, g_tuning.fwd_group));
  const int rows2 = std::min<int>(rows_for<T>(g_tuning.fwd_rows), sizeof(T) == 4 ? 8 : 4);
  if (fused) {
    std::memset(&a2, 0, sizeof(a2));
    a2.g = p.g;
    a2.pts = Points{dpt2.row_slot, dpt2.seg, dpt2.ent, dpt2.n_xt};
    a2.ca = p.ca;
    a2.cb = p.cb;
    a2.cq = p.cq;
    a2.n_shots = p.n_shots;
    a2.n_yt = (p.g.ny + rows2 - 1) / rows2;
    a2.split = make_split(p.g, dpt2.n_xt, a2.n_yt, rows2);
    groups2 = (p.n_shots + block2.y - 1) / block2.y;
  }
  int64_t t = start_t;
  LoopTimer timer(0, stream);
  int64_t launches = 0;
  if (kblk) {
    FwdKArgs<T> ak;
    std::memset(&ak, 0, sizeof(ak));
    ak.g = p.g;
    ak.pts = Points{dptk.row_slot, dptk.seg, dptk.ent, dptk.n_xt};
    ak.ca = p.ca;
    ak.cb = p.cb;
    ak.cq = p.cq;
    ak.n_shots = p.n_shots;
    const int rowsk = rows2 >= 8 ? 8 : (rows2 >= 4 ? 4 : 2);
    ak.n_yt = (p.g.ny + rowsk - 1) / rowsk;
    ak.split = make_split_k(p.g, dptk.n_xt, ak.n_yt, rowsk, kblk);
    for (; t + kblk - 1 < start_t + nt; t += kblk, ++launches) {
      ak.ey_i = cur[0]; ak.hx_i = cur[1]; ak.hz_i = cur[2];
      ak.mex_i = cur[3]; ak.mez_i = cur[4];
      ak.mhzx_i = cur_e[0]; ak.mhxz_i = cur_e[1];
      ak.ey_o = nxt[0]; ak.hx_o = nxt[1]; ak.hz_o = nxt[2];
      ak.mex_o = nxt[3]; ak.mez_o = nxt[4];
      ak.mhzx_o = nxt_e[0]; ak.mhxz_o = nxt_e[1];
      bool st = false;
      for (int j = 0; j < kblk; ++j) {
        ak.f[j] = p.n_src > 0 ? f + (t + j) * n_shots * p.n_src : nullptr;
        ak.r[j] = p.n_rec > 0 ? r + (t + j) * n_shots * p.n_rec : nullptr;
        ak.ey_store[j] = store_at(t + j, ey_store);
        ak.curl_store[j] = store_at(t + j, curl_store);
        st = st || ak.ey_store[j] || ak.curl_store[j];
      }
      if (!st) {
        launch_fwdk<T, 0>(kblk, rowsk, g_tuning.fwd_border_rows, groups2, block2, stream, ak);
      } else if (fmt == 2) {
        if constexpr (sizeof(T) == 4)
          launch_fwdk<T, 2>(kblk, rowsk, g_tuning.fwd_border_rows, groups2, block2, stream, ak);
      } else {
        launch_fwdk<T, 1>(kblk, rowsk, g_tuning.fwd_border_rows, groups2, block2, stream, ak);
      }
      for (int k = 0; k < 5; ++k) std::swap(cur[k], nxt[k]);
      for (int k = 0; k < 2; ++k) std::swap(cur_e[k], nxt_e[k]);
    }
  }
  for (; fused && t + 1 < start_t + nt; t += 2, ++launches) {
    a2.ey_i = cur[0]; a2.hx_i = cur[1]; a2.hz_i = cur[2];
    a2.mex_i = cur[3]; a2.mez_i = cur[4];
    a2.mhzx_i = cur_e[0]; a2.mhxz_i = cur_e[1];
    a2.ey_o = nxt[0]; a2.hx_o = nxt[1]; a2.hz_o = nxt[2];
    a2.mex_o = nxt[3]; a2.mez_o = nxt[4];
    a2.mhzx_o = nxt_e[0]; a2.mhxz_o = nxt_e[1];
    a2.f0 = p.n_src > 0 ? f + t * n_shots * p.n_src : nullptr;
    a2.f1 = p.n_src > 0 ? f + (t + 1) * n_shots * p.n_src : nullptr;
    a2.r0 = p.n_rec > 0 ? r + t * n_shots * p.n_rec : nullptr;
    a2.r1 = p.n_rec > 0 ? r + (t + 1) * n_shots * p.n_rec : nullptr;
    a2.ey_store0 = store_at(t, ey_store, ey_stage);
    a2.curl_store0 = store_at(t, curl_store, curl_stage);
    a2.ey_store1 = store_at(t + 1, ey_store, ey_stage);
    a2.curl_store1 = store_at(t + 1, curl_store, curl_stage);
    const bool st = a2.ey_store0 || a2.curl_store0 || a2.ey_store1 || a2.curl_store1;
    if (st) TIDE_CHECK(stage_steps(t, 2, false));
    if (!st) {
      launch_fwd2<T, 0>(rows2, g_tuning.fwd_border_rows, groups2, block2, stream, split2, a2);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_fwd2<T, 2>(rows2, g_tuning.fwd_border_rows, groups2, block2, stream, split2, a2);
    } else {
      launch_fwd2<T, 1>(rows2, g_tuning.fwd_border_rows, groups2, block2, stream, split2, a2);
    }
    if (st) TIDE_CHECK(stage_steps(t, 2, true));
    for (int k = 0; k < 5; ++k) std::swap(cur[k], nxt[k]);
    for (int k = 0; k < 2; ++k) std::swap(cur_e[k], nxt_e[k]);
  }
  a.mhzx = cur_e[0];
  a.mhxz = cur_e[1];
  for (; t < start_t + nt; ++t, ++launches) {
    a.ey_i = cur[0]; a.hx_i = cur[1]; a.hz_i = cur[2];
    a.mex_i = cur[3]; a.mez_i = cur[4];
    a.ey_o = nxt[0]; a.hx_o = nxt[1]; a.hz_o = nxt[2];
    a.mex_o = nxt[3]; a.mez_o = nxt[4];
    a.f_t = p.n_src > 0 ? f + t * n_shots * p.n_src : nullptr;
    a.r_t = p.n_rec > 0 ? r + t * n_shots * p.n_rec : nullptr;
    const bool st = storing && (t % step_ratio) == 0;
    a.ey_store = st ? store_at(t, ey_store, ey_stage) : nullptr;
    a.curl_store = st ? store_at(t, curl_store, curl_stage) : nullptr;
    if (st) TIDE_CHECK(stage_steps(t, 1, false));
    if (!st) {
      launch_fwd<T, 0>(rows, loop, grid, block, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_fwd<T, 2>(rows, loop, grid, block, stream, a);
    } else {
      launch_fwd<T, 1>(rows, loop, grid, block, stream, a);
    }
    if (st) TIDE_CHECK(stage_steps(t, 1, true));
    for (int k = 0; k < 5; ++k) std::swap(cur[k], nxt[k]);
  }
  timer.finish(launches);
  TIDE_HIP_TRY(hipGetLastError());
  if (cur[0] != ey) {
    TIDE_HIP_TRY(hipMemcpyAsync(ey, cur[0], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(hx, cur[1], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(hz, cur[2], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(m_ey_x, cur[3], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(m_ey_z, cur[4], bytes, hipMemcpyDeviceToDevice, stream));
  }
  if (cur_e[0] != m_hz_x) {
    TIDE_HIP_TRY(hipMemcpyAsync(m_hz_x, cur_e[0], bytes, hipMemcpyDeviceToDevice, stream));
    TIDE_HIP_TRY(hipMemcpyAsync(m_hx_z, cur_e[1], bytes, hipMemcpyDeviceToDevice, stream));
  }
  dpt.release(stream);
  if (fused) dpt2.release(stream);
  if (kblk) dptk.release(stream);
  return kOk;
}

// gca/gcb: FP64 accumulators ([ny,nx] summed over shots, or [n_shots,ny,nx]
// when the coefficient is per shot), added to; null when not required.
template <typename T>
int backward_impl(Problem<T> p, const T *grad_r, T *l_ey, T *l_hx, T *l_hz, T *ml_ey_x,
                  T *ml_ey_z, T *ml_hx_z, T *ml_hz_x, const void *ey_store,
                  const void *curl_store, int fmt, size_t store_step_bytes, T *grad_f,
                  double *gca, double *gcb, const int64_t *src, const int64_t *rec,
                  int64_t nt, int64_t start_t, int64_t step_ratio, hipStream_t stream,
                  SnapshotStage *ey_stage = nullptr, SnapshotStage *curl_stage = nullptr) {
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
  // One gradient layout serves both accumulators: per shot when a
  // coefficient is per shot.  A per-shot ca with a shared cb (or the
  // reverse) would index the shared [ny,nx] buffer per shot.
  if (gca && gcb && p.g.ca_b != p.g.cb_b) {
    set_error("ca and cb gradients need the same batching (ca_batched == cb_batched)");
    return kInvalidArgument;
  }
  const int64_t n_shots = p.n_shots, total = n_shots * p.g.numel;
  const bool staged = (ey_stage && ey_stage->active()) || (curl_stage && curl_stage->active());
  // K-step launches need K ring slots at once (see forward_impl)
  const int kblk = (g_tuning.adj_steps >= 3 && nt >= g_tuning.adj_steps && !staged)
                       ? g_tuning.adj_steps : 0;
  const bool fused = g_tuning.adj_steps >= 2 && nt >= 2;
  DevicePointTable dpt, dpt2, dptk;
  TIDE_CHECK(upload_points(p, src, rec, dpt, stream, fused ? &dpt2 : nullptr));
  if (kblk) TIDE_CHECK(upload_points_k(p, src, rec, kblk, dptk, stream));
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
  const bool loop = g_tuning.adj_loop != 0 && !fused;  // two-step launches group shots
  const int group = loop ? std::min<int>(g_tuning.adj_group, kMaxLoopTiles)
                         : balanced_group(p.n_shots, g_tuning.adj_group);
  a.n_yt = (p.g.ny + rows - 1) / rows;
  const int64_t n_tiles = (int64_t)dpt.n_xt * a.n_yt;
  const int64_t n_groups = loop ? 1 : (p.n_shots + group - 1) / group;
  // More shots than one workgroup holds: each shot group accumulates into
  // its own FP64 slice, folded into gca/gcb in group order at the end.
  a.group_stride = 0;
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
  const dim3 grid((unsigned)(loop ? (n_tiles + group - 1) / group : n_tiles * n_groups));
  const size_t lds = loop ? 0 : 2 * (size_t)group * rows * 64 * sizeof(double);

  // order: ley, lhx, lhz, mx (hz_x), my (hx_z), nx (ey_x), nz (ey_z)
  T *cur[7] = {l_ey, l_hx, l_hz, ml_hz_x, ml_hx_z, ml_ey_x, ml_ey_z};
  T *nxt[7] = {ley2, lhx2, lhz2, m2[0], m2[1], m2[2], m2[3]};
  auto snap = [&](int64_t t, const void *buf, const double *acc,
                  SnapshotStage *stage) -> const void * {
    if (acc == nullptr || (t % step_ratio) != 0) return nullptr;
    if (stage != nullptr && stage->active()) return stage->slot(t / step_ratio);
    return (const char *)buf + (size_t)(t / step_ratio) * store_step_bytes;
  };
  // Host-staged streams: the snapshots of steps t0-n+1 .. t0 are copied into
  // their ring slots before the launch that reads them; after it, the slots
  // are marked used and the next launch's snapshots are prefetched.
  SnapshotStage *stages[2] = {gca ? ey_stage : nullptr, gcb ? curl_stage : nullptr};
  auto stage_steps = [&](int64_t t0, int n, int phase) -> int {
    for (int k = 0; k < 2; ++k) {
      if (stages[k] == nullptr || !stages[k]->active()) continue;
      for (int64_t tt = t0; tt > t0 - n; --tt) {
        if (tt < start_t - nt || (tt % step_ratio) != 0) continue;
        const int64_t idx = tt / step_ratio;
        const int rc = phase == 0 ? stages[k]->before_read(idx)
                       : phase == 1 ? stages[k]->after_read(idx)
                                    : stages[k]->prefetch(idx);
        if (rc != kOk) return rc;
      }
    }
    return kOk;
  };
  int64_t t = start_t - 1;
  LoopTimer timer(1, stream);
  int64_t launches = 0;
  if (kblk) {
    const int rowsk = std::min<int>(rows_for<T>(g_tuning.adj_rows), 4) >= 4 ? 4 : 2;
    AdjKArgs<T> ak;
    std::memset(&ak, 0, sizeof(ak));
    ak.g = p.g;
    ak.pts = Points{dptk.row_slot, dptk.seg, dptk.ent, dptk.n_xt};
    ak.ca = p.ca;
    ak.cb = p.cb;
    ak.cq = p.cq;
    ak.per_shot_grad = a.per_shot_grad;
    ak.step_scale = a.step_scale;
    ak.n_shots = p.n_shots;
    ak.n_yt = (p.g.ny + rowsk - 1) / rowsk;
    ak.group_stride = a.group_stride;
    ak.gca = ga_k;
    ak.gcb = gb_k;
    ak.split = make_split_k(p.g, dptk.n_xt, ak.n_yt, rowsk, kblk);
    for (; t - (kblk - 1) >= start_t - nt; t -= kblk, ++launches) {
      ak.ley_i = cur[0]; ak.lhx_i = cur[1]; ak.lhz_i = cur[2];
      ak.mx_i = cur[3]; ak.my_i = cur[4]; ak.nx_i = cur[5]; ak.nz_i = cur[6];
      ak.ley_o = nxt[0]; ak.lhx_o = nxt[1]; ak.lhz_o = nxt[2];
      ak.mx_o = nxt[3]; ak.my_o = nxt[4]; ak.nx_o = nxt[5]; ak.nz_o = nxt[6];
      bool g_any = false;
      for (int j = 0; j < kblk; ++j) {
        ak.grad_r[j] = p.n_rec > 0 ? grad_r + (t - j) * n_shots * p.n_rec : nullptr;
        ak.grad_f[j] = (p.n_src > 0 && grad_f) ? grad_f + (t - j) * n_shots * p.n_src : nullptr;
        ak.ey_store[j] = snap(t - j, ey_store, gca, ey_stage);
        ak.curl_store[j] = snap(t - j, curl_store, gcb, curl_stage);
        g_any = g_any || ak.ey_store[j] || ak.curl_store[j];
      }
      if (!g_any) {
        launch_adjk<T, 0>(kblk, rowsk, g_tuning.adj_border_rows, n_groups, block, stream, ak);
      } else if (fmt == 2) {
        if constexpr (sizeof(T) == 4)
          launch_adjk<T, 2>(kblk, rowsk, g_tuning.adj_border_rows, n_groups, block, stream, ak);
      } else {
        launch_adjk<T, 1>(kblk, rowsk, g_tuning.adj_border_rows, n_groups, block, stream, ak);
      }
      for (int k = 0; k < 7; ++k) std::swap(cur[k], nxt[k]);
    }
  }
  if (fused) {
    const bool split2 = g_tuning.adj_split != 0;
    const int rows2 = std::min<int>(rows_for<T>(g_tuning.adj_rows), 4);
    Adj2Args<T> a2;
    std::memset(&a2, 0, sizeof(a2));
    a2.g = p.g;
    a2.pts = Points{dpt2.row_slot, dpt2.seg, dpt2.ent, dpt2.n_xt};
    a2.ca = p.ca;
    a2.cb = p.cb;
    a2.cq = p.cq;
    a2.per_shot_grad = a.per_shot_grad;
    a2.step_scale = a.step_scale;
    a2.n_shots = p.n_shots;
    a2.n_yt = (p.g.ny + rows2 - 1) / rows2;
    a2.group_stride = a.group_stride;
    a2.gca = ga_k;
    a2.gcb = gb_k;
    const dim3 block2 = block;  // same shot groups as the single-step launches
    a2.split = make_split(p.g, dpt2.n_xt, a2.n_yt, rows2);
    for (; t - 1 >= start_t - nt; t -= 2, ++launches) {
      a2.ley_i = cur[0]; a2.lhx_i = cur[1]; a2.lhz_i = cur[2];
      a2.mx_i = cur[3]; a2.my_i = cur[4]; a2.nx_i = cur[5]; a2.nz_i = cur[6];
      a2.ley_o = nxt[0]; a2.lhx_o = nxt[1]; a2.lhz_o = nxt[2];
      a2.mx_o = nxt[3]; a2.my_o = nxt[4]; a2.nx_o = nxt[5]; a2.nz_o = nxt[6];
      a2.grad_r0 = p.n_rec > 0 ? grad_r + t * n_shots * p.n_rec : nullptr;
      a2.grad_r1 = p.n_rec > 0 ? grad_r + (t - 1) * n_shots * p.n_rec : nullptr;
      a2.grad_f0 = (p.n_src > 0 && grad_f) ? grad_f + t * n_shots * p.n_src : nullptr;
      a2.grad_f1 = (p.n_src > 0 && grad_f) ? grad_f + (t - 1) * n_shots * p.n_src : nullptr;
      a2.ey_store0 = snap(t, ey_store, gca, ey_stage);
      a2.curl_store0 = snap(t, curl_store, gcb, curl_stage);
      a2.ey_store1 = snap(t - 1, ey_store, gca, ey_stage);
      a2.curl_store1 = snap(t - 1, curl_store, gcb, curl_stage);
      const bool g_any = a2.ey_store0 || a2.curl_store0 || a2.ey_store1 || a2.curl_store1;
      TIDE_CHECK(stage_steps(t, 2, 0));
      if (!g_any) {
        launch_adj2<T, 0>(rows2, g_tuning.adj_border_rows, n_groups, block2, stream, split2, a2);
      } else if (fmt == 2) {
        if constexpr (sizeof(T) == 4) launch_adj2<T, 2>(rows2, g_tuning.adj_border_rows, n_groups, block2, stream, split2, a2);
      } else {
        launch_adj2<T, 1>(rows2, g_tuning.adj_border_rows, n_groups, block2, stream, split2, a2);
      }
      TIDE_CHECK(stage_steps(t, 2, 1));
      TIDE_CHECK(stage_steps(t - 2, t - 3 >= start_t - nt ? 2 : 1, 2));
      for (int k = 0; k < 7; ++k) std::swap(cur[k], nxt[k]);
    }
  }
  for (; t >= start_t - nt; --t, ++launches) {
    a.ley_i = cur[0]; a.lhx_i = cur[1]; a.lhz_i = cur[2];
    a.mx_i = cur[3]; a.my_i = cur[4]; a.nx_i = cur[5]; a.nz_i = cur[6];
    a.ley_o = nxt[0]; a.lhx_o = nxt[1]; a.lhz_o = nxt[2];
    a.mx_o = nxt[3]; a.my_o = nxt[4]; a.nx_o = nxt[5]; a.nz_o = nxt[6];
    a.grad_r_t = p.n_rec > 0 ? grad_r + t * n_shots * p.n_rec : nullptr;
    a.grad_f_t = (p.n_src > 0 && grad_f) ? grad_f + t * n_shots * p.n_src : nullptr;
    const bool g_step = (gca || gcb) && (t % step_ratio) == 0;
    a.ey_store = g_step ? snap(t, ey_store, gca, ey_stage) : nullptr;
    a.curl_store = g_step ? snap(t, curl_store, gcb, curl_stage) : nullptr;
    a.gca = g_step ? ga_k : nullptr;
    a.gcb = g_step ? gb_k : nullptr;
    if (g_step) TIDE_CHECK(stage_steps(t, 1, 0));
    if (!g_step) {
      launch_adj<T, 0>(rows, loop, grid, block, lds, stream, a);
    } else if (fmt == 2) {
      if constexpr (sizeof(T) == 4) launch_adj<T, 2>(rows, loop, grid, block, lds, stream, a);
    } else {
      launch_adj<T, 1>(rows, loop, grid, block, lds, stream, a);
    }
    if (g_step) {
      TIDE_CHECK(stage_steps(t, 1, 1));
      TIDE_CHECK(stage_steps(t - 1, 1, 2));
    }
    for (int k = 0; k < 7; ++k) std::swap(cur[k], nxt[k]);
  }
  timer.finish(launches);
  TIDE_HIP_TRY(hipGetLastError());
  const unsigned fold_blocks = (unsigned)((numel + 255) / 256);
  if (gacc) fold_groups<<<fold_blocks, 256, 0, stream>>>(gca, gacc, numel, (int)n_groups);
  if (gbcc) fold_groups<<<fold_blocks, 256, 0, stream>>>(gcb, gbcc, numel, (int)n_groups);
  TIDE_HIP_TRY(hipGetLastError());
  T *orig[7] = {l_ey, l_hx, l_hz, ml_hz_x, ml_hx_z, ml_ey_x, ml_ey_z};
  for (int k = 0; k < 7; ++k)
    if (cur[k] != orig[k])
      TIDE_HIP_TRY(hipMemcpyAsync(orig[k], cur[k], bytes, hipMemcpyDeviceToDevice, stream));
  dpt.release(stream);
  if (fused) dpt2.release(stream);
  if (kblk) dptk.release(stream);
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

// Host-staged snapshot arguments of a call: the reference's store_3 buffers,
// file-name arrays and storage stream (mode 0: device history or none).
struct StageArgs {
  int64_t mode = 0;
  void *ey_host = nullptr, *curl_host = nullptr;
  char const *const *ey_files = nullptr;
  char const *const *curl_files = nullptr;
  void *storage_stream = nullptr;
};

// Opens the stages of the snapshot streams in use (store_1 non-null).
inline int open_stages(const StageArgs &sa, bool saving, size_t step_bytes, void *ey_ring,
                       void *curl_ring, int64_t first_store, hipStream_t stream,
                       SnapshotStage &ey_st, SnapshotStage &curl_st) {
  if (sa.mode != kStageCpu && sa.mode != kStageDisk) return kOk;
  const hipStream_t storage = (hipStream_t)sa.storage_stream;
  if (ey_ring != nullptr)
    TIDE_CHECK(ey_st.open(sa.mode, saving, step_bytes, ey_ring, sa.ey_host,
                          sa.ey_files ? sa.ey_files[0] : nullptr, first_store, stream, storage));
  if (curl_ring != nullptr)
    TIDE_CHECK(curl_st.open(sa.mode, saving, step_bytes, curl_ring, sa.curl_host,
                            sa.curl_files ? sa.curl_files[0] : nullptr, first_store, stream,
                            storage));
  return kOk;
}

// Closes both stages; the first failure (of the call or a stage) wins.
inline int close_stages(int rc, SnapshotStage &ey_st, SnapshotStage &curl_st) {
  const int r1 = ey_st.finish();
  const int r2 = curl_st.finish();
  if (rc != kOk) return rc;
  return r1 != kOk ? r1 : r2;
}

template <typename T>
tm2d_ho::Problem<T> ho_problem(const T *ca, const T *cb, const T *cq, const T *const prof[12],
                               T rdy, T rdx, int64_t n_shots, int64_t ny, int64_t nx,
                               int64_t n_src, int64_t n_rec, bool ca_b, bool cb_b, bool cq_b,
                               int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1) {
  tm2d_ho::Problem<T> hp;
  hp.ca = ca;
  hp.cb = cb;
  hp.cq = cq;
  for (int k = 0; k < 12; ++k) hp.prof[k] = prof[k];
  hp.rdy = rdy;
  hp.rdx = rdx;
  hp.n_shots = n_shots;
  hp.ny = ny;
  hp.nx = nx;
  hp.n_src = n_src;
  hp.n_rec = n_rec;
  hp.pml_y0 = pml_y0;
  hp.pml_x0 = pml_x0;
  hp.pml_y1 = pml_y1;
  hp.pml_x1 = pml_x1;
  hp.ca_b = ca_b;
  hp.cb_b = cb_b;
  hp.cq_b = cq_b;
  return hp;
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
                  void *curl_store, int fmt, size_t store_step_bytes,
                  char *(*arena_for)(int, size_t *) = nullptr,
                  const StageArgs &sa = StageArgs(), int stencil = 2) {
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  char *arena = nullptr;
  size_t arena_cap = 0;
  if (arena_for != nullptr && (arena = arena_for((int)device, &arena_cap)) == nullptr)
    return kHipError;  // the message is set by arena_for
  const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};
  Problem<T> p = make_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx, n_src, n_rec,
                                 ca_b, cb_b, cq_b, pml_y0, pml_x0, pml_y1, pml_x1);
  const hipStream_t stream = (hipStream_t)stream_handle;
  SnapshotStage ey_st, curl_st;
  if (step_ratio > 0)
    TIDE_CHECK(open_stages(sa, true, store_step_bytes, ey_store, curl_store,
                           (start_t + step_ratio - 1) / step_ratio, stream, ey_st, curl_st));
  int rc;
  if (stencil == 2) {
    rc = forward_impl<T>(p, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, f, r, sources_i,
                         receivers_i, nt, start_t, step_ratio, ey_store, curl_store, fmt,
                         store_step_bytes, stream, arena, arena_cap, &ey_st, &curl_st);
  } else {
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx,
                                                 n_src, n_rec, ca_b, cb_b, cq_b, pml_y0, pml_x0,
                                                 pml_y1, pml_x1);
    T *const fld[7] = {ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x};
    rc = step_ratio <= 0 ? (set_error("step_ratio must be positive"), (int)kInvalidArgument)
                         : tm2d_ho::forward<T>(stencil, hp, fld, f, r, sources_i, receivers_i, nt,
                                               start_t, step_ratio, ey_store, curl_store, fmt,
                                               store_step_bytes, stream, &ey_st, &curl_st);
  }
  return close_stages(rc, ey_st, curl_st);
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
                   int64_t device, void *stream_handle, const StageArgs &sa = StageArgs(),
                   int stencil = 2) {
  if (ca_rg && cb_rg && ca_b != cb_b) {  // checked before any device call
    set_error("ca and cb gradients need the same batching (ca_batched == cb_batched)");
    return kInvalidArgument;
  }
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  hipStream_t stream = (hipStream_t)stream_handle;
  const bool want = ca_rg || cb_rg;
  if (want && storage_mode == kStorageNone) {
    set_error("model gradients need stored snapshots (storage_mode 'none')");
    return kInvalidArgument;
  }
  if (storage_mode != kStorageDevice && storage_mode != kStorageNone &&
      storage_mode != kStageCpu && storage_mode != kStageDisk) {
    set_error("storage_mode must be 0 (device), 1 (cpu), 2 (disk) or 3 (none)");
    return kInvalidArgument;
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
  const size_t step_bytes = (size_t)shot_bytes_uncomp * (size_t)n_shots;
  SnapshotStage ey_st, curl_st;
  if (want && step_ratio > 0) {
    StageArgs s2 = sa;
    s2.mode = storage_mode;
    TIDE_CHECK(open_stages(s2, false, step_bytes, ca_rg ? ey_store_1 : nullptr,
                           cb_rg ? curl_store_1 : nullptr, 0, stream, ey_st, curl_st));
  }
  int rc;
  if (stencil == 2) {
    rc = backward_impl<T>(p, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,
                          m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, ey_store_1, curl_store_1,
                          fmt, step_bytes, grad_f, ca_rg ? ga : nullptr, cb_rg ? gb : nullptr,
                          sources_i, receivers_i, nt, start_t, step_ratio, stream, &ey_st,
                          &curl_st);
  } else if (step_ratio <= 0) {
    set_error("step_ratio must be positive");
    rc = kInvalidArgument;
  } else {
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, nx,
                                                 n_src, n_rec, ca_b, cb_b, cq_b, pml_y0, pml_x0,
                                                 pml_y1, pml_x1);
    T *const lam[7] = {lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x, m_lambda_ey_z,
                       m_lambda_hx_z, m_lambda_hz_x};
    rc = tm2d_ho::backward<T>(stencil, hp, lam, grad_r, ey_store_1, curl_store_1, fmt, step_bytes,
                              grad_f, ca_rg ? ga : nullptr, cb_rg ? gb : nullptr, sources_i,
                              receivers_i, nt, start_t, step_ratio, stream, &ey_st, &curl_st);
  }
  TIDE_CHECK(close_stages(rc, ey_st, curl_st));
  if (own) {
    if (ca_rg) TIDE_CHECK(fold_into<T>(grad_ca, ga, na, stream));
    if (cb_rg) TIDE_CHECK(fold_into<T>(grad_cb, gb, nb, stream));
  }
  return kOk;
}

}  // namespace

#define TIDE_TM2D_ENTRY_POINTS(S, T, NAME)                                                    \
  extern "C" int maxwell_tm_##S##_##NAME##_forward_cuda(                                      \
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
    (void)dt; (void)n_threads;                                                            \
    if (has_dispersion) {                                                                 \
      if (step_ratio != 1) {                                                              \
        set_error("Debye dispersion: step_ratio must be 1");                              \
        return kInvalidArgument;                                                          \
      }                                                                                   \
      if (hipSetDevice((int)device) != hipSuccess) {                                      \
        set_error("hipSetDevice failed");                                                 \
        return kHipError;                                                                 \
      }                                                                                   \
      const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};         \
      const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots,   \
                                                   ny, nx, n_src, n_rec, ca_b, cb_b,      \
                                                   cq_b, pml_y0, pml_x0, pml_y1, pml_x1); \
      T *const fld[7] = {ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x};                     \
      return tm2d_ho::forward_debye<T>(S, hp, fld, f, r, sources_i, receivers_i, nt,      \
                                       start_t, debye_a, debye_b, debye_cp, polarization, \
                                       ey_prev, n_poles, (hipStream_t)compute_stream);    \
    }                                                                                     \
    return forward_entry<T>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, \
                            ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh,         \
                            sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, \
                            n_rec, step_ratio, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, \
                            pml_y1, pml_x1, device, compute_stream, nullptr, nullptr, 0,  \
                            0, nullptr, StageArgs(), S);                                  \
  }                                                                                       \
                                                                                          \
  extern "C" int maxwell_tm_##S##_##NAME##_forward_with_storage_cuda(                         \
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
    (void)dt; (void)n_threads;                                                            \
    if (storage_mode < kStorageDevice || storage_mode > kStorageNone) {                   \
      set_error("storage_mode must be 0 (device), 1 (cpu), 2 (disk) or 3 (none)");       \
      return kInvalidArgument;                                                            \
    }                                                                                     \
    const bool st = storage_mode != kStorageNone;                                         \
    const int fmt = storage_format == kFormatBf16 ? 2 : 1;                                \
    StageArgs sa;                                                                         \
    if (storage_mode == kStageCpu || storage_mode == kStageDisk) {                        \
      sa.mode = storage_mode;                                                             \
      sa.ey_host = ey_store_3;                                                            \
      sa.curl_host = curl_store_3;                                                        \
      sa.ey_files = ey_filenames;                                                         \
      sa.curl_files = curl_filenames;                                                     \
      sa.storage_stream = storage_stream;                                                 \
    }                                                                                     \
    return forward_entry<T>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, \
                            ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh,         \
                            sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, \
                            n_rec, step_ratio, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, \
                            pml_y1, pml_x1, device, compute_stream,                       \
                            st && ca_rg ? ey_store_1 : nullptr,                           \
                            st && cb_rg ? curl_store_1 : nullptr, fmt,                    \
                            (size_t)shot_bytes_uncomp * (size_t)n_shots, nullptr, sa, S); \
  }                                                                                       \
                                                                                          \
  extern "C" int maxwell_tm_##S##_##NAME##_backward_cuda(                                     \
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
    (void)adjoint_memory_scratch; (void)grad_ca_shot;                                     \
    (void)grad_cb_shot; (void)dt; (void)n_threads;                                        \
    StageArgs sa;                                                                         \
    sa.ey_host = ey_store_3;                                                              \
    sa.curl_host = curl_store_3;                                                          \
    sa.ey_files = ey_filenames;                                                           \
    sa.curl_files = curl_filenames;                                                       \
    sa.storage_stream = storage_stream;                                                   \
    return backward_entry<T>(ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz,         \
                             m_lambda_ey_x, m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x,  \
                             ey_store_1, curl_store_1, grad_f, grad_ca, grad_cb, nullptr, \
                             nullptr, ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx,    \
                             kxh, sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx,  \
                             n_src, n_rec, step_ratio, storage_mode, storage_format,      \
                             shot_bytes_uncomp, ca_rg, cb_rg, ca_b, cb_b, cq_b, start_t,  \
                             pml_y0, pml_x0, pml_y1, pml_x1, device, compute_stream, sa,  \
                             S);                                                          \
  }

TIDE_TM2D_ENTRY_POINTS(2, float, float)
TIDE_TM2D_ENTRY_POINTS(2, double, double)
TIDE_TM2D_ENTRY_POINTS(4, float, float)
TIDE_TM2D_ENTRY_POINTS(4, double, double)
TIDE_TM2D_ENTRY_POINTS(6, float, float)
TIDE_TM2D_ENTRY_POINTS(6, double, double)
TIDE_TM2D_ENTRY_POINTS(8, float, float)
TIDE_TM2D_ENTRY_POINTS(8, double, double)

// Born forward entry points: reference ABI src/tide/backend_utils.py:336-349
// (_TM_BORN_FORWARD_SPEC); r receives the scattered receivers, background_r
// the background's.  dca / dcb null: no coefficient perturbation.
#define TIDE_TM2D_BORN_ENTRY(S, T, NAME)                                                   \
  extern "C" int maxwell_tm_##S##_##NAME##_born_forward_cuda(                             \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *f0,     \
      T const *df, T *ey, T *hx, T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x,       \
      T *dey, T *dhx, T *dhz, T *dm_ey_x, T *dm_ey_z, T *dm_hx_z, T *dm_hz_x, T *r,       \
      T *background_r, T const *ay, T const *by, T const *ayh, T const *byh, T const *ax, \
      T const *bx, T const *axh, T const *bxh, T const *ky, T const *kyh, T const *kx,    \
      T const *kxh, int64_t const *sources_i, int64_t const *receivers_i, T rdy, T rdx,   \
      T dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,           \
      int64_t n_rec, int64_t step_ratio, bool ca_b, bool cb_b, bool cq_b,                 \
      int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1,    \
      int64_t n_threads, int64_t device, void *compute_stream) {                          \
    (void)dt; (void)n_threads;                                                            \
    if (step_ratio != 1) {                                                                \
      set_error("born_forward: step_ratio must be 1 (resample the sources beforehand)");  \
      return kInvalidArgument;                                                            \
    }                                                                                     \
    if (hipSetDevice((int)device) != hipSuccess) {                                        \
      set_error("hipSetDevice failed");                                                   \
      return kHipError;                                                                   \
    }                                                                                     \
    const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};           \
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, \
                                                 nx, n_src, n_rec, ca_b, cb_b, cq_b,      \
                                                 pml_y0, pml_x0, pml_y1, pml_x1);         \
    T *const bg[7] = {ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x};                        \
    T *const sc[7] = {dey, dhx, dhz, dm_ey_x, dm_ey_z, dm_hx_z, dm_hz_x};                 \
    return tm2d_ho::born_forward<T>(S, hp, bg, sc, dca, dcb, f0, df, r, background_r,     \
                                    sources_i, receivers_i, nt, start_t,                  \
                                    (hipStream_t)compute_stream, nullptr);                \
  }                                                                                       \
                                                                                          \
  /* reference ABI src/tide/backend_utils.py:351-370 (_TM_BORN_FORWARD_WITH_STORAGE_SPEC): \
     device storage of full-precision snapshots; ey_store_1 / curl_store_1 keep the    \
     background's Ey^n / curl^n and dey_store / dcurl_store the scattered field's, one  \
     [n_shots, ny, nx] slot per step (slot t for step t) */                             \
  extern "C" int maxwell_tm_##S##_##NAME##_born_forward_with_storage_cuda(                \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *f0,     \
      T const *df, T *ey, T *hx, T *hz, T *m_ey_x, T *m_ey_z, T *m_hx_z, T *m_hz_x,       \
      T *dey, T *dhx, T *dhz, T *dm_ey_x, T *dm_ey_z, T *dm_hx_z, T *dm_hz_x, T *r,       \
      T *background_r, void *ey_store_1, void *ey_store_3, void *ey_filenames,            \
      void *curl_store_1, void *curl_store_3, void *curl_filenames, void *dey_store,      \
      void *dcurl_store, T const *ay, T const *by, T const *ayh, T const *byh,            \
      T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky, T const *kyh,    \
      T const *kx, T const *kxh, int64_t const *sources_i, int64_t const *receivers_i,    \
      T rdy, T rdx, T dt, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx,            \
      int64_t n_src, int64_t n_rec, int64_t step_ratio, int64_t storage_mode,             \
      int64_t storage_format, int64_t shot_bytes_uncomp, bool ca_rg, bool cb_rg,          \
      bool ca_b, bool cb_b, bool cq_b, int64_t start_t, int64_t pml_y0, int64_t pml_x0,   \
      int64_t pml_y1, int64_t pml_x1, int64_t n_threads, int64_t device,                  \
      void *compute_stream, void *storage_stream) {                                       \
    (void)dt; (void)n_threads; (void)ca_rg; (void)cb_rg; (void)storage_stream;            \
    if (step_ratio != 1) {                                                                \
      set_error("born_forward: step_ratio must be 1 (resample the sources beforehand)");  \
      return kInvalidArgument;                                                            \
    }                                                                                     \
    const bool any = ey_store_1 || curl_store_1 || dey_store || dcurl_store;              \
    const bool all = ey_store_1 && curl_store_1 && dey_store && dcurl_store;              \
    if (any && (!all || storage_mode != kStorageDevice || storage_format == kFormatBf16 ||\
                ey_store_3 || curl_store_3 || ey_filenames || curl_filenames ||           \
                shot_bytes_uncomp != ny * nx * (int64_t)sizeof(T))) {                     \
      set_error("born_forward_with_storage: the four snapshot stores are kept on the "    \
                "device in full precision, all four or none, shot_bytes_uncomp = "        \
                "ny * nx * sizeof(field)");                                               \
      return kUnsupported;                                                                \
    }                                                                                     \
    if (hipSetDevice((int)device) != hipSuccess) {                                        \
      set_error("hipSetDevice failed");                                                   \
      return kHipError;                                                                   \
    }                                                                                     \
    const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};           \
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, \
                                                 nx, n_src, n_rec, ca_b, cb_b, cq_b,      \
                                                 pml_y0, pml_x0, pml_y1, pml_x1);         \
    T *const bg[7] = {ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x};                        \
    T *const sc[7] = {dey, dhx, dhz, dm_ey_x, dm_ey_z, dm_hx_z, dm_hz_x};                 \
    T *const hist[4] = {(T *)ey_store_1, (T *)curl_store_1, (T *)dey_store,               \
                        (T *)dcurl_store};                                                \
    return tm2d_ho::born_forward<T>(S, hp, bg, sc, dca, dcb, f0, df, r, background_r,     \
                                    sources_i, receivers_i, nt, start_t,                  \
                                    (hipStream_t)compute_stream, any ? hist : nullptr);   \
  }                                                                                       \
                                                                                          \
  /* reference ABI src/tide/backend_utils.py:407-444 (_TM_BORN_BACKWARD_BGGRAD_SPEC):   \
     ey/hx/hz + m_lambda_* the scattered adjoint (cotangent grad_r), dey/dhx/dhz +      \
     m_eta_* the background adjoint (cotangent grad_background_r); gradients are       \
     accumulated in FP64 and added to grad_ca, grad_cb, grad_dca, grad_dcb; grad_f0 /   \
     grad_df are written per step.  Device storage, unbatched coefficients;          \
     lambda_store, eta_source_old, work_eta_*, grad_*_shot are not used. */            \
  extern "C" int maxwell_tm_##S##_##NAME##_born_backward_bggrad_cuda(                     \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *f0,     \
      T const *df, T const *grad_r, T const *grad_background_r, void *ey_store_1,         \
      void *ey_store_3, void *ey_filenames, void *curl_store_1, void *curl_store_3,       \
      void *curl_filenames, void *dey_store, void *dcurl_store, void *lambda_store,       \
      int64_t background_n_shots, T *ey, T *hx, T *hz, T *dey, T *dhx, T *dhz,            \
      T *grad_f0, T *grad_df, T *grad_ca, T *grad_cb, T *grad_dca, T *grad_dcb,           \
      T *m_lambda_ey_x, T *m_lambda_ey_z, T *m_lambda_hx_z, T *m_lambda_hz_x,             \
      T *m_eta_ey_x, T *m_eta_ey_z, T *m_eta_hx_z, T *m_eta_hz_x, T *eta_source_old,      \
      T *work_eta_x, T *work_eta_z, T *grad_ca_shot, T *grad_cb_shot, T *grad_dca_shot,   \
      T *grad_dcb_shot, T *adjoint_memory_scratch, T const *ay, T const *by,              \
      T const *ayh, T const *byh, T const *ax, T const *bx, T const *axh, T const *bxh,   \
      T const *ky, T const *kyh, T const *kx, T const *kxh, int64_t const *sources_i,     \
      int64_t const *receivers_i, T rdy, T rdx, T dt, int64_t nt, int64_t n_shots,        \
      int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec, int64_t step_ratio,           \
      int64_t storage_mode, int64_t storage_format, int64_t shot_bytes_uncomp,            \
      bool ca_rg, bool cb_rg, bool capture_background_adjoint,                            \
      bool reuse_background_adjoint, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,    \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t n_threads,  \
      int64_t device, void *compute_stream, void *storage_stream) {                       \
    (void)f0; (void)df; (void)lambda_store; (void)eta_source_old; (void)work_eta_x;       \
    (void)work_eta_z; (void)grad_ca_shot; (void)grad_cb_shot; (void)grad_dca_shot;        \
    (void)grad_dcb_shot; (void)dt; (void)n_threads; (void)storage_stream;                 \
    if (step_ratio != 1 || storage_mode != kStorageDevice ||                              \
        storage_format == kFormatBf16 || capture_background_adjoint ||                    \
        reuse_background_adjoint || ey_store_3 || curl_store_3 || ey_filenames ||         \
        curl_filenames || (background_n_shots != 0 && background_n_shots != n_shots) ||   \
        shot_bytes_uncomp != ny * nx * (int64_t)sizeof(T)) {                              \
      set_error("born_backward_bggrad: device storage of full-precision snapshots, "      \
                "step_ratio 1, no background-adjoint capture or reuse");                  \
      return kUnsupported;                                                                \
    }                                                                                     \
    if (hipSetDevice((int)device) != hipSuccess) {                                        \
      set_error("hipSetDevice failed");                                                   \
      return kHipError;                                                                   \
    }                                                                                     \
    const hipStream_t stream = (hipStream_t)compute_stream;                               \
    const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};           \
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, \
                                                 nx, n_src, n_rec, ca_b, cb_b, cq_b,      \
                                                 pml_y0, pml_x0, pml_y1, pml_x1);         \
    T *const lam[7] = {ey, hx, hz, m_lambda_ey_x, m_lambda_ey_z, m_lambda_hx_z,           \
                       m_lambda_hz_x};                                                    \
    T *const eta[7] = {dey, dhx, dhz, m_eta_ey_x, m_eta_ey_z, m_eta_hx_z, m_eta_hz_x};    \
    const T *const hist[4] = {(const T *)ey_store_1, (const T *)curl_store_1,             \
                              (const T *)dey_store, (const T *)dcurl_store};              \
    const int64_t numel = ny * nx;                                                        \
    T *const outs[4] = {ca_rg ? grad_ca : nullptr, cb_rg ? grad_cb : nullptr, grad_dca,   \
                        grad_dcb};                                                        \
    Scratch scr(stream);                                                                  \
    double *acc[4] = {nullptr, nullptr, nullptr, nullptr};                                \
    for (int k = 0; k < 4; ++k)                                                           \
      if (outs[k] != nullptr) {                                                           \
        TIDE_CHECK(scr.alloc(&acc[k], (size_t)numel));                                    \
        TIDE_HIP_TRY(hipMemsetAsync(acc[k], 0, (size_t)numel * sizeof(double), stream));  \
      }                                                                                   \
    TIDE_CHECK(tm2d_ho::born_backward<T>(S, hp, lam, eta, dca, dcb, grad_r,               \
                                         grad_background_r, hist, grad_df, grad_f0,       \
                                         acc[0], acc[1], acc[2], acc[3], sources_i,       \
                                         receivers_i, nt, start_t, stream,                \
                                         adjoint_memory_scratch));                        \
    for (int k = 0; k < 4; ++k)                                                           \
      if (outs[k] != nullptr) TIDE_CHECK(fold_into<T>(outs[k], acc[k], numel, stream));   \
    return kOk;                                                                           \
  }                                                                                       \
                                                                                          \
  /* this library's form of the Born adjoint: FP64 accumulators (added to), the four   \
     snapshot histories and both adjoint states as plain arrays */                      \
  extern "C" int tide_tm2d_##S##_##NAME##_born_backward_f64acc(                           \
      T const *ca, T const *cb, T const *cq, T const *dca, T const *dcb, T const *w,      \
      T const *w_bg, T const *bg_ey_hist, T const *bg_curl_hist, T const *sc_ey_hist,     \
      T const *sc_curl_hist, T *l_ey, T *l_hx, T *l_hz, T *ml_ey_x, T *ml_ey_z,           \
      T *ml_hx_z, T *ml_hz_x, T *e_ey, T *e_hx, T *e_hz, T *me_ey_x, T *me_ey_z,          \
      T *me_hx_z, T *me_hz_x, T *grad_df, T *grad_f0, double *gca, double *gcb,           \
      double *gdca, double *gdcb, T const *ay, T const *by, T const *ayh, T const *byh,   \
      T const *ax, T const *bx, T const *axh, T const *bxh, T const *ky, T const *kyh,    \
      T const *kx, T const *kxh, int64_t const *sources_i, int64_t const *receivers_i,    \
      T rdy, T rdx, int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src,   \
      int64_t n_rec, int64_t start_t, int64_t pml_y0, int64_t pml_x0, int64_t pml_y1,     \
      int64_t pml_x1, int64_t device, void *compute_stream) {                             \
    if (hipSetDevice((int)device) != hipSuccess) {                                        \
      set_error("hipSetDevice failed");                                                   \
      return kHipError;                                                                   \
    }                                                                                     \
    const T *prof[12] = {ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh};           \
    const tm2d_ho::Problem<T> hp = ho_problem<T>(ca, cb, cq, prof, rdy, rdx, n_shots, ny, \
                                                 nx, n_src, n_rec, false, false, false,   \
                                                 pml_y0, pml_x0, pml_y1, pml_x1);         \
    T *const lam[7] = {l_ey, l_hx, l_hz, ml_ey_x, ml_ey_z, ml_hx_z, ml_hz_x};             \
    T *const eta[7] = {e_ey, e_hx, e_hz, me_ey_x, me_ey_z, me_hx_z, me_hz_x};             \
    const T *const hist[4] = {bg_ey_hist, bg_curl_hist, sc_ey_hist, sc_curl_hist};        \
    return tm2d_ho::born_backward<T>(S, hp, lam, eta, dca, dcb, w, w_bg, hist, grad_df,   \
                                     grad_f0, gca, gcb, gdca, gdcb, sources_i,            \
                                     receivers_i, nt, start_t, (hipStream_t)compute_stream,\
                                     nullptr);                                            \
  }

TIDE_TM2D_BORN_ENTRY(2, float, float)
TIDE_TM2D_BORN_ENTRY(2, double, double)
TIDE_TM2D_BORN_ENTRY(4, float, float)
TIDE_TM2D_BORN_ENTRY(4, double, double)
TIDE_TM2D_BORN_ENTRY(6, float, float)
TIDE_TM2D_BORN_ENTRY(6, double, double)
TIDE_TM2D_BORN_ENTRY(8, float, float)
TIDE_TM2D_BORN_ENTRY(8, double, double)

// ------------------------------------------- allocation-free forward entry
// The forward's scratch (the ping-pong copies of the fields and CPML
// memories, the packed profiles and the point tables) lives in this fixed
// device array instead of stream-ordered allocations.  Each device holds its
// own copy for the life of the process; its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current (the
// array's name in host code is not a device address).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<char *> g_scratch_ptr;  // per device; null until looked up

// The device address of `scratch` on `device` (the current device), looked
// up on the first call for that device.  Null (with the error set) on failure.
char *scratch_on(int device, size_t *cap) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0) {
    set_error("invalid device index");
    return nullptr;
  }
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device] == nullptr) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[(size_t)device] = (char *)ptr;
  }
  *cap = kNoAllocScratchBytes;
  return g_scratch_ptr[(size_t)device];
}
}  // namespace

extern "C" int tide_tm2d_float_forward_noalloc(
    float const *ca, float const *cb, float const *cq, float const *f, float *ey, float *hx,
    float *hz, float *m_ey_x, float *m_ey_z, float *m_hx_z, float *m_hz_x,
    float const *debye_a, float const *debye_b, float const *debye_cp, float *polarization,
    float *ey_prev, float *r, int64_t n_poles, float const *ay, float const *by,
    float const *ayh, float const *byh, float const *ax, float const *bx, float const *axh,
    float const *bxh, float const *ky, float const *kyh, float const *kx, float const *kxh,
    int64_t const *sources_i, int64_t const *receivers_i, float rdy, float rdx, float dt,
    int64_t nt, int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,
    int64_t step_ratio, bool has_dispersion, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,
    int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t n_threads,
    int64_t device, void *compute_stream) {
  (void)debye_a; (void)debye_b; (void)debye_cp; (void)polarization;
  (void)ey_prev; (void)n_poles; (void)dt; (void)n_threads;
  if (has_dispersion) {
    set_error("Debye dispersion is not part of this build");
    return kUnsupported;
  }
  // The field copies dominate the scratch; checked before any device work
  // (the remaining pieces are checked as they are placed, also before any
  // launch).
  if (n_shots > 0 && ny > 0 && nx > 0 &&
      (size_t)7 * (size_t)n_shots * (size_t)ny * (size_t)nx * sizeof(float) +
              (size_t)12 * (size_t)(ny + nx) * sizeof(float) >
          kNoAllocScratchBytes) {
    set_error("tide_tm2d_float_forward_noalloc: 7 x n_shots x ny x nx fp32 scratch values exceed "
              "the fixed " + std::to_string(kNoAllocScratchBytes >> 20) +
              " MiB scratch array; use maxwell_tm_2_float_forward_cuda");
    return kInvalidArgument;
  }
  return forward_entry<float>(ca, cb, cq, f, ey, hx, hz, m_ey_x, m_ey_z, m_hx_z, m_hz_x, r, ay,
                              by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh, sources_i,
                              receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,
                              step_ratio, ca_b, cb_b, cq_b, start_t, pml_y0, pml_x0, pml_y1,
                              pml_x1, device, compute_stream, nullptr, nullptr, 0, 0,
                              scratch_on);
}

// FP64-accumulator backward used by the Python operator (and any caller that
// runs shot chunks and wants one FP64 sum across them); see tide_hip.h.
#define TIDE_TM2D_BACKWARD_F64(T, NAME)                                                    \
  extern "C" int tide_tm2d_##NAME##_backward_f64acc(                                      \
      T const *ca, T const *cb, T const *cq, T const *grad_r, T *lambda_ey, T *lambda_hx, \
      T *lambda_hz, T *m_lambda_ey_x, T *m_lambda_ey_z, T *m_lambda_hx_z,                 \
      T *m_lambda_hz_x, void *ey_store_1, void *ey_store_3,                               \
      char const *const *ey_filenames, void *curl_store_1, void *curl_store_3,            \
      char const *const *curl_filenames, T *grad_f, double *grad_ca, double *grad_cb,     \
      T const *ay, T const *by, T const *ayh, T const *byh, T const *ax, T const *bx,     \
      T const *axh, T const *bxh, T const *ky, T const *kyh, T const *kx, T const *kxh,   \
      int64_t const *sources_i, int64_t const *receivers_i, T rdy, T rdx, int64_t nt,     \
      int64_t n_shots, int64_t ny, int64_t nx, int64_t n_src, int64_t n_rec,              \
      int64_t step_ratio, int64_t storage_mode, int64_t storage_format,                   \
      int64_t shot_bytes_uncomp, bool ca_b, bool cb_b, bool cq_b, int64_t start_t,        \
      int64_t pml_y0, int64_t pml_x0, int64_t pml_y1, int64_t pml_x1, int64_t stencil,    \
      int64_t device, void *compute_stream, void *storage_stream) {                       \
    if (stencil != 2 && stencil != 4 && stencil != 6 && stencil != 8) {                   \
      set_error("stencil must be 2, 4, 6 or 8");                                          \
      return kInvalidArgument;                                                            \
    }                                                                                     \
    const bool want = grad_ca != nullptr || grad_cb != nullptr;                           \
    StageArgs sa;                                                                         \
    sa.ey_host = ey_store_3;                                                              \
    sa.curl_host = curl_store_3;                                                          \
    sa.ey_files = ey_filenames;                                                           \
    sa.curl_files = curl_filenames;                                                       \
    sa.storage_stream = storage_stream;                                                   \
    return backward_entry<T>(                                                             \
        ca, cb, cq, grad_r, lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x,               \
        m_lambda_ey_z, m_lambda_hx_z, m_lambda_hz_x, ey_store_1, curl_store_1, grad_f,    \
        nullptr, nullptr, grad_ca, grad_cb, ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh,  \
        kx, kxh, sources_i, receivers_i, rdy, rdx, nt, n_shots, ny, nx, n_src, n_rec,     \
        step_ratio, want ? storage_mode : kStorageNone, storage_format, shot_bytes_uncomp, \
        grad_ca != nullptr, grad_cb != nullptr, ca_b, cb_b, cq_b, start_t, pml_y0,        \
        pml_x0, pml_y1, pml_x1, device, compute_stream, sa, (int)stencil);                \
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
  g_tuning.fwd_rows = fwd_rows;
  g_tuning.fwd_group = fwd_group;
  g_tuning.adj_rows = adj_rows;
  g_tuning.adj_group = adj_group;
  return kOk;
}

extern "C" int tide_tm2d_set_shot_loop(int fwd_loop, int adj_loop) {
  if ((fwd_loop != 0 && fwd_loop != 1) || (adj_loop != 0 && adj_loop != 1)) {
    set_error("shot loop flags must be 0 or 1");
    return kInvalidArgument;
  }
  g_tuning.fwd_loop = fwd_loop;
  g_tuning.adj_loop = adj_loop;
  return kOk;
}

extern "C" int tide_tm2d_get_shot_loop(int *out2) {
  if (out2 == nullptr) {
    set_error("tuning: null output");
    return kInvalidArgument;
  }
  out2[0] = g_tuning.fwd_loop;
  out2[1] = g_tuning.adj_loop;
  return kOk;
}

extern "C" int tide_tm2d_set_adjoint_waves(int waves) {
  if (waves != 0 && waves != 4) {
    set_error("adjoint waves: 0 (compiler's register count) or 4 (128 VGPRs)");
    return kInvalidArgument;
  }
  g_tuning.adj_waves = waves;
  return kOk;
}

extern "C" int tide_tm2d_get_adjoint_waves(int *out) {
  if (out == nullptr) {
    set_error("get_adjoint_waves: null output");
    return kInvalidArgument;
  }
  *out = g_tuning.adj_waves;
  return kOk;
}

extern "C" int tide_tm2d_set_time_block(int fwd_steps, int adj_steps) {
  if (fwd_steps < 1 || fwd_steps > kMaxBlockSteps || adj_steps < 1 ||
      adj_steps > kMaxBlockSteps) {
    set_error("time steps per launch must be 1, 2, 3 or 4");
    return kInvalidArgument;
  }
  g_tuning.fwd_steps = fwd_steps;
  g_tuning.adj_steps = adj_steps;
  return kOk;
}

extern "C" int tide_tm2d_get_time_block(int *out2) {
  if (out2 == nullptr) {
    set_error("tuning: null output");
    return kInvalidArgument;
  }
  out2[0] = g_tuning.fwd_steps;
  out2[1] = g_tuning.adj_steps;
  return kOk;
}

extern "C" int tide_tm2d_set_tile_split(int fwd_split, int adj_split) {
  if ((fwd_split != 0 && fwd_split != 1) || (adj_split != 0 && adj_split != 1)) {
    set_error("tile split flags must be 0 or 1");
    return kInvalidArgument;
  }
  g_tuning.fwd_split = fwd_split;
  g_tuning.adj_split = adj_split;
  return kOk;
}

extern "C" int tide_tm2d_set_border_rows(int fwd_rows, int adj_rows) {
  if (fwd_rows < 0 || fwd_rows > 2 || adj_rows < 0 || adj_rows > 2) {
    set_error("border tile height must be 0 (the interior height), 1 or 2");
    return kInvalidArgument;
  }
  g_tuning.fwd_border_rows = fwd_rows;
  g_tuning.adj_border_rows = adj_rows;
  return kOk;
}

extern "C" int tide_tm2d_get_border_rows(int *out2) {
  if (out2 == nullptr) {
    set_error("null output");
    return kInvalidArgument;
  }
  out2[0] = g_tuning.fwd_border_rows;
  out2[1] = g_tuning.adj_border_rows;
  return kOk;
}

extern "C" int tide_tm2d_get_tile_split(int *out2) {
  if (out2 == nullptr) {
    set_error("tuning: null output");
    return kInvalidArgument;
  }
  out2[0] = g_tuning.fwd_split;
  out2[1] = g_tuning.adj_split;
  return kOk;
}

extern "C" int tide_tm2d_set_loop_timing(int enable) {
  g_loop_times.enabled = enable != 0;
  return kOk;
}

extern "C" int tide_tm2d_take_loop_timing(int pass, double *ms, int64_t *launches, int64_t *loops) {
  if (pass < 0 || pass > 1 || ms == nullptr || launches == nullptr || loops == nullptr) {
    set_error("loop timing: pass 0 or 1 and three outputs are required");
    return kInvalidArgument;
  }
  double total = 0.0;
  int64_t n = 0;
  auto &recs = g_loop_times.recs[pass];
  int rc = kOk;
  for (auto &r : recs) {
    float e = 0.0f;
    if (rc == kOk && hipEventSynchronize(r.stop) == hipSuccess &&
        hipEventElapsedTime(&e, r.start, r.stop) == hipSuccess) {
      total += e;
      n += r.launches;
    } else if (rc == kOk) {
      set_error("loop timing: event query failed");
      rc = kHipError;
    }
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  *ms = total;
  *launches = n;
  *loops = (int64_t)recs.size();
  recs.clear();
  return rc;
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
