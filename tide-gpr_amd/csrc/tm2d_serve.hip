// Serving engine for 2-D TM solves: continuous batching of independent
// problems that arrive and leave at any time.
//
// The engine keeps one descriptor per admitted job (its geometry, its
// arrays, its current phase and step).  A tick advances every running job
// by up to K steps of its phase with one launch per step: the jobs are
// sorted by the steps they take in this tick (longest first), so the jobs
// still running at step j of the tick own a prefix of the flat tile range,
// and a workgroup finds its job by a binary search over the jobs' first
// tiles.  With two steps per launch (the default time block), each launch
// advances the jobs by two steps of their phase with the two-step tiles
// (tm2d_fused.h); a job whose steps in the tick are odd takes its last one
// in a one-step launch over its own tiles.  The host tracks which state
// buffer of each job is current.  The job's phase picks the forward step or the adjoint iteration
// (the single-problem tile code), so forward-only solves, forward passes of
// gradient solves and adjoints all share the same launches.
//
// Host uploads go through two pinned staging slots, each reused only after
// the copies last issued from it have completed (an event per slot), so the
// host never waits on the device except to recycle a slot.  Per-job device
// memory (packed CPML profiles and the source/receiver table) comes from the
// stream-ordered allocator and is released in stream order.

#include "tide_hip.h"
#include "tm2d_items.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

namespace tide_hip {
namespace tm2d {

struct ServeItem {
  BatchItem<float> b;
  int32_t phase;  // 0 forward, 2 adjoint
  int32_t fmt;    // snapshot payload: 0 none, 1 fp32, 2 BF16
  int32_t k0;     // step of the phase at the start of the tick
  int32_t par0;   // two-step ticks: current state buffer at the start of the tick
  int32_t epar0;  // two-step ticks: current E-phase memory copy (forward)
  int32_t pad;
};

// n: jobs active at this step (a prefix of the tick's job list); the
// workgroups cover flat tiles [first, first + gridDim.x).  State parity:
// pair_mode 1, the single step after j / 2 two-step launches of this tick
// (each flips the state and the E-phase memory copy once); pair_mode 2, the
// j-th of consecutive single steps in a two-step engine (each flips the
// state, none the memory copy); pair_mode 0, a one-step engine, where it
// follows the step index.
template <int TYF, int TYA>
__global__ __launch_bounds__(512) void serve_step(const ServeItem *__restrict__ items,
                                                  const int32_t *__restrict__ start, int n, int j,
                                                  int first, int pair_mode) {
  extern __shared__ double red[];
  const int64_t flat = first + xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  int lo = 0, hi = n - 1;  // largest i with start[i] <= flat
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= flat)
      lo = mid;
    else
      hi = mid - 1;
  }
  const ServeItem &it = items[lo];
  const int k = it.k0 + j;
  const int local = (int)(flat - start[lo]);
  const int flip = (j >> 1) & 1;
  const int par = pair_mode == 1 ? it.par0 ^ flip : pair_mode == 2 ? it.par0 ^ (j & 1) : (k & 1);
  const int epar = pair_mode == 1 ? it.epar0 ^ flip : pair_mode == 2 ? it.epar0 : 0;
  if (it.b.n_shots == 1 && blockDim.y > 1) {
    // A single-shot job gives each wavefront of the workgroup its own tile
    // (no idle wavefronts, no cross-wavefront gradient sum).
    const int n_xt = it.b.pts.n_xt;
    const int tile = local * (int)blockDim.y + (int)threadIdx.y;
    if (it.phase == 0) {
      if (tile >= n_xt * it.b.n_yt) return;  // uniform over the wavefront; no barriers here
      const int yt = tile / n_xt, xt = tile - yt * n_xt;
      const FwdArgs<float> a = item_fwd_args(it.b, k, par, epar);
      const bool inner = interior_tile<float, TYF>(a.g, xt, yt);
      if (it.fmt == 0)
        inner ? fwd_tile<float, TYF, 0, false>(a, 0, xt, yt) : fwd_tile<float, TYF, 0, true>(a, 0, xt, yt);
      else if (it.fmt == 2)
        inner ? fwd_tile<float, TYF, 2, false>(a, 0, xt, yt) : fwd_tile<float, TYF, 2, true>(a, 0, xt, yt);
      else
        inner ? fwd_tile<float, TYF, 1, false>(a, 0, xt, yt) : fwd_tile<float, TYF, 1, true>(a, 0, xt, yt);
    } else {
      if (tile >= n_xt * it.b.n_yt) return;
      const int yt = tile / n_xt, xt = tile - yt * n_xt;
      const AdjArgs<float> a = item_adj_args(it.b, k, par);
      const bool inner = interior_tile<float, TYA>(a.g, xt, yt);
      if (it.fmt == 2) {
        if (inner)
          adj_tile<float, TYA, 2, false>(a, 0, xt, yt, nullptr, nullptr, false);
        else
          adj_tile<float, TYA, 2, true>(a, 0, xt, yt, nullptr, nullptr, false);
      } else {
        if (inner)
          adj_tile<float, TYA, 1, false>(a, 0, xt, yt, nullptr, nullptr, false);
        else
          adj_tile<float, TYA, 1, true>(a, 0, xt, yt, nullptr, nullptr, false);
      }
    }
    return;
  }
  const TileCoord tc = batch_tile(local, it.b.pts.n_xt, it.b.n_yt);
  if (it.phase == 0) {
    const FwdArgs<float> a = item_fwd_args(it.b, k, par, epar);
    if (it.fmt == 0)
      fwd_workgroup<float, TYF, 0>(a, tc);
    else if (it.fmt == 2)
      fwd_workgroup<float, TYF, 2>(a, tc);
    else
      fwd_workgroup<float, TYF, 1>(a, tc);
  } else {
    const AdjArgs<float> a = item_adj_args(it.b, k, par);
    if (it.fmt == 2)
      adj_workgroup<float, TYA, 2>(a, tc, red);
    else
      adj_workgroup<float, TYA, 1>(a, tc, red);
  }
}

constexpr int kServeRows2 = 4;  // tile rows of the two-step engine launches

// One forward tile of a two-step engine launch: edge 0 interior, 1 x checks
// only, 3 x and y checks; fmt is the snapshot payload (0 none, 1 fp32, 2
// BF16).
template <int TY>
__device__ __forceinline__ void fwd2_serve_tile(const Fwd2Args<float> &a, int s, const TileCoord &tc,
                                                int edge, int fmt) {
  auto run = [&](auto f) {
    constexpr int F = decltype(f)::value;
    if (edge == 0)
      fwd2_tile<float, TY, F, 0>(a, s, tc.xt, tc.yt);
    else if (edge == 1)
      fwd2_tile<float, TY, F, 1>(a, s, tc.xt, tc.yt);
    else
      fwd2_tile<float, TY, F, 3>(a, s, tc.xt, tc.yt);
  };
  if (fmt == 0)
    run(std::integral_constant<int, 0>{});
  else if (fmt == 2)
    run(std::integral_constant<int, 2>{});
  else
    run(std::integral_constant<int, 1>{});
}

// Two-step launch of pair q of the tick over the tiles [first, first +
// gridDim.x) of jobs [0, n): each advances steps k0 + 2q and k0 + 2q + 1.
// PH 0 runs forward jobs, PH 1 adjoint jobs (the tick orders its jobs by
// phase, so each launch covers one phase): a kernel holding both phases
// would run the forward tiles at the adjoint's register count (172 VGPRs,
// two wavefronts per SIMD).
template <int TY, int PH>
__global__ __launch_bounds__(512) void serve_pair(const ServeItem *__restrict__ items,
                                                  const int32_t *__restrict__ start, int n, int q,
                                                  int first) {
  extern __shared__ double red[];
  const int64_t flat = first + xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  int lo = 0, hi = n - 1;  // largest i with start[i] <= flat
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= flat)
      lo = mid;
    else
      hi = mid - 1;
  }
  const ServeItem &it = items[lo];
  const int k = it.k0 + 2 * q;
  const int par = it.par0 ^ (q & 1);
  const int local = (int)(flat - start[lo]);
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (it.b.n_shots == 1 && blockDim.y > 1) {
    // single-shot job: each wavefront its own tile; the gradient goes
    // straight to the job's array (its tiles own disjoint cells)
    const int n_xt = it.b.pts2.n_xt;
    const int tile = local * (int)blockDim.y + w;
    if constexpr (PH == 0) {
      if (tile >= n_xt * it.b.n_yt2) return;  // uniform over the wavefront; no barriers here
      const Fwd2Args<float> a = item_fwd2_args(it.b, k, par);
      int edge = 3;
      const TileCoord tc = ordered_tile<TY, TY>(a.split, n_xt, a.g.ny, 1, tile, false, &edge);
      fwd2_serve_tile<TY>(a, 0, tc, edge, it.fmt);
    } else {
      // ordered tiles: 2-row border tiles, TY-row interior tiles
      const Adj2Args<float> a = item_adj2_args(it.b, k, par);
      if (tile >= ordered_tiles<TY, 2>(a.split, n_xt, a.g.ny)) return;  // uniform over the wavefront
      int edge = 3;
      const TileCoord tc = ordered_tile<TY, 2>(a.split, n_xt, a.g.ny, 1, tile, false, &edge);
      if (it.fmt == 2) {
        if (edge == 0)
          adj2_tile<float, TY, 2, 0>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
        else if (edge == 1)
          adj2_tile<float, 2, 2, 1>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
        else
          adj2_tile<float, 2, 2, 3>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
      } else {
        if (edge == 0)
          adj2_tile<float, TY, 1, 0>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
        else if (edge == 1)
          adj2_tile<float, 2, 1, 1>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
        else
          adj2_tile<float, 2, 1, 3>(a, 0, tc.xt, tc.yt, nullptr, nullptr, false);
      }
    }
    return;
  }
  if constexpr (PH == 0) {
    // ordered tiles with border tiles of the interior height (the same tile
    // count as the TY-row grid), as the single-problem forward
    const Fwd2Args<float> a = item_fwd2_args(it.b, k, par);
    int edge = 3;
    const TileCoord tc = ordered_tile<TY, TY>(a.split, it.b.pts2.n_xt, a.g.ny, it.b.n_groups, local, false, &edge);
    const int s = tc.grp * (int)blockDim.y + w;
    if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
    fwd2_serve_tile<TY>(a, s, tc, edge, it.fmt);
  } else {
    // ordered tiles, as the single-problem adjoint: 2-row border tiles keep
    // the register count of the CPML paths down
    const Adj2Args<float> a = item_adj2_args(it.b, k, par);
    int edge = 3;
    const TileCoord tc = ordered_tile<TY, 2>(a.split, it.b.pts2.n_xt, a.g.ny, it.b.n_groups, local, false, &edge);
    // edge is uniform over the workgroup
    if (it.fmt == 2) {
      if (edge == 0)
        adj2_workgroup<float, TY, 2, 1>(a, tc, 0, red);
      else
        adj2_workgroup<float, 2, 2, 2 | 8>(a, tc, edge, red);
    } else {
      if (edge == 0)
        adj2_workgroup<float, TY, 1, 1>(a, tc, 0, red);
      else
        adj2_workgroup<float, 2, 1, 2 | 8>(a, tc, edge, red);
    }
  }
}

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct EJob {
  int64_t id = 0;
  tide_tm2d_job spec;
  BatchItem<float> item;
  PointTable table, table2;  // one-step tiles; two-step tiles (with halo copies)
  int par = 0, epar = 0;     // current state buffer / E-phase memory copy (two-step engines)
  char *block = nullptr;  // packed profiles + point table (device)
  size_t block_bytes = 0;
  bool uploaded = false;
  bool grad = false;
  int fmt = 1;            // snapshot payload when storing
  int phase = 0;          // 0 forward, 1 forward done, 2 adjoint, 3 adjoint done
  int64_t k = 0;
  int64_t nt = 0;
  int n_groups = 1;
  int64_t weight = 1;  // step multiplier (tide_tm2d_engine_set_weight)
};

struct Slot {
  char *host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};

struct Engine {
  int device = 0;
  int group = 4;
  int rows_f = 4, rows_a = 2;
  bool two_step = false;  // launches of two steps (set from the time block at creation)
  int64_t max_jobs = 64;
  // most steps a small job takes per tick, as a multiple of the largest
  // job's (tide_tm2d_engine_set_boost)
  int64_t boost = 4;
  std::map<int64_t, EJob> jobs;
  int64_t next_id = 1;
  ServeItem *d_items = nullptr;
  int32_t *d_start = nullptr, *d_start2 = nullptr;
  Slot slots[2];
  int cur = 0;
  std::vector<int64_t> done;  // jobs whose phase ended since the last report
};

// The (forward rows, adjoint rows) pairs the step kernel is built for; the
// tile counts of a tick use exactly the pair the launch uses.
void pick_rows(int rows_f, int rows_a, int *out_f, int *out_a) {
  if (rows_f >= 8) {
    *out_f = 8;
    *out_a = 4;
  } else if (rows_f <= 2) {
    *out_f = 2;
    *out_a = 2;
  } else {
    *out_f = 4;
    *out_a = rows_a >= 4 ? 4 : 2;
  }
}

void launch_step(const Engine &e, int64_t grid, int n, int j, hipStream_t s, int first = 0,
                 int pair_mode = 0) {
  const dim3 block(64, (unsigned)e.group);
  const size_t lds = 2 * (size_t)e.group * e.rows_a * 64 * sizeof(double);
  if (e.rows_f == 8 && e.rows_a == 4)
    serve_step<8, 4><<<(unsigned)grid, block, lds, s>>>(e.d_items, e.d_start, n, j, first, pair_mode);
  else if (e.rows_f == 2 && e.rows_a == 2)
    serve_step<2, 2><<<(unsigned)grid, block, lds, s>>>(e.d_items, e.d_start, n, j, first, pair_mode);
  else if (e.rows_f == 4 && e.rows_a == 4)
    serve_step<4, 4><<<(unsigned)grid, block, lds, s>>>(e.d_items, e.d_start, n, j, first, pair_mode);
  else
    serve_step<4, 2><<<(unsigned)grid, block, lds, s>>>(e.d_items, e.d_start, n, j, first, pair_mode);
}

// Tiles [first, first + grid) of jobs [0, n), all of one phase (0 forward,
// 1 adjoint).
void launch_pair(const Engine &e, int phase, int64_t grid, int n, int q, int64_t first,
                 hipStream_t s) {
  const dim3 block(64, (unsigned)e.group);
  if (phase == 0) {
    serve_pair<kServeRows2, 0><<<(unsigned)grid, block, 0, s>>>(e.d_items, e.d_start2, n, q, (int)first);
  } else {
    const size_t lds = 2 * (size_t)e.group * kServeRows2 * 64 * sizeof(double);
    serve_pair<kServeRows2, 1><<<(unsigned)grid, block, lds, s>>>(e.d_items, e.d_start2, n, q, (int)first);
  }
}

int slot_reserve(Slot &sl, size_t bytes) {
  if (sl.pending) {
    TIDE_HIP_TRY(hipEventSynchronize(sl.ev));
    sl.pending = false;
  }
  if (bytes <= sl.cap) return kOk;
  if (sl.host) TIDE_HIP_TRY(hipHostFree(sl.host));
  sl.host = nullptr;
  sl.cap = 0;
  const size_t cap = std::max<size_t>(bytes * 2, (size_t)1 << 20);
  TIDE_HIP_TRY(hipHostMalloc((void **)&sl.host, cap, hipHostMallocDefault));
  sl.cap = cap;
  return kOk;
}

size_t point_bytes(const PointTable &t) {
  return align_up(t.row_slot.size() * sizeof(int32_t)) + align_up(t.seg.size() * sizeof(int32_t)) +
         align_up(std::max<size_t>(1, t.ent.size()) * sizeof(int2));
}

size_t table_bytes(const EJob &j) {
  const int64_t ny = j.spec.p.ny, nx = j.spec.p.nx;
  return align_up(6 * (size_t)(ny + nx) * sizeof(float)) + point_bytes(j.table) +
         (j.table2.row_slot.empty() ? 0 : point_bytes(j.table2));
}

// Allocates the job's device block and enqueues its profile and table
// copies; host bytes are staged at `stage`.  Returns the staged bytes.
int upload_job(EJob &j, char *stage, size_t *staged, hipStream_t s) {
  const tide_tm2d_problem &q = j.spec.p;
  j.block_bytes = table_bytes(j);
  TIDE_HIP_TRY(hipMallocAsync((void **)&j.block, j.block_bytes, s));
  char *d = j.block;
  float *prof = (float *)d;
  const int ysrc[6] = {0, 1, 8, 2, 3, 9}, xsrc[6] = {4, 5, 10, 6, 7, 11};
  for (int k = 0; k < 6; ++k) {
    TIDE_HIP_TRY(hipMemcpyAsync(prof + (size_t)k * q.ny, q.prof[ysrc[k]], (size_t)q.ny * sizeof(float),
                                hipMemcpyDeviceToDevice, s));
    TIDE_HIP_TRY(hipMemcpyAsync(prof + (size_t)6 * q.ny + (size_t)k * q.nx, q.prof[xsrc[k]],
                                (size_t)q.nx * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  d += align_up(6 * (size_t)(q.ny + q.nx) * sizeof(float));
  size_t off = 0;
  auto put = [&](const void *src, size_t bytes, char **dst) -> int {
    std::memcpy(stage + off, src, bytes);
    const hipError_t e = hipMemcpyAsync(d, stage + off, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      set_error(std::string("serve upload: ") + hipGetErrorString(e));
      return kHipError;
    }
    *dst = d;
    d += align_up(bytes);
    off += align_up(bytes);
    return kOk;
  };
  const int2 zero = make_int2(0, 0);
  auto put_table = [&](const PointTable &t, Points *out) -> int {
    char *row_slot, *seg, *ent;
    int rc = put(t.row_slot.data(), t.row_slot.size() * sizeof(int32_t), &row_slot);
    if (rc == kOk) rc = put(t.seg.data(), t.seg.size() * sizeof(int32_t), &seg);
    if (rc == kOk)
      rc = t.ent.empty() ? put(&zero, sizeof(int2), &ent) : put(t.ent.data(), t.ent.size() * sizeof(int2), &ent);
    *out = Points{(const int32_t *)row_slot, (const int32_t *)seg, (const int2 *)ent, t.n_xt};
    return rc;
  };
  if (int rc = put_table(j.table, &j.item.pts); rc != kOk) return rc;
  if (!j.table2.row_slot.empty()) {
    if (int rc = put_table(j.table2, &j.item.pts2); rc != kOk) return rc;
    j.item.n_yt2 = (int32_t)((q.ny + kServeRows2 - 1) / kServeRows2);
    j.item.split2 = make_split(j.item.g, j.item.pts2.n_xt, j.item.n_yt2, kServeRows2);
  }
  j.item.g.py = prof;
  j.item.g.px = prof + (size_t)6 * q.ny;
  j.uploaded = true;
  j.table = PointTable();  // host copies no longer needed
  j.table2 = PointTable();
  *staged = off;
  return kOk;
}

int validate(const tide_tm2d_job &job, int group, EJob &out) {
  const tide_tm2d_problem &q = job.p;
  if (q.n_shots <= 0 || q.ny < 3 || q.nx < 3 || q.nt < 0 || q.interval <= 0 ||
      q.ny * q.nx >= ((int64_t)1 << 29) || q.n_shots > 4096 || q.nt >= ((int64_t)1 << 30)) {
    set_error("engine: a job needs n_shots in 1..4096, 3 <= ny, nx, ny*nx < 2^29, 0 <= nt < 2^30, interval >= 1");
    return kInvalidArgument;
  }
  if (!q.ca || !q.cb || !q.cq) {
    set_error("engine: coefficients are required");
    return kInvalidArgument;
  }
  for (int c = 0; c < 12; ++c)
    if (!q.prof[c]) {
      set_error("engine: all 12 CPML profiles are required");
      return kInvalidArgument;
    }
  for (int c = 0; c < 7; ++c)
    if (!q.fields[c] || !job.fields_b[c]) {
      set_error("engine: forward state needs fields[7] and fields_b[7]");
      return kInvalidArgument;
    }
  if ((q.n_src > 0 && (!q.f || !q.sources_i)) || (q.n_rec > 0 && (!q.r || !q.receivers_i))) {
    set_error("engine: sources need f and sources_i, receivers need r and receivers_i");
    return kInvalidArgument;
  }
  if (job.storage_format != 0 && job.storage_format != 1) {
    set_error("engine: storage_format must be 0 (fp32) or 1 (BF16)");
    return kInvalidArgument;
  }
  out.grad = q.lambda[0] != nullptr;
  out.n_groups = (int)((q.n_shots + group - 1) / group);
  if (out.grad) {
    for (int c = 0; c < 7; ++c)
      if (!q.lambda[c] || !job.lambda_b[c]) {
        set_error("engine: a gradient job needs lambda[7] and lambda_b[7]");
        return kInvalidArgument;
      }
    if (!q.ey_store || !q.curl_store || !q.grad_ca || !q.grad_cb || (q.n_rec > 0 && !q.grad_r)) {
      set_error("engine: a gradient job needs both snapshot buffers, grad_ca, grad_cb and grad_r");
      return kInvalidArgument;
    }
    if (out.n_groups > 1 && (!job.grad_ca_groups || !job.grad_cb_groups)) {
      set_error("engine: a gradient job with more shots than a workgroup holds needs grad_*_groups");
      return kInvalidArgument;
    }
  }
  return kOk;
}

void fill_item(EJob &j) {
  const tide_tm2d_problem &q = j.spec.p;
  BatchItem<float> &it = j.item;
  std::memset(&it, 0, sizeof(it));
  Geom<float> &g = it.g;
  g.rdy = q.rdy;
  g.rdx = q.rdx;
  g.ny = (int32_t)q.ny;
  g.nx = (int32_t)q.nx;
  g.numel = (int32_t)(q.ny * q.nx);
  g.pml_y0 = (int32_t)q.pml[0];
  g.pml_x0 = (int32_t)q.pml[1];
  g.pml_y1 = (int32_t)q.pml[2];
  g.pml_x1 = (int32_t)q.pml[3];
  g.pml_y1h = (int32_t)(q.pml[2] > q.pml[0] ? q.pml[2] - 1 : q.pml[0]);
  g.pml_x1h = (int32_t)(q.pml[3] > q.pml[1] ? q.pml[3] - 1 : q.pml[1]);
  it.ca = q.ca;
  it.cb = q.cb;
  it.cq = q.cq;
  for (int c = 0; c < 5; ++c) {
    it.fwd[0][c] = q.fields[c];
    it.fwd[1][c] = j.spec.fields_b[c];
  }
  it.mhzx = q.fields[6];
  it.mhxz = q.fields[5];
  it.mhzx_b = j.spec.fields_b[6];
  it.mhxz_b = j.spec.fields_b[5];
  it.f = q.f;
  it.r = q.r;
  j.fmt = j.spec.storage_format == 1 ? 2 : 1;
  const int64_t total = q.n_shots * q.ny * q.nx;
  it.store_step_bytes = total * (j.fmt == 2 ? 2 : (int64_t)sizeof(float));
  if (j.grad) {
    it.ey_store = q.ey_store;
    it.curl_store = q.curl_store;
    // caller's order ley lhx lhz ml_ey_x ml_ey_z ml_hx_z ml_hz_x;
    // kernel order ley lhx lhz m_hz_x m_hx_z m_ey_x m_ey_z
    const int perm[7] = {0, 1, 2, 6, 5, 3, 4};
    for (int c = 0; c < 7; ++c) {
      it.lam[0][c] = q.lambda[perm[c]];
      it.lam[1][c] = j.spec.lambda_b[perm[c]];
    }
    it.grad_r = q.grad_r;
    it.grad_f = q.grad_f;
    if (j.n_groups > 1) {
      it.gca = j.spec.grad_ca_groups;
      it.gcb = j.spec.grad_cb_groups;
      it.group_stride = q.ny * q.nx;
    } else {
      it.gca = q.grad_ca;
      it.gcb = q.grad_cb;
      it.group_stride = 0;
    }
  }
  it.n_shots = (int32_t)q.n_shots;
  it.n_src = (int32_t)q.n_src;
  it.n_rec = (int32_t)q.n_rec;
  it.nt = (int32_t)q.nt;
  it.interval = (int32_t)q.interval;
  it.n_groups = (int32_t)j.n_groups;
}

int engine_tick(Engine &e, int64_t max_steps, hipStream_t s, int64_t *n_running) {
  // two-step engines advance an even number of steps per tick (one step
  // when max_steps is 1); jobs with an odd remainder end on a one-step launch
  const bool pairs = e.two_step && max_steps >= 2;
  if (pairs) max_steps &= ~(int64_t)1;
  // Steps per job this tick: the largest running job (cells x shots per
  // step) takes max_steps, a smaller one proportionally more, up to
  // e.boost times: in lock-step launches a small job would otherwise move
  // at the large one's pace.  The host's per-job weight multiplies the
  // result (a priority: the server gives its oldest request the most).
  // Allotments are multiples of max_steps, so every non-final tick of a job
  // is still an even number of steps.
  auto cells = [](const EJob &j) {
    return (double)j.spec.p.n_shots * (double)j.spec.p.ny * (double)j.spec.p.nx;
  };
  double biggest = 0.0;
  for (auto &kv : e.jobs) {
    const EJob &j = kv.second;
    if ((j.phase == 0 || j.phase == 2) && j.nt > j.k) biggest = std::max(biggest, cells(j));
  }
  std::vector<std::pair<int64_t, EJob *>> run;  // (steps this tick, job)
  for (auto &kv : e.jobs) {
    EJob &j = kv.second;
    if (j.phase != 0 && j.phase != 2) continue;
    const int64_t left = j.nt - j.k;
    if (left <= 0) {
      j.phase += 1;
      e.done.push_back(j.id);
      continue;
    }
    const double ratio = biggest / std::max(cells(j), 1.0);
    const int64_t boost = std::max<int64_t>(1, std::min<int64_t>(e.boost, (int64_t)ratio));
    run.emplace_back(std::min<int64_t>(left, max_steps * boost * j.weight), &j);
  }
  if (run.empty()) {
    *n_running = 0;
    return kOk;
  }
  // Longest first; with two-step launches, forward jobs before adjoint jobs
  // (longest first within each), so a pair launch covers one phase.
  std::stable_sort(run.begin(), run.end(),
                   [pairs](const std::pair<int64_t, EJob *> &a, const std::pair<int64_t, EJob *> &b) {
                     if (pairs && a.second->phase != b.second->phase)
                       return a.second->phase < b.second->phase;
                     return a.first > b.first;
                   });
  const int n = (int)run.size();
  size_t need = align_up((size_t)n * sizeof(ServeItem)) + 2 * align_up((size_t)(n + 1) * sizeof(int32_t));
  for (auto &r : run)
    if (!r.second->uploaded) need += table_bytes(*r.second);
  Slot &sl = e.slots[e.cur];
  if (int rc = slot_reserve(sl, need); rc != kOk) return rc;
  size_t off = 0;
  for (auto &r : run) {
    if (r.second->uploaded) continue;
    size_t used = 0;
    if (int rc = upload_job(*r.second, sl.host + off, &used, s); rc != kOk) return rc;
    off += used;
  }
  ServeItem *items = (ServeItem *)(sl.host + off);
  int32_t *start = (int32_t *)(sl.host + off + align_up((size_t)n * sizeof(ServeItem)));
  int32_t *start2 = (int32_t *)((char *)start + align_up((size_t)(n + 1) * sizeof(int32_t)));
  int64_t tile = 0, tile2 = 0;
  for (int i = 0; i < n; ++i) {
    EJob &j = *run[(size_t)i].second;
    ServeItem &si = items[i];
    si.b = j.item;
    const int rows = j.phase == 0 ? e.rows_f : e.rows_a;
    si.b.n_yt = (int32_t)((j.spec.p.ny + rows - 1) / rows);
    si.b.tile0 = (int32_t)tile;
    si.phase = j.phase;
    si.fmt = j.phase == 0 ? (j.grad ? j.fmt : 0) : j.fmt;
    si.k0 = (int32_t)j.k;
    si.par0 = j.par;
    si.epar0 = j.epar;
    si.pad = 0;
    start[i] = (int32_t)tile;
    start2[i] = (int32_t)tile2;
    const int64_t tiles = (int64_t)si.b.pts.n_xt * si.b.n_yt;
    // single-shot jobs pack `group` tiles per workgroup (see serve_step)
    tile += j.spec.p.n_shots == 1 ? (tiles + e.group - 1) / e.group : tiles * j.n_groups;
    if (pairs) {
      // forward: TY-row tiles; adjoint: the ordered set (2-row border tiles)
      const int64_t tiles2 = j.phase == 0 ? (int64_t)si.b.pts2.n_xt * si.b.n_yt2
                                          : ordered_tiles<kServeRows2, 2>(si.b.split2, si.b.pts2.n_xt,
                                                                           (int)j.spec.p.ny);
      tile2 += j.spec.p.n_shots == 1 ? (tiles2 + e.group - 1) / e.group : tiles2 * j.n_groups;
    }
    if (tile >= ((int64_t)1 << 31) || tile2 >= ((int64_t)1 << 31)) {
      set_error("engine: too many tiles in one tick");
      return kInvalidArgument;
    }
  }
  start[n] = (int32_t)tile;
  start2[n] = (int32_t)tile2;
  TIDE_HIP_TRY(hipMemcpyAsync(e.d_items, items, (size_t)n * sizeof(ServeItem), hipMemcpyHostToDevice, s));
  TIDE_HIP_TRY(hipMemcpyAsync(e.d_start, start, (size_t)(n + 1) * sizeof(int32_t),
                              hipMemcpyHostToDevice, s));
  if (pairs)
    TIDE_HIP_TRY(hipMemcpyAsync(e.d_start2, start2, (size_t)(n + 1) * sizeof(int32_t),
                                hipMemcpyHostToDevice, s));
  TIDE_HIP_TRY(hipEventRecord(sl.ev, s));
  sl.pending = true;
  e.cur ^= 1;
  const int64_t steps = run[0].first;
  if (pairs) {
    // jobs [0, nf) run forward, [nf, n) adjoint, each segment longest first
    int nf = 0;
    while (nf < n && run[(size_t)nf].second->phase == 0) ++nf;
    // end of the jobs in [b, e) taking at least v steps
    auto at_least = [&](int b, int e, int64_t v) {
      int c = b;
      while (c < e && run[(size_t)c].first >= v) ++c;
      return c;
    };
    int64_t longest = 0;
    for (auto &r : run) longest = std::max(longest, r.first);
    for (int64_t q = 0; 2 * q < longest; ++q) {
      for (int ph = 0; ph < 2; ++ph) {
        const int b = ph == 0 ? 0 : nf, en = ph == 0 ? nf : n;
        const int m2 = at_least(b, en, 2 * q + 2), m1 = at_least(b, en, 2 * q + 1);
        if (m2 > b) launch_pair(e, ph, start2[m2] - start2[b], m2, (int)q, start2[b], s);
        if (m1 > m2) launch_step(e, start[m1] - start[m2], m1, (int)(2 * q), s, start[m2], 1);
      }
    }
  } else {
    int m = n;
    for (int64_t jstep = 0; jstep < steps; ++jstep) {
      while (m > 0 && run[(size_t)m - 1].first <= jstep) --m;
      launch_step(e, start[m], m, (int)jstep, s, 0, e.two_step ? 2 : 0);
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  int64_t still = 0;
  for (auto &r : run) {
    EJob &j = *r.second;
    if (e.two_step) {  // state buffers after this tick's launches
      const int64_t n_pairs = pairs ? r.first / 2 : 0, single = pairs ? (r.first & 1) : r.first;
      j.par ^= (int)((n_pairs + single) & 1);
      j.epar ^= (int)(n_pairs & 1);
    }
    j.k += r.first;
    if (j.k >= j.nt) {
      j.phase += 1;
      e.done.push_back(j.id);
    } else {
      ++still;
    }
  }
  *n_running = still;
  return kOk;
}

void engine_destroy(Engine *e) {
  if (e == nullptr) return;
  (void)hipSetDevice(e->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : e->jobs)
    if (kv.second.block) (void)hipFreeAsync(kv.second.block, nullptr);
  (void)hipDeviceSynchronize();
  if (e->d_items) (void)hipFree(e->d_items);
  if (e->d_start) (void)hipFree(e->d_start);
  if (e->d_start2) (void)hipFree(e->d_start2);
  for (Slot &sl : e->slots) {
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
  }
  delete e;
}

Engine *as_engine(void *p) {
  if (p == nullptr) set_error("engine: null handle");
  return (Engine *)p;
}

}  // namespace
}  // namespace tm2d
}  // namespace tide_hip

using namespace tide_hip;
using namespace tide_hip::tm2d;

extern "C" int tide_tm2d_engine_create(int64_t device, int64_t group, int64_t max_jobs, void **engine) {
  if (engine == nullptr || group < 1 || group > kMaxGroup || max_jobs < 1 || max_jobs > 65536) {
    set_error("engine: need an output handle, group in 1..8 and max_jobs in 1..65536");
    return kInvalidArgument;
  }
  *engine = nullptr;
  TIDE_HIP_TRY(hipSetDevice((int)device));
  Engine *e = new Engine();
  e->device = (int)device;
  e->group = (int)group;
  pick_rows(rows_for<float>(g_tuning.fwd_rows), rows_for<float>(g_tuning.adj_rows), &e->rows_f, &e->rows_a);
  e->two_step = g_tuning.fwd_steps >= 2 && g_tuning.adj_steps >= 2;  // K-step settings run at 2
  e->max_jobs = max_jobs;
  hipError_t err = hipMalloc((void **)&e->d_items, (size_t)max_jobs * sizeof(ServeItem));
  if (err == hipSuccess) err = hipMalloc((void **)&e->d_start, (size_t)(max_jobs + 1) * sizeof(int32_t));
  if (err == hipSuccess) err = hipMalloc((void **)&e->d_start2, (size_t)(max_jobs + 1) * sizeof(int32_t));
  for (Slot &sl : e->slots)
    if (err == hipSuccess) err = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
  if (err != hipSuccess) {
    set_error(std::string("engine: ") + hipGetErrorString(err));
    engine_destroy(e);
    return kHipError;
  }
  *engine = e;
  return kOk;
}

extern "C" int tide_tm2d_engine_set_boost(void *engine, int64_t boost) {
  Engine *e = as_engine(engine);
  if (e == nullptr || boost < 1 || boost > 64) {
    set_error("engine boost: 1..64 (most steps of a small job per tick, in largest-job steps)");
    return kInvalidArgument;
  }
  e->boost = boost;
  return kOk;
}

extern "C" int tide_tm2d_engine_set_weight(void *engine, int64_t job, int64_t weight) {
  Engine *e = as_engine(engine);
  if (e == nullptr || weight < 1 || weight > 64) {
    set_error("engine weight: 1..64 (step multiplier of one job)");
    return kInvalidArgument;
  }
  auto it = e->jobs.find(job);
  if (it == e->jobs.end()) {
    set_error("engine weight: unknown job");
    return kInvalidArgument;
  }
  it->second.weight = weight;
  return kOk;
}

extern "C" int tide_tm2d_engine_group(void *engine, int64_t *group) {
  Engine *e = as_engine(engine);
  if (e == nullptr || group == nullptr) return kInvalidArgument;
  *group = e->group;
  return kOk;
}

extern "C" int tide_tm2d_engine_add(void *engine, const tide_tm2d_job *job, int64_t *job_id) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  if (job == nullptr || job_id == nullptr) {
    set_error("engine: job and id output are required");
    return kInvalidArgument;
  }
  if ((int64_t)e->jobs.size() >= e->max_jobs) {
    set_error("engine: job table full (max_jobs)");
    return kInvalidArgument;
  }
  EJob j;
  j.spec = *job;
  if (int rc = validate(*job, e->group, j); rc != kOk) return rc;
  const tide_tm2d_problem &q = job->p;
  std::vector<int64_t> hs((size_t)(q.n_shots * q.n_src)), hr((size_t)(q.n_shots * q.n_rec));
  // the index arrays must be host memory: a device read here would order
  // this call behind all queued work
  hipPointerAttribute_t attr;
  for (const int64_t *p : {q.sources_i, q.receivers_i}) {
    if (p == nullptr) continue;
    std::memset(&attr, 0, sizeof(attr));
    if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeDevice) {
      set_error("engine: sources_i / receivers_i must be host arrays");
      return kInvalidArgument;
    }
    (void)hipGetLastError();
  }
  if (!hs.empty()) std::memcpy(hs.data(), q.sources_i, hs.size() * sizeof(int64_t));
  if (!hr.empty()) std::memcpy(hr.data(), q.receivers_i, hr.size() * sizeof(int64_t));
  if (!build_point_table(hs.data(), q.n_src, hr.data(), q.n_rec, q.n_shots, q.ny, q.nx, kOwn, 1, j.table))
    return kInvalidArgument;
  if (e->two_step &&
      !build_point_table_halo(hs.data(), q.n_src, hr.data(), q.n_rec, q.n_shots, q.ny, q.nx, kOwn2, 2,
                              j.table2))
    return kInvalidArgument;
  fill_item(j);
  j.nt = q.nt;
  j.phase = 0;
  j.k = 0;
  const int64_t id = e->next_id++;
  j.id = id;
  e->jobs.emplace(id, std::move(j));
  *job_id = id;
  return kOk;
}

extern "C" int tide_tm2d_engine_tick(void *engine, int64_t max_steps, void *stream, int64_t *n_running) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  if (max_steps < 1 || n_running == nullptr) {
    set_error("engine: max_steps >= 1 and a running-count output are required");
    return kInvalidArgument;
  }
  TIDE_HIP_TRY(hipSetDevice(e->device));
  return engine_tick(*e, max_steps, (hipStream_t)stream, n_running);
}

extern "C" int tide_tm2d_engine_finished(void *engine, int64_t *ids, int64_t cap, int64_t *n) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  if (n == nullptr || (cap > 0 && ids == nullptr)) {
    set_error("engine: output array and count are required");
    return kInvalidArgument;
  }
  const int64_t m = std::min<int64_t>(cap, (int64_t)e->done.size());
  for (int64_t i = 0; i < m; ++i) ids[i] = e->done[(size_t)i];
  e->done.erase(e->done.begin(), e->done.begin() + m);
  *n = m;
  return kOk;
}

extern "C" int tide_tm2d_engine_query(void *engine, int64_t job_id, int64_t *phase, int64_t *step) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  auto it = e->jobs.find(job_id);
  if (it == e->jobs.end()) {
    set_error("engine: unknown job id");
    return kInvalidArgument;
  }
  if (phase) *phase = it->second.phase;
  if (step) *step = it->second.k;
  return kOk;
}

extern "C" int tide_tm2d_engine_begin_adjoint(void *engine, int64_t job_id) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  auto it = e->jobs.find(job_id);
  if (it == e->jobs.end()) {
    set_error("engine: unknown job id");
    return kInvalidArgument;
  }
  EJob &j = it->second;
  if (!j.grad || j.phase != 1) {
    set_error("engine: the adjoint starts after the forward of a gradient job");
    return kInvalidArgument;
  }
  j.phase = 2;
  j.k = 0;
  j.par = j.epar = 0;  // the adjoint state starts in the parity-0 buffers
  return kOk;
}

extern "C" int tide_tm2d_engine_remove(void *engine, int64_t job_id, void *stream) {
  Engine *e = as_engine(engine);
  if (e == nullptr) return kInvalidArgument;
  auto it = e->jobs.find(job_id);
  if (it == e->jobs.end()) {
    set_error("engine: unknown job id");
    return kInvalidArgument;
  }
  TIDE_HIP_TRY(hipSetDevice(e->device));
  EJob &j = it->second;
  hipStream_t s = (hipStream_t)stream;
  if (j.phase == 3 && j.n_groups > 1) {
    const int64_t numel = j.spec.p.ny * j.spec.p.nx;
    const unsigned blocks = (unsigned)((numel + 255) / 256);
    fold_groups<<<blocks, 256, 0, s>>>(j.spec.p.grad_ca, j.spec.grad_ca_groups, numel, j.n_groups);
    fold_groups<<<blocks, 256, 0, s>>>(j.spec.p.grad_cb, j.spec.grad_cb_groups, numel, j.n_groups);
    TIDE_HIP_TRY(hipGetLastError());
  }
  if (j.block) TIDE_HIP_TRY(hipFreeAsync(j.block, s));
  e->jobs.erase(it);
  return kOk;
}

extern "C" int tide_tm2d_engine_destroy(void *engine) {
  engine_destroy((Engine *)engine);
  return kOk;
}
