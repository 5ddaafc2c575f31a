// Batched 2-D TM plans: many independent problems (different grid sizes,
// shot counts and nt) advanced together, one launch per time step, or (the
// default, as in the single-problem path) one launch per two time steps
// with the two-step kernels of tm2d_fused.h; a problem whose nt is odd takes
// its last step in a one-step launch over its own tiles.
//
// Every problem keeps its own arrays; a plan holds a device table with one
// descriptor per problem and a tile -> problem map.  Problems are sorted by
// nt (longest first), so the tiles still running at step k form a prefix
// of the tile list: step k launches exactly that prefix.  Forward step k
// advances every problem with nt > k by its step k; adjoint iteration k
// applies every such problem's step nt-1-k.  A workgroup is a tile of one
// problem, so the tile kernels are the single-problem ones (tm2d_device.h)
// fed from the descriptor.
//
// The time loops are captured once into HIP graphs when the plan is built
// and replayed by every run: the per-run resets are fill kernels and
// device copies (no memset nodes), the per-step kernels read the
// problems' inputs through fixed pointers, so a caller updates
// coefficients or sources in place between runs.  With graphs off, the same
// sequence is launched eagerly.

#include "tide_hip.h"
#include "tm2d_items.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace tide_hip {
namespace tm2d {

// One-step launches over the flat tiles [first, first + gridDim.x) of the
// one-step tile map; `par` / `epar` select the state buffers (see
// item_fwd_args).
template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void fwd_batch(const BatchItem<T> *__restrict__ items,
                                                 const int32_t *__restrict__ tile_item, int k,
                                                 int first, int par, int epar) {
  const int64_t flat = first + xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  const BatchItem<T> &it = items[tile_item[flat]];
  const TileCoord tc = batch_tile((int)(flat - it.tile0), it.pts.n_xt, it.n_yt);
  fwd_workgroup<T, TY, FMT>(item_fwd_args(it, k, par, epar), tc);
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void adj_batch(const BatchItem<T> *__restrict__ items,
                                                 const int32_t *__restrict__ tile_item, int k,
                                                 int first, int par) {
  extern __shared__ double red[];
  const int64_t flat = first + xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  const BatchItem<T> &it = items[tile_item[flat]];
  const TileCoord tc = batch_tile((int)(flat - it.tile0), it.pts.n_xt, it.n_yt);
  adj_workgroup<T, TY, FMT>(item_adj_args(it, k, par), tc, red);
}

// Two-step launches over the first gridDim.x tiles of the two-step tile map
// (the problems with at least two steps left at step k).
template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void fwd2_batch(const BatchItem<T> *__restrict__ items,
                                                  const int32_t *__restrict__ tile_item, int k,
                                                  int par) {
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  const BatchItem<T> &it = items[tile_item[flat]];
  const Fwd2Args<T> a = item_fwd2_args(it, k, par);
  // ordered tiles with border tiles of the interior height (the same tile
  // count as the TY-row grid): border tiles first, x-only checks beside the
  // interior rows, as the single-problem forward
  int edge = 3;
  const TileCoord tc = ordered_tile<TY, TY>(a.split, it.pts2.n_xt, a.g.ny, it.n_groups,
                                            flat - it.tile0_2, false, &edge);
  const int s = tc.grp * (int)blockDim.y + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (s >= a.n_shots) return;  // uniform over the wavefront; no barriers here
  if (edge == 0)
    fwd2_tile<T, TY, FMT, 0>(a, s, tc.xt, tc.yt);
  else if (edge == 1)
    fwd2_tile<T, TY, FMT, 1>(a, s, tc.xt, tc.yt);
  else
    fwd2_tile<T, TY, FMT, 3>(a, s, tc.xt, tc.yt);
}

template <typename T, int TY, int FMT>
__global__ __launch_bounds__(512) void adj2_batch(const BatchItem<T> *__restrict__ items,
                                                  const int32_t *__restrict__ tile_item, int k,
                                                  int par) {
  extern __shared__ double red[];
  const int64_t flat = xcd_remap((int64_t)blockIdx.x, (int64_t)gridDim.x);
  const BatchItem<T> &it = items[tile_item[flat]];
  const Adj2Args<T> a = item_adj2_args(it, k, par);
  // ordered tiles, as the single-problem adjoint: 2-row border tiles keep
  // the register count of the CPML paths down; edge is uniform over the
  // workgroup
  int edge = 3;
  const TileCoord tc = ordered_tile<TY, 2>(a.split, it.pts2.n_xt, a.g.ny, it.n_groups,
                                           flat - it.tile0_2, false, &edge);
  if (edge == 0)
    adj2_workgroup<T, TY, FMT, 1>(a, tc, 0, red);
  else
    adj2_workgroup<T, 2, FMT, 2 | 8>(a, tc, edge, red);
}

template <typename T>
__global__ void fill_zero(T *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

// ------------------------------------------------------------------ plan

struct Plan {
  int device = 0;
  int64_t n = 0;
  int fmt_fwd = 0, fmt_adj = 0;  // 0 none, 1 T payload, 2 BF16
  int rows_f = 8, rows_a = 4, group = 8;
  int steps_f = 1, steps_a = 1;  // time steps per launch of each pass
  int max_nt = 0;
  std::vector<tide_tm2d_problem> probs;  // sorted by nt, descending
  std::vector<void *> allocs;            // plan-owned device memory
  BatchItem<float> *d_fwd = nullptr, *d_adj = nullptr;
  int32_t *d_tiles_fwd = nullptr, *d_tiles_adj = nullptr;
  std::vector<int64_t> active_fwd, active_adj;  // tiles running at step k
  // two-step launches: tiles with >= 2 steps left at step 2j (active2[j]),
  // and the one-step tile range [rem_*[j], rem_*[j+1]...) of the problems
  // taking their single last step at step 2j (first, end per j)
  int32_t *d_tiles2 = nullptr;   // forward two-step tile -> problem
  int32_t *d_tiles2a = nullptr;  // adjoint two-step (ordered) tile -> problem
  std::vector<int64_t> active2a;  // adjoint tiles with >= 2 steps left at step 2j
  std::vector<int64_t> active2, rem_f_first, rem_f_end, rem_a_first, rem_a_end;
  std::vector<BatchItem<float>> h_fwd, h_adj;
  std::vector<DevicePointTable> dpts, dpts2;
  hipStream_t capture = nullptr;
  hipGraphExec_t g_fwd = nullptr, g_adj = nullptr;
  hipGraph_t gr_fwd = nullptr, gr_adj = nullptr;
  bool use_graph = false;
};

namespace {

template <typename U> int plan_alloc(Plan &p, U **out, size_t count) {
  void *ptr = nullptr;
  TIDE_HIP_TRY(hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(U)));
  p.allocs.push_back(ptr);
  *out = (U *)ptr;
  return kOk;
}

int fetch_host(const int64_t *p, int64_t n, std::vector<int64_t> &out) {
  out.assign((size_t)std::max<int64_t>(n, 0), 0);
  if (n <= 0 || p == nullptr) return kOk;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeDevice) {
    TIDE_HIP_TRY(hipMemcpy(out.data(), p, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost));
  } else {
    (void)hipGetLastError();
    std::memcpy(out.data(), p, (size_t)n * sizeof(int64_t));
  }
  return kOk;
}

#define PLAN_CHECK(expr)                    \
  do {                                      \
    const int _rc = (expr);                 \
    if (_rc != ::tide_hip::kOk) return _rc; \
  } while (0)

template <typename T>
void launch_zero(T *p, int64_t n, hipStream_t s) {
  if (p == nullptr || n <= 0) return;
  const int threads = 256;
  const int64_t blocks = std::min<int64_t>((n + threads - 1) / threads, 4096);
  fill_zero<T><<<(unsigned)blocks, threads, 0, s>>>(p, n);
}

template <int FMT>
void launch_fwd_b(int rows, int64_t grid, int group, hipStream_t s, const BatchItem<float> *d,
                  const int32_t *tiles, int k, int first = 0, int par = -1, int epar = 0) {
  const dim3 block(64, (unsigned)group);
  if (par < 0) par = k & 1;
  switch (rows) {
    case 2: fwd_batch<float, 2, FMT><<<(unsigned)grid, block, 0, s>>>(d, tiles, k, first, par, epar); break;
    case 4: fwd_batch<float, 4, FMT><<<(unsigned)grid, block, 0, s>>>(d, tiles, k, first, par, epar); break;
    case 16: fwd_batch<float, 16, FMT><<<(unsigned)grid, block, 0, s>>>(d, tiles, k, first, par, epar); break;
    default: fwd_batch<float, 8, FMT><<<(unsigned)grid, block, 0, s>>>(d, tiles, k, first, par, epar); break;
  }
}

template <int FMT>
void launch_adj_b(int rows, int64_t grid, int group, hipStream_t s, const BatchItem<float> *d,
                  const int32_t *tiles, int k, int first = 0, int par = -1) {
  const dim3 block(64, (unsigned)group);
  const size_t lds = 2 * (size_t)group * rows * 64 * sizeof(double);
  if (par < 0) par = k & 1;
  switch (rows) {
    case 2: adj_batch<float, 2, FMT><<<(unsigned)grid, block, lds, s>>>(d, tiles, k, first, par); break;
    case 8: adj_batch<float, 8, FMT><<<(unsigned)grid, block, lds, s>>>(d, tiles, k, first, par); break;
    default: adj_batch<float, 4, FMT><<<(unsigned)grid, block, lds, s>>>(d, tiles, k, first, par); break;
  }
}

constexpr int kBatchRows2 = 4;  // tile rows of the batched two-step kernels

void launch_step_f(const Plan &p, int fmt, int64_t grid, hipStream_t s, int k, int first, int par,
                   int epar) {
  if (fmt == 0) launch_fwd_b<0>(p.rows_f, grid, p.group, s, p.d_fwd, p.d_tiles_fwd, k, first, par, epar);
  else if (fmt == 2) launch_fwd_b<2>(p.rows_f, grid, p.group, s, p.d_fwd, p.d_tiles_fwd, k, first, par, epar);
  else launch_fwd_b<1>(p.rows_f, grid, p.group, s, p.d_fwd, p.d_tiles_fwd, k, first, par, epar);
}

void launch_step_a(const Plan &p, int fmt, int64_t grid, hipStream_t s, int k, int first, int par) {
  if (fmt == 0) launch_adj_b<0>(p.rows_a, grid, p.group, s, p.d_adj, p.d_tiles_adj, k, first, par);
  else if (fmt == 2) launch_adj_b<2>(p.rows_a, grid, p.group, s, p.d_adj, p.d_tiles_adj, k, first, par);
  else launch_adj_b<1>(p.rows_a, grid, p.group, s, p.d_adj, p.d_tiles_adj, k, first, par);
}

void launch_pair_f(const Plan &p, int fmt, int64_t grid, hipStream_t s, int k, int par) {
  const dim3 block(64, (unsigned)p.group);
  constexpr int TY = kBatchRows2;
  if (fmt == 0) fwd2_batch<float, TY, 0><<<(unsigned)grid, block, 0, s>>>(p.d_fwd, p.d_tiles2, k, par);
  else if (fmt == 2) fwd2_batch<float, TY, 2><<<(unsigned)grid, block, 0, s>>>(p.d_fwd, p.d_tiles2, k, par);
  else fwd2_batch<float, TY, 1><<<(unsigned)grid, block, 0, s>>>(p.d_fwd, p.d_tiles2, k, par);
}

void launch_pair_a(const Plan &p, int fmt, int64_t grid, hipStream_t s, int k, int par) {
  const dim3 block(64, (unsigned)p.group);
  constexpr int TY = kBatchRows2;
  const size_t lds = 2 * (size_t)p.group * TY * 64 * sizeof(double);
  if (fmt == 0) adj2_batch<float, TY, 0><<<(unsigned)grid, block, lds, s>>>(p.d_adj, p.d_tiles2a, k, par);
  else if (fmt == 2) adj2_batch<float, TY, 2><<<(unsigned)grid, block, lds, s>>>(p.d_adj, p.d_tiles2a, k, par);
  else adj2_batch<float, TY, 1><<<(unsigned)grid, block, lds, s>>>(p.d_adj, p.d_tiles2a, k, par);
}

// Enqueues one whole forward run on `s`.
int enqueue_forward(Plan &p, hipStream_t s) {
  const size_t fb = sizeof(float);
  for (size_t m = 0; m < p.probs.size(); ++m) {
    const tide_tm2d_problem &q = p.probs[m];
    const BatchItem<float> &it = p.h_fwd[m];
    const int64_t total = q.n_shots * q.ny * q.nx;
    // the second copy of the H-phase memories starts equal to the first
    TIDE_HIP_TRY(hipMemcpyAsync(it.fwd[1][3], it.fwd[0][3], total * fb, hipMemcpyDeviceToDevice, s));
    TIDE_HIP_TRY(hipMemcpyAsync(it.fwd[1][4], it.fwd[0][4], total * fb, hipMemcpyDeviceToDevice, s));
    if (p.steps_f == 2) {  // and of the E-phase memories, which the two-step kernels ping-pong
      TIDE_HIP_TRY(hipMemcpyAsync(it.mhzx_b, it.mhzx, total * fb, hipMemcpyDeviceToDevice, s));
      TIDE_HIP_TRY(hipMemcpyAsync(it.mhxz_b, it.mhxz, total * fb, hipMemcpyDeviceToDevice, s));
    }
    if (q.n_rec > 0) launch_zero(q.r, q.nt * q.n_shots * q.n_rec, s);
  }
  if (p.steps_f == 2) {
    for (int k = 0; k < p.max_nt; k += 2) {
      const size_t j = (size_t)k / 2;
      const int par = (int)(j & 1);
      const int64_t grid2 = p.active2[j];
      const int64_t first = p.rem_f_first[j], end = p.rem_f_end[j];
      if (grid2 > 0) launch_pair_f(p, p.fmt_fwd, grid2, s, k, par);
      if (end > first) launch_step_f(p, p.fmt_fwd, end - first, s, k, (int)first, par, par);
    }
  } else {
    for (int k = 0; k < p.max_nt; ++k) {
      const int64_t grid = p.active_fwd[(size_t)k];
      if (grid <= 0) break;
      launch_step_f(p, p.fmt_fwd, grid, s, k, 0, k & 1, 0);
    }
  }
  // final state back into the caller's arrays
  for (size_t m = 0; m < p.probs.size(); ++m) {
    const tide_tm2d_problem &q = p.probs[m];
    const BatchItem<float> &it = p.h_fwd[m];
    const int64_t total = q.n_shots * q.ny * q.nx;
    const bool two = p.steps_f == 2;
    const int fpar = (int)((two ? (q.nt + 1) / 2 : q.nt) & 1), epar = two ? (int)((q.nt / 2) & 1) : 0;
    if (fpar)
      for (int c = 0; c < 5; ++c)
        TIDE_HIP_TRY(hipMemcpyAsync(it.fwd[0][c], it.fwd[1][c], total * fb, hipMemcpyDeviceToDevice, s));
    if (epar) {
      TIDE_HIP_TRY(hipMemcpyAsync(it.mhzx, it.mhzx_b, total * fb, hipMemcpyDeviceToDevice, s));
      TIDE_HIP_TRY(hipMemcpyAsync(it.mhxz, it.mhxz_b, total * fb, hipMemcpyDeviceToDevice, s));
    }
  }
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

// Enqueues one whole adjoint run on `s` (adjoint state and gradients start at zero).
int enqueue_backward(Plan &p, hipStream_t s) {
  const size_t fb = sizeof(float);
  for (size_t m = 0; m < p.probs.size(); ++m) {
    const tide_tm2d_problem &q = p.probs[m];
    const BatchItem<float> &it = p.h_adj[m];
    const int64_t total = q.n_shots * q.ny * q.nx;
    for (int par = 0; par < 2; ++par)
      for (int c = 0; c < 7; ++c) launch_zero(it.lam[par][c], total, s);
    launch_zero(q.grad_ca, q.ny * q.nx, s);
    launch_zero(q.grad_cb, q.ny * q.nx, s);
    if (it.group_stride > 0) {
      if (it.gca) launch_zero(it.gca, it.n_groups * it.group_stride, s);
      if (it.gcb) launch_zero(it.gcb, it.n_groups * it.group_stride, s);
    }
    if (q.grad_f && q.n_src > 0) launch_zero(q.grad_f, q.nt * q.n_shots * q.n_src, s);
  }
  if (p.steps_a == 2) {
    for (int k = 0; k < p.max_nt; k += 2) {
      const size_t j = (size_t)k / 2;
      const int par = (int)(j & 1);
      const int64_t grid2 = p.active2a[j];
      const int64_t first = p.rem_a_first[j], end = p.rem_a_end[j];
      if (grid2 > 0) launch_pair_a(p, p.fmt_adj, grid2, s, k, par);
      if (end > first) launch_step_a(p, p.fmt_adj, end - first, s, k, (int)first, par);
    }
  } else {
    for (int k = 0; k < p.max_nt; ++k) {
      const int64_t grid = p.active_adj[(size_t)k];
      if (grid <= 0) break;
      launch_step_a(p, p.fmt_adj, grid, s, k, 0, k & 1);
    }
  }
  for (size_t m = 0; m < p.probs.size(); ++m) {
    const BatchItem<float> &it = p.h_adj[m];
    if (it.group_stride == 0) continue;
    const tide_tm2d_problem &q = p.probs[m];
    const int64_t numel = q.ny * q.nx;
    const unsigned blocks = (unsigned)((numel + 255) / 256);
    if (it.gca) fold_groups<<<blocks, 256, 0, s>>>(q.grad_ca, it.gca, numel, it.n_groups);
    if (it.gcb) fold_groups<<<blocks, 256, 0, s>>>(q.grad_cb, it.gcb, numel, it.n_groups);
  }
  for (size_t m = 0; m < p.probs.size(); ++m) {
    const tide_tm2d_problem &q = p.probs[m];
    if ((((p.steps_a == 2) ? (q.nt + 1) / 2 : q.nt) & 1) == 0) continue;
    const BatchItem<float> &it = p.h_adj[m];
    const int64_t total = q.n_shots * q.ny * q.nx;
    for (int c = 0; c < 7; ++c)
      TIDE_HIP_TRY(hipMemcpyAsync(it.lam[0][c], it.lam[1][c], total * fb, hipMemcpyDeviceToDevice, s));
  }
  TIDE_HIP_TRY(hipGetLastError());
  return kOk;
}

int capture(Plan &p, bool forward, hipGraph_t *graph, hipGraphExec_t *exec) {
  TIDE_HIP_TRY(hipStreamBeginCapture(p.capture, hipStreamCaptureModeRelaxed));
  const int rc = forward ? enqueue_forward(p, p.capture) : enqueue_backward(p, p.capture);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p.capture, &g);
  if (rc != kOk) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  TIDE_HIP_TRY(e);
  *graph = g;
  TIDE_HIP_TRY(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return kOk;
}

void destroy(Plan *p) {
  if (p == nullptr) return;
  (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  if (p->g_fwd) (void)hipGraphExecDestroy(p->g_fwd);
  if (p->g_adj) (void)hipGraphExecDestroy(p->g_adj);
  if (p->gr_fwd) (void)hipGraphDestroy(p->gr_fwd);
  if (p->gr_adj) (void)hipGraphDestroy(p->gr_adj);
  for (DevicePointTable &d : p->dpts) d.release(nullptr);
  for (DevicePointTable &d : p->dpts2) d.release(nullptr);
  (void)hipDeviceSynchronize();
  for (void *a : p->allocs) (void)hipFree(a);
  if (p->capture) (void)hipStreamDestroy(p->capture);
  delete p;
}

// Builds descriptors for one tile shape; returns the tile -> item map.
int build_items(Plan &p, int rows, std::vector<BatchItem<float>> &items,
                std::vector<int32_t> &tile_item, std::vector<int64_t> &active,
                const std::vector<BatchItem<float>> &proto) {
  items = proto;
  tile_item.clear();
  int64_t tile0 = 0;
  std::vector<int64_t> tiles_of(p.probs.size());
  for (size_t m = 0; m < p.probs.size(); ++m) {
    BatchItem<float> &it = items[m];
    it.n_yt = (int32_t)((p.probs[m].ny + rows - 1) / rows);
    const int64_t n_groups = (it.n_shots + p.group - 1) / p.group;
    it.n_groups = (int32_t)n_groups;
    const int64_t tiles = (int64_t)it.pts.n_xt * it.n_yt * n_groups;
    if (tile0 + tiles >= ((int64_t)1 << 31)) {
      set_error("batch: too many tiles for one launch");
      return kInvalidArgument;
    }
    it.tile0 = (int32_t)tile0;
    tiles_of[m] = tiles;
    tile_item.insert(tile_item.end(), (size_t)tiles, (int32_t)m);
    tile0 += tiles;
  }
  active.assign((size_t)p.max_nt, 0);
  // problems are sorted by nt descending: those with nt > k are a prefix
  int64_t run = 0;
  size_t m = p.probs.size();
  std::vector<int64_t> prefix(p.probs.size() + 1, 0);
  for (size_t i = 0; i < p.probs.size(); ++i) prefix[i + 1] = prefix[i] + tiles_of[i];
  for (int k = 0; k < p.max_nt; ++k) {
    while (m > 0 && p.probs[m - 1].nt <= k) --m;
    run = prefix[m];
    active[(size_t)k] = run;
  }
  return kOk;
}

// Two-step tiles (kOwn2 columns x kBatchRows2 rows, the same shot groups as
// the one-step tiles): fills n_yt2 / tile0_2 / split2 of both passes' items
// and returns the two-step tile -> item maps (forward: TY-row tiles; adjoint:
// the ordered set with 2-row border tiles); also the per-pair launch sizes.
int build_items2(Plan &p, std::vector<int32_t> &tile_item, std::vector<int32_t> &tile_item_a) {
  constexpr int TY = kBatchRows2;
  const size_t n = p.probs.size();
  tile_item.clear();
  tile_item_a.clear();
  std::vector<int64_t> prefix(n + 1, 0), prefix_a(n + 1, 0);
  for (size_t m = 0; m < n; ++m) {
    BatchItem<float> &f = p.h_fwd[m];
    const int32_t n_yt2 = (int32_t)((p.probs[m].ny + TY - 1) / TY);
    const int64_t tiles = (int64_t)f.pts2.n_xt * n_yt2 * f.n_groups;
    if (prefix[m] + tiles >= ((int64_t)1 << 31)) {
      set_error("batch: too many tiles for one launch");
      return kInvalidArgument;
    }
    const TileSplit sp = make_split(f.g, f.pts2.n_xt, n_yt2, TY);
    const int64_t tiles_a = ordered_tiles<TY, 2>(sp, f.pts2.n_xt, (int)p.probs[m].ny) * p.h_adj[m].n_groups;
    if (prefix_a[m] + tiles_a >= ((int64_t)1 << 31)) {
      set_error("batch: too many tiles for one launch");
      return kInvalidArgument;
    }
    for (BatchItem<float> *it : {&p.h_fwd[m], &p.h_adj[m]}) {
      it->n_yt2 = n_yt2;
      it->split2 = sp;
    }
    p.h_fwd[m].tile0_2 = (int32_t)prefix[m];
    p.h_adj[m].tile0_2 = (int32_t)prefix_a[m];
    tile_item.insert(tile_item.end(), (size_t)tiles, (int32_t)m);
    tile_item_a.insert(tile_item_a.end(), (size_t)tiles_a, (int32_t)m);
    prefix[m + 1] = prefix[m] + tiles;
    prefix_a[m + 1] = prefix_a[m] + tiles_a;
  }
  // problems are sorted by nt descending: count those with nt >= v
  auto count_at_least = [&](int64_t v) {
    size_t c = 0;
    while (c < n && p.probs[c].nt >= v) ++c;
    return c;
  };
  const size_t pairs = (size_t)(p.max_nt + 1) / 2;
  p.active2.assign(pairs, 0);
  p.active2a.assign(pairs, 0);
  p.rem_f_first.assign(pairs, 0);
  p.rem_f_end.assign(pairs, 0);
  p.rem_a_first.assign(pairs, 0);
  p.rem_a_end.assign(pairs, 0);
  auto tile_start = [&](const std::vector<BatchItem<float>> &items, size_t m, int64_t end) {
    return m < n ? (int64_t)items[m].tile0 : end;
  };
  const int64_t end_f = n ? p.h_fwd[n - 1].tile0 + (int64_t)p.h_fwd[n - 1].pts.n_xt *
                                                      p.h_fwd[n - 1].n_yt * p.h_fwd[n - 1].n_groups
                          : 0;
  const int64_t end_a = n ? p.h_adj[n - 1].tile0 + (int64_t)p.h_adj[n - 1].pts.n_xt *
                                                      p.h_adj[n - 1].n_yt * p.h_adj[n - 1].n_groups
                          : 0;
  for (size_t j = 0; j < pairs; ++j) {
    const int64_t k = 2 * (int64_t)j;
    const size_t two = count_at_least(k + 2), one = count_at_least(k + 1);
    p.active2[j] = prefix[two];
    p.active2a[j] = prefix_a[two];
    p.rem_f_first[j] = tile_start(p.h_fwd, two, end_f);
    p.rem_f_end[j] = tile_start(p.h_fwd, one, end_f);
    p.rem_a_first[j] = tile_start(p.h_adj, two, end_a);
    p.rem_a_end[j] = tile_start(p.h_adj, one, end_a);
  }
  return kOk;
}

}  // namespace
}  // namespace tm2d
}  // namespace tide_hip

using namespace tide_hip;
using namespace tide_hip::tm2d;

namespace {
int build_plan(Plan *p, const tide_tm2d_problem *problems, int64_t n, int64_t storage_format,
               int64_t use_graph, int64_t device);
}  // namespace

extern "C" int tide_tm2d_plan_create(const tide_tm2d_problem *problems, int64_t n,
                                     int64_t storage_format, int64_t use_graph, int64_t device,
                                     void **plan_out) {
  if (plan_out == nullptr || problems == nullptr || n <= 0) {
    set_error("plan: problems and output handle are required");
    return kInvalidArgument;
  }
  *plan_out = nullptr;
  if (hipSetDevice((int)device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return kHipError;
  }
  Plan *p = new Plan();
  const int rc = build_plan(p, problems, n, storage_format, use_graph, device);
  if (rc != kOk) {
    destroy(p);
    return rc;
  }
  *plan_out = p;
  return kOk;
}

namespace {
int build_plan(Plan *p, const tide_tm2d_problem *problems, int64_t n, int64_t storage_format,
               int64_t use_graph, int64_t device) {
  p->device = (int)device;
  p->n = n;
  p->use_graph = use_graph != 0;
  p->rows_f = rows_for<float>(g_tuning.fwd_rows);
  p->rows_a = rows_for<float>(g_tuning.adj_rows);
  p->group = g_tuning.adj_group;  // one group size for both passes
  // plans run one- or two-step launches (a K-step setting runs at 2)
  p->steps_f = g_tuning.fwd_steps >= 2 ? 2 : 1;
  p->steps_a = g_tuning.adj_steps >= 2 ? 2 : 1;
  const bool two = p->steps_f == 2 || p->steps_a == 2;
  p->probs.assign(problems, problems + n);
  std::stable_sort(p->probs.begin(), p->probs.end(),
                   [](const tide_tm2d_problem &a, const tide_tm2d_problem &b) { return a.nt > b.nt; });
  const int fmt = storage_format == 1 ? 2 : 1;
  bool any_store = false, any_grad = false;
  auto fail = [](int rc) { return rc; };
  std::vector<BatchItem<float>> proto(p->probs.size());
  p->dpts.reserve(p->probs.size());
  p->dpts2.reserve(p->probs.size());
  for (size_t m = 0; m < p->probs.size(); ++m) {
    const tide_tm2d_problem &q = p->probs[m];
    if (q.n_shots <= 0 || q.ny < 3 || q.nx < 3 || q.nt < 0 || q.interval <= 0 ||
        q.ny * q.nx >= ((int64_t)1 << 29) || q.n_shots > 4096) {
      set_error("plan: each problem needs n_shots in 1..4096, ny, nx >= 3, nt >= 0, interval >= 1");
      return fail(kInvalidArgument);
    }
    if (q.nt >= ((int64_t)1 << 30)) {
      set_error("plan: nt too large");
      return fail(kInvalidArgument);
    }
    p->max_nt = std::max<int>(p->max_nt, (int)q.nt);
    any_store = any_store || q.ey_store || q.curl_store;
    any_grad = any_grad || q.grad_ca || q.grad_cb;
    if ((q.grad_ca && !q.ey_store) || (q.grad_cb && !q.curl_store)) {
      set_error("plan: coefficient gradients require the snapshot buffers");
      return fail(kInvalidArgument);
    }
    BatchItem<float> &it = proto[m];
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
    g.ca_b = g.cb_b = g.cq_b = 0;
    // profiles packed [a, b, k, ah, bh, kh] per axis
    float *pbuf;
    PLAN_CHECK(plan_alloc(*p, &pbuf, (size_t)6 * (q.ny + q.nx)));
    const int ysrc[6] = {0, 1, 8, 2, 3, 9}, xsrc[6] = {4, 5, 10, 6, 7, 11};
    for (int k = 0; k < 6; ++k) {
      TIDE_HIP_TRY(hipMemcpy(pbuf + (size_t)k * q.ny, q.prof[ysrc[k]], (size_t)q.ny * sizeof(float),
                             hipMemcpyDefault));
      TIDE_HIP_TRY(hipMemcpy(pbuf + (size_t)6 * q.ny + (size_t)k * q.nx, q.prof[xsrc[k]],
                             (size_t)q.nx * sizeof(float), hipMemcpyDefault));
    }
    g.py = pbuf;
    g.px = pbuf + (size_t)6 * q.ny;
    // point table
    std::vector<int64_t> hs, hr;
    PLAN_CHECK(fetch_host(q.sources_i, q.n_shots * q.n_src, hs));
    PLAN_CHECK(fetch_host(q.receivers_i, q.n_shots * q.n_rec, hr));
    PointTable table;
    if (!build_point_table(hs.data(), q.n_src, hr.data(), q.n_rec, q.n_shots, q.ny, q.nx, kOwn,
                           1, table))
      return fail(kInvalidArgument);
    p->dpts.emplace_back();
    DevicePointTable &dpt = p->dpts.back();
    PLAN_CHECK(dpt.upload(table, nullptr));
    it.pts = Points{dpt.row_slot, dpt.seg, dpt.ent, dpt.n_xt};
    if (two) {  // point table of the two-step tiles (halo copies included)
      PointTable table2;
      if (!build_point_table_halo(hs.data(), q.n_src, hr.data(), q.n_rec, q.n_shots, q.ny, q.nx,
                                  kOwn2, 2, table2))
        return fail(kInvalidArgument);
      p->dpts2.emplace_back();
      DevicePointTable &dpt2 = p->dpts2.back();
      PLAN_CHECK(dpt2.upload(table2, nullptr));
      it.pts2 = Points{dpt2.row_slot, dpt2.seg, dpt2.ent, dpt2.n_xt};
    }
    it.ca = q.ca;
    it.cb = q.cb;
    it.cq = q.cq;
    const int64_t total = q.n_shots * q.ny * q.nx;
    it.fwd[0][0] = q.fields[0];
    it.fwd[0][1] = q.fields[1];
    it.fwd[0][2] = q.fields[2];
    it.fwd[0][3] = q.fields[3];
    it.fwd[0][4] = q.fields[4];
    for (int c = 0; c < 5; ++c) PLAN_CHECK(plan_alloc(*p, &it.fwd[1][c], (size_t)total));
    for (int c = 3; c < 5; ++c) TIDE_HIP_TRY(hipMemset(it.fwd[1][c], 0, (size_t)total * sizeof(float)));
    it.mhzx = q.fields[6];
    it.mhxz = q.fields[5];
    if (p->steps_f == 2) {
      PLAN_CHECK(plan_alloc(*p, &it.mhzx_b, (size_t)total));
      PLAN_CHECK(plan_alloc(*p, &it.mhxz_b, (size_t)total));
    }
    it.f = q.f;
    it.r = q.r;
    it.ey_store = q.ey_store;
    it.curl_store = q.curl_store;
    it.store_step_bytes = total * (fmt == 2 ? 2 : (int64_t)sizeof(float));
    // adjoint state: caller's order ley lhx lhz ml_ey_x ml_ey_z ml_hx_z ml_hz_x;
    // kernel order ley lhx lhz m_hz_x m_hx_z m_ey_x m_ey_z
    if (q.lambda[0] != nullptr) {
      const int perm[7] = {0, 1, 2, 6, 5, 3, 4};
      for (int c = 0; c < 7; ++c) {
        it.lam[0][c] = q.lambda[perm[c]];
        PLAN_CHECK(plan_alloc(*p, &it.lam[1][c], (size_t)total));
      }
    }
    it.grad_r = q.grad_r;
    it.grad_f = q.grad_f;
    it.gca = q.grad_ca;
    it.gcb = q.grad_cb;
    it.n_shots = (int32_t)q.n_shots;
    it.n_src = (int32_t)q.n_src;
    it.n_rec = (int32_t)q.n_rec;
    it.nt = (int32_t)q.nt;
    it.interval = (int32_t)q.interval;
    if (q.n_src > 0 && q.f == nullptr) {
      set_error("plan: a problem with sources needs its source term f");
      return fail(kInvalidArgument);
    }
    if (q.n_rec > 0 && q.r == nullptr) {
      set_error("plan: a problem with receivers needs its receiver buffer r");
      return fail(kInvalidArgument);
    }
  }
  p->fmt_fwd = any_store ? fmt : 0;
  p->fmt_adj = any_grad ? fmt : 0;
  std::vector<int32_t> tiles_f, tiles_a;
  PLAN_CHECK(build_items(*p, p->rows_f, p->h_fwd, tiles_f, p->active_fwd, proto));
  PLAN_CHECK(build_items(*p, p->rows_a, p->h_adj, tiles_a, p->active_adj, proto));
  std::vector<int32_t> tiles2, tiles2a;
  if (two) PLAN_CHECK(build_items2(*p, tiles2, tiles2a));
  // problems with several shot groups accumulate into per-group FP64 slices
  for (size_t m = 0; m < p->probs.size(); ++m) {
    BatchItem<float> &it = p->h_adj[m];
    it.group_stride = 0;
    if (it.n_groups > 1 && (it.gca || it.gcb)) {
      const int64_t numel = p->probs[m].ny * p->probs[m].nx;
      it.group_stride = numel;
      if (it.gca) PLAN_CHECK(plan_alloc(*p, &it.gca, (size_t)(it.n_groups * numel)));
      if (it.gcb) PLAN_CHECK(plan_alloc(*p, &it.gcb, (size_t)(it.n_groups * numel)));
    }
  }
  PLAN_CHECK(plan_alloc(*p, &p->d_fwd, p->h_fwd.size()));
  PLAN_CHECK(plan_alloc(*p, &p->d_adj, p->h_adj.size()));
  PLAN_CHECK(plan_alloc(*p, &p->d_tiles_fwd, tiles_f.size()));
  PLAN_CHECK(plan_alloc(*p, &p->d_tiles_adj, tiles_a.size()));
  if (two) {
    PLAN_CHECK(plan_alloc(*p, &p->d_tiles2, tiles2.size()));
    TIDE_HIP_TRY(hipMemcpy(p->d_tiles2, tiles2.data(), tiles2.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));
    PLAN_CHECK(plan_alloc(*p, &p->d_tiles2a, tiles2a.size()));
    TIDE_HIP_TRY(hipMemcpy(p->d_tiles2a, tiles2a.data(), tiles2a.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));
  }
  TIDE_HIP_TRY(hipMemcpy(p->d_fwd, p->h_fwd.data(), p->h_fwd.size() * sizeof(BatchItem<float>),
                         hipMemcpyHostToDevice));
  TIDE_HIP_TRY(hipMemcpy(p->d_adj, p->h_adj.data(), p->h_adj.size() * sizeof(BatchItem<float>),
                         hipMemcpyHostToDevice));
  TIDE_HIP_TRY(hipMemcpy(p->d_tiles_fwd, tiles_f.data(), tiles_f.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  TIDE_HIP_TRY(hipMemcpy(p->d_tiles_adj, tiles_a.data(), tiles_a.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  TIDE_HIP_TRY(hipDeviceSynchronize());
  if (p->use_graph) {
    if (hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking) != hipSuccess) {
      set_error("plan: stream creation failed");
      return fail(kHipError);
    }
    PLAN_CHECK(capture(*p, true, &p->gr_fwd, &p->g_fwd));
    const bool any_adj = p->probs[0].lambda[0] != nullptr;
    if (any_adj) PLAN_CHECK(capture(*p, false, &p->gr_adj, &p->g_adj));
  }
  return kOk;
}
}  // namespace

extern "C" int tide_tm2d_plan_forward(void *plan, void *stream) {
  Plan *p = (Plan *)plan;
  if (p == nullptr) {
    set_error("plan: null handle");
    return kInvalidArgument;
  }
  TIDE_HIP_TRY(hipSetDevice(p->device));
  if (p->use_graph) {
    TIDE_HIP_TRY(hipGraphLaunch(p->g_fwd, (hipStream_t)stream));
    return kOk;
  }
  return enqueue_forward(*p, (hipStream_t)stream);
}

extern "C" int tide_tm2d_plan_backward(void *plan, void *stream) {
  Plan *p = (Plan *)plan;
  if (p == nullptr) {
    set_error("plan: null handle");
    return kInvalidArgument;
  }
  for (const tide_tm2d_problem &q : p->probs)
    if (q.lambda[0] == nullptr) {
      set_error("plan: built without adjoint state; pass lambda arrays to run the adjoint");
      return kInvalidArgument;
    }
  TIDE_HIP_TRY(hipSetDevice(p->device));
  if (p->use_graph) {
    TIDE_HIP_TRY(hipGraphLaunch(p->g_adj, (hipStream_t)stream));
    return kOk;
  }
  return enqueue_backward(*p, (hipStream_t)stream);
}

extern "C" int tide_tm2d_plan_destroy(void *plan) {
  destroy((Plan *)plan);
  return kOk;
}
