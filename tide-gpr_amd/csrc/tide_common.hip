// Error reporting and point-table construction shared by the propagators.
#include "tide_common.h"

#include <algorithm>
#include <cstdio>
#include <tuple>

namespace tide_hip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

namespace {
bool build_table(const int64_t *src, int64_t n_src, const int64_t *rec, int64_t n_rec,
                 int64_t n_shots, int64_t ny, int64_t nx, int tile_w, int lane_offset,
                 bool halo, PointTable &out) {
  const int64_t numel = ny * nx;
  const int64_t n_xt = (nx + tile_w - 1) / tile_w;
  // (y, shot, xt, kind, k, xloc, q); kind bit 1 marks a halo copy
  struct Entry {
    int64_t y, shot, xt, kind, k;
    int32_t xloc, q;
  };
  std::vector<Entry> entries;
  entries.reserve((size_t)(n_shots * (n_src + n_rec)));
  auto add = [&](const int64_t *idx, int64_t n, int kind) -> bool {
    for (int64_t s = 0; s < n_shots; ++s) {
      for (int64_t k = 0; k < n; ++k) {
        const int64_t v = idx[s * n + k];
        if (v < 0) continue;
        if (v >= numel) {
          char buf[160];
          std::snprintf(buf, sizeof(buf),
                        "%s index %lld of shot %lld is outside the padded "
                        "%lldx%lld grid",
                        kind == 0 ? "source" : "receiver", (long long)v,
                        (long long)s, (long long)ny, (long long)nx);
          set_error(buf);
          return false;
        }
        const int64_t y = v / nx, x = v % nx;
        const int64_t xt = x / tile_w;
        const int32_t lane = (int32_t)(x % tile_w + lane_offset);
        entries.push_back({y, s, xt, kind, k, lane, (int32_t)(s * n + k)});
        if (halo) {
          if (lane + tile_w <= 63 && xt > 0)
            entries.push_back({y, s, xt - 1, kind | 2, k, lane + tile_w, (int32_t)(s * n + k)});
          if (lane - tile_w >= 0 && xt + 1 < n_xt)
            entries.push_back({y, s, xt + 1, kind | 2, k, lane - tile_w, (int32_t)(s * n + k)});
        }
      }
    }
    return true;
  };
  if (!add(src, n_src, 0) || !add(rec, n_rec, 1)) return false;
  std::sort(entries.begin(), entries.end(), [](const Entry &a, const Entry &b) {
    return std::tie(a.y, a.shot, a.xt, a.kind, a.k) <
           std::tie(b.y, b.shot, b.xt, b.kind, b.k);
  });
  out.n_xt = (int32_t)n_xt;
  out.row_slot.assign((size_t)ny, -1);
  std::vector<int64_t> rows;
  for (const Entry &e : entries)
    if (rows.empty() || rows.back() != e.y) rows.push_back(e.y);
  out.n_slots = (int32_t)rows.size();
  for (size_t i = 0; i < rows.size(); ++i) out.row_slot[(size_t)rows[i]] = (int32_t)i;
  const int64_t n_seg = (int64_t)out.n_slots * n_shots * out.n_xt;
  out.seg.assign((size_t)n_seg + 1, 0);
  std::vector<int32_t> counts((size_t)n_seg, 0);
  for (const Entry &e : entries) {
    const int64_t slot = out.row_slot[(size_t)e.y];
    counts[(size_t)((slot * n_shots + e.shot) * out.n_xt + e.xt)]++;
  }
  for (int64_t i = 0; i < n_seg; ++i) out.seg[(size_t)i + 1] = out.seg[(size_t)i] + counts[(size_t)i];
  out.ent.resize(entries.size());
  // entries are already in (row, shot, tile, kind, k) order == segment order.
  for (size_t i = 0; i < entries.size(); ++i)
    out.ent[i] = make_int2(entries[i].xloc | ((int32_t)(entries[i].kind & 1) << 8) |
                               ((int32_t)(entries[i].kind >> 1) << 9),
                           entries[i].q);
  return true;
}
}  // namespace

bool build_point_table(const int64_t *src, int64_t n_src, const int64_t *rec,
                       int64_t n_rec, int64_t n_shots, int64_t ny, int64_t nx,
                       int tile_w, int lane_offset, PointTable &out) {
  return build_table(src, n_src, rec, n_rec, n_shots, ny, nx, tile_w, lane_offset, false, out);
}

bool build_point_table_halo(const int64_t *src, int64_t n_src, const int64_t *rec,
                            int64_t n_rec, int64_t n_shots, int64_t ny, int64_t nx,
                            int tile_w, int lane_offset, PointTable &out) {
  return build_table(src, n_src, rec, n_rec, n_shots, ny, nx, tile_w, lane_offset, true, out);
}

int DevicePointTable::upload(const PointTable &t, hipStream_t stream) {
  n_xt = t.n_xt;
  const size_t nr = t.row_slot.size() * sizeof(int32_t);
  const size_t ns = t.seg.size() * sizeof(int32_t);
  const size_t ne = std::max<size_t>(1, t.ent.size()) * sizeof(int2);
  TIDE_HIP_TRY(hipMallocAsync((void **)&row_slot, nr, stream));
  TIDE_HIP_TRY(hipMallocAsync((void **)&seg, ns, stream));
  TIDE_HIP_TRY(hipMallocAsync((void **)&ent, ne, stream));
  TIDE_HIP_TRY(hipMemcpyAsync(row_slot, t.row_slot.data(), nr, hipMemcpyHostToDevice, stream));
  TIDE_HIP_TRY(hipMemcpyAsync(seg, t.seg.data(), ns, hipMemcpyHostToDevice, stream));
  if (!t.ent.empty())
    TIDE_HIP_TRY(hipMemcpyAsync(ent, t.ent.data(), t.ent.size() * sizeof(int2),
                                hipMemcpyHostToDevice, stream));
  // The host vectors are pageable and may be freed when the caller returns.
  TIDE_HIP_TRY(hipStreamSynchronize(stream));
  return kOk;
}

int DevicePointTable::place(const PointTable &t, int32_t *row_slot_mem, int32_t *seg_mem,
                            int2 *ent_mem, hipStream_t stream) {
  n_xt = t.n_xt;
  owned = false;
  row_slot = row_slot_mem;
  seg = seg_mem;
  ent = ent_mem;
  TIDE_HIP_TRY(hipMemcpyAsync(row_slot, t.row_slot.data(), t.row_slot.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, stream));
  TIDE_HIP_TRY(hipMemcpyAsync(seg, t.seg.data(), t.seg.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, stream));
  if (!t.ent.empty())
    TIDE_HIP_TRY(hipMemcpyAsync(ent, t.ent.data(), t.ent.size() * sizeof(int2),
                                hipMemcpyHostToDevice, stream));
  // The host vectors are pageable and may be freed when the caller returns.
  TIDE_HIP_TRY(hipStreamSynchronize(stream));
  return kOk;
}

void DevicePointTable::release(hipStream_t stream) {
  if (owned) {
    if (row_slot) (void)hipFreeAsync(row_slot, stream);
    if (seg) (void)hipFreeAsync(seg, stream);
    if (ent) (void)hipFreeAsync(ent, stream);
  }
  row_slot = nullptr;
  seg = nullptr;
  ent = nullptr;
}

}  // namespace tide_hip

extern "C" const char *tide_last_error(void) { return tide_hip::last_error(); }
