// Shared host/device helpers for the MI355X (gfx950) TIDE propagators.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace tide_hip {

// ------------------------------------------------------------------ errors
// Every C-ABI entry point returns 0 on success or a nonzero code; the text
// of the last failure on the calling thread is kept for tide_last_error().
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kUnsupported = 2,
  kHipError = 3,
  kOutOfMemory = 4,
};

void set_error(const std::string &msg);
const char *last_error();

#define TIDE_HIP_TRY(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tide_hip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return _e == hipErrorOutOfMemory ? ::tide_hip::kOutOfMemory            \
                                       : ::tide_hip::kHipError;              \
    }                                                                        \
  } while (0)

// --------------------------------------------------------------- BF16 codec
// Round-to-nearest-even on the upper 16 bits, identical on host and device.
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float v) {
  uint32_t u = __builtin_bit_cast(uint32_t, v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, ((uint32_t)h) << 16);
}

// -------------------------------------------------------------- point table
// Sources and receivers are applied inside the step kernels by the thread
// that owns the cell.  The table groups the (shot, point) entries by padded
// row, shot and column tile (tile_w columns, tile xt starting at column
// xt*tile_w) so that a wavefront walking row y reads only the handful of
// entries that fall in its columns:
//   row_slot[y]                      -1 or the index of row y among point rows
//   seg[(slot*n_shots+shot)*n_xt+xt] first entry of that (row, shot, tile);
//                                     seg[last+1] is the end
//   ent[e] = {lane | kind<<8, q}      lane = column - xt*tile_w + lane_offset;
//                                     kind 0 = source (q indexes f[t]),
//                                     kind 1 = receiver (q indexes r[t]);
//                                     sources precede receivers in a segment,
//                                     each kind in ascending point order.
struct PointTable {
  std::vector<int32_t> row_slot;
  std::vector<int32_t> seg;
  std::vector<int2> ent;
  int32_t n_slots = 0;
  int32_t n_xt = 0;
};

// Builds the table from host copies of sources_i / receivers_i ([n_shots, n]
// flat indices into one shot's [ny, nx] plane; negative entries are
// skipped).  Returns false (with set_error) on an index outside the plane.
bool build_point_table(const int64_t *src, int64_t n_src, const int64_t *rec,
                       int64_t n_rec, int64_t n_shots, int64_t ny, int64_t nx,
                       int tile_w, int lane_offset, PointTable &out);

// As above, and additionally every point is listed in the neighbouring
// tiles whose 64-lane window (columns xt*tile_w - lane_offset ...
// + 63) holds its column, as a copy: ent.x = lane | kind<<8 | 1<<9.  The
// multi-step kernels recompute halo cells and apply sources there too.
bool build_point_table_halo(const int64_t *src, int64_t n_src, const int64_t *rec,
                            int64_t n_rec, int64_t n_shots, int64_t ny, int64_t nx,
                            int tile_w, int lane_offset, PointTable &out);

// Device copy of a PointTable; owns three device allocations.
struct DevicePointTable {
  int32_t *row_slot = nullptr;
  int32_t *seg = nullptr;
  int2 *ent = nullptr;
  int32_t n_xt = 0;
  int upload(const PointTable &t, hipStream_t stream);
  void release(hipStream_t stream);
};

// ----------------------------------------------------------- lane moves
// Cross-lane moves by one lane through full-wave DPP shifts (a VALU move, no
// LDS round trip): shfl_dn1 gives lane i the value of lane i+1, shfl_up1 that
// of lane i-1; the value at the wave edge (lane 63 resp. 0) is unspecified
// and never used (those lanes are halo lanes).
__device__ __forceinline__ int dpp_dn1_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}
__device__ __forceinline__ int dpp_up1_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}
__device__ __forceinline__ float shfl_dn1(float v) {
  return __builtin_bit_cast(float, dpp_dn1_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float shfl_up1(float v) {
  return __builtin_bit_cast(float, dpp_up1_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double shfl_dn1(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = dpp_dn1_i((int)u), hi = dpp_dn1_i((int)(u >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shfl_up1(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = dpp_up1_i((int)u), hi = dpp_up1_i((int)(u >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// ----------------------------------------------------------- tile mapping
// Maps a flat workgroup id onto (tile, group) so that the 8 XCDs of the
// MI355X each receive a contiguous band of tiles.  Workgroups are dispatched
// to XCDs round-robin, so id b runs on XCD b % 8; the band of XCD k starts at
// k*per + min(k, rem).  The map is a bijection on [0, n).
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n) {
  const int64_t per = n >> 3, rem = n & 7;
  const int64_t xcd = b & 7, idx = b >> 3;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

}  // namespace tide_hip
