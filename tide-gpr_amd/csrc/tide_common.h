// Shared host/device helpers for the MI355X (gfx950) TIDE propagators.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
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
  kIoError = 5,
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

// Device copy of a PointTable: either three device allocations of its own
// (upload) or three caller-provided arrays of the sizes below (place).
struct DevicePointTable {
  int32_t *row_slot = nullptr;
  int32_t *seg = nullptr;
  int2 *ent = nullptr;
  int32_t n_xt = 0;
  bool owned = true;
  int upload(const PointTable &t, hipStream_t stream);
  int place(const PointTable &t, int32_t *row_slot_mem, int32_t *seg_mem, int2 *ent_mem,
            hipStream_t stream);
  static size_t row_slot_count(const PointTable &t) { return t.row_slot.size(); }
  static size_t seg_count(const PointTable &t) { return t.seg.size(); }
  static size_t ent_count(const PointTable &t) { return t.ent.empty() ? 1 : t.ent.size(); }
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

// -------------------------------------------------- addressing and loads
// Element i of an array through a base pointer that is the same in every
// lane of the wavefront (a shot's plane, say).  The byte offset is formed in
// 32 bits, so the access compiles to the scalar-base plus 32-bit
// vector-offset address form instead of 64-bit address arithmetic per access.
// Callers keep i * sizeof(T) below 2^32: a 2-D shot plane holds fewer than
// 2^29 cells (fp64 included); 3-D code indexes within a row whose start is
// a wave-uniform 64-bit pointer.
template <typename T>
__device__ __forceinline__ T *elem32(T *p, uint32_t i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T *>(reinterpret_cast<B *>(p) + i * (uint32_t)sizeof(T));
}

// Shot plane of a snapshot buffer (payload T, or BF16 for FMT 2) starting at
// element `base`; null stays null.
template <typename T, int FMT, typename V>
__device__ __forceinline__ V *snap_plane(V *p, int64_t base) {
  using B = typename std::conditional<std::is_const<V>::value, const char, char>::type;
  constexpr int64_t bytes = FMT == 2 ? 2 : (int64_t)sizeof(T);
  return p == nullptr ? nullptr : reinterpret_cast<V *>(reinterpret_cast<B *>(p) + base * bytes);
}

// put_store / get_store through a snapshot plane (snap_plane) and a 32-bit
// in-plane index.
template <typename T, int FMT>
__device__ __forceinline__ void put_store32(void *plane, uint32_t j, T v) {
  if constexpr (FMT == 2) {
    __builtin_nontemporal_store(f32_to_bf16((float)v), elem32((uint16_t *)plane, j));
  } else {
    __builtin_nontemporal_store(v, elem32((T *)plane, j));
  }
}

template <typename T, int FMT>
__device__ __forceinline__ double get_store32(const void *plane, uint32_t j) {
  if constexpr (FMT == 2) {
    return (double)bf16_to_f32(__builtin_nontemporal_load(elem32((const uint16_t *)plane, j)));
  } else {
    return (double)__builtin_nontemporal_load(elem32((const T *)plane, j));
  }
}

// v where `keep`, else +0, without a select: a select with a loaded operand
// may be turned into a branch around the load, and the wavefront then waits
// for that load on the spot instead of with the other loads in flight.  The
// empty asm hides where the mask comes from, so the AND stays an AND.
template <typename T>
__device__ __forceinline__ T keep_or_zero(T v, bool keep) {
  using U = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
  U m = keep ? ~U(0) : U(0);
  asm("" : "+v"(m));
  return __builtin_bit_cast(T, __builtin_bit_cast(U, v) & m);
}

// Value of v in lane `lane` (wave-uniform index), as a scalar.
template <typename T>
__device__ __forceinline__ T lane_read(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  }
}

}  // namespace tide_hip
