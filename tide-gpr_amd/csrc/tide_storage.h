// Host-staged snapshot storage: the storage modes that keep the forward
// snapshots outside HBM (reference storage_mode 1 "cpu" and 2 "disk";
// src/tide/csrc/storage/storage_utils.{c,cu},
// src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:286-315, 1407-1571, 2458-2623).
//
// The kernels read and write snapshots in a device ring of kRingSlots
// stored steps (the caller's store_1 buffer).  A stored step's slot is
// (store index % kRingSlots); the host side is
//   cpu:  store_3 = [n_store][step_bytes] pinned host history, step i at i;
//   disk: store_3 = [kRingSlots][step_bytes] pinned host ring, and a file
//         (filenames[0]) that holds step i at byte offset i * step_bytes.
// Copies run on the storage stream (the compute stream when none is given
// or both are the same).  Events order every reuse of a slot: the compute
// stream waits for the copy that last used a ring slot before a kernel
// writes (forward) or reads (adjoint) it, and the copy into or out of a slot
// waits for the kernels that last used it.  A worker thread per file does
// the pread / pwrite after the copy that feeds it (or the copy that last
// drained its host slot) has completed.  Nothing aborts: a failed copy or
// file operation is reported by the call's status.
#pragma once

#include "tide_common.h"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace tide_hip {

constexpr int kRingSlots = 3;  // reference NUM_BUFFERS / _CPU_STORAGE_BUFFERS
constexpr int64_t kStageCpu = 1;
constexpr int64_t kStageDisk = 2;

class DiskWorker;

// One snapshot stream (Ey or curl H) of one call.
class SnapshotStage {
 public:
  SnapshotStage();
  ~SnapshotStage();
  SnapshotStage(const SnapshotStage &) = delete;
  SnapshotStage &operator=(const SnapshotStage &) = delete;

  // mode kStageCpu / kStageDisk; saving = forward (device -> host).  `ring`
  // is the device ring, `host` the pinned host buffer (layout above),
  // `path` the disk file (disk only).  `first_store` is the first store
  // index this call touches (a forward starting at t = 0 truncates the file).
  int open(int64_t mode, bool saving, size_t step_bytes, void *ring, void *host,
           const char *path, int64_t first_store, hipStream_t compute, hipStream_t storage);

  bool active() const { return mode_ != 0; }

  // Device address of the ring slot of store index `idx`.
  void *slot(int64_t idx) const {
    return ring_ + (size_t)(idx % kRingSlots) * step_bytes_;
  }

  // Forward: before a kernel writes store `idx`, make the compute stream
  // wait for the copy that last drained its slot.
  int before_write(int64_t idx);
  // Forward: after the kernel that wrote `idx` is enqueued, copy the slot
  // out (host history, or host ring + file).
  int after_write(int64_t idx);

  // Adjoint: enqueue the copy of store `idx` into its slot (once; a later
  // call for the same index is a no-op), for a kernel launched later.
  int prefetch(int64_t idx);
  // Adjoint: before a kernel reads store `idx` (prefetches it if needed),
  // make the compute stream wait for its copy.
  int before_read(int64_t idx);
  // Adjoint: after the kernel that read `idx` is enqueued, mark the slot's
  // use so the next copy into it waits for that kernel.
  int after_read(int64_t idx);

  // Waits for the file worker (disk) and reports a failed copy or file
  // operation; the stage is closed afterwards.  Copies already enqueued on
  // the streams stay ordered after the call's kernels.
  int finish();

 private:
  int record_use(int64_t idx);

  int64_t mode_ = 0;
  bool saving_ = true;
  size_t step_bytes_ = 0;
  char *ring_ = nullptr;
  char *host_ = nullptr;
  hipStream_t compute_ = nullptr, storage_ = nullptr;
  bool split_ = false;  // separate storage stream
  hipEvent_t copied_[kRingSlots] = {};  // last copy into / out of the slot
  hipEvent_t used_[kRingSlots] = {};    // last kernel that used the slot
  bool has_copied_[kRingSlots] = {};
  bool has_used_[kRingSlots] = {};
  int64_t loaded_[kRingSlots] = {-1, -1, -1};  // adjoint: store index now in the slot
  std::unique_ptr<DiskWorker> disk_;
};

}  // namespace tide_hip
