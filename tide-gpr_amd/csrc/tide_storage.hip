// Host-staged snapshot storage (see tide_storage.h).
//
// The reference does the same staging with a device ring of three steps,
// pinned host buffers and an asynchronous file per snapshot stream
// (src/tide/csrc/storage/storage_utils.cu:26-239); it aborts the process on
// a failed copy or file operation, this library returns a status instead.

#include "tide_storage.h"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>

namespace tide_hip {

// One file and one thread doing its pread / pwrite in submission order.  A
// task first waits for its event (the copy that fills its host slot, or the
// last copy that drained it), then moves the bytes.
class DiskWorker {
 public:
  struct Task {
    int slot;
    char *host;
    size_t bytes;
    int64_t offset;
    hipEvent_t after;  // owned; null: no wait
    bool write;
  };

  int open(const char *path, bool write, bool truncate) {
    if (path == nullptr || path[0] == '\0') {
      set_error("disk snapshot storage needs a file name");
      return kInvalidArgument;
    }
    const int flags = write ? (O_CREAT | O_WRONLY | (truncate ? O_TRUNC : 0)) : O_RDONLY;
    fd_ = ::open(path, flags, 0600);
    if (fd_ < 0) {
      set_error(std::string("open(") + path + ") failed: " + std::strerror(errno));
      return kIoError;
    }
    thread_ = std::thread([this] { run(); });
    return kOk;
  }

  ~DiskWorker() { (void)close(); }

  // Blocks until the last task on `slot` is done.
  int wait_slot(int slot) {
    std::unique_lock<std::mutex> lock(mu_);
    cv_.wait(lock, [&] { return !busy_[slot] || !error_.empty(); });
    return check_locked();
  }

  int enqueue(const Task &t) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!error_.empty()) {
      if (t.after) (void)hipEventDestroy(t.after);
      return check_locked();
    }
    busy_[t.slot] = true;
    tasks_.push_back(t);
    cv_.notify_all();
    return kOk;
  }

  // Drains the queue, joins the thread and closes the file.
  int close() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      closing_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    if (fd_ >= 0) {
      ::close(fd_);
      fd_ = -1;
    }
    std::lock_guard<std::mutex> lock(mu_);
    return check_locked();
  }

 private:
  int check_locked() {
    if (error_.empty()) return kOk;
    set_error("disk snapshot storage: " + error_);
    return kIoError;
  }

  static bool move_bytes(int fd, char *p, size_t n, int64_t off, bool write) {
    size_t done = 0;
    while (done < n) {
      const ssize_t k = write ? ::pwrite(fd, p + done, n - done, (off_t)(off + (int64_t)done))
                              : ::pread(fd, p + done, n - done, (off_t)(off + (int64_t)done));
      if (k <= 0) {
        if (k < 0 && errno == EINTR) continue;
        if (k == 0) errno = write ? ENOSPC : ENODATA;
        return false;
      }
      done += (size_t)k;
    }
    return true;
  }

  void run() {
    for (;;) {
      Task t;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return closing_ || !tasks_.empty(); });
        if (tasks_.empty()) return;  // closing and drained
        t = tasks_.front();
        tasks_.pop_front();
      }
      std::string err;
      if (t.after != nullptr) {
        const hipError_t e = hipEventSynchronize(t.after);
        (void)hipEventDestroy(t.after);
        if (e != hipSuccess) err = std::string("copy before file I/O failed: ") + hipGetErrorString(e);
      }
      if (err.empty() && !move_bytes(fd_, t.host, t.bytes, t.offset, t.write))
        err = std::string(t.write ? "pwrite" : "pread") + " failed at offset " +
              std::to_string(t.offset) + ": " + std::strerror(errno);
      {
        std::lock_guard<std::mutex> lock(mu_);
        busy_[t.slot] = false;
        if (!err.empty() && error_.empty()) error_ = err;
      }
      cv_.notify_all();
    }
  }

  int fd_ = -1;
  std::thread thread_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Task> tasks_;
  bool busy_[kRingSlots] = {};
  bool closing_ = false;
  std::string error_;
};

SnapshotStage::SnapshotStage() = default;

SnapshotStage::~SnapshotStage() {
  if (disk_) (void)disk_->close();
  for (int j = 0; j < kRingSlots; ++j) {
    if (copied_[j]) (void)hipEventDestroy(copied_[j]);
    if (used_[j]) (void)hipEventDestroy(used_[j]);
  }
}

int SnapshotStage::open(int64_t mode, bool saving, size_t step_bytes, void *ring, void *host,
                        const char *path, int64_t first_store, hipStream_t compute,
                        hipStream_t storage) {
  if (mode != kStageCpu && mode != kStageDisk) {
    set_error("host-staged storage mode must be 1 (cpu) or 2 (disk)");
    return kInvalidArgument;
  }
  if (ring == nullptr || host == nullptr || step_bytes == 0) {
    set_error("host-staged storage needs the device ring (store_1) and the pinned host "
              "buffer (store_3)");
    return kInvalidArgument;
  }
  mode_ = mode;
  saving_ = saving;
  step_bytes_ = step_bytes;
  ring_ = (char *)ring;
  host_ = (char *)host;
  compute_ = compute;
  storage_ = storage;
  split_ = storage != nullptr && storage != compute;
  for (int j = 0; j < kRingSlots; ++j) {
    TIDE_HIP_TRY(hipEventCreateWithFlags(&copied_[j], hipEventDisableTiming));
    TIDE_HIP_TRY(hipEventCreateWithFlags(&used_[j], hipEventDisableTiming));
  }
  if (mode == kStageDisk) {
    disk_.reset(new DiskWorker());
    const int rc = disk_->open(path, saving, saving && first_store == 0);
    if (rc != kOk) {
      disk_.reset();
      mode_ = 0;
      return rc;
    }
  }
  return kOk;
}

int SnapshotStage::record_use(int64_t idx) {
  const int j = (int)(idx % kRingSlots);
  TIDE_HIP_TRY(hipEventRecord(used_[j], compute_));
  has_used_[j] = true;
  return kOk;
}

int SnapshotStage::before_write(int64_t idx) {
  const int j = (int)(idx % kRingSlots);
  if (split_ && has_copied_[j]) TIDE_HIP_TRY(hipStreamWaitEvent(compute_, copied_[j], 0));
  return kOk;
}

int SnapshotStage::after_write(int64_t idx) {
  const int j = (int)(idx % kRingSlots);
  hipStream_t s = compute_;
  if (split_) {
    const int rc = record_use(idx);
    if (rc != kOk) return rc;
    TIDE_HIP_TRY(hipStreamWaitEvent(storage_, used_[j], 0));
    s = storage_;
  }
  if (mode_ == kStageCpu) {
    TIDE_HIP_TRY(hipMemcpyAsync(host_ + (size_t)idx * step_bytes_, slot(idx), step_bytes_,
                                hipMemcpyDeviceToHost, s));
  } else {
    char *h = host_ + (size_t)j * step_bytes_;
    int rc = disk_->wait_slot(j);  // the previous write from this host slot
    if (rc != kOk) return rc;
    TIDE_HIP_TRY(hipMemcpyAsync(h, slot(idx), step_bytes_, hipMemcpyDeviceToHost, s));
    hipEvent_t ready = nullptr;
    TIDE_HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    const hipError_t e = hipEventRecord(ready, s);
    if (e != hipSuccess) {
      (void)hipEventDestroy(ready);
      set_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
      return kHipError;
    }
    rc = disk_->enqueue({j, h, step_bytes_, idx * (int64_t)step_bytes_, ready, true});
    if (rc != kOk) return rc;
  }
  TIDE_HIP_TRY(hipEventRecord(copied_[j], s));
  has_copied_[j] = true;
  return kOk;
}

int SnapshotStage::prefetch(int64_t idx) {
  if (idx < 0) return kOk;
  const int j = (int)(idx % kRingSlots);
  if (loaded_[j] == idx) return kOk;
  hipStream_t s = compute_;
  if (split_) {
    if (has_used_[j]) TIDE_HIP_TRY(hipStreamWaitEvent(storage_, used_[j], 0));
    s = storage_;
  }
  const char *src = host_ + (size_t)idx * step_bytes_;
  if (mode_ == kStageDisk) {
    // the read into host slot j follows the last copy out of it (every
    // copy so far on s)
    char *h = host_ + (size_t)j * step_bytes_;
    hipEvent_t drained = nullptr;
    TIDE_HIP_TRY(hipEventCreateWithFlags(&drained, hipEventDisableTiming));
    const hipError_t e = hipEventRecord(drained, s);
    if (e != hipSuccess) {
      (void)hipEventDestroy(drained);
      set_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
      return kHipError;
    }
    int rc = disk_->enqueue({j, h, step_bytes_, idx * (int64_t)step_bytes_, drained, false});
    if (rc == kOk) rc = disk_->wait_slot(j);
    if (rc != kOk) return rc;
    src = h;
  }
  TIDE_HIP_TRY(hipMemcpyAsync(slot(idx), src, step_bytes_, hipMemcpyHostToDevice, s));
  TIDE_HIP_TRY(hipEventRecord(copied_[j], s));
  has_copied_[j] = true;
  loaded_[j] = idx;
  return kOk;
}

int SnapshotStage::before_read(int64_t idx) {
  const int rc = prefetch(idx);
  if (rc != kOk) return rc;
  const int j = (int)(idx % kRingSlots);
  if (split_) TIDE_HIP_TRY(hipStreamWaitEvent(compute_, copied_[j], 0));
  return kOk;
}

int SnapshotStage::after_read(int64_t idx) { return split_ ? record_use(idx) : kOk; }

int SnapshotStage::finish() {
  int rc = kOk;
  if (disk_) {
    rc = disk_->close();
    disk_.reset();
  }
  mode_ = 0;
  return rc;
}

}  // namespace tide_hip
