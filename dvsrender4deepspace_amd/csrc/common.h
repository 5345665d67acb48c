// Host-side helpers shared by the C ABI translation units: error reporting in
// the reference's exception wording, owning device arrays and HIP-event timers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dvs_hip.h"

namespace dvsamd {

// Message of the calling thread's last failure (dvs_last_error).
std::string &last_error();

struct Failure {
  int code;
  std::string message;
};
[[noreturn]] inline void fail(int code, std::string message) { throw Failure{code, std::move(message)}; }
inline void hip_check(hipError_t e, const char *where) {
  // Message format of CUDA_CHECK (include/dvs/math_types.hpp:6-14).
  if (e != hipSuccess) fail(DVS_E_HIP, std::string("[HIP Error] in ") + where + " : " + hipGetErrorString(e));
}
// Runs f, mapping exceptions to the C ABI status codes.
template <class F> int guarded(F &&f) {
  try {
    f();
    last_error().clear();
    return DVS_OK;
  } catch (const Failure &x) {
    last_error() = x.message;
    return x.code;
  } catch (const std::invalid_argument &x) {
    last_error() = x.what();
    return DVS_E_ARGUMENT;
  } catch (const std::bad_alloc &) {
    last_error() = "host allocation failed";
    return DVS_E_STATE;
  } catch (const std::exception &x) {
    last_error() = x.what();
    return DVS_E_STATE;
  }
}

// Waits for `st` by polling it.  A blocking synchronize can put the host
// thread to sleep and wake it tens of microseconds after the stream drains;
// between the small read-backs and launches of one step that wake-up is
// most of the host's share of the step.
inline void stream_wait(hipStream_t st, const char *what) {
  for (;;) {
    const hipError_t e = hipStreamQuery(st);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) hip_check(e, what);
  }
}

// Staged copies of one scatter launch: copy k moves bytes[k] from byte
// offset src[k] of a staging mirror to dst[k].
constexpr int kScatterCopies = 16;
struct ScatterArgs {
  unsigned long long src[kScatterCopies];
  char *dst[kScatterCopies];
  unsigned bytes[kScatterCopies];
  int count;
};
// Queues the copies of `a` from device memory `mirror` on st (one block per
// copy).
hipError_t launch_staged_scatter(const char *mirror, const ScatterArgs &a, hipStream_t st);

// Waits for a queued device-to-host copy into page-locked memory by reading
// its destination: the caller set the n words to `unset` (a value the copy
// never writes) before queuing the copy, and the wait ends when none of them
// holds it.  This sees the data several microseconds before the stream
// reports the copy complete.  Copies on one stream run in order, so the
// words of the last copy queued also tell that the earlier ones landed.
// The stream is polled every 1024 reads: an error ends the wait, and so
// does a drained stream whose copy left a word unset.
inline void wait_words(const volatile int *host, size_t n, int unset, hipStream_t st, const char *what) {
  for (unsigned reads = 1;; ++reads) {
    bool all = true;
    for (size_t i = 0; i < n; ++i) all = all && host[i] != unset;
    if (all) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return;
    }
    if (reads % 1024 == 0) {
      const hipError_t e = hipStreamQuery(st);
      if (e == hipSuccess) {
        for (size_t i = 0; i < n; ++i)
          if (host[i] == unset) throw std::runtime_error(std::string(what) + ": read-back left a word unwritten");
        std::atomic_thread_fence(std::memory_order_acquire);
        return;
      }
      if (e != hipErrorNotReady) hip_check(e, what);
    }
  }
}

// Page-locked staging for the small host-to-device copies of one engine
// step.  A copy from pageable memory blocks the host until the stream has
// drained up to it; from page-locked memory it is queued and returns, but
// each queued copy still costs the host ten to twenty microseconds.  So
// upload() only records a copy: flush() moves every recorded region with one
// host-to-device copy into a device mirror of the staging block and one
// scatter kernel per kScatterCopies regions.  A kernel that reads an
// uploaded array must be queued after the flush.
//
// The staging is a bump allocator: regions stay untouched until reset(),
// which the engine calls when a step starts (the previous step's copies have
// completed by then, its results having been read back).
class Staging {
 public:
  Staging() = default;
  Staging(const Staging &) = delete;
  Staging &operator=(const Staging &) = delete;
  ~Staging() { free_blocks(); }
  void reset() {
    if (blocks_.size() > 1) {  // one block of the combined size from now on
      size_t total = 0;
      for (auto &b : blocks_) total += b.bytes;
      free_blocks();
      grow(total);
    }
    used_ = 0;
    pending_.clear();
  }
  // Copies n elements to staging and records their copy to dst.
  template <class T> void upload(T *dst, const T *src, size_t n, const char *what, hipStream_t st) {
    if (!n) return;
    const size_t bytes = n * sizeof(T), at = (used_ + 255) & ~size_t(255);
    if (bytes > 0xFFFFFFFFu) throw std::runtime_error(std::string(what) + ": staged copy too large");
    if (blocks_.empty() || at + bytes > blocks_.back().bytes) {
      flush(st);  // the recorded regions live in the current block
      grow(std::max(bytes, blocks_.empty() ? size_t(1) << 20 : 2 * blocks_.back().bytes));
      return upload(dst, src, n, what, st);
    }
    std::memcpy(blocks_.back().host + at, src, bytes);
    used_ = at + bytes;
    pending_.push_back(Region{at, reinterpret_cast<char *>(dst), bytes});
  }
  // Queues every recorded copy on st.
  void flush(hipStream_t st) {
    if (pending_.empty()) return;
    const Block &b = blocks_.back();
    const size_t lo = pending_.front().at, hi = pending_.back().at + pending_.back().bytes;
    hip_check(hipMemcpyAsync(b.device + lo, b.host + lo, hi - lo, hipMemcpyHostToDevice, st), "staged upload");
    for (size_t k = 0; k < pending_.size(); k += kScatterCopies) {
      ScatterArgs a{};
      a.count = (int)std::min<size_t>(kScatterCopies, pending_.size() - k);
      for (int j = 0; j < a.count; ++j) {
        a.src[j] = pending_[k + j].at;
        a.dst[j] = pending_[k + j].dst;
        a.bytes[j] = (unsigned)pending_[k + j].bytes;
      }
      hip_check(launch_staged_scatter(b.device, a, st), "staged upload");
    }
    pending_.clear();
  }

 private:
  struct Block {
    char *host, *device;
    size_t bytes;
  };
  struct Region {
    size_t at;
    char *dst;
    size_t bytes;
  };
  void grow(size_t bytes) {
    Block b{nullptr, nullptr, bytes};
    hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.host), bytes, hipHostMallocDefault), "page-locked staging");
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&b.device), bytes);
    if (e != hipSuccess) {
      (void)hipHostFree(b.host);
      hip_check(e, "staging mirror");
    }
    blocks_.push_back(b);
    used_ = 0;
  }
  void free_blocks() {
    for (auto &b : blocks_) {
      (void)hipHostFree(b.host);
      (void)hipFree(b.device);
    }
    blocks_.clear();
  }
  std::vector<Block> blocks_;
  std::vector<Region> pending_;
  size_t used_ = 0;
};

// Owning device allocation that grows on demand (contents are not kept).
template <class T> struct DeviceArray {
  T *ptr = nullptr;
  size_t count = 0;
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void reserve(size_t n, const char *what) {
    if (n <= count) return;
    release();
    hip_check(hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T)), what);
    count = n;
  }
  // reserve() with room to spare when it has to grow (at least half again
  // as much as asked and twice the old size), for sizes that drift from call
  // to call: a free and an allocation cost a device-wide synchronisation each.
  void reserve_ahead(size_t n, const char *what) {
    if (n > count) reserve(std::max(n + n / 2, 2 * count), what);
  }
  void upload(const T *host, size_t n, const char *what, hipStream_t st = nullptr) {
    reserve(n, what);
    if (n) hip_check(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, st), what);
  }
  void upload(const T *host, size_t n, const char *what, hipStream_t st, Staging &staging) {
    reserve(n, what);
    staging.upload(ptr, host, n, what, st);
  }
  // Grows to at least n elements (with the room of reserve_ahead), keeping
  // the first `used` (stream ordered).
  void grow_keep(size_t n, size_t used, const char *what, hipStream_t st) {
    if (n <= count) return;
    n = std::max(n + n / 2, 2 * count);
    T *bigger = nullptr;
    hip_check(hipMalloc(&bigger, n * sizeof(T)), what);
    if (used) hip_check(hipMemcpyAsync(bigger, ptr, used * sizeof(T), hipMemcpyDeviceToDevice, st), what);
    hip_check(hipStreamSynchronize(st), what);
    release();
    ptr = bigger;
    count = n;
  }
  size_t bytes() const { return count * sizeof(T); }
};

// Page-locked host buffer for read-backs, grown with room to spare.
template <class T> struct PinnedArray {
  T *ptr = nullptr;
  size_t count = 0;
  PinnedArray() = default;
  PinnedArray(const PinnedArray &) = delete;
  PinnedArray &operator=(const PinnedArray &) = delete;
  ~PinnedArray() {
    if (ptr) (void)hipHostFree(ptr);
  }
  void reserve(size_t n, const char *what) {
    if (n <= count) return;
    const size_t want = std::max<size_t>(std::max(n + n / 2, 2 * count), 1);
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    count = 0;
    hip_check(hipHostMalloc(reinterpret_cast<void **>(&ptr), want * sizeof(T), hipHostMallocDefault), what);
    count = want;
  }
};

// Pair of HIP events bracketing work on one stream.
struct Timer {
  hipEvent_t start = nullptr, stop = nullptr;
  Timer() {
    hip_check(hipEventCreate(&start), "timer event");
    hip_check(hipEventCreate(&stop), "timer event");
  }
  ~Timer() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
  Timer(const Timer &) = delete;
  Timer &operator=(const Timer &) = delete;
  double ms() const {
    float v = 0;
    // a wait that read its copy's words may return before the runtime has
    // marked the stream's events complete
    hip_check(hipEventSynchronize(stop), "timer elapsed");
    hip_check(hipEventElapsedTime(&v, start, stop), "timer elapsed");
    return v;
  }
};

}  // namespace dvsamd
