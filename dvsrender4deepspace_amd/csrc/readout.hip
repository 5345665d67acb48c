// The readout FIFO of each segment on the device (EventReadout::submit,
// src/event_readout.cpp:13-29): one work-group (one wave) per segment.
//
// The recurrence is serial by nature (an event's completion time depends on
// the previous one's, and whether it is queued or dropped depends on the
// queue length after the deliveries before it), so lane 0 runs it in the
// reference's order and arithmetic, which keeps every delivered timestamp,
// drop and the peak queue length bit-identical to the host loop.  The other
// lanes stage its inputs in LDS: chunks of the segment's sorted events, and
// chunks of the queue's ready times from the device ring (the entries
// queued in this step are appended to the staged chunk by lane 0 itself as
// long as they extend it).  When lane 0 runs out of either it stops, the
// wave refills, and lane 0 resumes.  Finally the wave copies the delivered
// range of the ring to the output.
#include <hip/hip_runtime.h>

#include "readout.h"

namespace dvsamd {
namespace {

constexpr int kFifoThreads = 64;
constexpr int kEventChunk = 512;
constexpr int kReadyChunk = 1024;

enum : int { kNeedNothing = 0, kNeedEvents = 1, kNeedReady = 2 };

struct FifoShared {
  double ev_t[kEventChunk];
  float ev_depth[kEventChunk];
  int ev_xy[kEventChunk];
  signed char ev_pol[kEventChunk];
  double ready[kReadyChunk];
  long long head, count, k, ev_lo, ev_n, rb_lo, rb_n;
  double available;
  unsigned long long dropped, peak;
  int need;
};

// Lane 0: advances the recurrence until it finishes (kNeedNothing) or needs
// the next chunk of events or of queue ready times.
__device__ void fifo_serial(FifoShared &sh, const FifoTask &task, long long n) {
  long long head = sh.head, count = sh.count, k = sh.k;
  const long long ev_lo = sh.ev_lo, ev_end = sh.ev_lo + sh.ev_n, rb_lo = sh.rb_lo;
  long long rb_n = sh.rb_n;
  double available = sh.available;
  unsigned long long dropped = sh.dropped, peak = sh.peak;
  const unsigned long long capacity = (unsigned long long)task.capacity;
  int need = kNeedNothing;
  for (;;) {
    if (k == n) {  // deliveries up to `now` (event_readout.cpp:27-29)
      while (count > 0) {
        if (head >= rb_lo + rb_n) {
          need = kNeedReady;
          break;
        }
        if (sh.ready[head - rb_lo] <= task.now) {
          ++head;
          --count;
        } else {
          break;
        }
      }
      break;
    }
    if (k >= ev_end) {
      need = kNeedEvents;
      break;
    }
    const double t = sh.ev_t[k - ev_lo];
    // deliveries before the arrival (event_readout.cpp:15-17)
    bool stalled = false;
    while (count > 0) {
      if (head >= rb_lo + rb_n) {
        stalled = true;
        break;
      }
      if (sh.ready[head - rb_lo] <= t) {
        ++head;
        --count;
      } else {
        break;
      }
    }
    if (stalled) {
      need = kNeedReady;
      break;
    }
    if ((unsigned long long)count >= capacity) {  // (event_readout.cpp:19)
      ++dropped;
      ++k;
      continue;
    }
    // completion and delivery time (event_readout.cpp:20-24)
    const double done = (available < t ? t : available) + task.step;
    available = done;
    QueuedEvent rec;
    rec.timestamp = done + task.latency;
    rec.depth = sh.ev_depth[k - ev_lo];
    const int xy = sh.ev_xy[k - ev_lo];
    rec.x = (int16_t)(xy & 0xFFFF);
    rec.y = (int16_t)(xy >> 16);
    rec.polarity = sh.ev_pol[k - ev_lo];
    const long long pos = head + count;
    task.ring[pos & task.mask] = rec;
    if (pos == rb_lo + rb_n && rb_n < kReadyChunk) sh.ready[rb_n++] = rec.timestamp;
    ++count;
    if ((unsigned long long)count > peak) peak = (unsigned long long)count;
    ++k;
  }
  sh.head = head;
  sh.count = count;
  sh.k = k;
  sh.rb_n = rb_n;
  sh.available = available;
  sh.dropped = dropped;
  sh.peak = peak;
  sh.need = need;
}

__global__ __launch_bounds__(kFifoThreads) void readout_fifo_kernel(const FifoTask *tasks, const int *ends,
                                                                    const dvs_event *sorted, const int *skip,
                                                                    dvs_event *out, FifoResult *results) {
  __shared__ FifoShared sh;
  if (skip && *skip) return;  // the events are to be sorted again first
  const FifoTask task = tasks[blockIdx.x];
  const int lane = threadIdx.x;
  const long long begin = task.pixel_begin > 0 ? ends[task.pixel_begin - 1] : 0;
  const long long n = ends[task.pixel_end - 1] - begin;
  const long long head0 = task.head;
  if (lane == 0) {
    sh.head = task.head;
    sh.count = task.count;
    sh.k = 0;
    sh.ev_lo = 0;
    sh.ev_n = 0;
    sh.rb_lo = task.head;
    sh.rb_n = 0;
    sh.available = task.available;
    sh.dropped = task.dropped;
    sh.peak = task.peak;
  }
  int need = (n > 0 ? kNeedEvents : 0) | (task.count > 0 ? kNeedReady : 0);
  __syncthreads();
  while (need != kNeedNothing) {
    if (need & kNeedEvents) {
      const long long lo = sh.k;
      const long long m = n - lo < kEventChunk ? n - lo : kEventChunk;
      for (long long j = lane; j < m; j += kFifoThreads) {
        const dvs_event e = sorted[begin + lo + j];
        sh.ev_t[j] = e.timestamp;
        sh.ev_depth[j] = e.depth;
        sh.ev_xy[j] = (int)(uint16_t)e.x | ((int)e.y << 16);
        sh.ev_pol[j] = e.polarity;
      }
      if (lane == 0) {
        sh.ev_lo = lo;
        sh.ev_n = m;
      }
    }
    if (need & kNeedReady) {
      const long long lo = sh.head;
      const long long m = sh.count < kReadyChunk ? sh.count : kReadyChunk;
      for (long long j = lane; j < m; j += kFifoThreads) sh.ready[j] = task.ring[(lo + j) & task.mask].timestamp;
      if (lane == 0) {
        sh.rb_lo = lo;
        sh.rb_n = m;
      }
    }
    __syncthreads();
    if (lane == 0) fifo_serial(sh, task, n);
    __syncthreads();
    need = sh.need;
    if (need == kNeedNothing) break;
  }
  // Deliveries of the step: ring positions [head0, head); no slot was reused.
  const long long delivered = sh.head - head0;
  const long long copy = delivered < task.out_room ? delivered : task.out_room;
  for (long long j = lane; j < copy; j += kFifoThreads)
    out[task.out_offset + j] = unqueue(task.ring[(head0 + j) & task.mask]);
  if (lane == 0) {
    FifoResult r;
    r.head = sh.head;
    r.count = sh.count;
    r.delivered = delivered;
    r.available = sh.available;
    r.dropped = sh.dropped;
    r.peak = sh.peak;
    r.overflow = delivered > task.out_room ? 1 : 0;
    r.pad = 0;
    results[blockIdx.x] = r;
  }
}

}  // namespace

hipError_t launch_readout_fifo(const FifoTask *tasks, int task_count, const int *ends, const dvs_event *sorted,
                               const int *skip, dvs_event *out, FifoResult *results, hipStream_t st) {
  if (task_count <= 0) return hipSuccess;
  readout_fifo_kernel<<<task_count, kFifoThreads, 0, st>>>(tasks, ends, sorted, skip, out, results);
  return hipGetLastError();
}

}  // namespace dvsamd
