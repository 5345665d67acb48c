// The readout FIFO on the device (readout.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dvs_hip.h"

namespace dvsamd {

// A queued event: the packed dvs_event record (17 bytes) widened to an
// aligned 24 bytes for the ring.
struct QueuedEvent {
  double timestamp;
  float depth;
  int16_t x, y;
  int8_t polarity;
  int8_t pad[7];
};
static_assert(sizeof(QueuedEvent) == 24, "queued event layout");

__host__ __device__ inline dvs_event unqueue(const QueuedEvent &q) {
  dvs_event e;
  e.timestamp = q.timestamp;
  e.x = q.x;
  e.y = q.y;
  e.polarity = q.polarity;
  e.depth = q.depth;
  return e;
}

// One segment's FIFO for one step.  The queue holds `count` events in ring
// positions [head, head + count) (absolute indices, slot = index & mask);
// the ring has room for every event queued in the step without reusing a
// slot, so delivered events stay readable until the step ends.
struct FifoTask {
  QueuedEvent *ring;
  long long mask;
  long long head, count;
  double available, step, latency, now;
  long long capacity;
  unsigned long long dropped, peak;
  int pixel_begin, pixel_end;  // the segment's pixels in the step's compaction order
  long long out_offset, out_room;
  int force_serial;  // take the serial recurrence even where the scans apply
  int pad;
};

struct FifoResult {
  long long head, count, delivered;
  double available;
  unsigned long long dropped, peak;
  int overflow;  // more deliveries than out_room (the host bound was wrong)
  int serial;    // 1: the serial recurrence ran, 0: the scans did
};

// The readout of every task: the scan form (one 1,024-thread work-group
// per task), and the serial recurrence (one wave per task) for the tasks the
// scan form could not certify.  `accepted` (int) and `done` (double) are
// scratch of one entry per sorted event; `serial` one int per task.  Each
// task reads its slice of `sorted` (sorted by (segment, t, y, x); the slice
// is [ends[pixel_begin - 1], ends[pixel_end - 1])), writes its delivered
// events to out[out_offset ...] and its queue state to results[task].
// Nothing runs while *skip (if given) is set.
hipError_t launch_readout_fifo(const FifoTask *tasks, int task_count, const int *ends, const dvs_event *sorted,
                               const int *skip, int *accepted, double *done, int *serial, dvs_event *out,
                               FifoResult *results, hipStream_t st);

}  // namespace dvsamd
