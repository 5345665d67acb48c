// The readout FIFO of each segment on the device (EventReadout::submit,
// src/event_readout.cpp:13-29), bit-identical to the reference's serial
// loop, in two forms.
//
// Scan form (one 1,024-thread work-group per segment).  Within a step no
// event arrives at a full queue unless the queue could fill: with the queue
// holding c0 events and n arriving, c0 + n < capacity means every event is
// queued; otherwise the deliveries before event k come from the queue as it
// was (P_k = #ready <= t_k among the c0 entries) as long as the newly queued
// events are not yet due, and event k is queued iff A_<k < Q_k = capacity -
// c0 + P_k.  Q is non-decreasing, so the greedy count of queued events is
// the min-plus prefix scan A_<=k = min(k + 1, min_{j<=k}(Q_j - j) + k).  The
// completion times of the queued events follow done_a = fl(max(done_a-1,
// t_a) + 1/bandwidth).  Inside one binade, with every value on the grid of
// its ulp u, adding 1/bandwidth is adding sigma*u and fl(t_a + 1/bandwidth)
// is beta_a*u, so done_a/u - a*sigma is the max-plus prefix scan of beta_a -
// a*sigma: integers, exact.  Every step of the result is then checked
// against the reference's own floating-point expression, and every queue
// decision against the pops counted over the finished queue (old entries and
// the new ones); the peak is the maximum queue length after each push.
// Any step that fails (a binade crossed, a tie in the rounding, new entries
// due inside the step while the queue can fill, more deliveries than the
// staged window) hands the segment to the serial form.
//
// Serial form (one wave per segment).  All 64 lanes step through the
// recurrence together on wave-uniform values, with its inputs in registers:
// a window of 64 event times and one of 64 queue ready times, one per lane,
// read with v_readlane at a uniform index, refilled from LDS chunks (1,024
// events, 1,024 ready times) that refill from HBM.
#include <hip/hip_runtime.h>

#include "readout.h"

namespace dvsamd {
namespace {

constexpr int kLanes = 64;
constexpr int kEventChunk = 1024;
constexpr int kReadyChunk = 1024;

// The queue ring is global memory.  Said so (in the device pass), its
// accesses are global instructions; through a plain pointer they would be
// flat ones, and every LDS access after a flat store would wait for that
// store to land.
#if defined(__HIP_DEVICE_COMPILE__)
#define kGlobal __attribute__((address_space(1)))
#else
#define kGlobal
#endif

// v_readlane of a double at a wave-uniform lane.
__device__ inline double lane_value(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(kLanes) void readout_fifo_kernel(const FifoTask *tasks, const int *ends,
                                                              const dvs_event *sorted, const int *skip,
                                                              const int *serial, dvs_event *out,
                                                              FifoResult *results) {
  __shared__ double s_ev_t[kEventChunk];
  __shared__ float s_ev_depth[kEventChunk];
  __shared__ int s_ev_xy[kEventChunk];
  __shared__ signed char s_ev_pol[kEventChunk];
  __shared__ double s_ready[kReadyChunk];
  if (skip && *skip) return;  // the events are to be sorted again first
  if (!serial[blockIdx.x]) return;  // the scan form finished this task
  const FifoTask task = tasks[blockIdx.x];
  const int lane = threadIdx.x;
  const long long begin = task.pixel_begin > 0 ? ends[task.pixel_begin - 1] : 0;
  const long long n = ends[task.pixel_end - 1] - begin;
  const long long head0 = task.head;
  const unsigned long long capacity = (unsigned long long)task.capacity;
  QueuedEvent kGlobal *const ring = (QueuedEvent kGlobal *)task.ring;
  const long long mask = task.mask;
  // wave-uniform recurrence state
  long long head = task.head, count = task.count, k = 0;
  double available = task.available;
  unsigned long long dropped = task.dropped, peak = task.peak;
  // event chunk [ec_lo, ec_lo + ec_n) in LDS; window [ew_lo, ew_lo + ew_n)
  // in registers (lane j: event ew_lo + j)
  long long ec_lo = 0, ec_n = 0, ew_lo = 0, ew_n = 0;
  double my_t = 0;
  float my_depth = 0;
  int my_xy = 0, my_pol = 0;
  // ready-time chunk [rc_lo, rc_lo + rc_n) in LDS; window [rw_lo, rw_lo +
  // rw_n) in registers (lane j: queue position rw_lo + j)
  long long rc_lo = head, rc_n = 0, rw_lo = head, rw_n = 0;
  double my_ready = 0;

  // The next ready window, from head on (count > 0).
  auto refill_ready = [&]() {
    __syncthreads();  // the lanes' queue entries and LDS appends are visible
    const long long want = count < kLanes ? count : kLanes;
    if (head < rc_lo || head + want > rc_lo + rc_n) {
      rc_lo = head;
      rc_n = count < kReadyChunk ? count : kReadyChunk;
      for (long long j = lane; j < rc_n; j += kLanes) s_ready[j] = ring[(rc_lo + j) & mask].timestamp;
      __syncthreads();
    }
    const long long have = rc_lo + rc_n - head;
    rw_lo = head;
    rw_n = want < have ? want : have;
    my_ready = lane < rw_n ? s_ready[head - rc_lo + lane] : 0.0;
  };
  // The next event window, from k on (k < n).
  auto refill_events = [&]() {
    if (k >= ec_lo + ec_n) {
      __syncthreads();
      ec_lo = k;
      ec_n = n - k < kEventChunk ? n - k : kEventChunk;
      for (long long j = lane; j < ec_n; j += kLanes) {
        const dvs_event e = sorted[begin + ec_lo + j];
        s_ev_t[j] = e.timestamp;
        s_ev_depth[j] = e.depth;
        s_ev_xy[j] = (int)(uint16_t)e.x | ((int)e.y << 16);
        s_ev_pol[j] = e.polarity;
      }
      __syncthreads();
    }
    const long long have = ec_lo + ec_n - k;
    ew_lo = k;
    ew_n = have < kLanes ? have : kLanes;
    if (lane < ew_n) {
      const long long j = k - ec_lo + lane;
      my_t = s_ev_t[j];
      my_depth = s_ev_depth[j];
      my_xy = s_ev_xy[j];
      my_pol = s_ev_pol[j];
    }
  };

  for (;;) {
    if (k == n) {  // deliveries up to `now` (event_readout.cpp:27-29)
      while (count > 0) {
        if (head == rw_lo + rw_n) refill_ready();
        if (lane_value(my_ready, (int)(head - rw_lo)) <= task.now) {
          ++head;
          --count;
        } else {
          break;
        }
      }
      break;
    }
    if (k == ew_lo + ew_n) refill_events();
    const double t = lane_value(my_t, (int)(k - ew_lo));
    // deliveries before the arrival (event_readout.cpp:15-17)
    while (count > 0) {
      if (head == rw_lo + rw_n) refill_ready();
      if (lane_value(my_ready, (int)(head - rw_lo)) <= t) {
        ++head;
        --count;
      } else {
        break;
      }
    }
    if ((unsigned long long)count >= capacity) {  // (event_readout.cpp:19)
      ++dropped;
      ++k;
      continue;
    }
    // completion and delivery time (event_readout.cpp:20-24)
    const double done = (available < t ? t : available) + task.step;
    available = done;
    const double ready = done + task.latency;
    const long long pos = head + count;
    if (lane == (int)(k - ew_lo)) {
      QueuedEvent q;
      q.timestamp = ready;
      q.depth = my_depth;
      q.x = (int16_t)(my_xy & 0xFFFF);
      q.y = (int16_t)(my_xy >> 16);
      q.polarity = (int8_t)my_pol;
      ring[pos & mask] = q;
    }
    if (pos == rw_lo + rw_n && rw_n < kLanes) {
      if (lane == (int)rw_n) my_ready = ready;
      ++rw_n;
    }
    if (pos == rc_lo + rc_n && rc_n < kReadyChunk) {
      if (lane == 0) s_ready[rc_n] = ready;
      ++rc_n;
    }
    ++count;
    if ((unsigned long long)count > peak) peak = (unsigned long long)count;
    ++k;
  }
  __syncthreads();
  // Deliveries of the step: ring positions [head0, head); no slot was reused.
  const long long delivered = head - head0;
  const long long copy = delivered < task.out_room ? delivered : task.out_room;
  for (long long j = lane; j < copy; j += kLanes) out[task.out_offset + j] = unqueue(ring[(head0 + j) & mask]);
  if (lane == 0) {
    FifoResult r;
    r.head = head;
    r.count = count;
    r.delivered = delivered;
    r.available = available;
    r.dropped = dropped;
    r.peak = peak;
    r.overflow = delivered > task.out_room ? 1 : 0;
    r.serial = 1;
    results[blockIdx.x] = r;
  }
}


constexpr int kScanThreads = 1024;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kPer = 4;                          // events per thread per tile
constexpr int kScanTile = kScanThreads * kPer;   // 4,096 events
constexpr int kWindow = 8192;                    // queue ready times staged for the pop counts
constexpr long long kNoValue = -(1LL << 62);     // the identity of the max scan
constexpr long long kNoBound = (1LL << 62);      // the identity of the min scan

struct ScanShared {
  double win[kWindow];
  double tile_done[kScanTile];
  long long wave_ll[kScanWaves];
  long long value_ll[4];
  double value_d[2];
  long long win_lo, win_n;
  int fail;
  int first;
};

template <class T, class Op>
__device__ inline T block_exclusive(T v, T identity, Op op, T *wave_buf, T &total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const T y = __shfl_up(x, off, 64);
    if (lane >= off) x = op(x, y);
  }
  if (lane == 63) wave_buf[wave] = x;
  __syncthreads();
  T before = identity, all = identity;
#pragma unroll
  for (int w = 0; w < kScanWaves; ++w) {
    const T s = wave_buf[w];
    if (w < wave) before = op(before, s);
    all = op(all, s);
  }
  __syncthreads();
  T ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = identity;
  total = all;
  return op(before, ex);
}

__device__ inline long long min_ll(long long a, long long b) { return a < b ? a : b; }
__device__ inline long long max_ll(long long a, long long b) { return a < b ? b : a; }

__global__ __launch_bounds__(kScanThreads) void readout_scan_kernel(const FifoTask *tasks, const int *ends,
                                                                    const dvs_event *sorted, const int *skip,
                                                                    int *accepted, double *done_out, int *serial,
                                                                    dvs_event *out, FifoResult *results) {
  __shared__ ScanShared sh;
  if (skip && *skip) return;
  const FifoTask task = tasks[blockIdx.x];
  const int tid = threadIdx.x;
  if (task.force_serial) {
    if (tid == 0) serial[blockIdx.x] = 1;
    return;
  }
  QueuedEvent kGlobal *const ring = (QueuedEvent kGlobal *)task.ring;
  const long long mask = task.mask, head0 = task.head, count0 = task.count;
  const long long begin = task.pixel_begin > 0 ? ends[task.pixel_begin - 1] : 0;
  const long long n = ends[task.pixel_end - 1] - begin;
  const long long capacity = task.capacity;
  const double step = task.step;
  const bool all_queued = count0 + n < capacity;
  int *const acc = accepted + begin;
  double *const dval = done_out + begin;
  if (tid == 0) {
    sh.fail = 0;
    sh.first = 0x7FFFFFFF;
  }
  __syncthreads();
  // #{queue positions i < len with ready <= x}, from the staged window;
  // positions below the window are due by every x asked.
  auto count_le = [&](double x, long long len) -> long long {
    int lo = 0, hi = (int)sh.win_n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sh.win[mid] <= x) lo = mid + 1;
      else hi = mid;
    }
    if (lo == sh.win_n && sh.win_lo + sh.win_n < len) sh.fail = 1;
    return sh.win_lo + lo;
  };
  auto load_window = [&](long long from, long long len) {
    __syncthreads();
    const long long m = len - from < kWindow ? len - from : kWindow;
    for (long long j = tid; j < m; j += kScanThreads) sh.win[j] = ring[(head0 + from + j) & mask].timestamp;
    if (tid == 0) {
      sh.win_lo = from;
      sh.win_n = m;
    }
    __syncthreads();
  };
  auto a_incl = [&](long long k) -> long long { return k < 0 ? 0 : (all_queued ? k + 1 : (long long)acc[k]); };

  // 1. Queue decisions (min-plus scan); every event is queued if the queue cannot fill.
  if (!all_queued && n > 0) {
    long long carry = kNoBound, p_carry = 0;
    for (long long base = 0; base < n; base += kScanTile) {
      load_window(p_carry, count0);
      long long v[kPer], pops[kPer];
      long long agg = kNoBound;
      const long long k0 = base + (long long)tid * kPer;
      for (int j = 0; j < kPer; ++j) {
        const long long k = k0 + j;
        v[j] = kNoBound;
        pops[j] = 0;
        if (k < n) {
          pops[j] = count_le(sorted[begin + k].timestamp, count0);
          v[j] = capacity - count0 + pops[j] - k;
        }
        agg = min_ll(agg, v[j]);
      }
      long long tile_min;
      long long m = min_ll(carry, block_exclusive<long long>(agg, kNoBound, min_ll, sh.wave_ll, tile_min));
      long long prev = k0 > 0 ? min_ll(k0, k0 - 1 + m) : 0;  // A_<=k0-1
      for (int j = 0; j < kPer; ++j) {
        const long long k = k0 + j;
        if (k >= n) break;
        m = min_ll(m, v[j]);
        const long long a = min_ll(k + 1, k + m);
        acc[k] = (int)a;
        if (a == 1 && prev == 0) atomicMin(&sh.first, (int)k);
        prev = a;
      }
      carry = min_ll(carry, tile_min);
      const long long last = (n - base < kScanTile ? n - base : kScanTile) - 1;
      if (tid == (int)(last / kPer)) sh.value_ll[0] = pops[last % kPer];
      __syncthreads();
      p_carry = sh.value_ll[0];
      if (sh.fail) break;
    }
    __syncthreads();
  } else if (n > 0 && tid == 0) {
    sh.first = 0;
  }
  __syncthreads();
  const long long total = n > 0 ? a_incl(n - 1) : 0;  // events queued
  // 2. Completion times (max-plus scan on the ulp grid), checked step by step.
  double last_done = task.available;
  if (!sh.fail && total > 0) {
    const long long f = sh.first;
    const double t0 = sorted[begin + f].timestamp;
    const double d0 = (task.available < t0 ? t0 : task.available) + step;
    const long long efield = (__double_as_longlong(d0) >> 52) & 0x7FF;
    const double u = __longlong_as_double((efield - 52) << 52);
    const double binade = __longlong_as_double(efield << 52);
    const double S = (d0 + step) - d0;
    const long long sigma = (long long)(S / u);
    const long long e0 = (long long)(d0 / u);
    if (efield <= 53 || (double)sigma * u != S || sigma <= 0 || sigma > (1LL << 52) / (total + 1)) {
      if (tid == 0) sh.fail = 1;
    }
    __syncthreads();
    long long carry = kNoValue;
    double carry_done = task.available;
    for (long long base = 0; base < n && !sh.fail; base += kScanTile) {
      const long long k0 = base + (long long)tid * kPer;
      const long long a_before_tile = a_incl(base - 1);
      long long g[kPer], a_of[kPer];
      double t[kPer];
      bool q[kPer];
      long long agg = kNoValue;
      long long prev_a = a_incl(k0 - 1);
      for (int j = 0; j < kPer; ++j) {
        const long long k = k0 + j;
        g[j] = kNoValue;
        q[j] = false;
        a_of[j] = 0;
        t[j] = 0;
        if (k < n) {
          const long long a = a_incl(k);
          q[j] = a > prev_a;
          prev_a = a;
          a_of[j] = a - 1;
          if (q[j]) {
            t[j] = sorted[begin + k].timestamp;
            if (a_of[j] == 0) {
              g[j] = e0;
            } else {
              const double beta = t[j] + step;
              g[j] = beta < binade ? kNoValue : (long long)(beta / u) - a_of[j] * sigma;
            }
          }
        }
        agg = max_ll(agg, g[j]);
      }
      long long tile_max;
      long long m = max_ll(carry, block_exclusive<long long>(agg, kNoValue, max_ll, sh.wave_ll, tile_max));
      for (int j = 0; j < kPer; ++j) {
        if (!q[j]) continue;
        m = max_ll(m, g[j]);
        const double d = (double)(a_of[j] * sigma + m) * u;
        dval[a_of[j]] = d;
        sh.tile_done[a_of[j] - a_before_tile] = d;
        const dvs_event e = sorted[begin + k0 + j];
        QueuedEvent qe;
        qe.timestamp = d + task.latency;
        qe.depth = e.depth;
        qe.x = e.x;
        qe.y = e.y;
        qe.polarity = e.polarity;
        ring[(head0 + count0 + a_of[j]) & mask] = qe;
      }
      __syncthreads();
      for (int j = 0; j < kPer; ++j) {  // done_a == fl(max(done_a-1, t_a) + step), as the reference computes it
        if (!q[j]) continue;
        const long long local = a_of[j] - a_before_tile;
        const double prev = a_of[j] == 0 ? task.available : (local > 0 ? sh.tile_done[local - 1] : carry_done);
        const double want = (prev < t[j] ? t[j] : prev) + step;
        if (want != sh.tile_done[local]) sh.fail = 1;
      }
      const long long a_after_tile = a_incl((base + kScanTile < n ? base + kScanTile : n) - 1);
      carry = max_ll(carry, tile_max);
      if (a_after_tile > a_before_tile) carry_done = sh.tile_done[a_after_tile - 1 - a_before_tile];
      __syncthreads();
    }
    last_done = carry_done;
  }
  // 3. The queue decisions against the pops over the finished queue, and
  // the peak; then the deliveries up to `now`.
  const long long len = count0 + total;
  unsigned long long peak = task.peak;
  long long p_carry = 0;
  for (long long base = 0; base < n && !sh.fail; base += kScanTile) {
    load_window(p_carry, len);
    const long long k0 = base + (long long)tid * kPer;
    long long prev_a = a_incl(k0 - 1), pops_last = 0;
    unsigned long long top = 0;
    for (int j = 0; j < kPer; ++j) {
      const long long k = k0 + j;
      if (k >= n) break;
      const long long a = a_incl(k);
      const long long pops = count_le(sorted[begin + k].timestamp, len);
      const long long queued_before = count0 + prev_a - pops;
      const bool q = a > prev_a;
      if (q != (queued_before < capacity)) sh.fail = 1;
      if (q && (unsigned long long)(queued_before + 1) > top) top = (unsigned long long)(queued_before + 1);
      prev_a = a;
      pops_last = pops;
    }
    long long tile_top;
    block_exclusive<long long>((long long)top, 0, max_ll, sh.wave_ll, tile_top);
    if ((unsigned long long)tile_top > peak) peak = (unsigned long long)tile_top;
    const long long last = (n - base < kScanTile ? n - base : kScanTile) - 1;
    if (tid == (int)(last / kPer)) sh.value_ll[1] = pops_last;
    __syncthreads();
    p_carry = sh.value_ll[1];
  }
  long long delivered = 0;
  if (!sh.fail) {
    load_window(p_carry, len);
    if (tid == 0) sh.value_ll[2] = count_le(task.now, len);
    __syncthreads();
    delivered = sh.value_ll[2];
  }
  __syncthreads();
  if (sh.fail) {
    if (tid == 0) serial[blockIdx.x] = 1;
    return;
  }
  const long long copy = delivered < task.out_room ? delivered : task.out_room;
  for (long long j = tid; j < copy; j += kScanThreads) {
    QueuedEvent qe = ring[(head0 + j) & mask];
    out[task.out_offset + j] = unqueue(qe);
  }
  if (tid == 0) {
    serial[blockIdx.x] = 0;
    FifoResult r;
    r.head = head0 + delivered;
    r.count = len - delivered;
    r.delivered = delivered;
    r.available = last_done;
    r.dropped = task.dropped + (unsigned long long)(n - total);
    r.peak = peak;
    r.overflow = delivered > task.out_room ? 1 : 0;
    r.serial = 0;
    results[blockIdx.x] = r;
  }
}
}  // namespace

hipError_t launch_readout_fifo(const FifoTask *tasks, int task_count, const int *ends, const dvs_event *sorted,
                               const int *skip, int *accepted, double *done, int *serial, dvs_event *out,
                               FifoResult *results, hipStream_t st) {
  if (task_count <= 0) return hipSuccess;
  readout_scan_kernel<<<task_count, kScanThreads, 0, st>>>(tasks, ends, sorted, skip, accepted, done, serial, out,
                                                           results);
  readout_fifo_kernel<<<task_count, kLanes, 0, st>>>(tasks, ends, sorted, skip, serial, out, results);
  return hipGetLastError();
}

}  // namespace dvsamd
