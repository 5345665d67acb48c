// Multi-segment event engine: the EventSensor of the reference
// (src/cuda/event_sensor.cu:244-305) generalised to many sensors
// ("segments") that advance together in shared gfx950 launches.
//
// One step takes the next queued host horizons of every segment, builds the
// circuit-cell table of each (generate_pixels, event_sensor.cu:37-40),
// traces every (pixel, cell) of every segment, integrates every pixel's
// circuit, compacts the crossings by a device prefix sum, traces event-time
// depth, sorts the events by (segment, t, y, x) on the device and runs each
// segment's readout FIFO (src/event_readout.cpp:6-30) on the sorted stream,
// also on the device (readout.hip).
// The plain dvs_sensor_* API is a one-segment engine whose every
// process_many call is one step over exactly the call's intervals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <vector>

#include "adaptive.h"
#include "common.h"
#include "dvs_hip.h"
#include "kernels.h"
#include "readout.h"
#include "scan_sort.h"
#include "scene.h"
#include "trajectory_device.h"

namespace dvsamd {
namespace {

constexpr double kQ = 1.602176634e-19, kK = 1.380649e-23, kH = 6.62607015e-34, kC = 299792458;

// parameters / photo_parameters (src/cuda/event_sensor.cu:151-176), with the
// compile-time ENABLE_PHYSIC_EFFECTS switch as a runtime flag.
void derive(const dvs_physic &c, bool effects, CircuitParams &p, PhotoParams &photo) {
  p.threshold_pos = c.threshold_pos;
  p.threshold_neg = c.threshold_neg;
  const double voltage = kK * c.temperature_k / (kQ * c.subthreshold_kappa);
  p.charge_scale = c.receptor_capacitance_f * voltage / kQ;
  p.refractory_time = effects ? c.refractory_time : 0.0;
  p.reference_leak = effects ? c.reset_leakage_a / (c.reset_capacitance_f * voltage) : 0.0;
  p.reset_noise_std =
      effects && c.reset_thermal_noise ? std::sqrt(kK * c.temperature_k / c.reset_capacitance_f) / voltage : 0.0;
  const double pitch = c.pixel_pitch_um * 1e-6;
  photo.electrons_per_radiance = c.radiance_scale * 3.141592653589793 / (4 * c.f_number * c.f_number) *
                                 c.lens_transmission * pitch * pitch * c.fill_factor * c.quantum_efficiency *
                                 (c.wavelength_nm * 1e-9) / (kH * kC);
  photo.integration_step = c.circuit_step;
  photo.dark_electrons = effects ? c.dark_current_a / kQ : 0.0;
  photo.shot_noise = effects ? c.photon_shot_noise : 0;
  photo.pad = 0;
}

// Bounded FIFO readout (EventReadout::submit / drain, src/event_readout.cpp:6-34)
// on the host, for the EventReadout API on host events (dvs_readout_*),
// over events already in (t, y, x) order.  The pending events are a ring of
// {ready time, event} in submission order; their ready times ascend (each is
// max(previous, t) + 1 / bandwidth + latency).
struct Readout {
  double bandwidth = 0, latency = 0, available = 0;
  size_t capacity = 1, dropped = 0, peak = 0;
  struct Pending { double ready; dvs_event event; };
  std::vector<Pending> ring;  // power-of-two size
  size_t head = 0, count = 0;

  const Pending &at(size_t i) const { return ring[(head + i) & (ring.size() - 1)]; }
  void push(const Pending &x) {
    if (count == ring.size()) grow(count + 1);
    ring[(head + count) & (ring.size() - 1)] = x;
    ++count;
  }
  void grow(size_t need) {
    size_t n = std::max<size_t>(ring.size(), 4096);
    while (n < need) n *= 2;
    if (n == ring.size()) return;
    std::vector<Pending> bigger(n);
    for (size_t i = 0; i < count; ++i) bigger[i] = at(i);
    ring.swap(bigger);
    head = 0;
  }
  // Moves the first n pending events to `ready`.
  void deliver(size_t n, std::vector<dvs_event> &ready) {
    for (size_t i = 0; i < n; ++i) ready.push_back(at(i).event);
    head = (head + n) & (ring.size() - 1);
    count -= n;
  }

  void submit_sorted(const dvs_event *events, size_t n, double now, std::vector<dvs_event> &ready) {
    if (bandwidth <= 0) {  // no queue: every event is delivered as it comes
      ready.insert(ready.end(), events, events + n);
      return;
    }
    if (count + n < capacity) {
      submit_unbounded(events, n, now, ready);
      return;
    }
    for (size_t k = 0; k < n; ++k) {
      const dvs_event &e = events[k];
      size_t pops = 0;
      while (pops < count && at(pops).ready <= e.timestamp) ++pops;
      deliver(pops, ready);
      if (count >= capacity) { ++dropped; continue; }
      const double done = std::max(available, (double)e.timestamp) + 1.0 / bandwidth;
      available = done;
      dvs_event d = e;
      d.timestamp = done + latency;
      push(Pending{done + latency, d});
      peak = std::max(peak, count);
    }
    size_t pops = 0;
    while (pops < count && at(pops).ready <= now) ++pops;
    deliver(pops, ready);
  }
  // The same when no event can be dropped (fewer than `capacity` pending
  // even if nothing were delivered): all events are queued first, then one
  // pass over the merged order of ready times and event times replays the
  // deliveries before each event, which only the peak depends on.  The
  // merge is branch free: one step delivers or queues.
  void submit_unbounded(const dvs_event *events, size_t n, double now, std::vector<dvs_event> &ready) {
    const size_t old = count;
    grow(count + n);
    const double step = 1.0 / bandwidth;
    for (size_t k = 0; k < n; ++k) {
      const double done = std::max(available, (double)events[k].timestamp) + step;
      available = done;
      dvs_event d = events[k];
      d.timestamp = done + latency;
      ring[(head + count) & (ring.size() - 1)] = Pending{done + latency, d};
      ++count;
    }
    const size_t mask = ring.size() - 1;
    size_t p = 0, k = 0, size = old, top = peak;
    while (k < n) {
      const bool pop = (p < old + k) & (ring[(head + p) & mask].ready <= (double)events[k].timestamp);
      p += pop;
      k += !pop;
      size = pop ? size - 1 : size + 1;
      top = std::max(top, size);
    }
    peak = top;
    while (p < count && at(p).ready <= now) ++p;
    ready.reserve(ready.size() + p);
    deliver(p, ready);
  }
  void drain(std::vector<dvs_event> &ready) { deliver(count, ready); }
};

// Bits of the bucket the event sort orders by first (two radix passes).
constexpr int kTimeBucketBits = 16;

// The order-preserving 64-bit key of an IEEE double (as the compaction
// kernel builds it for event times).
inline unsigned long long time_key(double t) {
  unsigned long long bits;
  std::memcpy(&bits, &t, sizeof(bits));
  return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ULL);
}

// A segment's readout FIFO on the device (readout.hip): the queue lives in
// a device ring, the rest of its state here between steps.  Without a
// bandwidth every event is delivered as it comes (event_readout.cpp:18).
struct DeviceFifo {
  double bandwidth = 0, latency = 0, step = 0, available = 0;
  double last_now = -std::numeric_limits<double>::infinity();
  long long capacity = 1, head = 0, count = 0;
  unsigned long long dropped = 0, peak = 0;
  DeviceArray<QueuedEvent> ring;
  long long ring_size = 0;
  bool queued() const { return bandwidth > 0; }
};

struct Segment {
  int id = 0;
  dvs_scene scene = nullptr;
  int npix = 0, width = 0, max_events = 16;
  uint64_t seed = 0;
  CircuitParams params{};
  PhotoParams photo{};
  double threshold_std = 0, threshold_min = 0;
  double last_time = 0, fed_until = 0;
  bool initialized = false;
  DeviceArray<dvs_pixel_state> states;
  DeviceArray<CrossingEvent> slots;
  std::deque<dvs_optical_interval> queue;
  // per queued interval: the trajectory whose device evaluation supplies its
  // states when its step runs, or null (the states came with the interval)
  std::deque<dvs_trajectory> source;
  void clear_queue() {
    queue.clear();
    source.clear();
  }
  DeviceFifo fifo;
  std::vector<dvs_event> ready, raw;
  int64_t horizons_done = 0, events_total = 0;
  // adaptive optics: the host driver policy of apps/main_dvs.cpp:96-116 per
  // segment (window of `abatch` horizons, halved on a rejected check; a
  // window of one horizon is then integrated natively; the receptor check
  // runs at the first window and every check_period seconds after).
  uint64_t cells_done = 0;  // circuit cells integrated since initialisation (every pixel alike)
  bool adaptive = false;
  double tolerance = 1e-2, check_period = .25, next_check = 0;
  int abatch = 1, capacity = 0;
  dvs_refinement_stats last_refine{};
  dvs_receptor_probe last_probe{};
  bool last_accepted = true;
  // the last adaptive window: its engine step, its place in that step and
  // its bounds (verify_optics reads its knots while no other step ran)
  uint64_t window_step = ~0ull;
  int window_index = 0, window_pixel_offset = 0, window_ivs = 0;
  double window_start = 0, window_end = 0;
  // numerical diagnostics (EventSensor::enable_numerical_diagnostics): every
  // adaptive window also measures the receptor check, kept for
  // verify_receptor
  bool diagnostics = false;
  dvs_receptor_probe diag_probe{};
  uint64_t diag_step = ~0ull;
  ~Segment() { scene_release(scene); }
};

// Work of one segment inside one step.
struct StepSegment {
  Segment *seg;
  std::vector<dvs_optical_interval> ivs;
  std::vector<dvs_trajectory> src;  // Segment::source of each interval
  int cells = 0, cell_offset = 0, iv_offset = 0, pixel_offset = 0;
  double tolerance = 0;  // adaptive windows only
  bool check = false;
};

// A queued adaptive window setup (Engine::setup_adaptive): what the rest of
// the window needs, and for a setup queued ahead of its window, the window
// and the segment state it was made from.
struct AdaptSetup {
  Timer step_t, refine_t, probe_t;
  size_t lds = 0;
  int pcur = 0;
  std::vector<SegDev> tables;
  std::vector<AdaptDev> ad;
  std::vector<StepSegment> work;
  int cell_total = 0, capacity = 0;
  uint64_t cells_done = 0;
};

// The window a caller expects to run next on a segment.
struct Expected {
  int id = -1;
  dvs_trajectory trajectory = nullptr;
  std::vector<double> times;
  double tolerance = 0;
};

struct GraphKey {
  std::vector<long long> shape;
  bool operator==(const GraphKey &o) const { return shape == o.shape; }
};

}  // namespace

class Engine {
 public:
  explicit Engine(int horizons_per_step, bool use_graph) : horizons_(horizons_per_step), use_graph_(use_graph) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "engine stream");
    // DVS_READOUT_SERIAL=1 runs every readout through the serial recurrence
    // (the scan form's fallback), for comparing the two.
    const char *serial = std::getenv("DVS_READOUT_SERIAL");
    force_serial_readout_ = serial && serial[0] == '1';
  }
  ~Engine() {
    segments_.clear();
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int add(dvs_scene scene, const dvs_physic &c, bool effects, bool aer) {
    const int w = scene->view.width, h = scene->view.height;
    // EventSensor ctor checks (event_sensor.cu:251-252).
    if (w <= 0 || h <= 0 || w > 32767 || h > 32767 || c.max_events_per_step < 1 || c.max_events_per_step > 64 ||
        (long long)w * h * c.max_events_per_step > std::numeric_limits<int>::max())
      fail(DVS_E_ARGUMENT, "invalid sensor capacity");
    if (!segments_.empty() && segments_.begin()->second->scene->view.samples != scene->view.samples)
      fail(DVS_E_ARGUMENT, "all segments of one engine must use the same sub-ray grid");
    auto s = std::make_unique<Segment>();
    s->id = next_id_++;
    s->npix = w * h;
    s->width = w;
    s->max_events = c.max_events_per_step;
    s->seed = (uint64_t)(uint32_t)c.random_seed;
    derive(c, effects, s->params, s->photo);
    s->threshold_std = effects ? c.threshold_std : 0.0;
    s->threshold_min = c.threshold_min;
    s->fifo.bandwidth = aer ? (double)c.readout_bandwidth : 0.0;
    s->fifo.latency = aer ? (double)c.readout_latency : 0.0;
    s->fifo.step = s->fifo.bandwidth > 0 ? 1.0 / s->fifo.bandwidth : 0.0;
    s->fifo.capacity = std::max(1, c.readout_queue_capacity);
    s->states.reserve(s->npix, "pixel states");
    s->slots.reserve((size_t)s->npix * s->max_events, "events");
    scene_retain(scene);
    s->scene = scene;
    const int id = s->id;
    segments_[id] = std::move(s);
    return id;
  }

  Segment &segment(int id) {
    auto it = segments_.find(id);
    if (it == segments_.end()) fail(DVS_E_ARGUMENT, "unknown segment");
    return *it->second;
  }

  void remove(int id) { segments_.erase(id); }

  // EventSensor::initialize / initialize_pixels (event_sensor.cu:18-28,262-270).
  void initialize(int id, const dvs_optical_state &pose, double timestamp) {
    Segment &s = segment(id);
    if (!std::isfinite(timestamp) || s.initialized) fail(DVS_E_ARGUMENT, "invalid sensor initialization");
    // the poses, radiance and counters below are buffers a queued setup uses
    ahead_.reset();
    expected_ = Expected{};
    DeviceArray<dvs_optical_state> state;
    state.upload(&pose, 1, "initial pose", stream_);
    poses_.reserve(1, "initial pose");
    hip_check(launch_state_poses(1, state.ptr, poses_.ptr, stream_), "initialize circuit");
    radiance_.reserve(s.npix, "initial radiance");
    status_.reserve(1, "error flag");
    counter_.reserve(1, "ray counter");
    hip_check(hipMemsetAsync(status_.ptr, 0, sizeof(int), stream_), "clear initialization error");
    hip_check(hipMemsetAsync(counter_.ptr, 0, sizeof(unsigned long long), stream_), "clear ray counter");
    hip_check(launch_trace(dvsdev::MODE_FRAME, s.scene->view, poses_.ptr, 1, s.npix, s.npix, nullptr, radiance_.ptr,
                           nullptr, nullptr, status_.ptr, counter_.ptr, stream_),
              "initialize circuit");
    hip_check(launch_init(s.npix, radiance_.ptr, s.photo, s.params, s.threshold_std, s.threshold_min, s.seed,
                          s.states.ptr, status_.ptr, stream_),
              "initialize circuit");
    int status = 0;
    hip_check(hipMemcpyAsync(&status, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, stream_),
              "initialize circuit status");
    hip_check(hipStreamSynchronize(stream_), "initialize circuit status");
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample during initialization");
    s.last_time = s.fed_until = timestamp;
    s.initialized = true;
  }

  // Launch shapes of the adaptive kernels (results do not depend on them).
  void set_tuning(const dvs_tuning &t) {
    if (t.rate_cells < 8 || t.rate_cells > 1024 || t.rate_cells % 8 != 0 ||
        (t.trace_threads != 64 && t.trace_threads != 128 && t.trace_threads != 256) ||
        (t.traversal != 0 && t.traversal != 1))
      fail(DVS_E_ARGUMENT, "invalid launch tuning");
    tuning_ = t;
  }
  const dvs_tuning &tuning() const { return tuning_; }

  // Queues contiguous intervals (EventSensor::upload_intervals checks,
  // event_sensor.cu:274-282).
  void feed(int id, const dvs_optical_interval *ivs, int64_t n) {
    Segment &s = segment(id);
    if (!s.initialized || n <= 0 || !ivs) fail(DVS_E_ARGUMENT, "sensor not initialized or no intervals");
    double previous = s.fed_until;
    for (int64_t k = 0; k < n; ++k) {
      const dvs_optical_interval &iv = ivs[k];
      if (!std::isfinite(iv.t0) || !std::isfinite(iv.t1) || iv.t0 != previous || iv.t1 <= iv.t0)
        fail(DVS_E_ARGUMENT, "sensor intervals must be positive and contiguous");
      previous = iv.t1;
    }
    s.queue.insert(s.queue.end(), ivs, ivs + n);
    s.source.insert(s.source.end(), (size_t)n, nullptr);
    s.fed_until = previous;
  }

  // Queues the n - 1 intervals between consecutive times, their states to be
  // evaluated from `trajectory` on the device when their step runs.
  // The window the next adaptive-times call on segment `id` will most
  // likely run (n == 0 forgets it).  Kept until the end of the next
  // adaptive-times call.
  void expect_adaptive(int id, dvs_trajectory trajectory, const double *times, int64_t n, double tolerance) {
    (void)segment(id);
    expected_ = Expected{};
    if (n == 0) return;
    if (!trajectory || !times || n < 2 || !std::isfinite(tolerance) || tolerance <= 0)
      fail(DVS_E_ARGUMENT, "invalid expected adaptive window");
    for (int64_t k = 0; k + 1 < n; ++k)
      if (!std::isfinite(times[k]) || !std::isfinite(times[k + 1]) || times[k + 1] <= times[k])
        fail(DVS_E_ARGUMENT, "sensor intervals must be positive and contiguous");
    expected_.id = id;
    expected_.trajectory = trajectory;
    expected_.times.assign(times, times + n);
    expected_.tolerance = tolerance;
  }
  void forget_expected() { expected_ = Expected{}; }

  void feed_trajectory(int id, dvs_trajectory trajectory, const double *times, int64_t n) {
    Segment &s = segment(id);
    if (!s.initialized || n < 2 || !times || !trajectory) fail(DVS_E_ARGUMENT, "sensor not initialized or no intervals");
    double previous = s.fed_until;
    if (times[0] != previous) fail(DVS_E_ARGUMENT, "sensor intervals must be positive and contiguous");
    std::vector<dvs_optical_interval> ivs((size_t)n - 1);
    for (int64_t k = 0; k + 1 < n; ++k) {
      if (!std::isfinite(times[k + 1]) || times[k + 1] <= times[k])
        fail(DVS_E_ARGUMENT, "sensor intervals must be positive and contiguous");
      dvs_optical_interval &iv = ivs[(size_t)k];
      iv = dvs_optical_interval{};
      iv.t0 = times[k];
      iv.t1 = times[k + 1];
      previous = iv.t1;
    }
    s.queue.insert(s.queue.end(), ivs.begin(), ivs.end());
    s.source.insert(s.source.end(), ivs.size(), trajectory);
    s.fed_until = previous;
  }

  // One step over every segment with queued intervals; returns the number of
  // segments that still have queued intervals afterwards.
  int step(int only = -1) {
    struct Group {
      std::vector<StepSegment> work;
      int pixels = 0, cells = 0, ivs = 0;
      void add(StepSegment w) {
        w.pixel_offset = pixels;
        w.cell_offset = cells;
        w.iv_offset = ivs;
        pixels += w.seg->npix;
        ivs += (int)w.ivs.size();
        for (auto &iv : w.ivs) cells += (int)std::ceil((iv.t1 - iv.t0) / w.seg->photo.integration_step);
        work.push_back(std::move(w));
      }
    };
    Group native, adaptive;
    for (auto &kv : segments_) {
      Segment &s = *kv.second;
      if ((only >= 0 && s.id != only) || s.queue.empty()) continue;
      StepSegment w;
      w.seg = &s;
      const size_t take = std::min<size_t>(s.queue.size(), (size_t)(s.adaptive ? s.abatch : horizons_));
      w.ivs.assign(s.queue.begin(), s.queue.begin() + take);
      w.src.assign(s.source.begin(), s.source.begin() + take);
      if (s.adaptive) {
        w.tolerance = s.tolerance;
        w.check = s.check_period >= 0 && w.ivs.front().t0 >= s.next_check;
        check_adaptive_window(w);
        adaptive.add(std::move(w));
      } else {
        double cells_estimate = 0;
        for (auto &iv : w.ivs) cells_estimate += (iv.t1 - iv.t0) / s.photo.integration_step;
        if (cells_estimate > 1000000) fail(DVS_E_ARGUMENT, "optical horizon too long; reduce batch size");
        native.add(std::move(w));
      }
    }
    auto consume = [](StepSegment &w) {
      w.seg->queue.erase(w.seg->queue.begin(), w.seg->queue.begin() + w.ivs.size());
      w.seg->source.erase(w.seg->source.begin(), w.seg->source.begin() + w.ivs.size());
      w.seg->horizons_done += (int64_t)w.ivs.size();
    };
    if (!native.work.empty()) {
      run(native.work, native.pixels, native.cells, native.ivs);
      for (auto &w : native.work) consume(w);
    }
    if (!adaptive.work.empty()) {
      std::vector<char> accepted;
      run_adaptive(adaptive.work, adaptive.pixels, adaptive.cells, adaptive.ivs, accepted);
      Group retry;
      for (size_t i = 0; i < adaptive.work.size(); ++i) {
        StepSegment &w = adaptive.work[i];
        Segment &s = *w.seg;
        if (accepted[i]) {
          consume(w);
          if (w.check) s.next_check = w.ivs.back().t1 + s.check_period;
        } else if (s.abatch > 1) {
          s.abatch = std::max(1, s.abatch / 2);  // retried with the shorter window at the next step
        } else {
          StepSegment again;
          again.seg = &s;
          again.ivs = w.ivs;
          again.src = w.src;
          again.check = w.check;
          retry.add(std::move(again));
        }
      }
      if (!retry.work.empty()) {
        run(retry.work, retry.pixels, retry.cells, retry.ivs);
        for (auto &w : retry.work) {
          consume(w);
          if (w.check) w.seg->next_check = w.ivs.back().t1 + w.seg->check_period;
        }
      }
    }
    int pending = 0;
    for (auto &kv : segments_) pending += !kv.second->queue.empty();
    return pending;
  }

  // process_adaptive's argument checks on a window before it is queued.
  void validate_adaptive(int id, const dvs_optical_interval *ivs, int n, double tolerance) {
    if (!ivs || n <= 0) fail(DVS_E_ARGUMENT, "empty adaptive optical horizon");
    StepSegment w;
    w.seg = &segment(id);
    w.ivs.assign(ivs, ivs + n);
    w.tolerance = tolerance;
    check_adaptive_window(w);
  }

  int capacity_of(int id) {
    Segment &s = segment(id);
    if (!s.capacity) s.capacity = adaptive_capacity(s.npix);
    return s.capacity;
  }

  // Switches a segment to adaptive windows of `horizons` horizons.
  void set_adaptive(int id, double tolerance, int horizons, double check_period) {
    Segment &s = segment(id);
    if (!std::isfinite(tolerance) || tolerance <= 0 || horizons < 1 || s.params.refractory_time <= 0)
      fail(DVS_E_ARGUMENT, "invalid adaptive tolerance or event capacity bound");
    s.adaptive = true;
    s.tolerance = tolerance;
    s.abatch = horizons;
    s.check_period = check_period;
    s.next_check = s.last_time;
  }

  dvs_stats stats{};
  int horizons() const { return horizons_; }

  // process_adaptive over exactly the queued intervals of one segment, as
  // one window (event_sensor.cu:306-355); returns whether the convergence
  // gate accepted it.  The queue is consumed only when accepted.
  // EventSensor::verify_optics / OpticalSampler::verify (event_sensor.hpp:
  // 57-59, optical_sampler.cu:86-103,195-215): the knot interpolation of the
  // segment's last adaptive window against native traces of 1024 hashed
  // pixels at 17 times, while no other step has run since.
  dvs_optical_probe verify_optics(int id, const dvs_optical_interval *ivs, int n, double tolerance) {
    Segment &s = segment(id);
    if (s.window_step != step_serial_) fail(DVS_E_STATE, "refine before validating optical knots");
    if (!ivs || n != s.window_ivs || ivs[0].t0 != s.window_start || ivs[n - 1].t1 != s.window_end)
      fail(DVS_E_ARGUMENT, "intervals differ from the last adaptive window");
    constexpr int positions = 1024, times = 17, items = positions * times;
    hipStream_t st = stream_;
    std::vector<int> pix(items), pose_of(items);
    for (int q = 0; q < items; ++q) {
      pix[q] = s.window_pixel_offset + (int)(((uint64_t)(q % positions) * 2654435761ULL) % (uint64_t)s.npix);
      pose_of[q] = q / positions;
    }
    verify_pix_.upload(pix.data(), pix.size(), "optical convergence probes", st, staging_);
    verify_pose_.upload(pose_of.data(), pose_of.size(), "optical convergence probes", st, staging_);
    staging_.flush(st);
    verify_poses_.reserve(times, "optical convergence probes");
    verify_rad_.reserve(items, "optical convergence probes");
    verify_err_.reserve(items, "optical convergence probes");
    hip_check(launch_fill_int(status_.ptr, 0, 1, st), "clear probe status");
    const SegDev *seg = segtab_.ptr + s.window_index;
    hip_check(launch_adapt_verify_poses(seg, s.window_start, s.window_end, times, verify_poses_.ptr, st),
              "optical convergence probe launch");
    const int samples = s.scene->view.samples;
    hip_check(launch_adapt_trace(dvsdev::LAYOUT_CELL, segtab_.ptr, pbegin_.ptr, last_nseg_, samples, items,
                                 verify_pix_.ptr, verify_pose_.ptr, verify_poses_.ptr, last_lds_, verify_rad_.ptr,
                                 status_.ptr, counter_.ptr, st, nullptr, nullptr, 256, wave_),
              "optical convergence probe launch");
    AdaptBufs b{};
    b.flags = flags_.ptr;
    b.kcount = kcount_.ptr;
    b.koff = koff_.ptr;
    b.knots = knots_.ptr;
    hip_check(launch_adapt_verify_errors(items, positions, times, verify_pix_.ptr, verify_rad_.ptr, b, s.photo,
                                         verify_err_.ptr, st),
              "optical convergence probe launch");
    std::vector<double> values(items);
    int status = 0;
    hip_check(hipMemcpyAsync(values.data(), verify_err_.ptr, sizeof(double) * items, hipMemcpyDeviceToHost, st),
              "download optical convergence probes");
    hip_check(hipMemcpyAsync(&status, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "probe status");
    hip_check(hipStreamSynchronize(st), "download optical convergence probes");
    if (status) fail(DVS_E_OPTICAL, "invalid ray in optical convergence probe");
    dvs_optical_probe r{};
    r.probes = items;
    for (double v : values) {
      if (v < 0) { ++r.exact_fallback_probes; continue; }
      r.maximum_log_error = std::max(r.maximum_log_error, v);
      r.mean_log_error += v;
      if (v > tolerance) ++r.over_tolerance;
    }
    r.mean_log_error /= items - r.exact_fallback_probes > 0 ? items - r.exact_fallback_probes : 1;
    std::sort(values.begin(), values.end());
    r.p99_log_error = values[(size_t)(items * .99)];
    return r;
  }

  // EventSensor::verify_receptor (event_sensor.cu:357-374): with numerical
  // diagnostics on, the receptor check's measurement of the last adaptive
  // window.
  dvs_receptor_probe verify_receptor(int id, const dvs_optical_interval *ivs, int n) {
    Segment &s = segment(id);
    if (!s.diagnostics) fail(DVS_E_STATE, "enable numerical diagnostics before processing");
    if (s.diag_step != s.window_step || s.window_step == ~0ull) fail(DVS_E_STATE, "no optical knots for receptor diagnostic");
    if (!ivs || n != s.window_ivs || ivs[0].t0 != s.window_start || ivs[n - 1].t1 != s.window_end)
      fail(DVS_E_ARGUMENT, "intervals differ from the last adaptive window");
    return s.diag_probe;
  }

  bool step_adaptive_once(int id, double tolerance, bool check) {
    Segment &s = segment(id);
    if (s.queue.empty()) fail(DVS_E_ARGUMENT, "empty adaptive optical horizon");
    StepSegment w;
    w.seg = &s;
    w.ivs.assign(s.queue.begin(), s.queue.end());
    w.src.assign(s.source.begin(), s.source.end());
    w.tolerance = tolerance;
    w.check = check;
    check_adaptive_window(w);
    int cells = 0;
    for (auto &iv : w.ivs) cells += (int)std::ceil((iv.t1 - iv.t0) / s.photo.integration_step);
    std::vector<StepSegment> work;
    work.push_back(std::move(w));
    std::vector<char> accepted;
    run_adaptive(work, s.npix, cells, (int)work[0].ivs.size(), accepted);
    if (accepted[0]) {
      s.clear_queue();
      s.horizons_done += (int64_t)work[0].ivs.size();
    }
    return accepted[0] != 0;
  }

 private:
  // EventSensor::process_adaptive argument checks (event_sensor.cu:308-312)
  // and the 1e6-cell horizon bound of upload_intervals (:276-282).
  void check_adaptive_window(const StepSegment &w) {
    const Segment &s = *w.seg;
    const double start = w.ivs.front().t0, end = w.ivs.back().t1;
    if (!std::isfinite(w.tolerance) || w.tolerance <= 0 || s.params.refractory_time <= 0 ||
        (end - start) / s.params.refractory_time + 2 > s.max_events)
      fail(DVS_E_ARGUMENT, "invalid adaptive tolerance or event capacity bound");
    double cells = 0;
    for (auto &iv : w.ivs) cells += (iv.t1 - iv.t0) / s.photo.integration_step;
    if (cells > 1000000) fail(DVS_E_ARGUMENT, "optical horizon too long; reduce batch size");
  }

  // Cell tables of a step: every segment's intervals and circuit cells
  // (generate_pixels' cell split, event_sensor.cu:37-40), pixel and circuit
  // block offsets.  Returns the trace kernels' LDS bytes.
  // `staged_done`: the caller knows the previous step's staged copies have
  // run (a host read queued after them has returned), so the staging memory
  // is free without waiting for the stream.
  size_t prepare(std::vector<StepSegment> &work, int pixel_total, int cell_total, int iv_total,
                 bool staged_done = false) {
    hipStream_t st = stream_;
    ahead_.reset();  // a step overwrites the tables a queued setup left
    if (!staged_done) stream_wait(st, "step start");  // the previous step's staged copies are done
    staging_.reset();
    std::vector<dvs_optical_interval> ivs;
    std::vector<double> t0s, t1s;
    std::vector<int> civ;
    ivs.reserve(iv_total);
    t0s.reserve(cell_total);
    t1s.reserve(cell_total);
    civ.reserve(cell_total);
    for (auto &w : work) {
      for (size_t k = 0; k < w.ivs.size(); ++k) {
        const dvs_optical_interval &iv = w.ivs[k];
        const int steps = (int)std::ceil((iv.t1 - iv.t0) / w.seg->photo.integration_step);
        const double dt = (iv.t1 - iv.t0) / steps;
        for (int c = 0; c < steps; ++c) {
          t0s.push_back(iv.t0 + c * dt);
          t1s.push_back(c + 1 == steps ? iv.t1 : iv.t0 + (c + 1) * dt);
          civ.push_back(w.iv_offset + (int)k);
        }
        ivs.push_back(iv);
      }
      w.cells = (int)(t0s.size()) - w.cell_offset;
    }
    intervals_.upload(ivs.data(), ivs.size(), "trajectory interval upload", st, staging_);
    cell_t0_.upload(t0s.data(), t0s.size(), "cell times", st, staging_);
    cell_t1_.upload(t1s.data(), t1s.size(), "cell times", st, staging_);
    cell_iv_.upload(civ.data(), civ.size(), "cell intervals", st, staging_);
    poses_.reserve(cell_total, "cell poses");
    counts_.reserve(pixel_total, "event counts");
    ends_.reserve(pixel_total, "event offsets");
    status_.reserve(1, "error flag");
    counter_.reserve(1, "ray counter");
    // the status word and the ray counter start at zero: staged with the tables
    static const int kZeros[2] = {0, 0};
    staging_.upload(status_.ptr, kZeros, 1, "clear status", st);
    staging_.upload(reinterpret_cast<int *>(counter_.ptr), kZeros, 2, "clear ray counter", st);
    const int nseg = (int)work.size();
    std::vector<int> cblocks(nseg + 1), pbegin(nseg + 1);
    cblocks[0] = 0;
    pbegin[0] = 0;
    size_t lds = 0;
    wave_ = tuning_.traversal != 0;
    for (int i = 0; i < nseg; ++i) {
      cblocks[i + 1] = cblocks[i] + (work[i].seg->npix + 255) / 256;
      pbegin[i + 1] = pbegin[i] + work[i].seg->npix;
      const int depth = std::max(3, work[i].seg->scene->view.stack_depth);
      lds = std::max(lds, (size_t)depth * 256 * sizeof(int));
      wave_ = wave_ && work[i].seg->scene->view.packet != 0;
    }
    cblock_.upload(cblocks.data(), cblocks.size(), "segment blocks", st, staging_);
    pbegin_.upload(pbegin.data(), pbegin.size(), "segment pixels", st, staging_);
    circuit_blocks_ = cblocks[nseg];
    // whole-cell receptor maps of every cell, for each segment's charge scale
    std::vector<int> cell_begin(nseg + 1);
    std::vector<double> charge(nseg);
    for (int i = 0; i < nseg; ++i) {
      cell_begin[i] = work[i].cell_offset;
      charge[i] = work[i].seg->params.charge_scale;
    }
    cell_begin[nseg] = cell_total;
    cellbeg_.upload(cell_begin.data(), cell_begin.size(), "segment cells", st, staging_);
    charge_.upload(charge.data(), charge.size(), "segment charge", st, staging_);
    staging_.flush(st);
    // intervals fed as trajectory times: their states evaluated on the device
    // over the staged copies, which carry only the times
    std::vector<double> times;
    for (auto &w : work) {
      for (size_t k = 0; k < w.src.size();) {
        size_t e = k + 1;
        while (e < w.src.size() && w.src[e] == w.src[k]) ++e;
        if (w.src[k]) {
          times.assign(1, w.ivs[k].t0);
          for (size_t j = k; j < e; ++j) times.push_back(w.ivs[j].t1);
          trajectory_intervals_on(w.src[k], times.data(), (int64_t)times.size(), intervals_.ptr + w.iv_offset + k, st);
        }
        k = e;
      }
    }
    maps_.reserve(cell_total, "cell maps");
    hip_check(launch_cell_maps(cell_total, cell_t0_.ptr, cell_t1_.ptr, cellbeg_.ptr, nseg, charge_.ptr, maps_.ptr, st),
              "cell maps");
    return lds;
  }

  // Device table entry of segment i over its cells [c0, c1).
  SegDev table_entry(const StepSegment &w, int c0, int c1, double *rad) {
    const Segment &s = *w.seg;
    SegDev d{};
    d.scene = s.scene->view;
    d.ivs = intervals_.ptr + w.iv_offset;
    d.iv_count = (int)w.ivs.size();
    d.t0 = cell_t0_.ptr + w.cell_offset + c0;
    d.t1 = cell_t1_.ptr + w.cell_offset + c0;
    d.maps = maps_.ptr + w.cell_offset + c0;
    d.poses = poses_.ptr + w.cell_offset + c0;
    d.cells = c1 - c0;
    d.rad = rad;
    d.states = s.states.ptr;
    d.slots = s.slots.ptr;
    d.counts = counts_.ptr + w.pixel_offset;
    d.params = s.params;
    d.photo = s.photo;
    d.seed = s.seed;
    d.npix = s.npix;
    d.max_events = s.max_events;
    d.pixel_offset = w.pixel_offset;
    d.pad = 0;
    d.pixel_list = nullptr;
    return d;
  }

  void run(std::vector<StepSegment> &work, int pixel_total, int cell_total, int iv_total) {
    hipStream_t st = stream_;
    ++step_serial_;
    const size_t lds = prepare(work, pixel_total, cell_total, iv_total);
    // Cells are traced in chunks so the [cell][pixel] radiance scratch stays
    // within budget: chunk k covers cells [k*C, (k+1)*C) of every segment.
    int max_cells = 0;
    for (auto &w : work) max_cells = std::max(max_cells, w.cells);
    // Two bounds: radiance scratch bytes, and trace blocks per launch
    // (chunk * sum(npix / pixels_per_block) + one partial block per segment).
    double block_rate = 0.0;
    for (auto &w : work) {
      const int ss = w.seg->scene->view.samples;
      block_rate += (double)w.seg->npix / (256 / (ss * ss));
    }
    const double block_cap = (double)(dvsdev::kMaxBlocks256 - (long long)work.size()) / std::max(block_rate, 1e-30);
    const int chunk = (int)std::max<double>(
        1.0, std::min<double>({(double)max_cells, (double)(radiance_budget_ / (sizeof(double) * pixel_total)), block_cap,
                               65535.0}));
    const int chunks = (max_cells + chunk - 1) / chunk;
    radiance_.reserve((size_t)chunk * pixel_total, "cell radiance");
    // Segment tables of every chunk (device), then the launches.
    const int nseg = (int)work.size();
    std::vector<SegDev> tables((size_t)chunks * nseg);
    std::vector<long long> tblocks((size_t)chunks * (nseg + 1));
    std::vector<long long> shape{nseg, chunks, chunk};
    for (int k = 0; k < chunks; ++k) {
      long long *tb = &tblocks[(size_t)k * (nseg + 1)];
      tb[0] = 0;
      size_t rad_offset = 0;
      for (int i = 0; i < nseg; ++i) {
        StepSegment &w = work[i];
        const int c0 = std::min(w.cells, k * chunk), c1 = std::min(w.cells, (k + 1) * chunk);
        tables[(size_t)k * nseg + i] = table_entry(w, c0, c1, radiance_.ptr + rad_offset);
        rad_offset += (size_t)w.seg->npix * chunk;
        const int slots = 256 / (w.seg->scene->view.samples * w.seg->scene->view.samples);
        tb[i + 1] = tb[i] + ((long long)w.seg->npix * (c1 - c0) + slots - 1) / slots;
        shape.push_back(c1 - c0);
        shape.push_back(w.seg->npix);
      }
    }
    segtab_.upload(tables.data(), tables.size(), "segment table", st, staging_);
    tblock_.upload(tblocks.data(), tblocks.size(), "segment blocks", st, staging_);
    staging_.flush(st);
    const size_t rare_native = cell_rates_flags(pixel_total, chunk);
    rare_native_.reserve(rare_native, "photocurrent flags");
    // A captured graph replays only while every buffer it reads stays put.
    for (const void *ptr : {(const void *)segtab_.ptr, (const void *)tblock_.ptr, (const void *)cblock_.ptr,
                            (const void *)radiance_.ptr, (const void *)counts_.ptr, (const void *)poses_.ptr,
                            (const void *)intervals_.ptr, (const void *)cell_t0_.ptr, (const void *)cell_t1_.ptr,
                            (const void *)cell_iv_.ptr, (const void *)status_.ptr, (const void *)counter_.ptr,
                            (const void *)rare_native_.ptr})
      shape.push_back((long long)(uintptr_t)ptr);
    shape.push_back((long long)lds);
    shape.push_back(wave_ ? 1 : 0);
    shape.push_back(pixel_total);
    shape.push_back(cell_total);
    shape.push_back(iv_total);

    // Timers bracket every trace and circuit launch (outside graph replay,
    // where only the whole step is timed).
    Timer step_t;
    std::vector<std::unique_ptr<Timer>> trace_t, circuit_t;
    const bool graph = use_graph_;
    const int cblocks_total = circuit_blocks_;
    hip_check(hipEventRecord(step_t.start, st), "record");
    auto enqueue = [&]() {
      hip_check(launch_fill_int(counts_.ptr, 0, pixel_total, st), "clear events");
      hip_check(launch_cell_poses(cell_total, cell_t0_.ptr, cell_t1_.ptr, cell_iv_.ptr, intervals_.ptr, poses_.ptr, st),
                "cell poses");
      for (int k = 0; k < chunks; ++k) {
        const long long blocks = tblocks[(size_t)k * (nseg + 1) + nseg];
        if (!graph) {
          trace_t.emplace_back(new Timer());
          circuit_t.emplace_back(new Timer());
          hip_check(hipEventRecord(trace_t.back()->start, st), "record");
        }
        hip_check(launch_trace_cells_multi(segtab_.ptr + (size_t)k * nseg, tblock_.ptr + (size_t)k * (nseg + 1), nseg,
                                           blocks, lds, status_.ptr, counter_.ptr, st, wave_),
                  "trace and integrate circuits");
        if (!graph) {
          hip_check(hipEventRecord(trace_t.back()->stop, st), "record");
          hip_check(hipEventRecord(circuit_t.back()->start, st), "record");
        }
        hip_check(launch_cell_rates_multi(segtab_.ptr + (size_t)k * nseg, pbegin_.ptr, nseg, pixel_total, chunk,
                                          status_.ptr, rare_native_.ptr, rare_native, st),
                  "trace and integrate circuits");
        hip_check(launch_circuit_multi(segtab_.ptr + (size_t)k * nseg, cblock_.ptr, nseg, cblocks_total, status_.ptr, st),
                  "trace and integrate circuits");
        if (!graph) hip_check(hipEventRecord(circuit_t.back()->stop, st), "record");
      }
    };
    if (graph) {
      GraphKey key{shape};
      if (!graph_exec_ || !(key == graph_key_)) {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
        enqueue();
        hip_check(hipStreamEndCapture(st, &g), "graph capture");
        hip_check(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        graph_key_ = key;
      }
      hip_check(hipGraphLaunch(graph_exec_, st), "graph launch");
    } else {
      enqueue();
    }
    hip_check(hipEventRecord(step_t.stop, st), "record");
    const std::vector<char> all(nseg, 1);
    const double events_ms = collect(work, pixel_total, lds, true, all);
    const double step_ms = step_t.ms();
    double trace_ms = 0, circuit_ms = 0;
    for (auto &t : trace_t) trace_ms += t->ms();
    for (auto &t : circuit_t) circuit_ms += t->ms();
    if (graph) trace_ms = step_ms;  // graph replay: trace + circuit of the step
    stats.trace_ms += trace_ms;
    stats.circuit_ms += circuit_ms;
    stats.total_ms += step_ms + events_ms;
    stats.trace_launches += chunks;
    stats.circuit_launches += chunks;
    for (auto &w : work) {
      const int ss = w.seg->scene->view.samples;
      stats.primary_rays += (uint64_t)w.seg->npix * w.cells * ss * ss;
      stats.native_equivalent_rays += (uint64_t)w.seg->npix * w.cells * ss * ss;
    }
  }

  // Grows a FIFO's ring to hold `need` entries, moving the queue to the
  // front of the new ring.
  void ensure_ring(DeviceFifo &f, long long need) {
    if (f.ring_size >= need) return;
    long long size = std::max<long long>(1024, f.ring_size);
    while (size < need) size *= 2;
    DeviceArray<QueuedEvent> bigger;
    bigger.reserve((size_t)size, "readout queue");
    if (f.count > 0) {
      const long long slot = f.head & (f.ring_size - 1);
      const long long first = std::min(f.count, f.ring_size - slot);
      hip_check(hipMemcpyAsync(bigger.ptr, f.ring.ptr + slot, sizeof(QueuedEvent) * first, hipMemcpyDeviceToDevice,
                               stream_),
                "readout queue");
      if (first < f.count)
        hip_check(hipMemcpyAsync(bigger.ptr + first, f.ring.ptr, sizeof(QueuedEvent) * (f.count - first),
                                 hipMemcpyDeviceToDevice, stream_),
                  "readout queue");
      hip_check(hipStreamSynchronize(stream_), "readout queue");
    }
    std::swap(f.ring.ptr, bigger.ptr);
    std::swap(f.ring.count, bigger.count);
    f.ring_size = size;
    f.head = 0;
  }

  // The readout tasks of a step: every accepted segment with a bandwidth.
  // Its delivered events go to ready_out_ at out_offset: the queue holds
  // only events whose ready time is after the previous call's `now` (every
  // earlier one was delivered then), and ready times are at least 1 /
  // bandwidth apart, so at most (now - last now) x bandwidth + 1 of them
  // are delivered (the room also never exceeds the queue plus the step's
  // events).
  int build_fifo_tasks(std::vector<StepSegment> &work, const std::vector<char> &accepted, int total) {
    host_fifo_tasks_.reserve(work.size(), "readout tasks");
    int ntask = 0;
    long long room_total = 0;
    for (size_t i = 0; i < work.size(); ++i) {
      Segment &s = *work[i].seg;
      if (!accepted[i] || !s.fifo.queued()) continue;
      DeviceFifo &f = s.fifo;
      const double now = work[i].ivs.back().t1;
      const long long need = f.count + (long long)total;
      ensure_ring(f, need);
      long long room = need;
      if (std::isfinite(f.last_now)) {
        const double span = (now - f.last_now) * f.bandwidth;
        if (span >= 0 && span < 1e15) room = std::min<long long>(room, (long long)(span * (1 + 1e-9)) + 64);
      }
      FifoTask &t = host_fifo_tasks_.ptr[ntask++];
      t.ring = f.ring.ptr;
      t.mask = f.ring_size - 1;
      t.head = f.head;
      t.count = f.count;
      t.available = f.available;
      t.step = f.step;
      t.latency = f.latency;
      t.now = now;
      t.capacity = f.capacity;
      t.dropped = f.dropped;
      t.peak = f.peak;
      t.pixel_begin = work[i].pixel_offset;
      t.pixel_end = work[i].pixel_offset + s.npix;
      t.out_offset = room_total;
      t.out_room = room;
      t.force_serial = force_serial_readout_ ? 1 : 0;
      t.pad = 0;
      room_total += room;
    }
    if (ntask) {
      fifo_tasks_.reserve_ahead(ntask, "readout tasks");
      host_fifo_results_.reserve(ntask, "readout results");
      fifo_results_.reserve_ahead(ntask, "readout results");
      const size_t room = (size_t)std::max<long long>(room_total, 1);
      host_ready_.reserve(room, "delivered events");
      ready_out_.reserve_ahead(room, "delivered events");
      fifo_room_ = room;
      fifo_accepted_.reserve_ahead((size_t)std::max(total, 1), "readout scratch");
      fifo_done_.reserve_ahead((size_t)std::max(total, 1), "readout scratch");
      fifo_serial_.reserve_ahead(ntask, "readout scratch");
    }
    return ntask;
  }

 public:
  // EventReadout::drain (event_readout.cpp:31-34): every queued event.
  void drain(Segment &s) {
    DeviceFifo &f = s.fifo;
    if (!f.queued() || f.count == 0) return;
    std::vector<QueuedEvent> queued((size_t)f.count);
    const long long slot = f.head & (f.ring_size - 1), first = std::min(f.count, f.ring_size - slot);
    hip_check(hipMemcpyAsync(queued.data(), f.ring.ptr + slot, sizeof(QueuedEvent) * first, hipMemcpyDeviceToHost,
                             stream_),
              "drain readout");
    if (first < f.count)
      hip_check(hipMemcpyAsync(queued.data() + first, f.ring.ptr, sizeof(QueuedEvent) * (f.count - first),
                               hipMemcpyDeviceToHost, stream_),
                "drain readout");
    hip_check(hipStreamSynchronize(stream_), "drain readout");
    s.ready.reserve(s.ready.size() + queued.size());
    for (const QueuedEvent &q : queued) s.ready.push_back(unqueue(q));
    f.head += f.count;
    f.count = 0;
  }

 private:
  // Compaction by a device prefix sum over every segment's pixel counts,
  // event-time depth (native steps), device sort by (segment, t, y, x) and
  // each accepted segment's readout FIFO.  Returns the events time (ms).
  // `before_wait`: queues more work on the stream once the step's last
  // copies are queued, before the host waits for them (the next window's
  // setup); the collect's own host read has returned by then.
  double collect(std::vector<StepSegment> &work, int pixel_total, size_t lds, bool depth,
                 const std::vector<char> &accepted, const std::function<void()> &before_wait = nullptr) {
    hipStream_t st = stream_;
    const int nseg = (int)work.size();
    Timer events_t;
    hip_check(hipEventRecord(events_t.start, st), "record");
    size_t temp_bytes = 0;
    hip_check(inclusive_scan_int(nullptr, temp_bytes, counts_.ptr, ends_.ptr, pixel_total, st), "scan size");
    temp_.reserve_ahead(temp_bytes, "scan scratch");
    hip_check(inclusive_scan_int(temp_.ptr, temp_bytes, counts_.ptr, ends_.ptr, pixel_total, st), "event scan");
    int status = 0, total = 0;
    unsigned long long shadows = 0;
    host_words_.reserve(6, "status");
    host_words_.ptr[0] = host_words_.ptr[1] = -1;  // the status and the count are never negative
    hip_check(hipMemcpyAsync(host_words_.ptr, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
    hip_check(hipMemcpyAsync(host_words_.ptr + 1, ends_.ptr + pixel_total - 1, sizeof(int), hipMemcpyDeviceToHost, st),
              "copy count");
    wait_words(host_words_.ptr, 2, -1, st, "copy count");
    status = host_words_.ptr[0];
    total = host_words_.ptr[1];
    if (status & 1) fail(DVS_E_CAPACITY, "pixel event capacity exceeded; reduce numerical window");
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
    // read-backs land in page-locked memory: events, their segments (with
    // more than one segment), the status word and the shadow ray count
    // Event buffers start with room for one event per pixel: the count
    // grows over the first steps after initialisation (the sensor starts at
    // equilibrium), and every growth is a reallocation.
    const size_t event_room = std::max<size_t>((size_t)total, (size_t)pixel_total);
    host_events_.reserve(event_room, "events");
    if (nseg > 1) host_event_seg_.reserve(event_room, "event segments");
    const int samples = work[0].seg->scene->view.samples;
    // Every event time of the step lies in the windows' span [tmin, tmax], so
    // the order-preserving time keys share the bits above the highest bit in
    // which key(tmin) and key(tmax) differ: only the bits below it are sorted.
    // The compaction flags a key outside the span (never expected), and then
    // the sort is repeated on all 64 bits before the readout runs.
    double tmin = std::numeric_limits<double>::infinity(), tmax = -tmin;
    for (auto &w : work) {
      tmin = std::min(tmin, w.ivs.front().t0);
      tmax = std::max(tmax, w.ivs.back().t1);
    }
    const unsigned long long kmin = time_key(tmin), kmax = time_key(tmax);
    const int time_bits = kmin == kmax ? 1 : 64 - __builtin_clzll(kmin ^ kmax);
    int seg_bits = 1;
    while ((1 << seg_bits) < nseg) ++seg_bits;
    keyflag_.reserve(1, "event key flag");
    // Sort by time (stably, so ties keep the compaction's (segment, y, x)
    // order), then stably by segment: (segment, t, y, x); permute the
    // records and copy them (and their segments) to the host.  Over the
    // window's span the time sort takes two passes: by the top 16 of the
    // key bits that vary (a bucket is a few tens of nanoseconds), then
    // inside each bucket by the whole key; a bucket of more than 1024 events
    // raises the key flag, and the events are then sorted on all bits.
    auto sort_events = [&](int end_bit) {
      size_t bytes = 0, bucketed = 0;
      hip_check(sort_pairs_u64(nullptr, bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, event_room, 0, 64, st),
                "sort size");
      hip_check(sort_pairs_u64_bucketed(nullptr, bucketed, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, event_room,
                                        kmin, 64, kTimeBucketBits, keyflag_.ptr, st),
                "sort size");
      bytes = std::max(bytes, bucketed);
      temp_.reserve_ahead(bytes, "sort scratch");
      if (end_bit < 64 && end_bit > kTimeBucketBits)
        hip_check(sort_pairs_u64_bucketed(temp_.ptr, bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, total,
                                          kmin, end_bit, kTimeBucketBits, keyflag_.ptr, st),
                  "event sort");
      else
        hip_check(sort_pairs_u64(temp_.ptr, bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, total, 0, end_bit,
                                 st),
                  "event sort");
      const unsigned *final_order = order2_.ptr;
      if (nseg > 1) {
        // in buffers of their own: the time keys and the compaction order
        // stay for a sort on all bits after a raised key flag
        hip_check(launch_segment_keys(total, order2_.ptr, ev_seg_.ptr, seg_keys_.ptr, st), "segment keys");
        hip_check(sort_pairs_u64(temp_.ptr, bytes, seg_keys_.ptr, seg_keys2_.ptr, order2_.ptr, seg_order_.ptr, total,
                                 0, seg_bits, st),
                  "event sort");
        final_order = seg_order_.ptr;
      }
      hip_check(launch_permute_multi(segtab_.ptr, total, final_order, ev_time_.ptr, ev_seg_.ptr, ev_pixel_.ptr,
                                     ev_pol_.ptr, ev_depth_.ptr, events_out_.ptr, events_seg_.ptr, st),
                "event records");
      hip_check(hipMemcpyAsync(host_events_.ptr, events_out_.ptr, sizeof(dvs_event) * total, hipMemcpyDeviceToHost,
                               st),
                "copy events");
      if (nseg > 1)
        hip_check(hipMemcpyAsync(host_event_seg_.ptr, events_seg_.ptr, sizeof(int) * total, hipMemcpyDeviceToHost,
                                 st),
                  "copy events");
    };
    if (total > 0) {
      ev_time_.reserve_ahead(event_room, "event times");
      ev_seg_.reserve_ahead(event_room, "event segments");
      ev_pixel_.reserve_ahead(event_room, "event pixels");
      ev_pol_.reserve_ahead(event_room, "event polarity");
      ev_depth_.reserve_ahead(event_room, "event depth");
      ev_poses_.reserve_ahead(event_room, "event poses");
      keys_.reserve_ahead(event_room, "event keys");
      keys2_.reserve_ahead(event_room, "event keys");
      if (nseg > 1) {
        seg_keys_.reserve_ahead(event_room, "event segment keys");
        seg_keys2_.reserve_ahead(event_room, "event segment keys");
        seg_order_.reserve_ahead(event_room, "event order");
      }
      order_.reserve_ahead(event_room, "event order");
      order2_.reserve_ahead(event_room, "event order");
      events_out_.reserve_ahead(event_room, "events");
      events_seg_.reserve_ahead(event_room, "event segments");
      const SegDev *tab = segtab_.ptr;  // chunk 0 table: ivs, scene, counts, slots are chunk independent
      hip_check(launch_fill_int(keyflag_.ptr, 0, 1, st), "clear event key flag");
      hip_check(launch_gather_multi(tab, pbegin_.ptr, nseg, pixel_total, ends_.ptr, kmin, kmax, ev_time_.ptr,
                                    ev_seg_.ptr, ev_pixel_.ptr, ev_pol_.ptr, keys_.ptr, order_.ptr, keyflag_.ptr, st),
                "event compaction");
      if (depth) {
        hip_check(launch_event_poses_multi(tab, total, ev_time_.ptr, ev_seg_.ptr, ev_poses_.ptr, st), "event poses");
        hip_check(launch_event_depth_multi(tab, total, samples, lds, ev_seg_.ptr, ev_pixel_.ptr, ev_poses_.ptr,
                                           ev_depth_.ptr, status_.ptr, st, wave_),
                  "event depth rays");
      } else {
        // The interpolating path supplies no per-event depth (event_sensor.cu:111-113).
        hip_check(launch_fill_int(reinterpret_cast<int *>(ev_depth_.ptr), (int)0xBF800000u, total, st), "event depth");
      }
      sort_events(time_bits);
    }
    // The readout FIFO of every accepted segment with a bandwidth, on the
    // device, over its slice of the sorted events.
    const int ntask = build_fifo_tasks(work, accepted, total);
    auto run_fifo = [&](const int *flag) {
      if (ntask == 0) return;
      hip_check(hipMemcpyAsync(fifo_tasks_.ptr, host_fifo_tasks_.ptr, sizeof(FifoTask) * ntask, hipMemcpyHostToDevice,
                               st),
                "readout tasks");
      hip_check(launch_readout_fifo(fifo_tasks_.ptr, ntask, ends_.ptr, events_out_.ptr, flag, fifo_accepted_.ptr,
                                    fifo_done_.ptr, fifo_serial_.ptr, ready_out_.ptr, fifo_results_.ptr, st),
                "readout FIFO");
      // the deliveries (the whole room: a copy is cheaper than a wait for the
      // counts) and the queue states
      hip_check(hipMemcpyAsync(host_ready_.ptr, ready_out_.ptr, sizeof(dvs_event) * fifo_room_, hipMemcpyDeviceToHost,
                               st),
                "delivered events");
      hip_check(hipMemcpyAsync(host_fifo_results_.ptr, fifo_results_.ptr, sizeof(FifoResult) * ntask,
                               hipMemcpyDeviceToHost, st),
                "readout results");
    };
    run_fifo(total > 0 ? keyflag_.ptr : nullptr);
    hip_check(hipMemcpyAsync(host_words_.ptr + 2, counter_.ptr, sizeof(shadows), hipMemcpyDeviceToHost, st),
              "ray counter");
    host_words_.ptr[4] = 0;
    if (total > 0)
      hip_check(hipMemcpyAsync(host_words_.ptr + 4, keyflag_.ptr, sizeof(int), hipMemcpyDeviceToHost, st),
                "event key flag");
    hip_check(hipEventRecord(events_t.stop, st), "record");
    host_words_.ptr[0] = -1;
    hip_check(hipMemcpyAsync(host_words_.ptr, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
    if (before_wait) before_wait();
    wait_words(host_words_.ptr, 1, -1, st, "copy events");
    // The event records, their segments, the ray counter and the readout's
    // results arrive by other copies than the polled status word; only the
    // stream orders them, so their completion is taken from the event
    // recorded after them.
    hip_check(hipEventSynchronize(events_t.stop), "copy events");
    status = host_words_.ptr[0];
    std::memcpy(&shadows, host_words_.ptr + 2, sizeof(shadows));
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
    if (host_words_.ptr[4]) {  // a time outside the span: sort on all bits, then the readout
      sort_events(64);
      run_fifo(nullptr);
      stream_wait(st, "copy events");
    }
    const dvs_event *sorted = host_events_.ptr;
    const int *sorted_seg = host_event_seg_.ptr;
    stats.shadow_rays += shadows;
    stats.comparator_events += (uint64_t)total;
    if (depth) stats.event_rays += (uint64_t)total * samples * samples;
    // Per-segment raw events and readout results.
    size_t at = 0;
    int task = 0;
    for (int i = 0; i < nseg; ++i) {
      Segment &s = *work[i].seg;
      size_t end = at;
      if (nseg == 1) end = (size_t)total;
      while (end < (size_t)total && sorted_seg[end] == i) ++end;
      if (!accepted[i]) continue;  // a rejected window has no events and changes nothing
      s.raw.assign(sorted + at, sorted + end);
      s.events_total += (int64_t)(end - at);
      const size_t before = s.ready.size();
      if (s.fifo.queued()) {
        const FifoTask &t = host_fifo_tasks_.ptr[task];
        const FifoResult &r = host_fifo_results_.ptr[task];
        ++task;
        if (r.overflow) fail(DVS_E_STATE, "readout delivered more events than its output room");
        stats.readout_tasks += 1;
        stats.readout_serial_tasks += (uint64_t)r.serial;
        s.ready.insert(s.ready.end(), host_ready_.ptr + t.out_offset, host_ready_.ptr + t.out_offset + r.delivered);
        s.fifo.head = r.head;
        s.fifo.count = r.count;
        s.fifo.available = r.available;
        s.fifo.dropped = r.dropped;
        s.fifo.peak = r.peak;
        s.fifo.last_now = t.now;
      } else {
        s.ready.insert(s.ready.end(), s.raw.begin(), s.raw.end());  // no queue (event_readout.cpp:18)
      }
      stats.delivered_events += s.ready.size() - before;
      stats.pixel_cells += (uint64_t)s.npix * work[i].cells;
      s.cells_done += (uint64_t)work[i].cells;
      s.last_time = work[i].ivs.back().t1;
      at = end;
    }
    const double events_ms = events_t.ms();  // after the readout: by now the runtime has seen the stream drain
    stats.events_ms += events_ms;
    return events_ms;
  }

  // Knot and exact-cell capacity of OpticalSampler (optical_sampler.cu:141-150):
  // the largest of 512, 256, 128 whose caches fit the free device memory with
  // 1 GiB to spare, else 128.
  static int adaptive_capacity(int npix) {
    size_t avail = 0, total = 0;
    hip_check(hipMemGetInfo(&avail, &total), "query optical cache memory budget");
    for (int candidate : {512, 256, 128}) {
      const size_t needed = (size_t)npix * candidate * (16 + 32 + sizeof(double) + sizeof(int));
      if (avail > needed + (size_t(1) << 30)) return candidate;
    }
    return 128;
  }

  // Adaptive windows of every segment of `work` (OpticalSampler::refine,
  // prepare_exact_cells, sample_cells and EventSensor::process_adaptive).
  // accepted[i] == 0 when segment i's convergence gate rejected its window;
  // such a window changes no state.
  // The part of an adaptive window that reads no pixel state: cell tables,
  // the window's horizon states, cell poses and fractions, and the five
  // initial probes of every pixel.  Queued on the stream; the returned
  // record carries what the rest of the window needs.
  // `ahead`: queued for the window after the current one, before the current
  // one's last copies have run; cell_base is then the segment's cell count
  // once the current window is integrated.
  std::unique_ptr<AdaptSetup> setup_adaptive(std::vector<StepSegment> &work, int pixel_total, int cell_total,
                                             int iv_total, bool ahead = false, uint64_t cell_base = 0) {
    hipStream_t st = stream_;
    auto su = std::make_unique<AdaptSetup>();
    hip_check(hipEventRecord(su->step_t.start, st), "record");
    ++step_serial_;
    const size_t lds = prepare(work, pixel_total, cell_total, iv_total, ahead);
    last_lds_ = lds;
    last_nseg_ = (int)work.size();
    const int nseg = (int)work.size();
    const int P = pixel_total;
    const int samples = work[0].seg->scene->view.samples;
    su->lds = lds;
    su->tables.resize(nseg);
    su->ad.resize(nseg);
    for (int i = 0; i < nseg; ++i) {
      StepSegment &w = work[i];
      Segment &s = *w.seg;
      if (!s.capacity) s.capacity = adaptive_capacity(s.npix);
      su->tables[i] = table_entry(w, 0, w.cells, nullptr);
      AdaptDev &a = su->ad[i];
      a = AdaptDev{};
      a.start = w.ivs.front().t0;
      a.end = w.ivs.back().t1;
      a.tolerance = w.tolerance;
      a.cell_base = ahead ? cell_base : s.cells_done;
      a.capacity = s.capacity;
      a.accepted = 1;
      a.cell_offset = w.cell_offset;
    }
    std::vector<SegDev> &tables = su->tables;
    std::vector<AdaptDev> &ad = su->ad;
    segtab_.upload(tables.data(), tables.size(), "segment table", st, staging_);
    adapt_.upload(ad.data(), ad.size(), "adaptive windows", st, staging_);
    staging_.flush(st);
    hip_check(launch_cell_poses(cell_total, cell_t0_.ptr, cell_t1_.ptr, cell_iv_.ptr, intervals_.ptr, poses_.ptr, st),
              "cell poses");
    cell_frac_.reserve(cell_total, "cell fractions");
    hip_check(launch_adapt_cell_frac(cell_total, cell_t0_.ptr, cell_t1_.ptr, cellbeg_.ptr, nseg, adapt_.ptr,
                                     cell_frac_.ptr, st),
              "cell fractions");
    hip_check(hipEventRecord(su->refine_t.start, st), "record");
    // Five initial probes per pixel.  The probes alternate between two
    // buffers across windows: a start probe whose pose equals the previous
    // window's end pose takes that probe's radiance, provided the segments
    // and their pixels are laid out as in the previous adaptive step.
    std::vector<long long> layout;
    for (auto &w : work) {
      layout.push_back(w.seg->id);
      layout.push_back(w.seg->npix);
    }
    const int pprev = probe_cur_, pcur = 1 - pprev;
    const bool reuse_layout = layout == probe_layout_;
    probe_poses_.reserve((size_t)nseg * 5, "probe poses");
    end_pose_[pcur].reserve(nseg, "probe poses");
    probe_reuse_.reserve(nseg, "probe poses");
    hip_check(launch_adapt_probe_poses(nseg, segtab_.ptr, adapt_.ptr, probe_poses_.ptr,
                                       reuse_layout ? end_pose_[pprev].ptr : nullptr, end_pose_[pcur].ptr,
                                       probe_reuse_.ptr, st),
              "probe poses");
    probes_[pcur].reserve((size_t)P * 5, "optical probes");
    hip_check(hipEventRecord(su->probe_t.start, st), "record");
    hip_check(launch_adapt_trace(dvsdev::LAYOUT_PROBE, segtab_.ptr, pbegin_.ptr, nseg, samples, (long long)P * 5,
                                 nullptr, nullptr, probe_poses_.ptr, lds, probes_[pcur].ptr, status_.ptr, counter_.ptr,
                                 st, probe_reuse_.ptr, probes_[pprev].ptr, tuning_.trace_threads, wave_),
              "refine pixel optical launch");
    hip_check(hipEventRecord(su->probe_t.stop, st), "record");
    probe_layout_ = std::move(layout);
    probe_cur_ = pcur;
    su->pcur = pcur;
    return su;
  }

  // The setup of the window the caller expects next (expect_adaptive),
  // queued in the collect step of the accepted window before it, once that
  // window's last copies are queued and before the host waits for them: the
  // device goes on from one window to the next while the host finishes the
  // first and turns to the next call.  `end_time` and `cells_done` are the
  // segment's once the current window is integrated.  The next window uses
  // it when everything the setup read is unchanged (ahead_matches); any
  // other step replaces it.
  void setup_ahead(int id, double end_time, uint64_t cells_done) {
    Expected e = std::move(expected_);
    expected_ = Expected{};
    Segment &s = segment(id);
    if (e.id != id || e.times.size() < 2 || s.diagnostics || e.times[0] != end_time) return;
    StepSegment w;
    w.seg = &s;
    for (size_t k = 0; k + 1 < e.times.size(); ++k) {
      dvs_optical_interval iv{};
      iv.t0 = e.times[k];
      iv.t1 = e.times[k + 1];
      w.ivs.push_back(iv);
      w.src.push_back(e.trajectory);
    }
    w.tolerance = e.tolerance;
    try {
      check_adaptive_window(w);
    } catch (...) {
      return;  // the call that runs this window reports its error itself
    }
    int cells = 0;
    for (auto &iv : w.ivs) cells += (int)std::ceil((iv.t1 - iv.t0) / s.photo.integration_step);
    std::vector<StepSegment> work;
    work.push_back(std::move(w));
    std::unique_ptr<AdaptSetup> su = setup_adaptive(work, s.npix, cells, (int)work[0].ivs.size(), true, cells_done);
    su->work = std::move(work);
    su->cell_total = cells;
    su->cells_done = cells_done;
    su->capacity = s.capacity;
    ahead_ = std::move(su);  // after prepare(), which drops any earlier one
  }


  // Whether the queued setup `su` is that of window `work` now: the same
  // segment, intervals, trajectory and tolerance, and the segment's cell
  // count and capacity unchanged since it was queued.
  static bool ahead_matches(const AdaptSetup &su, const std::vector<StepSegment> &work, int cell_total) {
    if (work.size() != 1 || su.work.size() != 1 || su.cell_total != cell_total) return false;
    const StepSegment &a = su.work[0], &b = work[0];
    if (a.seg != b.seg || a.ivs.size() != b.ivs.size() || a.src != b.src || a.tolerance != b.tolerance) return false;
    if (su.cells_done != b.seg->cells_done || su.capacity != b.seg->capacity) return false;
    for (size_t k = 0; k < a.ivs.size(); ++k)
      if (a.ivs[k].t0 != b.ivs[k].t0 || a.ivs[k].t1 != b.ivs[k].t1 || b.src[k] == nullptr) return false;
    return true;
  }

  void run_adaptive(std::vector<StepSegment> &work, int pixel_total, int cell_total, int iv_total,
                    std::vector<char> &accepted) {
    hipStream_t st = stream_;
    std::unique_ptr<AdaptSetup> su;
    if (ahead_ && ahead_matches(*ahead_, work, cell_total)) {
      su = std::move(ahead_);
      work[0].cells = su->work[0].cells;
      stats.windows_set_up_ahead += 1;
    } else {
      su = setup_adaptive(work, pixel_total, cell_total, iv_total);
    }
    ahead_.reset();
    Timer &step_t = su->step_t, &refine_t = su->refine_t, &probe_t = su->probe_t;
    Timer circuit_t;
    const size_t lds = su->lds;
    for (int i = 0; i < (int)work.size(); ++i) {
      Segment &s = *work[i].seg;
      s.window_step = step_serial_;
      s.window_index = i;
      s.window_pixel_offset = work[i].pixel_offset;
      s.window_ivs = (int)work[i].ivs.size();
      s.window_start = work[i].ivs.front().t0;
      s.window_end = work[i].ivs.back().t1;
    }
    const int nseg = (int)work.size();
    const int P = pixel_total;
    const int samples = work[0].seg->scene->view.samples, R = samples * samples, slots = 256 / R;
    accepted.assign(nseg, 1);
    std::vector<AdaptDev> &ad = su->ad;
    for (int i = 0; i < nseg; ++i) ad[i].tolerance = work[i].tolerance;
    const int pcur = su->pcur;
    std::vector<std::unique_ptr<Timer>> query_t;
    for (auto *a : {&bound_, &flags_, &bisects_, &exact_terms_, &kcount_, &koff_, &ecount_, &eoff_, &fb_index_,
                    &fb_scan_})
      a->reserve(P, "optical refinement state");
    AdaptBufs b{};
    b.probes = probes_[pcur].ptr;
    b.bound = bound_.ptr;
    b.flags = flags_.ptr;
    b.bisects = bisects_.ptr;
    b.exact_terms = exact_terms_.ptr;
    b.kcount = kcount_.ptr;
    b.koff = koff_.ptr;
    b.ecount = ecount_.ptr;
    b.eoff = eoff_.ptr;
    b.fb_index = fb_index_.ptr;
    b.cell_frac = cell_frac_.ptr;
    ent_[0].reserve((size_t)2 * P, "refinement segments");
    ctr_.reserve(kRefineCounterWords, "refinement counters");
    // also clears the refinement counters and the event counts
    hip_check(launch_adapt_init(P, b, ent_[0].ptr, ctr_.ptr, kRefineCounterWords, counts_.ptr, st),
              "refine pixel optical launch");
    // Level-synchronous refinement.  A traced bisection also traces the
    // midpoints of its halves' halves, so the next level's bisections need
    // no trace: the levels alternate between one ray launch and none, and
    // every probe value is the one a bisection would trace itself.
    //
    // The host reads the counters after each level but one kind: when every
    // bisection of a level was traced, each entry of the next level carries
    // its halves' midpoints (radiance is never negative), so that level
    // traces nothing.  It is followed at once by the level after it, which
    // takes its entry count from the device; the host checks the claim at
    // the next read.
    long long n = 2LL * P;  // entries of the level, or an upper bound on them when !counted
    bool counted = true, untraced = false;
    long long terms_bound = 0;  // resolved segments written so far, at most
    int cur = 0, terms_total = 0, levels = 0, exact_terms = 0, over_capacity = 0;
    host_small_.reserve(kRefineCounterWords + 2, "refinement counters");
    int *host_ctr = host_small_.ptr;  // page-locked: the copy is queued, then the stream is polled
    while (n > 0) {
      if (++levels > kMaxRefineLevels) fail(DVS_E_STATE, "optical refinement did not terminate");
      int *level_ctr = ctr_.ptr + kRefineHeadWords + 2 * (levels - 1);
      terms_.grow_keep((size_t)(terms_bound + n), (size_t)terms_bound, "resolved segments", st);
      tkeys_.grow_keep((size_t)(terms_bound + n), (size_t)terms_bound, "resolved segment keys", st);
      bis_.reserve(n, "bisected segments");
      qpix_.reserve((size_t)dvsdev::kBisectQueries * n, "optical queries");
      qpose_.reserve((size_t)dvsdev::kBisectQueries * n, "optical queries");
      ent_[1 - cur].reserve(2 * n, "refinement segments");
      hip_check(launch_adapt_decide((int)n, counted ? nullptr : level_ctr - 2, ent_[cur].ptr, segtab_.ptr,
                                    adapt_.ptr, pbegin_.ptr, nseg, b, terms_.ptr, tkeys_.ptr, bis_.ptr, qpix_.ptr,
                                    qpose_.ptr, ent_[1 - cur].ptr, ctr_.ptr, level_ctr, st),
                "refine pixel optical launch");
      terms_bound += n;
      if (untraced && counted) {
        cur = 1 - cur;
        n = 2 * n;
        counted = untraced = false;
        continue;
      }
      const size_t words = kRefineHeadWords + 2 * (size_t)levels;
      for (size_t k = 0; k < words; ++k) host_ctr[k] = -1;  // counters are never negative
      hip_check(hipMemcpyAsync(host_ctr, ctr_.ptr, sizeof(int) * words, hipMemcpyDeviceToHost, st), "refinement count");
      wait_words(host_ctr, words, -1, st, "refinement count");
      terms_total = host_ctr[0];
      terms_bound = terms_total;
      exact_terms = host_ctr[1];
      over_capacity = host_ctr[2];
      // a traced bisection at the level read without a wait: one of the
      // level's midpoint radiances was not a number
      if (!counted && host_ctr[words - 4] != 0) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
      const long long nb = host_ctr[words - 2], nd = host_ctr[words - 1];
      if (nb + nd == 0) break;
      if (nb > 0) {
        const long long queries = (long long)dvsdev::kBisectQueries * nb;
        qrad_.reserve_ahead(queries, "optical queries");
        query_t.emplace_back(new Timer());
        hip_check(hipEventRecord(query_t.back()->start, st), "record");
        hip_check(launch_adapt_trace(dvsdev::LAYOUT_QUERY, segtab_.ptr, pbegin_.ptr, nseg, samples, queries,
                                     qpix_.ptr, nullptr, qpose_.ptr, lds, qrad_.ptr, status_.ptr, counter_.ptr, st,
                                     nullptr, nullptr, tuning_.trace_threads, wave_),
                  "refine pixel optical launch");
        hip_check(hipEventRecord(query_t.back()->stop, st), "record");
        stats.query_samples += (uint64_t)queries;
        hip_check(launch_adapt_children((int)nb, bis_.ptr, qrad_.ptr, ent_[1 - cur].ptr + 2 * nd, st),
                  "refine pixel optical launch");
      }
      untraced = nd == 0;  // every entry of the next level is a traced bisection's half
      counted = true;
      cur = 1 - cur;
      n = 2 * (nb + nd);
    }
    // Without exact segments there are no exact cells, and without them and
    // without a pixel past its capacity no pixel falls back to native
    // tracing: the exact-cell and fallback lists, and their read-backs, are
    // skipped.
    const bool any_exact = exact_terms > 0, any_fallback = any_exact || over_capacity > 0;
    // Knot counts, offsets and depth-first knot order; adapt_counts also
    // clears the knot fills, the statistics and (without exact segments)
    // the exact counts.
    kfill_.reserve(P, "knot order");
    seg_stats_.reserve(nseg, "refinement statistics");
    hip_check(launch_adapt_counts(P, adapt_.ptr, pbegin_.ptr, nseg, b, kfill_.ptr, !any_exact, seg_stats_.ptr, st),
              "pixel optical knots");
    size_t temp_bytes = 0;
    hip_check(inclusive_scan_int(nullptr, temp_bytes, kcount_.ptr, koff_.ptr, P, st), "scan size");
    temp_.reserve_ahead(temp_bytes, "scan scratch");
    hip_check(inclusive_scan_int(temp_.ptr, temp_bytes, kcount_.ptr, koff_.ptr, P, st), "knot offsets");
    // A knot pixel has 1 + 2 (accepted) + (exact) knots: P + 2 terms_total
    // bounds the total without reading it back.
    const size_t knots_bound = (size_t)P + 2 * (size_t)terms_total;
    knots_.reserve_ahead(knots_bound, "pixel optical knots");
    b.knots = knots_.ptr;
    if (terms_total > 0) {
      kbucket_.reserve_ahead(knots_bound, "knot order");
      hip_check(launch_adapt_knots(terms_total, tkeys_.ptr, terms_.ptr, b, kfill_.ptr, kbucket_.ptr, st),
                "pixel optical knots");
    }
    // Local exact cells (identify_exact_cells / sample_exact_queries).
    int exact_total = 0;
    if (any_exact) {
      hip_check(launch_adapt_exact_count(P, segtab_.ptr, adapt_.ptr, pbegin_.ptr, nseg, b, st),
                "identify local exact optical cells");
      temp_bytes = 0;
      hip_check(inclusive_scan_int(nullptr, temp_bytes, ecount_.ptr, eoff_.ptr, P, st), "scan size");
      temp_.reserve_ahead(temp_bytes, "scan scratch");
      hip_check(inclusive_scan_int(temp_.ptr, temp_bytes, ecount_.ptr, eoff_.ptr, P, st), "local exact query count");
      hip_check(hipMemcpyAsync(host_ctr + kRefineCounterWords, eoff_.ptr + P - 1, sizeof(int), hipMemcpyDeviceToHost,
                               st),
                "local exact query count");
      stream_wait(st, "local exact query count");
      exact_total = host_ctr[kRefineCounterWords];
    }
    ecell_.reserve(std::max(exact_total, 1), "local exact optical cells");
    erad_.reserve(std::max(exact_total, 1), "local exact radiance");
    eitem_pix_.reserve(std::max(exact_total, 1), "compact exact ray queue");
    eitem_pose_.reserve(std::max(exact_total, 1), "compact exact ray queue");
    b.ecell = ecell_.ptr;
    b.erad = erad_.ptr;
    if (exact_total > 0) {
      hip_check(launch_adapt_exact_list(P, segtab_.ptr, adapt_.ptr, pbegin_.ptr, nseg, b, eitem_pix_.ptr,
                                        eitem_pose_.ptr, st),
                "identify local exact optical cells");
      hip_check(launch_adapt_trace(dvsdev::LAYOUT_CELL, segtab_.ptr, pbegin_.ptr, nseg, samples, exact_total,
                                   eitem_pix_.ptr, eitem_pose_.ptr, poses_.ptr, lds, erad_.ptr, status_.ptr,
                                   counter_.ptr, st, nullptr, nullptr, tuning_.trace_threads, wave_),
                "trace local exact ray cells");
    }
    // Whole-pixel fallback list (compact_fallback, event_sensor.cu:116-118).
    // fb_index is read for fallback pixels only.
    std::vector<int> fb_ends(nseg, 0);
    if (any_fallback) {
      hip_check(launch_adapt_fallback_flags(P, flags_.ptr, fb_scan_.ptr, st), "compact exact fallback pixels");
      temp_bytes = 0;
      hip_check(inclusive_scan_int(nullptr, temp_bytes, fb_scan_.ptr, fb_scan_.ptr, P, st), "scan size");
      temp_.reserve_ahead(temp_bytes, "scan scratch");
      hip_check(inclusive_scan_int(temp_.ptr, temp_bytes, fb_scan_.ptr, fb_scan_.ptr, P, st),
                "compact exact fallback pixels");
      fb_ends_.reserve(nseg, "fallback counts");
      hip_check(launch_adapt_segment_ends(nseg, pbegin_.ptr, fb_scan_.ptr, fb_ends_.ptr, st), "fallback counts");
      host_fb_ends_.reserve(nseg, "fallback counts");
      hip_check(hipMemcpyAsync(host_fb_ends_.ptr, fb_ends_.ptr, sizeof(int) * nseg, hipMemcpyDeviceToHost, st),
                "fallback counts");
      stream_wait(st, "fallback counts");
      fb_ends.assign(host_fb_ends_.ptr, host_fb_ends_.ptr + nseg);
    }
    const int fb_total = nseg ? fb_ends[nseg - 1] : 0;
    fb_list_.reserve(std::max(fb_total, 1), "exact fallback pixel IDs");
    if (any_fallback)
      hip_check(launch_adapt_fallback_list(P, adapt_.ptr, pbegin_.ptr, nseg, fb_scan_.ptr, b, fb_list_.ptr, st),
                "compact exact fallback pixels");
    // Refinement statistics of every segment.
    hip_check(launch_adapt_segment_stats(nseg, pbegin_.ptr, b, seg_stats_.ptr, st), "refinement statistics");
    host_seg_stats_.reserve(nseg, "refinement statistics");
    const AdaptSegStats *seg_stats = host_seg_stats_.ptr;  // read after the step's final wait
    hip_check(hipMemcpyAsync(host_seg_stats_.ptr, seg_stats_.ptr, sizeof(AdaptSegStats) * nseg, hipMemcpyDeviceToHost, st),
              "refinement statistics");
    host_reuse_.reserve(nseg, "refinement statistics");
    const int *probe_reused = host_reuse_.ptr;  // read after the step's final wait
    hip_check(hipMemcpyAsync(host_reuse_.ptr, probe_reuse_.ptr, sizeof(int) * nseg, hipMemcpyDeviceToHost, st),
              "refinement statistics");
    // Convergence gate (verify_receptor, event_sensor.cu:357-374) of the
    // segments that check this window.
    std::vector<int> gate_seg;
    for (int i = 0; i < nseg; ++i)
      if (work[i].check || work[i].seg->diagnostics) gate_seg.push_back(i);
    const int ngate = (int)gate_seg.size();
    uint64_t gate_samples = 0;
    if (ngate > 0) {
      // Probe pixel i of a segment is (i * 2654435761) mod npix.  They are
      // traced in ascending pixel order, so that the rays in flight together
      // meet nearby parts of the scene; slots[] takes each error back to
      // position i, where the statistics read it.
      std::vector<int> ids((size_t)ngate * 1024), slots_of((size_t)ngate * 1024);
      std::vector<SegDev> gtab(ngate);
      std::vector<long long> gblocks(ngate + 1);
      size_t rad_total = 0;
      for (int g = 0; g < ngate; ++g) {
        const StepSegment &w = work[gate_seg[g]];
        const int np = std::min(1024, w.seg->npix);
        std::vector<std::pair<int, int>> order(np);
        for (int i = 0; i < np; ++i) order[i] = {(int)(((uint64_t)i * 2654435761ULL) % (uint64_t)w.seg->npix), i};
        std::sort(order.begin(), order.end());
        for (int j = 0; j < np; ++j) {
          ids[(size_t)g * 1024 + j] = order[j].first;
          slots_of[(size_t)g * 1024 + j] = order[j].second;
        }
        rad_total += (size_t)np * w.cells;
      }
      gate_ids_.upload(ids.data(), ids.size(), "mean receptor probe IDs", st, staging_);
      gate_slots_.upload(slots_of.data(), slots_of.size(), "mean receptor probe IDs", st, staging_);
      gate_rad_.reserve(std::max<size_t>(rad_total, 1), "mean receptor probe radiance");
      size_t off = 0;
      gblocks[0] = 0;
      for (int g = 0; g < ngate; ++g) {
        const StepSegment &w = work[gate_seg[g]];
        const int np = std::min(1024, w.seg->npix);
        gtab[g] = table_entry(w, 0, w.cells, gate_rad_.ptr + off);
        gtab[g].npix = np;
        gtab[g].pixel_list = gate_ids_.ptr + (size_t)g * 1024;
        off += (size_t)np * w.cells;
        gblocks[g + 1] = gblocks[g] + ((long long)np * w.cells + slots - 1) / slots;
        gate_samples += (uint64_t)np * w.cells;
      }
      gate_tab_.upload(gtab.data(), gtab.size(), "mean receptor probes", st, staging_);
      gate_blocks_.upload(gblocks.data(), gblocks.size(), "mean receptor probes", st, staging_);
      gate_idx_.upload(gate_seg.data(), gate_seg.size(), "mean receptor probes", st, staging_);
      staging_.flush(st);
      hip_check(launch_trace_cells_multi(gate_tab_.ptr, gate_blocks_.ptr, ngate, gblocks[ngate], lds, status_.ptr,
                                         counter_.ptr, st, wave_),
                "mean receptor probe optics");
      gate_err_.reserve((size_t)ngate * 1024, "mean receptor probe errors");
      hip_check(launch_adapt_gate(gate_tab_.ptr, gate_idx_.ptr, gate_slots_.ptr, ngate, adapt_.ptr, b, gate_err_.ptr,
                                  st),
                "mean receptor probe integration");
      std::vector<double> errors((size_t)ngate * 1024);
      hip_check(hipMemcpyAsync(errors.data(), gate_err_.ptr, sizeof(double) * errors.size(), hipMemcpyDeviceToHost, st),
                "download mean receptor convergence");
      stream_wait(st, "download mean receptor convergence");
      bool changed = false;
      for (int g = 0; g < ngate; ++g) {
        StepSegment &w = work[gate_seg[g]];
        Segment &s = *w.seg;
        const int np = std::min(1024, s.npix);
        std::vector<double> e(errors.begin() + (size_t)g * 1024, errors.begin() + (size_t)g * 1024 + np);
        dvs_receptor_probe r{};
        r.pixels = np;
        r.circuit_cells = w.cells;
        for (double v : e) r.mean_log_voltage_error += v;
        std::sort(e.begin(), e.end());
        r.maximum_log_voltage_error = e.back();
        r.p99_log_voltage_error = e[(size_t)(np * .99)];
        r.mean_log_voltage_error /= np;
        if (s.diagnostics) {
          s.diag_probe = r;
          s.diag_step = step_serial_;
        }
        if (!w.check) continue;  // measured for the diagnostics only
        s.last_probe = r;
        const double threshold = std::min(s.params.threshold_pos, s.params.threshold_neg);
        if (r.maximum_log_voltage_error > .02 * threshold || r.p99_log_voltage_error > .005 * threshold) {
          accepted[gate_seg[g]] = 0;
          ad[gate_seg[g]].accepted = 0;
          changed = true;
        }
      }
      if (changed) {
        adapt_.upload(ad.data(), ad.size(), "adaptive windows", st, staging_);
        staging_.flush(st);
      }
    }
    hip_check(hipEventRecord(refine_t.stop, st), "record");
    // Integration, with the fallback pixels' native radiance traced in cell
    // chunks (sample_cells / integrate_fallback_tile).
    hip_check(hipEventRecord(circuit_t.start, st), "record");
    std::vector<std::unique_ptr<Timer>> integrate_t;
    int max_cells = 0;
    for (auto &w : work) max_cells = std::max(max_cells, w.cells);
    // Photocurrent of every pixel for a chunk of cells (knot pixels in
    // rates_ [cell][P], fallback pixels in radiance_ [cell][fallback list]),
    // then the circuit over the chunk, all on one stream.  The chunk is the
    // whole window unless the photocurrent would pass the radiance budget.
    // (Computing the next chunk's photocurrent on a second stream beside the
    // circuit gains nothing: the circuit is slowed by as much as the
    // photocurrent would take alone, and each cross-stream wait costs about
    // ten microseconds.)
    size_t chunk_cap = std::min<size_t>((size_t)max_cells, 65535);
    chunk_cap = std::min<size_t>(chunk_cap, radiance_budget_ / (sizeof(double) * ((size_t)P + (size_t)fb_total)));
    const int chunk = (int)std::max<size_t>(1, chunk_cap);
    rates_.reserve((size_t)chunk * P, "adaptive photocurrent");
    b.rates = rates_.ptr;
    codes_.reserve((size_t)((chunk + dvsdev::kCodeCells - 1) / dvsdev::kCodeCells) * P, "adaptive photocurrent");
    b.codes = codes_.ptr;
    // per-wave flags of the knot pixels' and the fallback pixels' photocurrent
    const size_t rare_knots = adapt_rates_flags(P, chunk, tuning_.rate_cells), rare_fallback = cell_rates_flags(fb_total, chunk);
    rare_adapt_.reserve(rare_knots, "photocurrent flags");
    rare_native_.reserve(rare_fallback, "photocurrent flags");
    uint64_t fallback_samples = 0;
    std::vector<long long> fblocks(nseg + 1);
    std::vector<SegDev> ftab(nseg);
    if (fb_total > 0) {
      radiance_.reserve((size_t)chunk * fb_total, "fallback optical tile");
      std::vector<int> fb_items(nseg + 1, 0);
      for (int i = 0; i < nseg; ++i) fb_items[i + 1] = fb_ends[i];
      fb_items_.upload(fb_items.data(), fb_items.size(), "fallback optical tile", st, staging_);
      size_t off = 0;
      for (int i = 0; i < nseg; ++i) {
        const int begin = i ? fb_ends[i - 1] : 0, count = fb_ends[i] - begin;
        ad[i].fb_begin = begin;
        ad[i].fb_count = count;
        ad[i].fb_rad = radiance_.ptr + off;
        off += (size_t)count * chunk;
        if (accepted[i]) fallback_samples += (uint64_t)count * work[i].cells;
      }
      adapt_.upload(ad.data(), ad.size(), "adaptive windows", st, staging_);
      staging_.flush(st);
    }
    for (int c0 = 0; c0 < max_cells; c0 += chunk) {
      const int c1 = std::min(max_cells, c0 + chunk);
      if (fb_total > 0) {
        fblocks[0] = 0;
        for (int i = 0; i < nseg; ++i) {
          StepSegment &w = work[i];
          const int a0 = std::min(w.cells, c0), a1 = std::min(w.cells, c1);
          ftab[i] = table_entry(w, a0, a1, const_cast<double *>(ad[i].fb_rad));
          ftab[i].npix = accepted[i] ? ad[i].fb_count : 0;
          ftab[i].pixel_list = fb_list_.ptr + ad[i].fb_begin;
          fblocks[i + 1] = fblocks[i] + ((long long)ftab[i].npix * (a1 - a0) + slots - 1) / slots;
        }
        fb_tab_.upload(ftab.data(), ftab.size(), "fallback optical tile", st, staging_);
        fb_blocks_.upload(fblocks.data(), fblocks.size(), "fallback optical tile", st, staging_);
        staging_.flush(st);
        if (fblocks[nseg] > 0) {
          hip_check(launch_trace_cells_multi(fb_tab_.ptr, fb_blocks_.ptr, nseg, fblocks[nseg], lds, status_.ptr,
                                             counter_.ptr, st, wave_),
                    "parallel exact cell optics launch");
          hip_check(launch_cell_rates_multi(fb_tab_.ptr, fb_items_.ptr, nseg, fb_total, c1 - c0, status_.ptr,
                                            rare_native_.ptr, rare_fallback, st),
                    "parallel exact cell optics launch");
        }
      }
      hip_check(launch_adapt_rates(segtab_.ptr, adapt_.ptr, pbegin_.ptr, nseg, P, b, c0, c1, status_.ptr,
                                   rare_adapt_.ptr, rare_knots, st, tuning_.rate_cells),
                "adaptive photocurrent launch");
      integrate_t.emplace_back(new Timer());
      hip_check(hipEventRecord(integrate_t.back()->start, st), "record");
      hip_check(launch_adapt_integrate(segtab_.ptr, adapt_.ptr, cblock_.ptr, nseg, circuit_blocks_, P, b, c0, c1,
                                       status_.ptr, st),
                "adaptive circuit launch");
      hip_check(hipEventRecord(integrate_t.back()->stop, st), "record");
      if (fb_total > 0) stream_wait(st, "exact fallback circuit tile");  // tables are reused
    }
    hip_check(hipEventRecord(circuit_t.stop, st), "record");
    hip_check(hipEventRecord(step_t.stop, st), "record");
    std::function<void()> ahead;
    if (nseg == 1 && accepted[0] && expected_.id == work[0].seg->id) {
      const int id = work[0].seg->id;
      const double end_time = work[0].ivs.back().t1;
      const uint64_t cells_after = work[0].seg->cells_done + (uint64_t)work[0].cells;
      ahead = [this, id, end_time, cells_after] { setup_ahead(id, end_time, cells_after); };
    }
    const double events_ms = collect(work, pixel_total, lds, false, accepted, ahead);
    // Statistics.
    const double refine_ms = refine_t.ms(), photo_circuit_ms = circuit_t.ms(), step_ms = step_t.ms();
    double circuit_ms = 0;
    for (auto &t : integrate_t) circuit_ms += t->ms();
    stats.trace_ms += refine_ms;
    stats.circuit_ms += circuit_ms;
    stats.rates_ms += photo_circuit_ms - circuit_ms;
    stats.probe_ms += probe_t.ms();
    stats.probe_launches += 1;
    for (int i = 0; i < nseg; ++i) stats.probe_samples += (uint64_t)work[i].seg->npix * (probe_reused[i] ? 4 : 5);
    for (auto &t : query_t) stats.query_ms += t->ms();
    stats.query_launches += (double)query_t.size();
    stats.circuit_launches += (double)integrate_t.size();
    stats.total_ms += step_ms + events_ms;
    uint64_t samples_traced = gate_samples + fallback_samples;
    for (int i = 0; i < nseg; ++i) {
      Segment &s = *work[i].seg;
      const AdaptSegStats &g = seg_stats[i];
      samples_traced += g.traced_pixel_samples + (uint64_t)g.native_cell_queries;
      s.last_refine = dvs_refinement_stats{g.traced_pixel_samples, g.fallback_pixels, g.maximum_knots,
                                           g.exact_segments, g.native_cell_queries};
      s.last_accepted = accepted[i] != 0;
      if (accepted[i]) {
        stats.native_equivalent_rays += (uint64_t)s.npix * work[i].cells * R;
        stats.adaptive_windows += 1;
      } else {
        stats.rejected_windows += 1;
      }
    }
    stats.primary_rays += samples_traced * R;
    stats.trace_launches += 1;
  }

  int horizons_;
  bool use_graph_;
  hipStream_t stream_ = nullptr;
  std::map<int, std::unique_ptr<Segment>> segments_;
  int next_id_ = 0;
  size_t radiance_budget_ = size_t(4) << 30;
  DeviceArray<SegDev> segtab_;
  DeviceArray<long long> tblock_;
  DeviceArray<int> cblock_, pbegin_, counts_, ends_, status_, cell_iv_, ev_seg_, ev_pixel_, ev_pol_, events_seg_;
  DeviceArray<dvs_optical_interval> intervals_;
  DeviceArray<double> cell_t0_, cell_t1_, radiance_, ev_time_, charge_;
  DeviceArray<CellMaps> maps_;
  DeviceArray<Pose> poses_, ev_poses_;
  DeviceArray<float> ev_depth_;
  DeviceArray<unsigned long long> keys_, keys2_, seg_keys_, seg_keys2_, counter_;
  DeviceArray<unsigned> order_, order2_, seg_order_;
  DeviceArray<dvs_event> events_out_;
  DeviceArray<unsigned char> temp_;
  hipGraphExec_t graph_exec_ = nullptr;
  GraphKey graph_key_;
  int circuit_blocks_ = 0;
  // adaptive step state
  DeviceArray<AdaptDev> adapt_;
  DeviceArray<int> cellbeg_, bound_, flags_, bisects_, exact_terms_, kcount_, koff_, ecount_, eoff_, fb_index_,
      fb_scan_, qpix_, kfill_, ecell_, eitem_pix_, eitem_pose_, fb_ends_, fb_list_, gate_ids_, gate_slots_, gate_idx_,
      fb_items_;
  DeviceArray<unsigned> kbucket_;  // per-pixel segment words of the knot order
  DeviceArray<double> cell_frac_, qrad_, erad_, gate_rad_, gate_err_, rates_;
  DeviceArray<uint64_t> codes_;
  // the setup of an expected window queued ahead of it, and the expectation
  std::unique_ptr<AdaptSetup> ahead_;
  Expected expected_;
  // probe radiance and end poses of the last two adaptive windows, and the
  // segment layout (id, pixels) of the last one
  DeviceArray<double> probes_[2];
  DeviceArray<Pose> end_pose_[2];
  DeviceArray<int> probe_reuse_;
  PinnedArray<int> host_reuse_;
  // event key flag and the readout FIFO's tasks, results and deliveries
  // (written on the device, copied to page-locked memory)
  DeviceArray<int> keyflag_;
  DeviceArray<FifoTask> fifo_tasks_;
  PinnedArray<FifoTask> host_fifo_tasks_;
  PinnedArray<FifoResult> host_fifo_results_;
  PinnedArray<dvs_event> host_ready_;
  DeviceArray<FifoResult> fifo_results_;
  DeviceArray<dvs_event> ready_out_;
  DeviceArray<int> fifo_accepted_, fifo_serial_;
  DeviceArray<double> fifo_done_;
  size_t fifo_room_ = 0;
  bool force_serial_readout_ = false;
  std::vector<long long> probe_layout_;
  int probe_cur_ = 0;
  DeviceArray<int> rare_native_, rare_adapt_;  // per-wave flags of the photocurrent's second phase
  DeviceArray<Pose> probe_poses_, qpose_;
  DeviceArray<AdaptEntry> ent_[2], bis_;
  DeviceArray<int> ctr_;  // refinement counters (kRefineCounterWords)
  Staging staging_;       // page-locked source of the step's uploads
  DeviceArray<AdaptTerm> terms_;
  DeviceArray<unsigned long long> tkeys_;
  PinnedArray<dvs_event> host_events_;
  PinnedArray<int> host_event_seg_, host_words_, host_small_, host_fb_ends_;
  PinnedArray<AdaptSegStats> host_seg_stats_;
  DeviceArray<double2> knots_;
  // engine steps run so far, and the last adaptive step's trace LDS bytes and segments
  uint64_t step_serial_ = 0;
  size_t last_lds_ = 0;
  // measured on the DAVIS346 request's first windows: 128-lane trace blocks
  // and 128-cell photocurrent runs were 1.5% faster per window than 256 / 32
  dvs_tuning tuning_{128, 128, 1};
  bool wave_ = false;  // this step's traces take the wave-packet traversal
  int last_nseg_ = 0;
  DeviceArray<int> verify_pix_, verify_pose_;
  DeviceArray<Pose> verify_poses_;
  DeviceArray<double> verify_rad_, verify_err_;
  DeviceArray<AdaptSegStats> seg_stats_;
  DeviceArray<SegDev> gate_tab_, fb_tab_;
  DeviceArray<long long> gate_blocks_, fb_blocks_;
 public:
  size_t segment_bytes(int id) {
    Segment &s = segment(id);
    return s.states.bytes() + s.slots.bytes() + s.fifo.ring.bytes() + (size_t)s.npix * sizeof(int) * 2;
  }
  std::map<int, std::unique_ptr<Segment>> &segments() { return segments_; }
};

}  // namespace dvsamd

using namespace dvsamd;

struct dvs_sensor_s {
  std::unique_ptr<Engine> engine;
  int id = 0;
};
struct dvs_batch_s {
  std::unique_ptr<Engine> engine;
};

namespace {
void copy_out(std::vector<dvs_event> &v, dvs_event *out, int64_t capacity, int64_t *count, bool clear) {
  *count = (int64_t)v.size();
  if ((int64_t)v.size() > capacity) fail(DVS_E_ARGUMENT, "event buffer too small");
  if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(dvs_event));
  if (clear) v.clear();
}
}  // namespace

extern "C" {

// ---- EventSensor (one-segment engine) -----------------------------------
int dvs_sensor_create(dvs_scene scene, const dvs_physic *c, int32_t effects, int32_t aer, dvs_sensor *out) {
  return guarded([&] {
    if (!scene || !c || !out) fail(DVS_E_ARGUMENT, "invalid sensor arguments");
    *out = nullptr;
    std::unique_ptr<dvs_sensor_s> s(new dvs_sensor_s());
    s->engine.reset(new Engine(INT_MAX, false));
    s->id = s->engine->add(scene, *c, effects != 0, aer != 0);
    *out = s.release();
  });
}

int dvs_sensor_destroy(dvs_sensor s) {
  delete s;
  return DVS_OK;
}

int dvs_sensor_initialize(dvs_sensor s, const dvs_optical_state *pose, double timestamp) {
  return guarded([&] {
    if (!s || !pose) fail(DVS_E_ARGUMENT, "invalid sensor initialization");
    s->engine->initialize(s->id, *pose, timestamp);
  });
}

int dvs_sensor_process(dvs_sensor s, const dvs_optical_interval *ivs, int32_t n, int64_t *ready_count) {
  return guarded([&] {
    if (!s || !ivs) fail(DVS_E_ARGUMENT, "sensor not initialized or no intervals");
    Segment &seg = s->engine->segment(s->id);
    seg.clear_queue();
    seg.fed_until = seg.last_time;
    s->engine->feed(s->id, ivs, n);
    try {
      s->engine->step(s->id);
    } catch (...) {
      seg.clear_queue();
      seg.fed_until = seg.last_time;
      throw;
    }
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
  });
}

int dvs_sensor_take(dvs_sensor s, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!s || !count) fail(DVS_E_ARGUMENT, "null sensor");
    copy_out(s->engine->segment(s->id).ready, out, capacity, count, true);
  });
}

int dvs_sensor_raw_events(dvs_sensor s, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!s || !count) fail(DVS_E_ARGUMENT, "null sensor");
    copy_out(s->engine->segment(s->id).raw, out, capacity, count, false);
  });
}

struct dvs_readout_s {
  Readout fifo;
  std::vector<dvs_event> ready;
};

int dvs_readout_create(double bandwidth, double latency, int64_t capacity, dvs_readout *out) {
  return guarded([&] {
    if (!out) fail(DVS_E_ARGUMENT, "null readout");
    *out = nullptr;
    if (!std::isfinite(bandwidth) || !std::isfinite(latency) || capacity < 0)
      fail(DVS_E_ARGUMENT, "invalid readout bandwidth, latency or capacity");
    auto r = std::make_unique<dvs_readout_s>();
    r->fifo.bandwidth = bandwidth;
    r->fifo.latency = latency;
    r->fifo.capacity = (size_t)capacity;
    *out = r.release();
  });
}

int dvs_readout_destroy(dvs_readout r) {
  delete r;
  return DVS_OK;
}

int dvs_readout_submit(dvs_readout r, const dvs_event *events, int64_t n, double now, int64_t *ready_count) {
  return guarded([&] {
    if (!r || n < 0 || (n > 0 && !events)) fail(DVS_E_ARGUMENT, "invalid readout submission");
    // stable order by (t, y, x) (event_readout.cpp:8-12), then the FIFO
    std::vector<dvs_event> sorted(events, events + n);
    std::stable_sort(sorted.begin(), sorted.end(), [](const dvs_event &a, const dvs_event &b) {
      if (a.timestamp != b.timestamp) return a.timestamp < b.timestamp;
      if (a.y != b.y) return a.y < b.y;
      return a.x < b.x;
    });
    r->fifo.submit_sorted(sorted.data(), sorted.size(), now, r->ready);
    if (ready_count) *ready_count = (int64_t)r->ready.size();
  });
}

int dvs_readout_drain(dvs_readout r, int64_t *ready_count) {
  return guarded([&] {
    if (!r) fail(DVS_E_ARGUMENT, "null readout");
    r->fifo.drain(r->ready);
    if (ready_count) *ready_count = (int64_t)r->ready.size();
  });
}

int dvs_readout_take(dvs_readout r, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!r || !count || (capacity > 0 && !out)) fail(DVS_E_ARGUMENT, "null readout");
    copy_out(r->ready, out, capacity, count, true);
  });
}

int dvs_readout_stats(dvs_readout r, int64_t *dropped, int64_t *peak_queue) {
  return guarded([&] {
    if (!r) fail(DVS_E_ARGUMENT, "null readout");
    if (dropped) *dropped = (int64_t)r->fifo.dropped;
    if (peak_queue) *peak_queue = (int64_t)r->fifo.peak;
  });
}

int dvs_sensor_drain(dvs_sensor s, int64_t *ready_count) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    Segment &seg = s->engine->segment(s->id);
    s->engine->drain(seg);
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
  });
}

int dvs_sensor_readout_stats(dvs_sensor s, int64_t *dropped, int64_t *peak) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    Segment &seg = s->engine->segment(s->id);
    if (dropped) *dropped = (int64_t)seg.fifo.dropped;
    if (peak) *peak = (int64_t)seg.fifo.peak;
  });
}

int dvs_sensor_states(dvs_sensor s, dvs_pixel_state *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    Segment &seg = s->engine->segment(s->id);
    hip_check(hipMemcpy(out, seg.states.ptr, sizeof(dvs_pixel_state) * seg.npix, hipMemcpyDeviceToHost),
              "download pixel states");
  });
}

int dvs_sensor_stats(dvs_sensor s, dvs_stats *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->engine->stats;
  });
}

int dvs_sensor_reset_stats(dvs_sensor s) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    s->engine->stats = dvs_stats{};
  });
}

int dvs_sensor_process_adaptive(dvs_sensor s, const dvs_optical_interval *ivs, int32_t n, double tolerance,
                                int32_t check_convergence, int32_t *accepted, int64_t *ready_count) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    if (accepted) *accepted = 0;
    s->engine->validate_adaptive(s->id, ivs, n, tolerance);
    Segment &seg = s->engine->segment(s->id);
    seg.clear_queue();
    seg.fed_until = seg.last_time;
    bool ok = false;
    try {
      s->engine->feed(s->id, ivs, n);
      ok = s->engine->step_adaptive_once(s->id, tolerance, check_convergence != 0);
    } catch (...) {
      seg.clear_queue();
      seg.fed_until = seg.last_time;
      throw;
    }
    seg.clear_queue();
    seg.fed_until = seg.last_time;
    if (accepted) *accepted = ok ? 1 : 0;
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
  });
}

int dvs_sensor_process_adaptive_times(dvs_sensor s, dvs_trajectory trajectory, const double *times, int32_t n,
                                      double tolerance, int32_t check_convergence, int32_t *accepted,
                                      int64_t *ready_count) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    if (accepted) *accepted = 0;
    if (!times || n < 2) fail(DVS_E_ARGUMENT, "empty adaptive optical horizon");
    Segment &seg = s->engine->segment(s->id);
    seg.clear_queue();
    seg.fed_until = seg.last_time;
    bool ok = false;
    try {
      s->engine->feed_trajectory(s->id, trajectory, times, n);
      ok = s->engine->step_adaptive_once(s->id, tolerance, check_convergence != 0);
    } catch (...) {
      seg.clear_queue();
      seg.fed_until = seg.last_time;
      s->engine->forget_expected();
      throw;
    }
    seg.clear_queue();
    seg.fed_until = seg.last_time;
    if (accepted) *accepted = ok ? 1 : 0;
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
    // an expected next window was set up in the window's collect step, or not at all
    s->engine->forget_expected();
  });
}

int dvs_sensor_expect_adaptive_times(dvs_sensor s, dvs_trajectory trajectory, const double *times, int32_t n,
                                     double tolerance) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    if (n < 0) fail(DVS_E_ARGUMENT, "invalid expected adaptive window");
    s->engine->expect_adaptive(s->id, trajectory, times, n, tolerance);
  });
}

int dvs_sensor_enable_diagnostics(dvs_sensor s) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    s->engine->segment(s->id).diagnostics = true;
  });
}

int dvs_sensor_verify_receptor(dvs_sensor s, const dvs_optical_interval *ivs, int32_t n, dvs_receptor_probe *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->engine->verify_receptor(s->id, ivs, n);
  });
}

int dvs_sensor_verify_optics(dvs_sensor s, const dvs_optical_interval *ivs, int32_t n, double tolerance,
                             dvs_optical_probe *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->engine->verify_optics(s->id, ivs, n, tolerance);
  });
}

int dvs_sensor_adaptive_capacity(dvs_sensor s, int32_t *capacity) {
  return guarded([&] {
    if (!s || !capacity) fail(DVS_E_ARGUMENT, "null sensor");
    *capacity = s->engine->capacity_of(s->id);
  });
}

int dvs_sensor_refinement_stats(dvs_sensor s, dvs_refinement_stats *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->engine->segment(s->id).last_refine;
  });
}

int dvs_sensor_last_convergence(dvs_sensor s, dvs_receptor_probe *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->engine->segment(s->id).last_probe;
  });
}

// ---- Multi-segment batch engine -------------------------------------------
int dvs_batch_create(int32_t horizons_per_step, int32_t use_graph, dvs_batch *out) {
  return guarded([&] {
    if (!out || horizons_per_step < 1) fail(DVS_E_ARGUMENT, "invalid batch arguments");
    std::unique_ptr<dvs_batch_s> b(new dvs_batch_s());
    b->engine.reset(new Engine(horizons_per_step, use_graph != 0));
    *out = b.release();
  });
}

int dvs_batch_destroy(dvs_batch b) {
  delete b;
  return DVS_OK;
}

int dvs_batch_add(dvs_batch b, dvs_scene scene, const dvs_physic *c, int32_t effects, int32_t aer,
                  const dvs_optical_state *initial, double t0, int32_t *segment) {
  return guarded([&] {
    if (!b || !scene || !c || !initial || !segment) fail(DVS_E_ARGUMENT, "invalid batch segment");
    const int id = b->engine->add(scene, *c, effects != 0, aer != 0);
    try {
      b->engine->initialize(id, *initial, t0);
    } catch (...) {
      b->engine->remove(id);
      throw;
    }
    *segment = id;
  });
}

int dvs_batch_feed(dvs_batch b, int32_t segment, const dvs_optical_interval *ivs, int64_t n) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    b->engine->feed(segment, ivs, n);
  });
}

int dvs_batch_feed_trajectory(dvs_batch b, int32_t segment, dvs_trajectory trajectory, const double *times,
                              int64_t n) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    b->engine->feed_trajectory(segment, trajectory, times, n);
  });
}

int dvs_batch_set_tuning(dvs_batch b, const dvs_tuning *tuning) {
  return guarded([&] {
    if (!b || !tuning) fail(DVS_E_ARGUMENT, "null batch or tuning");
    b->engine->set_tuning(*tuning);
  });
}

int dvs_batch_get_tuning(dvs_batch b, dvs_tuning *tuning) {
  return guarded([&] {
    if (!b || !tuning) fail(DVS_E_ARGUMENT, "null batch or tuning");
    *tuning = b->engine->tuning();
  });
}

int dvs_sensor_set_tuning(dvs_sensor s, const dvs_tuning *tuning) {
  return guarded([&] {
    if (!s || !tuning) fail(DVS_E_ARGUMENT, "null sensor or tuning");
    s->engine->set_tuning(*tuning);
  });
}

int dvs_batch_step(dvs_batch b, int32_t *pending_segments) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    const int pending = b->engine->step();
    if (pending_segments) *pending_segments = pending;
  });
}

int dvs_batch_take(dvs_batch b, int32_t segment, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!b || !count) fail(DVS_E_ARGUMENT, "null batch");
    copy_out(b->engine->segment(segment).ready, out, capacity, count, true);
  });
}

int dvs_batch_drain(dvs_batch b, int32_t segment, int64_t *ready_count) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    Segment &seg = b->engine->segment(segment);
    b->engine->drain(seg);
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
  });
}

int dvs_batch_segment_info(dvs_batch b, int32_t segment, int64_t *queued_horizons, int64_t *done_horizons,
                           int64_t *events, int64_t *device_bytes) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    Segment &seg = b->engine->segment(segment);
    if (queued_horizons) *queued_horizons = (int64_t)seg.queue.size();
    if (done_horizons) *done_horizons = seg.horizons_done;
    if (events) *events = seg.events_total;
    if (device_bytes) *device_bytes = (int64_t)b->engine->segment_bytes(segment);
  });
}

int dvs_batch_remove(dvs_batch b, int32_t segment) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    b->engine->segment(segment);
    b->engine->remove(segment);
  });
}

int dvs_batch_set_adaptive(dvs_batch b, int32_t segment, double tolerance, int32_t horizons, double check_period) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    b->engine->set_adaptive(segment, tolerance, horizons, check_period);
  });
}

int dvs_batch_refinement_stats(dvs_batch b, int32_t segment, dvs_refinement_stats *out) {
  return guarded([&] {
    if (!b || !out) fail(DVS_E_ARGUMENT, "null batch");
    *out = b->engine->segment(segment).last_refine;
  });
}

int dvs_batch_stats(dvs_batch b, dvs_stats *out) {
  return guarded([&] {
    if (!b || !out) fail(DVS_E_ARGUMENT, "null batch");
    *out = b->engine->stats;
  });
}

int dvs_batch_reset_stats(dvs_batch b) {
  return guarded([&] {
    if (!b) fail(DVS_E_ARGUMENT, "null batch");
    b->engine->stats = dvs_stats{};
  });
}

}  // extern "C"
