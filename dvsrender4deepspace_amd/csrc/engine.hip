// Multi-segment event engine: the EventSensor of the reference
// (src/cuda/event_sensor.cu:244-305) generalised to many sensors
// ("segments") that advance together in shared gfx950 launches.
//
// One step takes the next queued host horizons of every segment, builds the
// circuit-cell table of each (generate_pixels, event_sensor.cu:37-40),
// traces every (pixel, cell) of every segment, integrates every pixel's
// circuit, compacts the crossings by a device prefix sum, traces event-time
// depth, sorts the events by (segment, t, y, x) on the device and runs each
// segment's readout FIFO (src/event_readout.cpp:6-30) on the sorted stream.
// The plain dvs_sensor_* API is a one-segment engine whose every
// process_many call is one step over exactly the call's intervals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <vector>

#include "common.h"
#include "dvs_hip.h"
#include "events_sort.h"
#include "kernels.h"
#include "scene.h"

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
// over events already in (t, y, x) order.
struct Readout {
  double bandwidth = 0, latency = 0, available = 0;
  size_t capacity = 1, dropped = 0, peak = 0;
  struct Pending { double ready; dvs_event event; };
  std::deque<Pending> pending;
  void submit_sorted(const dvs_event *events, size_t n, double now, std::vector<dvs_event> &ready) {
    for (size_t k = 0; k < n; ++k) {
      const dvs_event &e = events[k];
      while (!pending.empty() && pending.front().ready <= e.timestamp) {
        ready.push_back(pending.front().event);
        pending.pop_front();
      }
      if (bandwidth <= 0) { ready.push_back(e); continue; }
      if (pending.size() >= capacity) { ++dropped; continue; }
      const double done = std::max(available, (double)e.timestamp) + 1.0 / bandwidth;
      available = done;
      dvs_event d = e;
      d.timestamp = done + latency;
      pending.push_back({done + latency, d});
      peak = std::max(peak, pending.size());
    }
    while (!pending.empty() && pending.front().ready <= now) {
      ready.push_back(pending.front().event);
      pending.pop_front();
    }
  }
  void drain(std::vector<dvs_event> &ready) {
    while (!pending.empty()) {
      ready.push_back(pending.front().event);
      pending.pop_front();
    }
  }
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
  Readout readout;
  std::vector<dvs_event> ready, raw;
  int64_t horizons_done = 0, events_total = 0;
  ~Segment() { scene_release(scene); }
};

// Work of one segment inside one step.
struct StepSegment {
  Segment *seg;
  std::vector<dvs_optical_interval> ivs;
  int cells = 0, cell_offset = 0, iv_offset = 0, pixel_offset = 0;
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
    s->readout.bandwidth = aer ? (double)c.readout_bandwidth : 0.0;
    s->readout.latency = aer ? (double)c.readout_latency : 0.0;
    s->readout.capacity = (size_t)std::max(1, c.readout_queue_capacity);
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
    s.fed_until = previous;
  }

  // One step over every segment with queued intervals; returns the number of
  // segments that still have queued intervals afterwards.
  int step(int only = -1) {
    std::vector<StepSegment> work;
    int pixel_total = 0, cell_total = 0, iv_total = 0;
    for (auto &kv : segments_) {
      Segment &s = *kv.second;
      if ((only >= 0 && s.id != only) || s.queue.empty()) continue;
      StepSegment w;
      w.seg = &s;
      const size_t take = std::min<size_t>(s.queue.size(), (size_t)horizons_);
      w.ivs.assign(s.queue.begin(), s.queue.begin() + take);
      double cells_estimate = 0;
      for (auto &iv : w.ivs) cells_estimate += (iv.t1 - iv.t0) / s.photo.integration_step;
      if (cells_estimate > 1000000) fail(DVS_E_ARGUMENT, "optical horizon too long; reduce batch size");
      w.pixel_offset = pixel_total;
      w.cell_offset = cell_total;
      w.iv_offset = iv_total;
      pixel_total += s.npix;
      iv_total += (int)w.ivs.size();
      for (auto &iv : w.ivs) cell_total += (int)std::ceil((iv.t1 - iv.t0) / s.photo.integration_step);
      work.push_back(std::move(w));
    }
    if (work.empty()) return 0;
    run(work, pixel_total, cell_total, iv_total);
    int pending = 0;
    for (auto &w : work) {
      w.seg->queue.erase(w.seg->queue.begin(), w.seg->queue.begin() + w.ivs.size());
      w.seg->horizons_done += (int64_t)w.ivs.size();
    }
    for (auto &kv : segments_) pending += !kv.second->queue.empty();
    return pending;
  }

  dvs_stats stats{};
  int horizons() const { return horizons_; }

 private:
  void run(std::vector<StepSegment> &work, int pixel_total, int cell_total, int iv_total) {
    hipStream_t st = stream_;
    // Host tables: intervals and circuit cells of every segment
    // (generate_pixels' cell split, event_sensor.cu:37-40).
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
    intervals_.upload(ivs.data(), ivs.size(), "trajectory interval upload", st);
    cell_t0_.upload(t0s.data(), t0s.size(), "cell times", st);
    cell_t1_.upload(t1s.data(), t1s.size(), "cell times", st);
    cell_iv_.upload(civ.data(), civ.size(), "cell intervals", st);
    poses_.reserve(cell_total, "cell poses");
    counts_.reserve(pixel_total, "event counts");
    ends_.reserve(pixel_total, "event offsets");
    status_.reserve(1, "error flag");
    counter_.reserve(1, "ray counter");
    hip_check(hipMemsetAsync(status_.ptr, 0, sizeof(int), st), "clear status");
    hip_check(hipMemsetAsync(counter_.ptr, 0, sizeof(unsigned long long), st), "clear ray counter");
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
        1.0, std::min<double>({(double)max_cells, (double)(radiance_budget_ / (sizeof(double) * pixel_total)), block_cap}));
    const int chunks = (max_cells + chunk - 1) / chunk;
    radiance_.reserve((size_t)chunk * pixel_total, "cell radiance");
    size_t lds = 0;
    for (auto &w : work) {
      const int depth = std::max(3, w.seg->scene->view.stack_depth);
      lds = std::max(lds, (size_t)depth * 256 * sizeof(int));
    }
    // Segment tables of every chunk (device), then the launches.
    const int nseg = (int)work.size();
    std::vector<SegDev> tables((size_t)chunks * nseg);
    std::vector<long long> tblocks((size_t)chunks * (nseg + 1));
    std::vector<int> cblocks(nseg + 1), pbegin(nseg + 1);
    std::vector<long long> shape{nseg, chunks, chunk};
    cblocks[0] = 0;
    pbegin[0] = 0;
    for (int i = 0; i < nseg; ++i) {
      cblocks[i + 1] = cblocks[i] + (work[i].seg->npix + 255) / 256;
      pbegin[i + 1] = pbegin[i] + work[i].seg->npix;
    }
    size_t rad_offset = 0;
    for (int k = 0; k < chunks; ++k) {
      long long *tb = &tblocks[(size_t)k * (nseg + 1)];
      tb[0] = 0;
      rad_offset = 0;
      for (int i = 0; i < nseg; ++i) {
        StepSegment &w = work[i];
        Segment &s = *w.seg;
        const int c0 = std::min(w.cells, k * chunk), c1 = std::min(w.cells, (k + 1) * chunk);
        SegDev &d = tables[(size_t)k * nseg + i];
        d.scene = s.scene->view;
        d.ivs = intervals_.ptr + w.iv_offset;
        d.iv_count = (int)w.ivs.size();
        d.t0 = cell_t0_.ptr + w.cell_offset + c0;
        d.t1 = cell_t1_.ptr + w.cell_offset + c0;
        d.poses = poses_.ptr + w.cell_offset + c0;
        d.cells = c1 - c0;
        d.rad = radiance_.ptr + rad_offset;
        rad_offset += (size_t)s.npix * chunk;
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
        const int slots = 256 / (s.scene->view.samples * s.scene->view.samples);
        tb[i + 1] = tb[i] + ((long long)s.npix * d.cells + slots - 1) / slots;
        shape.push_back(d.cells);
        shape.push_back(s.npix);
      }
    }
    segtab_.upload(tables.data(), tables.size(), "segment table", st);
    tblock_.upload(tblocks.data(), tblocks.size(), "segment blocks", st);
    cblock_.upload(cblocks.data(), cblocks.size(), "segment blocks", st);
    pbegin_.upload(pbegin.data(), pbegin.size(), "segment pixels", st);
    // A captured graph replays only while every buffer it reads stays put.
    for (const void *ptr : {(const void *)segtab_.ptr, (const void *)tblock_.ptr, (const void *)cblock_.ptr,
                            (const void *)radiance_.ptr, (const void *)counts_.ptr, (const void *)poses_.ptr,
                            (const void *)intervals_.ptr, (const void *)cell_t0_.ptr, (const void *)cell_t1_.ptr,
                            (const void *)cell_iv_.ptr, (const void *)status_.ptr, (const void *)counter_.ptr})
      shape.push_back((long long)(uintptr_t)ptr);
    shape.push_back((long long)lds);
    shape.push_back(pixel_total);
    shape.push_back(cell_total);
    shape.push_back(iv_total);

    // Timers bracket every trace and circuit launch (outside graph replay,
    // where only the whole step is timed).
    Timer step_t, events_t;
    std::vector<std::unique_ptr<Timer>> trace_t, circuit_t;
    const bool graph = use_graph_;
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
                                           blocks, lds, status_.ptr, counter_.ptr, st),
                  "trace and integrate circuits");
        if (!graph) {
          hip_check(hipEventRecord(trace_t.back()->stop, st), "record");
          hip_check(hipEventRecord(circuit_t.back()->start, st), "record");
        }
        hip_check(launch_circuit_multi(segtab_.ptr + (size_t)k * nseg, cblock_.ptr, nseg, cblocks[nseg], status_.ptr, st),
                  "trace and integrate circuits");
        if (!graph) hip_check(hipEventRecord(circuit_t.back()->stop, st), "record");
      }
    };
    if (use_graph_) {
      GraphKey key{shape};
      if (!graph_exec_ || !(key == graph_key_)) {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
        enqueue();
        hip_check(hipStreamEndCapture(st, &graph), "graph capture");
        hip_check(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(graph);
        graph_key_ = key;
      }
      hip_check(hipGraphLaunch(graph_exec_, st), "graph launch");
    } else {
      enqueue();
    }
    hip_check(hipEventRecord(step_t.stop, st), "record");
    // Compaction by a device prefix sum over every segment's pixel counts.
    hip_check(hipEventRecord(events_t.start, st), "record");
    size_t temp_bytes = 0;
    hip_check(inclusive_scan_int(nullptr, temp_bytes, counts_.ptr, ends_.ptr, pixel_total, st), "scan size");
    temp_.reserve(temp_bytes, "scan scratch");
    hip_check(inclusive_scan_int(temp_.ptr, temp_bytes, counts_.ptr, ends_.ptr, pixel_total, st), "event scan");
    int status = 0, total = 0;
    unsigned long long shadows = 0;
    hip_check(hipMemcpyAsync(&status, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
    hip_check(hipMemcpyAsync(&total, ends_.ptr + pixel_total - 1, sizeof(int), hipMemcpyDeviceToHost, st), "copy count");
    hip_check(hipStreamSynchronize(st), "copy count");
    if (status & 1) fail(DVS_E_CAPACITY, "pixel event capacity exceeded; reduce numerical window");
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
    std::vector<dvs_event> sorted(total);
    std::vector<int> sorted_seg(total);
    const int samples = work[0].seg->scene->view.samples;
    if (total > 0) {
      ev_time_.reserve(total, "event times");
      ev_seg_.reserve(total, "event segments");
      ev_pixel_.reserve(total, "event pixels");
      ev_pol_.reserve(total, "event polarity");
      ev_depth_.reserve(total, "event depth");
      ev_poses_.reserve(total, "event poses");
      keys_.reserve(total, "event keys");
      keys2_.reserve(total, "event keys");
      order_.reserve(total, "event order");
      order2_.reserve(total, "event order");
      events_out_.reserve(total, "events");
      events_seg_.reserve(total, "event segments");
      const SegDev *tab = segtab_.ptr;  // chunk 0 table: ivs, scene, counts, slots are chunk independent
      hip_check(launch_gather_multi(tab, pbegin_.ptr, nseg, pixel_total, ends_.ptr, ev_time_.ptr, ev_seg_.ptr,
                                    ev_pixel_.ptr, ev_pol_.ptr, keys_.ptr, order_.ptr, st),
                "event compaction");
      hip_check(launch_event_poses_multi(tab, total, ev_time_.ptr, ev_seg_.ptr, ev_poses_.ptr, st), "event poses");
      hip_check(launch_event_depth_multi(tab, total, samples, lds, ev_seg_.ptr, ev_pixel_.ptr, ev_poses_.ptr,
                                         ev_depth_.ptr, status_.ptr, st),
                "event depth rays");
      // Stable sort by time, then stable by segment: (segment, t, y, x),
      // since compaction order is (segment, y, x).
      temp_bytes = 0;
      hip_check(sort_pairs_u64(nullptr, temp_bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, total, 0, 64, st),
                "sort size");
      size_t need = temp_bytes;
      int seg_bits = 1;
      while ((1 << seg_bits) < nseg) ++seg_bits;
      hip_check(sort_pairs_u64(nullptr, temp_bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, total, 0, seg_bits,
                               st),
                "sort size");
      need = std::max(need, temp_bytes);
      temp_.reserve(need, "sort scratch");
      temp_bytes = need;
      hip_check(sort_pairs_u64(temp_.ptr, temp_bytes, keys_.ptr, keys2_.ptr, order_.ptr, order2_.ptr, total, 0, 64, st),
                "event sort");
      const unsigned *by_time = order2_.ptr;
      const unsigned *final_order = by_time;
      if (nseg > 1) {
        hip_check(launch_segment_keys(total, by_time, ev_seg_.ptr, keys_.ptr, st), "segment keys");
        temp_bytes = need;
        hip_check(sort_pairs_u64(temp_.ptr, temp_bytes, keys_.ptr, keys2_.ptr, by_time, order_.ptr, total, 0, seg_bits,
                                 st),
                  "event sort");
        final_order = order_.ptr;
      }
      hip_check(launch_permute_multi(tab, total, final_order, ev_time_.ptr, ev_seg_.ptr, ev_pixel_.ptr, ev_pol_.ptr,
                                     ev_depth_.ptr, events_out_.ptr, events_seg_.ptr, st),
                "event records");
      hip_check(hipMemcpyAsync(sorted.data(), events_out_.ptr, sizeof(dvs_event) * total, hipMemcpyDeviceToHost, st),
                "copy events");
      hip_check(hipMemcpyAsync(sorted_seg.data(), events_seg_.ptr, sizeof(int) * total, hipMemcpyDeviceToHost, st),
                "copy events");
    }
    hip_check(hipEventRecord(events_t.stop, st), "record");
    hip_check(hipMemcpyAsync(&status, status_.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
    hip_check(hipMemcpyAsync(&shadows, counter_.ptr, sizeof(shadows), hipMemcpyDeviceToHost, st), "ray counter");
    hip_check(hipStreamSynchronize(st), "copy events");
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
    const double step_ms = step_t.ms(), events_ms = events_t.ms();
    double trace_ms = 0, circuit_ms = 0;
    for (auto &t : trace_t) trace_ms += t->ms();
    for (auto &t : circuit_t) circuit_ms += t->ms();
    if (graph) trace_ms = step_ms;  // graph replay: trace + circuit of the step
    stats.trace_ms += trace_ms;
    stats.circuit_ms += circuit_ms;
    stats.events_ms += events_ms;
    stats.total_ms += step_ms + events_ms;
    stats.trace_launches += chunks;
    stats.shadow_rays += shadows;
    stats.comparator_events += (uint64_t)total;
    stats.event_rays += (uint64_t)total * samples * samples;
    // Per-segment readout on its (t, y, x)-ordered slice.
    size_t at = 0;
    for (int i = 0; i < nseg; ++i) {
      Segment &s = *work[i].seg;
      size_t end = at;
      while (end < (size_t)total && sorted_seg[end] == i) ++end;
      s.raw.assign(sorted.begin() + at, sorted.begin() + end);
      s.events_total += (int64_t)(end - at);
      const size_t before = s.ready.size();
      s.readout.submit_sorted(s.raw.data(), s.raw.size(), work[i].ivs.back().t1, s.ready);
      stats.delivered_events += s.ready.size() - before;
      stats.primary_rays += (uint64_t)s.npix * work[i].cells * samples * samples;
      stats.pixel_cells += (uint64_t)s.npix * work[i].cells;
      s.last_time = work[i].ivs.back().t1;
      at = end;
    }
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
  DeviceArray<double> cell_t0_, cell_t1_, radiance_, ev_time_;
  DeviceArray<Pose> poses_, ev_poses_;
  DeviceArray<float> ev_depth_;
  DeviceArray<unsigned long long> keys_, keys2_, counter_;
  DeviceArray<unsigned> order_, order2_;
  DeviceArray<dvs_event> events_out_;
  DeviceArray<unsigned char> temp_;
  hipGraphExec_t graph_exec_ = nullptr;
  GraphKey graph_key_;
 public:
  size_t segment_bytes(int id) {
    Segment &s = segment(id);
    return s.states.bytes() + s.slots.bytes() + (size_t)s.npix * sizeof(int) * 2;
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
    seg.queue.clear();
    seg.fed_until = seg.last_time;
    s->engine->feed(s->id, ivs, n);
    try {
      s->engine->step(s->id);
    } catch (...) {
      seg.queue.clear();
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

int dvs_sensor_drain(dvs_sensor s, int64_t *ready_count) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    Segment &seg = s->engine->segment(s->id);
    seg.readout.drain(seg.ready);
    if (ready_count) *ready_count = (int64_t)seg.ready.size();
  });
}

int dvs_sensor_readout_stats(dvs_sensor s, int64_t *dropped, int64_t *peak) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    Segment &seg = s->engine->segment(s->id);
    if (dropped) *dropped = (int64_t)seg.readout.dropped;
    if (peak) *peak = (int64_t)seg.readout.peak;
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
    seg.readout.drain(seg.ready);
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
