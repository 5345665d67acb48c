// C ABI of the gfx950 DVS library: scene upload, frame render, event sensor.
// Host-side orchestration only; every per-ray and per-pixel computation runs
// in the kernels of dvs_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_host.h"
#include "dvs_hip.h"
#include "events_sort.h"
#include "kernels.h"

namespace dvsamd {
namespace {

thread_local std::string g_last_error;

struct Failure {
  int code;
  std::string message;
};
[[noreturn]] void fail(int code, std::string message) { throw Failure{code, std::move(message)}; }
void hip_check(hipError_t e, const char *where) {
  // Message format of CUDA_CHECK (include/dvs/math_types.hpp:6-14).
  if (e != hipSuccess) fail(DVS_E_HIP, std::string("[HIP Error] in ") + where + " : " + hipGetErrorString(e));
}
template <class F> int guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return DVS_OK;
  } catch (const Failure &x) {
    g_last_error = x.message;
    return x.code;
  } catch (const std::invalid_argument &x) {
    g_last_error = x.what();
    return DVS_E_ARGUMENT;
  } catch (const std::exception &x) {
    g_last_error = x.what();
    return DVS_E_STATE;
  }
}

// Owning device allocation that grows on demand.
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
  void upload(const T *host, size_t n, const char *what, hipStream_t st = nullptr) {
    reserve(n, what);
    if (n) hip_check(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, st), what);
  }
};

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
  double ms() const {
    float v = 0;
    hip_check(hipEventElapsedTime(&v, start, stop), "timer elapsed");
    return v;
  }
};

}  // namespace
}  // namespace dvsamd

using namespace dvsamd;

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
struct dvs_scene_s {
  std::atomic<int> refs{1};
  DeviceArray<float4> pairs, triangles;
  DeviceArray<int2> leaves;
  DeviceArray<uint32_t> indices;
  DeviceArray<float> normals, albedos;
  DevScene view{};
  int64_t internal_nodes = 0, leaf_count = 0, device_bytes = 0, triangle_count = 0;
  hipStream_t stream = nullptr;
  DeviceArray<Pose> poses;
  DeviceArray<float> image, depth;
  DeviceArray<double> radiance;
  DeviceArray<int> status;
  DeviceArray<unsigned long long> counter;
  dvs_stats stats{};
  ~dvs_scene_s() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void scene_release(dvs_scene s) {
  if (s && s->refs.fetch_sub(1) == 1) delete s;
}

namespace {
// Renders `n` device poses of `s` into device image/depth/radiance buffers.
void frame_trace(dvs_scene s, const Pose *poses, int n, double *rad, float *img, float *dep, hipStream_t st) {
  const int npix = s->view.width * s->view.height;
  s->status.reserve(1, "optical status");
  s->counter.reserve(1, "ray counter");
  hip_check(hipMemsetAsync(s->status.ptr, 0, sizeof(int), st), "clear optical status");
  hip_check(hipMemsetAsync(s->counter.ptr, 0, sizeof(unsigned long long), st), "clear ray counter");
  Timer timer;
  hip_check(hipEventRecord(timer.start, st), "record render start");
  hip_check(launch_trace(dvsdev::MODE_FRAME, s->view, poses, n, npix, (long long)npix * n, nullptr, rad, img, dep,
                         s->status.ptr, s->counter.ptr, st),
            "parallel optical launch");
  hip_check(hipEventRecord(timer.stop, st), "record render stop");
  int status = 0;
  unsigned long long shadows = 0;
  hip_check(hipMemcpyAsync(&status, s->status.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "parallel optical status");
  hip_check(hipMemcpyAsync(&shadows, s->counter.ptr, sizeof(shadows), hipMemcpyDeviceToHost, st), "ray counter");
  hip_check(hipStreamSynchronize(st), "parallel optical wait");
  const double ms = timer.ms();
  s->stats.trace_ms += ms;
  s->stats.total_ms += ms;
  s->stats.trace_launches += 1;
  s->stats.primary_rays += (uint64_t)npix * n * s->view.samples * s->view.samples;
  s->stats.shadow_rays += shadows;
  if (status) fail(DVS_E_OPTICAL, "invalid optical sample/BVH stack overflow");
}
}  // namespace

extern "C" {

const char *dvs_last_error(void) { return g_last_error.c_str(); }

int dvs_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int dvs_bvh_build(int max_leaf_size, float sah_node_cost, float sah_primitive_cost, const float *vertices,
                  int64_t vertex_count, uint32_t *indices, int64_t triangle_count, dvs_bvh_node *nodes,
                  int64_t *node_count, int32_t *root) {
  return guarded([&] {
    if (!vertices || !indices || !nodes || !node_count || !root || vertex_count < 0 || triangle_count < 0)
      fail(DVS_E_ARGUMENT, "invalid BVH build arguments");
    BuildOptions opt;
    opt.max_leaf_size = max_leaf_size;
    opt.node_cost = sah_node_cost;
    opt.primitive_cost = sah_primitive_cost;
    std::vector<dvs_bvh_node> out;
    *root = build_bvh(opt, vertices, vertex_count, indices, triangle_count, out);
    *node_count = (int64_t)out.size();
    if (!out.empty()) std::memcpy(nodes, out.data(), out.size() * sizeof(dvs_bvh_node));
  });
}

int dvs_bvh_validate(const dvs_bvh_node *nodes, int64_t node_count, int32_t root) {
  return guarded([&] {
    const std::string why = validate_bvh(nodes, node_count, root);
    if (!why.empty()) fail(DVS_E_ARGUMENT, why);
  });
}

int dvs_scene_create(const dvs_bvh_node *nodes, int64_t node_count, int32_t root, const float *vertices,
                     const float *normals, int64_t vertex_count, const uint32_t *indices, int64_t triangle_count,
                     const float *albedos, const dvs_camera *camera, const dvs_hapke *hapke, dvs_scene *out) {
  return dvs_scene_create_ex(nodes, node_count, root, vertices, normals, vertex_count, indices, triangle_count,
                             albedos, camera, hapke, DVS_DEFAULT_DEVICE_LEAF_SIZE, out);
}

int dvs_scene_create_ex(const dvs_bvh_node *nodes, int64_t node_count, int32_t root, const float *vertices,
                        const float *normals, int64_t vertex_count, const uint32_t *indices, int64_t triangle_count,
                        const float *albedos, const dvs_camera *camera, const dvs_hapke *hapke,
                        int32_t device_leaf_size, dvs_scene *out) {
  return guarded([&] {
    if (device_leaf_size < 0) fail(DVS_E_ARGUMENT, "device leaf size must be >= 0");
    if (!out || !camera || !hapke) fail(DVS_E_ARGUMENT, "invalid scene arguments");
    *out = nullptr;
    // DeviceScene checks (src/cuda/device_scene.cu:16-22).
    if (!nodes || node_count <= 0 || !indices || triangle_count <= 0 || !vertices || vertex_count <= 0 ||
        !normals || root < 0 || root >= node_count)
      fail(DVS_E_ARGUMENT, "invalid indexed mesh/BVH/material dimensions");
    const std::string why = validate_bvh(nodes, node_count, root);
    if (!why.empty()) fail(DVS_E_ARGUMENT, why);
    for (int64_t i = 0; i < 3 * triangle_count; ++i)
      if (indices[i] >= (uint64_t)vertex_count) fail(DVS_E_ARGUMENT, "mesh triangle index outside vertex array");
    if (camera->width <= 0 || camera->height <= 0 || camera->samples < 1 || camera->samples > 16)
      fail(DVS_E_ARGUMENT, "camera grid must be positive with 1..16 sub-rays per axis");
    SoaBvh soa = relayout(nodes, node_count, root, vertices, indices, triangle_count, device_leaf_size);
    if (soa.max_depth > 64) fail(DVS_E_ARGUMENT, "BVH deeper than the 64-entry device traversal stack");
    std::unique_ptr<dvs_scene_s> s(new dvs_scene_s());
    hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "scene stream");
    s->pairs.upload(reinterpret_cast<const float4 *>(soa.pairs.data()), soa.pairs.size() / 4, "BVH node pairs", s->stream);
    s->leaves.upload(reinterpret_cast<const int2 *>(soa.leaf_ranges.data()), soa.leaf_ranges.size() / 2, "BVH leaves", s->stream);
    s->triangles.upload(reinterpret_cast<const float4 *>(soa.triangles.data()), soa.triangles.size() / 4, "triangle records", s->stream);
    s->indices.upload(indices, 3 * (size_t)triangle_count, "triangle indices", s->stream);
    s->normals.upload(normals, 3 * (size_t)vertex_count, "vertex normals", s->stream);
    if (albedos) s->albedos.upload(albedos, (size_t)vertex_count, "vertex albedo", s->stream);
    hip_check(hipStreamSynchronize(s->stream), "scene upload");
    DevScene &v = s->view;
    v.pairs = s->pairs.ptr;
    v.leaves = s->leaves.ptr;
    v.triangles = s->triangles.ptr;
    v.indices = s->indices.ptr;
    v.normals = s->normals.ptr;
    v.albedos = albedos ? s->albedos.ptr : nullptr;
    v.root_ref = soa.root_ref;
    v.stack_depth = std::max(1, soa.max_depth);
    v.width = camera->width;
    v.height = camera->height;
    v.samples = camera->samples;
    v.fx = camera->fx; v.fy = camera->fy; v.cx = camera->cx; v.cy = camera->cy;
    v.hapke = *hapke;
    s->internal_nodes = soa.internal_nodes;
    s->leaf_count = soa.leaf_count;
    s->triangle_count = triangle_count;
    s->device_bytes = (int64_t)(soa.pairs.size() * 4 + soa.leaf_ranges.size() * 4 + soa.triangles.size() * 4 +
                                12 * triangle_count + 12 * vertex_count + (albedos ? 4 * vertex_count : 0));
    *out = s.release();
  });
}

int dvs_scene_destroy(dvs_scene scene) {
  scene_release(scene);
  return DVS_OK;
}

int dvs_scene_info(dvs_scene s, int32_t *max_depth, int64_t *internal_nodes, int64_t *leaves, int64_t *device_bytes) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null scene");
    if (max_depth) *max_depth = s->view.stack_depth;
    if (internal_nodes) *internal_nodes = s->internal_nodes;
    if (leaves) *leaves = s->leaf_count;
    if (device_bytes) *device_bytes = s->device_bytes;
  });
}

int dvs_frame_render(dvs_scene s, const dvs_optical_state *poses, int32_t n, float *image, float *depth,
                     double *radiance) {
  return guarded([&] {
    if (!s || !poses || n <= 0) fail(DVS_E_ARGUMENT, "invalid parallel optical request");
    const size_t npix = (size_t)s->view.width * s->view.height, total = npix * n;
    hipStream_t st = s->stream;
    DeviceArray<dvs_optical_state> states;
    states.upload(poses, n, "upload optical poses", st);
    s->poses.reserve(n, "optical poses");
    hip_check(launch_state_poses(n, states.ptr, s->poses.ptr, st), "pose conversion");
    s->image.reserve(total, "optical image");
    s->depth.reserve(total, "optical depth");
    if (radiance) s->radiance.reserve(total, "optical radiance");
    frame_trace(s, s->poses.ptr, n, radiance ? s->radiance.ptr : nullptr, s->image.ptr, s->depth.ptr, st);
    if (image) hip_check(hipMemcpyAsync(image, s->image.ptr, total * sizeof(float), hipMemcpyDeviceToHost, st), "download optical values");
    if (depth) hip_check(hipMemcpyAsync(depth, s->depth.ptr, total * sizeof(float), hipMemcpyDeviceToHost, st), "download optical depth");
    if (radiance) hip_check(hipMemcpyAsync(radiance, s->radiance.ptr, total * sizeof(double), hipMemcpyDeviceToHost, st), "download optical radiance");
    hip_check(hipStreamSynchronize(st), "download optical values");
  });
}

int dvs_frame_render_device(dvs_scene s, const dvs_optical_state *device_poses, int32_t n, float *device_image,
                            float *device_depth) {
  return guarded([&] {
    if (!s || !device_poses || n <= 0) fail(DVS_E_ARGUMENT, "invalid parallel optical request");
    s->poses.reserve(n, "optical poses");
    hip_check(launch_state_poses(n, device_poses, s->poses.ptr, s->stream), "pose conversion");
    frame_trace(s, s->poses.ptr, n, nullptr, device_image, device_depth, s->stream);
  });
}

int dvs_frame_traversal_stats(dvs_scene s, const dvs_optical_state *pose, uint64_t totals[8]) {
  return guarded([&] {
    if (!s || !pose || !totals) fail(DVS_E_ARGUMENT, "invalid traversal statistics request");
    hipStream_t st = s->stream;
    DeviceArray<dvs_optical_state> state;
    state.upload(pose, 1, "upload optical pose", st);
    s->poses.reserve(1, "optical poses");
    hip_check(launch_state_poses(1, state.ptr, s->poses.ptr, st), "pose conversion");
    DeviceArray<unsigned long long> acc;
    acc.reserve(8, "traversal counters");
    hip_check(hipMemsetAsync(acc.ptr, 0, 8 * sizeof(unsigned long long), st), "clear traversal counters");
    hip_check(launch_traversal_stats(s->view, s->poses.ptr, s->view.width * s->view.height, acc.ptr, st),
              "traversal statistics");
    hip_check(hipMemcpyAsync(totals, acc.ptr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
              "traversal statistics");
    hip_check(hipStreamSynchronize(st), "traversal statistics");
  });
}

int dvs_scene_stats(dvs_scene s, dvs_stats *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null scene");
    *out = s->stats;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Event sensor
// ---------------------------------------------------------------------------
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
  p.reset_noise_std = effects && c.reset_thermal_noise ? std::sqrt(kK * c.temperature_k / c.reset_capacitance_f) / voltage : 0.0;
  const double pitch = c.pixel_pitch_um * 1e-6;
  photo.electrons_per_radiance = c.radiance_scale * 3.141592653589793 / (4 * c.f_number * c.f_number) *
                                 c.lens_transmission * pitch * pitch * c.fill_factor * c.quantum_efficiency *
                                 (c.wavelength_nm * 1e-9) / (kH * kC);
  photo.integration_step = c.circuit_step;
  photo.dark_electrons = effects ? c.dark_current_a / kQ : 0.0;
  photo.shot_noise = effects ? c.photon_shot_noise : 0;
  photo.pad = 0;
}

// Bounded FIFO readout (EventReadout::submit / drain, src/event_readout.cpp:6-34).
// Events arrive already in (t, y, x) order from the device sort.
struct Readout {
  double bandwidth, latency, available = 0;
  size_t capacity, dropped = 0, peak = 0;
  struct Pending { double ready; dvs_event event; };
  std::deque<Pending> pending;
  Readout(double b, double l, size_t c) : bandwidth(b), latency(l), capacity(c) {}
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
}  // namespace

struct dvs_sensor_s {
  dvs_scene scene = nullptr;
  int width = 0, height = 0, npix = 0, max_events = 16;
  uint64_t seed = 0;
  CircuitParams params{};
  PhotoParams photo{};
  double threshold_std = 0, threshold_min = 0, last_time = 0;
  bool initialized = false;
  hipStream_t stream = nullptr;
  Readout readout{0, 0, 1};
  // device state and work buffers
  DeviceArray<dvs_pixel_state> states;
  DeviceArray<CrossingEvent> slots;
  DeviceArray<int> counts, ends, status, cell_iv, ev_pixel, ev_polarity;
  DeviceArray<double> cell_t0, cell_t1, radiance, ev_time;
  DeviceArray<dvs_optical_interval> intervals;
  DeviceArray<Pose> poses, ev_poses;
  DeviceArray<float> ev_depth;
  DeviceArray<unsigned long long> keys, keys_sorted, counter;
  DeviceArray<unsigned> order, order_sorted;
  DeviceArray<dvs_event> events_out;
  DeviceArray<unsigned char> temp;
  std::vector<dvs_event> raw_host, ready_host;
  dvs_stats stats{};
  ~dvs_sensor_s() {
    if (stream) (void)hipStreamDestroy(stream);
    scene_release(scene);
  }
};

namespace {
// Radiance scratch is [cells x pixels] doubles; cells are processed in chunks
// so the buffer stays under this many bytes.
constexpr size_t kRadianceBudget = size_t(1) << 30;

void process_intervals(dvs_sensor s, const dvs_optical_interval *ivs, int n) {
  // EventSensor::upload_intervals (event_sensor.cu:274-289).
  if (!s->initialized || n <= 0) fail(DVS_E_ARGUMENT, "sensor not initialized or no intervals");
  double previous = s->last_time, cells_estimate = 0;
  for (int k = 0; k < n; ++k) {
    const dvs_optical_interval &iv = ivs[k];
    if (!std::isfinite(iv.t0) || !std::isfinite(iv.t1) || iv.t0 != previous || iv.t1 <= iv.t0)
      fail(DVS_E_ARGUMENT, "sensor intervals must be positive and contiguous");
    cells_estimate += (iv.t1 - iv.t0) / s->photo.integration_step;
    previous = iv.t1;
  }
  if (cells_estimate > 1000000) fail(DVS_E_ARGUMENT, "optical horizon too long; reduce batch size");
  // Circuit cells of each interval (generate_pixels, event_sensor.cu:37-40).
  std::vector<double> t0s, t1s;
  std::vector<int> ivx;
  for (int k = 0; k < n; ++k) {
    const dvs_optical_interval &iv = ivs[k];
    const int steps = (int)std::ceil((iv.t1 - iv.t0) / s->photo.integration_step);
    const double dt = (iv.t1 - iv.t0) / steps;
    for (int step = 0; step < steps; ++step) {
      t0s.push_back(iv.t0 + step * dt);
      t1s.push_back(step + 1 == steps ? iv.t1 : iv.t0 + (step + 1) * dt);
      ivx.push_back(k);
    }
  }
  const int cells = (int)t0s.size();
  hipStream_t st = s->stream;
  const DevScene &view = s->scene->view;
  const int npix = s->npix, R = view.samples * view.samples;
  s->intervals.upload(ivs, n, "trajectory interval upload", st);
  s->cell_t0.upload(t0s.data(), cells, "cell times", st);
  s->cell_t1.upload(t1s.data(), cells, "cell times", st);
  s->cell_iv.upload(ivx.data(), cells, "cell intervals", st);
  s->poses.reserve(cells, "cell poses");
  s->counts.reserve(npix, "event counts");
  s->status.reserve(1, "error flag");
  s->counter.reserve(1, "ray counter");
  hip_check(hipMemsetAsync(s->counts.ptr, 0, npix * sizeof(int), st), "clear events");
  hip_check(hipMemsetAsync(s->status.ptr, 0, sizeof(int), st), "clear status");
  hip_check(hipMemsetAsync(s->counter.ptr, 0, sizeof(unsigned long long), st), "clear ray counter");
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(cells, kRadianceBudget / (sizeof(double) * npix)));
  s->radiance.reserve((size_t)chunk * npix, "cell radiance");
  s->slots.reserve((size_t)npix * s->max_events, "events");
  hip_check(launch_cell_poses(cells, s->cell_t0.ptr, s->cell_t1.ptr, s->cell_iv.ptr, s->intervals.ptr, s->poses.ptr, st),
            "cell poses");
  const int chunks = (cells + chunk - 1) / chunk;
  std::vector<std::unique_ptr<Timer>> trace_t(chunks), circuit_t(chunks);
  for (int c = 0, k = 0; c < cells; c += chunk, ++k) {
    const int count = std::min(chunk, cells - c);
    trace_t[k].reset(new Timer());
    circuit_t[k].reset(new Timer());
    hip_check(hipEventRecord(trace_t[k]->start, st), "record");
    hip_check(launch_trace(dvsdev::MODE_CELLS, view, s->poses.ptr + c, count, npix, (long long)npix * count, nullptr,
                           s->radiance.ptr, nullptr, nullptr, s->status.ptr, s->counter.ptr, st),
              "trace and integrate circuits");
    hip_check(hipEventRecord(trace_t[k]->stop, st), "record");
    hip_check(hipEventRecord(circuit_t[k]->start, st), "record");
    hip_check(launch_circuit(npix, c, count, s->cell_t0.ptr, s->cell_t1.ptr, s->radiance.ptr, s->states.ptr, s->params,
                             s->photo, s->seed, s->max_events, s->slots.ptr, s->counts.ptr, s->status.ptr, st),
              "trace and integrate circuits");
    hip_check(hipEventRecord(circuit_t[k]->stop, st), "record");
  }
  // Compaction by scan of per-pixel counts.
  Timer ev_t;
  hip_check(hipEventRecord(ev_t.start, st), "record");
  s->ends.reserve(npix, "event offsets");
  size_t temp_bytes = 0;
  hip_check(inclusive_scan_int(nullptr, temp_bytes, s->counts.ptr, s->ends.ptr, npix, st), "scan size");
  s->temp.reserve(temp_bytes, "scan scratch");
  hip_check(inclusive_scan_int(s->temp.ptr, temp_bytes, s->counts.ptr, s->ends.ptr, npix, st), "event scan");
  int status = 0, total = 0;
  unsigned long long shadows = 0;
  hip_check(hipMemcpyAsync(&status, s->status.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
  hip_check(hipMemcpyAsync(&total, s->ends.ptr + npix - 1, sizeof(int), hipMemcpyDeviceToHost, st), "copy count");
  hip_check(hipStreamSynchronize(st), "copy count");
  if (status & 1) fail(DVS_E_CAPACITY, "pixel event capacity exceeded; reduce numerical window");
  if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
  s->raw_host.resize(total);
  if (total > 0) {
    s->ev_time.reserve(total, "event times");
    s->ev_pixel.reserve(total, "event pixels");
    s->ev_polarity.reserve(total, "event polarity");
    s->ev_depth.reserve(total, "event depth");
    s->ev_poses.reserve(total, "event poses");
    s->keys.reserve(total, "event keys");
    s->keys_sorted.reserve(total, "event keys");
    s->order.reserve(total, "event order");
    s->order_sorted.reserve(total, "event order");
    s->events_out.reserve(total, "events");
    hip_check(launch_gather(npix, s->ends.ptr, s->counts.ptr, s->max_events, s->slots.ptr, s->ev_time.ptr,
                            s->ev_pixel.ptr, s->ev_polarity.ptr, s->keys.ptr, s->order.ptr, st),
              "event compaction");
    hip_check(launch_event_poses(total, s->ev_time.ptr, s->intervals.ptr, n, s->ev_poses.ptr, st), "event poses");
    hip_check(launch_trace(dvsdev::MODE_DEPTH, view, s->ev_poses.ptr, total, npix, total, s->ev_pixel.ptr, nullptr,
                           nullptr, s->ev_depth.ptr, s->status.ptr, s->counter.ptr, st),
              "event depth rays");
    temp_bytes = 0;
    hip_check(sort_pairs_u64(nullptr, temp_bytes, s->keys.ptr, s->keys_sorted.ptr, s->order.ptr, s->order_sorted.ptr,
                             total, 0, 64, st),
              "sort size");
    s->temp.reserve(temp_bytes, "sort scratch");
    hip_check(sort_pairs_u64(s->temp.ptr, temp_bytes, s->keys.ptr, s->keys_sorted.ptr, s->order.ptr,
                             s->order_sorted.ptr, total, 0, 64, st),
              "event sort");
    hip_check(launch_permute(total, s->width, s->order_sorted.ptr, s->ev_time.ptr, s->ev_pixel.ptr,
                             s->ev_polarity.ptr, s->ev_depth.ptr, s->events_out.ptr, st),
              "event records");
    hip_check(hipMemcpyAsync(s->raw_host.data(), s->events_out.ptr, sizeof(dvs_event) * total, hipMemcpyDeviceToHost, st),
              "copy events");
  }
  hip_check(hipEventRecord(ev_t.stop, st), "record");
  hip_check(hipMemcpyAsync(&status, s->status.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "copy status");
  hip_check(hipMemcpyAsync(&shadows, s->counter.ptr, sizeof(shadows), hipMemcpyDeviceToHost, st), "ray counter");
  hip_check(hipStreamSynchronize(st), "copy events");
  if (status) fail(DVS_E_OPTICAL, "invalid optical sample or BVH traversal overflow");
  double trace_ms = 0, circuit_ms = 0;
  for (int k = 0; k < chunks; ++k) { trace_ms += trace_t[k]->ms(); circuit_ms += circuit_t[k]->ms(); }
  const double events_ms = ev_t.ms();
  s->stats.trace_ms += trace_ms;
  s->stats.circuit_ms += circuit_ms;
  s->stats.events_ms += events_ms;
  s->stats.total_ms += trace_ms + circuit_ms + events_ms;
  s->stats.trace_launches += chunks;
  s->stats.primary_rays += (uint64_t)npix * cells * R;
  s->stats.event_rays += (uint64_t)total * R;
  s->stats.shadow_rays += shadows;
  s->stats.pixel_cells += (uint64_t)npix * cells;
  s->stats.comparator_events += (uint64_t)total;
  const size_t before = s->ready_host.size();
  s->readout.submit_sorted(s->raw_host.data(), s->raw_host.size(), ivs[n - 1].t1, s->ready_host);
  s->stats.delivered_events += s->ready_host.size() - before;
  s->last_time = ivs[n - 1].t1;
}
}  // namespace

extern "C" {

int dvs_sensor_create(dvs_scene scene, const dvs_physic *c, int32_t effects, int32_t aer, dvs_sensor *out) {
  return guarded([&] {
    if (!scene || !c || !out) fail(DVS_E_ARGUMENT, "invalid sensor arguments");
    *out = nullptr;
    const int w = scene->view.width, h = scene->view.height;
    // EventSensor ctor checks (event_sensor.cu:251-252).
    if (w <= 0 || h <= 0 || w > 32767 || h > 32767 || c->max_events_per_step < 1 || c->max_events_per_step > 64 ||
        (long long)w * h * c->max_events_per_step > std::numeric_limits<int>::max())
      fail(DVS_E_ARGUMENT, "invalid sensor capacity");
    std::unique_ptr<dvs_sensor_s> s(new dvs_sensor_s());
    s->width = w; s->height = h; s->npix = w * h;
    s->max_events = c->max_events_per_step;
    s->seed = (uint64_t)(uint32_t)c->random_seed;
    derive(*c, effects != 0, s->params, s->photo);
    s->threshold_std = effects ? c->threshold_std : 0.0;
    s->threshold_min = c->threshold_min;
    s->readout = Readout(aer ? (double)c->readout_bandwidth : 0.0, aer ? (double)c->readout_latency : 0.0,
                         (size_t)std::max(1, c->readout_queue_capacity));
    hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "sensor stream");
    s->states.reserve(s->npix, "pixel states");
    s->slots.reserve((size_t)s->npix * s->max_events, "events");
    scene->refs.fetch_add(1);
    s->scene = scene;
    *out = s.release();
  });
}

int dvs_sensor_destroy(dvs_sensor s) {
  delete s;
  return DVS_OK;
}

int dvs_sensor_initialize(dvs_sensor s, const dvs_optical_state *pose, double timestamp) {
  return guarded([&] {
    if (!s || !pose || !std::isfinite(timestamp) || s->initialized) fail(DVS_E_ARGUMENT, "invalid sensor initialization");
    hipStream_t st = s->stream;
    DeviceArray<dvs_optical_state> state;
    state.upload(pose, 1, "initial pose", st);
    s->poses.reserve(1, "initial pose");
    hip_check(launch_state_poses(1, state.ptr, s->poses.ptr, st), "initial pose");
    s->radiance.reserve(s->npix, "initial radiance");
    s->status.reserve(1, "error flag");
    s->counter.reserve(1, "ray counter");
    hip_check(hipMemsetAsync(s->status.ptr, 0, sizeof(int), st), "clear initialization error");
    hip_check(hipMemsetAsync(s->counter.ptr, 0, sizeof(unsigned long long), st), "clear ray counter");
    hip_check(launch_trace(dvsdev::MODE_FRAME, s->scene->view, s->poses.ptr, 1, s->npix, s->npix, nullptr,
                           s->radiance.ptr, nullptr, nullptr, s->status.ptr, s->counter.ptr, st),
              "initialize circuit");
    hip_check(launch_init(s->npix, s->radiance.ptr, s->photo, s->params, s->threshold_std, s->threshold_min, s->seed,
                          s->states.ptr, s->status.ptr, st),
              "initialize circuit");
    int status = 0;
    hip_check(hipMemcpyAsync(&status, s->status.ptr, sizeof(int), hipMemcpyDeviceToHost, st), "initialize circuit status");
    hip_check(hipStreamSynchronize(st), "initialize circuit status");
    if (status) fail(DVS_E_OPTICAL, "invalid optical sample during initialization");
    s->last_time = timestamp;
    s->initialized = true;
  });
}

int dvs_sensor_process(dvs_sensor s, const dvs_optical_interval *ivs, int32_t n, int64_t *ready_count) {
  return guarded([&] {
    if (!s || !ivs) fail(DVS_E_ARGUMENT, "sensor not initialized or no intervals");
    process_intervals(s, ivs, n);
    if (ready_count) *ready_count = (int64_t)s->ready_host.size();
  });
}

int dvs_sensor_take(dvs_sensor s, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!s || !count) fail(DVS_E_ARGUMENT, "null sensor");
    *count = (int64_t)s->ready_host.size();
    if ((int64_t)s->ready_host.size() > capacity) fail(DVS_E_ARGUMENT, "event buffer too small");
    if (!s->ready_host.empty()) std::memcpy(out, s->ready_host.data(), s->ready_host.size() * sizeof(dvs_event));
    s->ready_host.clear();
  });
}

int dvs_sensor_raw_events(dvs_sensor s, dvs_event *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!s || !count) fail(DVS_E_ARGUMENT, "null sensor");
    *count = (int64_t)s->raw_host.size();
    if ((int64_t)s->raw_host.size() > capacity) fail(DVS_E_ARGUMENT, "event buffer too small");
    if (!s->raw_host.empty()) std::memcpy(out, s->raw_host.data(), s->raw_host.size() * sizeof(dvs_event));
  });
}

int dvs_sensor_drain(dvs_sensor s, int64_t *ready_count) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    s->readout.drain(s->ready_host);
    if (ready_count) *ready_count = (int64_t)s->ready_host.size();
  });
}

int dvs_sensor_readout_stats(dvs_sensor s, int64_t *dropped, int64_t *peak) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    if (dropped) *dropped = (int64_t)s->readout.dropped;
    if (peak) *peak = (int64_t)s->readout.peak;
  });
}

int dvs_sensor_states(dvs_sensor s, dvs_pixel_state *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    hip_check(hipMemcpyAsync(out, s->states.ptr, sizeof(dvs_pixel_state) * s->npix, hipMemcpyDeviceToHost, s->stream),
              "download pixel states");
    hip_check(hipStreamSynchronize(s->stream), "download pixel states");
  });
}

int dvs_sensor_stats(dvs_sensor s, dvs_stats *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null sensor");
    *out = s->stats;
  });
}

int dvs_sensor_reset_stats(dvs_sensor s) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null sensor");
    s->stats = dvs_stats{};
  });
}

// ---- unit entry points ----------------------------------------------------
int dvs_hapke_eval(const float *ci, const float *ce, const float *cg, const float *alb, int64_t n,
                   const dvs_hapke *h, float *out) {
  return guarded([&] {
    if (!ci || !ce || !cg || !alb || !h || !out || n < 0) fail(DVS_E_ARGUMENT, "invalid Hapke arguments");
    DeviceArray<float> a, b, c, d, o;
    a.upload(ci, n, "hapke input"); b.upload(ce, n, "hapke input"); c.upload(cg, n, "hapke input");
    d.upload(alb, n, "hapke input"); o.reserve(n, "hapke output");
    hip_check(launch_hapke(n, a.ptr, b.ptr, c.ptr, d.ptr, *h, o.ptr, nullptr), "hapke kernel");
    hip_check(hipMemcpy(out, o.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "hapke output");
  });
}

int dvs_receptor_eval(const double *u, const double *rate, const double *dt, double charge, int64_t n, double *out) {
  return guarded([&] {
    if (!u || !rate || !dt || !out || n < 0) fail(DVS_E_ARGUMENT, "invalid receptor arguments");
    DeviceArray<double> a, b, c, o;
    a.upload(u, n, "receptor input"); b.upload(rate, n, "receptor input"); c.upload(dt, n, "receptor input");
    o.reserve(n, "receptor output");
    hip_check(launch_receptor(n, a.ptr, b.ptr, c.ptr, charge, o.ptr, nullptr), "receptor kernel");
    hip_check(hipMemcpy(out, o.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "receptor output");
  });
}

int dvs_noise_eval(uint64_t seed, int64_t pixels, double lambda, int32_t draws, double *poisson, double *normal) {
  return guarded([&] {
    if (pixels < 0 || draws < 0) fail(DVS_E_ARGUMENT, "invalid noise arguments");
    DeviceArray<double> p, q;
    const size_t n = (size_t)pixels * draws;
    if (poisson) p.reserve(n, "poisson output");
    if (normal) q.reserve(n, "normal output");
    hip_check(launch_noise(seed, pixels, lambda, draws, poisson ? p.ptr : nullptr, normal ? q.ptr : nullptr, nullptr),
              "noise kernel");
    if (poisson) hip_check(hipMemcpy(poisson, p.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "poisson output");
    if (normal) hip_check(hipMemcpy(normal, q.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "normal output");
  });
}

static int trace_rays_common(dvs_scene s, const float *o, const float *d, const int32_t *self, int64_t n, int shadow,
                             int32_t *hit, int32_t *tri, float *t, float *u, float *v) {
  return guarded([&] {
    if (!s || !o || !d || !hit || n < 0 || (shadow && !self) || (!shadow && (!tri || !t || !u || !v)))
      fail(DVS_E_ARGUMENT, "invalid ray query");
    DeviceArray<float> dO, dD, dT, dU, dV;
    DeviceArray<int> dS, dH, dTri;
    dO.upload(o, 3 * n, "rays"); dD.upload(d, 3 * n, "rays");
    dH.reserve(n, "ray output");
    if (shadow) dS.upload(self, n, "rays");
    else { dTri.reserve(n, "ray output"); dT.reserve(n, "ray output"); dU.reserve(n, "ray output"); dV.reserve(n, "ray output"); }
    hip_check(launch_rays(s->view, n, dO.ptr, dD.ptr, dS.ptr, shadow, dH.ptr, dTri.ptr, dT.ptr, dU.ptr, dV.ptr, nullptr),
              "ray kernel");
    hip_check(hipMemcpy(hit, dH.ptr, n * sizeof(int), hipMemcpyDeviceToHost), "ray output");
    if (!shadow) {
      hip_check(hipMemcpy(tri, dTri.ptr, n * sizeof(int), hipMemcpyDeviceToHost), "ray output");
      hip_check(hipMemcpy(t, dT.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "ray output");
      hip_check(hipMemcpy(u, dU.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "ray output");
      hip_check(hipMemcpy(v, dV.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "ray output");
    }
  });
}

int dvs_trace_rays(dvs_scene s, const float *o, const float *d, int64_t n, int32_t *hit, int32_t *tri, float *t,
                   float *u, float *v) {
  return trace_rays_common(s, o, d, nullptr, n, 0, hit, tri, t, u, v);
}

int dvs_trace_shadow_rays(dvs_scene s, const float *o, const float *d, const int32_t *self, int64_t n,
                          int32_t *occluded) {
  return trace_rays_common(s, o, d, self, n, 1, occluded, nullptr, nullptr, nullptr, nullptr);
}

}  // extern "C"
