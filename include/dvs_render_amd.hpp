// C++17 API of the DVS path on MI355X: the classes of the reference's
// include/dvs_render.hpp (DeviceScene, DvsRayCaster, CamRayCaster,
// BVHBuilder, Config, the event stream writers) with the same names,
// signatures and error behaviour, implemented over the C ABIs of
// libdvs_hip.so (dvs_hip.h) and libdvs_io.so (dvs_io.h).  Types are re-homed
// off CUDA's vector types and Eigen: float3/float4/uint3 here are plain
// structs of the same layout, and any trajectory state with Eigen-style
// accessors (x(), y(), z(), w()) converts through to_optical_state.
//
// Errors: an invalid argument throws std::invalid_argument, every other
// failure std::runtime_error, with the reference's message text.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cerrno>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include "dvs_hip.h"
#include "dvs_io.h"

namespace dvs {

struct float3 { float x = 0, y = 0, z = 0; };
struct float4 { float x = 0, y = 0, z = 0, w = 0; };
struct uint3 { uint32_t x = 0, y = 0, z = 0; };
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }

using Event = dvs_event;  // packed 17 bytes (include/dvs/dvs_base.hpp:9-18)

// OpticalState / OpticalInterval (include/dvs/optical_sensor.hpp:15-20).
struct OpticalState {
  float3 position, velocity, light, light_rate;
  float4 rotation;
  float intensity = 0, intensity_rate = 0;
};
struct OpticalInterval {
  OpticalState start, end;
  double t0 = 0, t1 = 0;
};
static_assert(sizeof(OpticalState) == sizeof(dvs_optical_state), "OpticalState layout");
static_assert(sizeof(OpticalInterval) == sizeof(dvs_optical_interval), "OpticalInterval layout");

// BVHNode<float3> (include/dvs/scene_types.hpp:29-35): 40 bytes.
template <class F3> struct AABB { F3 min, max; };
template <class F3> struct BVHNode {
  AABB<F3> box;
  int left = -1, right = -1, start = 0, count = 0;
};
static_assert(sizeof(BVHNode<float3>) == sizeof(dvs_bvh_node), "BVHNode layout");

// Configuration (include/dvs/scene_types.hpp:59-109), PhysicConfig and
// HapkeConfig with the reference's defaults.
struct CameraConfig {
  int width = 0, height = 0;
  float fx = 0, fy = 0, cx = 0, cy = 0;
};
struct PhysicConfig {
  float threshold_pos = .3f, threshold_neg = .3f, refractory_time = .0002f;
  float readout_bandwidth = 5e6f, readout_latency = 0;
  float threshold_std = 0.03f, threshold_min = 0.01f;
  float wavelength_nm = 550, quantum_efficiency = .3f, pixel_pitch_um = 15;
  float fill_factor = .2f, f_number = 2, lens_transmission = .8f;
  float radiance_scale = .001f;
  float temperature_k = 300, subthreshold_kappa = .8f;
  float receptor_capacitance_f = 1e-15f;
  float reset_capacitance_f = 20e-15f, dark_current_a = 1e-18f;
  float reset_leakage_a = 0, circuit_step = 1e-6f;
  int photon_shot_noise = 1, reset_thermal_noise = 0;
  int random_seed = 114514, max_events_per_step = 16;
  int readout_queue_capacity = 1000000;
};
static_assert(sizeof(PhysicConfig) == sizeof(dvs_physic), "PhysicConfig layout");
struct SceneConfig {
  int frame_count = 0;
  float fps = 0, imu_fps = 0, cam_fps = 0;
  float time_resolution = 0, blur_resolution = 0;
  int supersample_factor = 1;
  std::vector<float> translation{0, 0, 0}, rotation{0, 0, 0};
};
struct HapkeConfig {
  float b = 0.24f, c = 0.305f, hS = 0.08f, hC = 1.0f;
  float thetap = 26.0f * 3.14159265358979323846f / 180.0f;
  float w = 0.324f, B_S0 = 2.42f, B_C0 = 0.0f, phi = 0.3f;
};
static_assert(sizeof(HapkeConfig) == sizeof(dvs_hapke), "HapkeConfig layout");

namespace detail {
[[noreturn]] inline void raise(int code, const char *message) {
  if (code == DVS_E_ARGUMENT) throw std::invalid_argument(message);
  throw std::runtime_error(message);
}
inline void check(int code) {
  if (code != DVS_OK) raise(code, dvs_last_error());
}
// The diagnostics' state errors are std::logic_error, as in the reference
// (raycast.hpp:47-53, optical_sampler.cu:198).
inline void check_logic(int code) {
  if (code == DVS_E_STATE) throw std::logic_error(dvs_last_error());
  check(code);
}
inline void check_io(int code) {
  if (code != DVS_OK) raise(code, dvs_io_last_error());
}
inline dvs_physic to_c(const PhysicConfig &p) {
  dvs_physic c;
  static_assert(sizeof c == sizeof p, "layout");
  std::memcpy(&c, &p, sizeof c);
  return c;
}
inline dvs_hapke to_c(const HapkeConfig &h) {
  dvs_hapke c;
  std::memcpy(&c, &h, sizeof c);
  return c;
}
}  // namespace detail

class Config {
 public:
  CameraConfig camera;
  PhysicConfig physic;
  SceneConfig scene;
  HapkeConfig hapke;
  // Config(path, err) (src/scene_types.cpp): on a malformed file err holds
  // the message; an unparsable optional physic value throws
  // std::invalid_argument, as in the reference.
  Config(const std::string &config_file, std::string &err) {
    dvs_config c{};
    const int code = dvs_config_load(config_file.c_str(), &c);
    if (code != DVS_OK) {
      err = dvs_io_last_error();
      if (err.rfind("invalid physic parameter", 0) == 0) throw std::invalid_argument(err);
      return;
    }
    camera = {c.camera.width, c.camera.height, c.camera.fx, c.camera.fy, c.camera.cx, c.camera.cy};
    // same fields in the same order (see the static_asserts on their sizes)
    std::memcpy(static_cast<void *>(&physic), &c.physic, sizeof physic);
    std::memcpy(static_cast<void *>(&hapke), &c.hapke, sizeof hapke);
    scene.frame_count = c.scene.frame_count;
    scene.fps = c.scene.fps;
    scene.imu_fps = c.scene.imu_fps;
    scene.cam_fps = c.scene.cam_fps;
    scene.time_resolution = c.scene.time_resolution;
    scene.blur_resolution = c.scene.blur_resolution;
    scene.supersample_factor = c.scene.supersample_factor;
    scene.translation.assign(c.scene.translation, c.scene.translation + 3);
    scene.rotation.assign(c.scene.rotation, c.scene.rotation + 3);
  }
};

template <class F3, class U3> struct SceneBundle {
  std::vector<BVHNode<F3>> bvh_nodes;
  std::vector<F3> vertices, normals;
  std::vector<U3> bvh_triangles;
  std::vector<float> albedos;
  bool use_albedo = false;
};

// BVHBuilder (include/bvh/build.h:8-40).
struct BVHBuildParams {
  int max_leaf_size = 8;
  float sah_node_cost = 1.0f;
  float sah_primitive_cost = 1.0f;
};
class BVHBuilder {
 public:
  // Indexed build: nodes are replaced, indices reordered to leaf order;
  // returns the root (the last node).
  static int build(const BVHBuildParams &params, std::vector<BVHNode<float3>> &out_nodes,
                   std::vector<uint3> &indices, std::vector<float3> &vertices) {
    out_nodes.assign(std::max<size_t>(2 * indices.size(), 1), BVHNode<float3>{});
    int64_t count = 0;
    int32_t root = -1;
    detail::check(dvs_bvh_build(params.max_leaf_size, params.sah_node_cost, params.sah_primitive_cost,
                                reinterpret_cast<const float *>(vertices.data()), (int64_t)vertices.size(),
                                reinterpret_cast<uint32_t *>(indices.data()), (int64_t)indices.size(),
                                reinterpret_cast<dvs_bvh_node *>(out_nodes.data()), &count, &root));
    out_nodes.resize((size_t)count);
    return root;
  }
  static void validate(const std::vector<BVHNode<float3>> &nodes, int root) {
    detail::check(dvs_bvh_validate(reinterpret_cast<const dvs_bvh_node *>(nodes.data()), (int64_t)nodes.size(),
                                   root));
  }
};

// DeviceScene (include/dvs/device_scene.hpp:5-15): one owner of the
// uploaded BVH and mesh, shared by the DVS and frame casters.
class DeviceScene {
 public:
  DeviceScene(const SceneBundle<float3, uint3> &mesh, const Config &config, int root) {
    dvs_camera cam{config.camera.width, config.camera.height, config.camera.fx, config.camera.fy,
                   config.camera.cx,    config.camera.cy,     config.scene.supersample_factor};
    const dvs_hapke hk = detail::to_c(config.hapke);
    if (mesh.use_albedo && mesh.albedos.size() != mesh.vertices.size())
      throw std::invalid_argument("invalid indexed mesh/BVH/material dimensions");
    detail::check(dvs_scene_create(reinterpret_cast<const dvs_bvh_node *>(mesh.bvh_nodes.data()),
                                   (int64_t)mesh.bvh_nodes.size(), root,
                                   reinterpret_cast<const float *>(mesh.vertices.data()),
                                   reinterpret_cast<const float *>(mesh.normals.data()),
                                   (int64_t)mesh.vertices.size(),
                                   reinterpret_cast<const uint32_t *>(mesh.bvh_triangles.data()),
                                   (int64_t)mesh.bvh_triangles.size(),
                                   mesh.use_albedo ? mesh.albedos.data() : nullptr, &cam, &hk, &scene_));
    width_ = cam.width;
    height_ = cam.height;
  }
  ~DeviceScene() { dvs_scene_destroy(scene_); }
  DeviceScene(const DeviceScene &) = delete;
  DeviceScene &operator=(const DeviceScene &) = delete;
  dvs_scene view() const { return scene_; }
  int width() const { return width_; }
  int height() const { return height_; }

 private:
  dvs_scene scene_ = nullptr;
  int width_ = 0, height_ = 0;
};

// Vec3 / Mat3 / Quaternionf stand-ins with Eigen-style accessors, and
// CameraTrajectoryState (include/dvs/trajectory.hpp:44-60).
struct Vec3 {
  float v[3] = {0, 0, 0};
  Vec3() = default;
  Vec3(float x, float y, float z) : v{x, y, z} {}
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float z() const { return v[2]; }
  float operator()(int i) const { return v[i]; }
  float *data() { return v; }
  const float *data() const { return v; }
};
// Column major, (row, col) access as Eigen's Matrix3f.
struct Mat3 {
  float m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  static Mat3 Identity() { return Mat3{}; }
  float operator()(int r, int c) const { return m[3 * c + r]; }
  float &operator()(int r, int c) { return m[3 * c + r]; }
  float *data() { return m; }
  const float *data() const { return m; }
};
struct Quaternion {
  float q[4] = {0, 0, 0, 1};  // x, y, z, w
  float x() const { return q[0]; }
  float y() const { return q[1]; }
  float z() const { return q[2]; }
  float w() const { return q[3]; }
};
struct CameraTrajectoryState {
  float timestamp = 0;
  Vec3 position;
  Mat3 orientation;
  Quaternion quaternion;
  Vec3 linear_velocity, angular_velocity, light_direction{0, 0, 1}, light_direction_rate;
  float light_intensity = 0, light_intensity_rate = 0;
};

// FrameData (include/dvs/scene_types.hpp:43-50): a NaN timestamp means
// frame_index / fps.
template <class V, class M> struct FrameData {
  int frame_index = 0;
  double timestamp_seconds = std::numeric_limits<double>::quiet_NaN();
  V camera_origin;
  M camera_world;
  V light_dir;
  float light_intensity = 0.0f;
};

// to_optical_state (include/dvs/raycast.hpp:17-22) for any state type with
// these members and Eigen-style accessors.
template <class State>
OpticalState to_optical_state(const State &s) {
  OpticalState p{};
  p.position = {s.position.x(), s.position.y(), s.position.z()};
  p.velocity = {s.linear_velocity.x(), s.linear_velocity.y(), s.linear_velocity.z()};
  p.light = {s.light_direction.x(), s.light_direction.y(), s.light_direction.z()};
  p.light_rate = {s.light_direction_rate.x(), s.light_direction_rate.y(), s.light_direction_rate.z()};
  p.intensity = s.light_intensity;
  p.intensity_rate = s.light_intensity_rate;
  p.rotation = {s.quaternion.x(), s.quaternion.y(), s.quaternion.z(), s.quaternion.w()};
  return p;
}
// An optical state is already one (callers that build intervals directly).
inline OpticalState to_optical_state(const OpticalState &s) { return s; }

// CameraTrajectory (include/dvs/trajectory.hpp:81-116, src/trajectory.cpp):
// the keys live in a dvs_trajectory; single evaluations run on the host,
// intervals() evaluates every endpoint of a window on the device.  Copies
// share the keys.
class CameraTrajectory {
 public:
  CameraTrajectory() = default;
  CameraTrajectory(const std::vector<FrameData<Vec3, Mat3>> &frames, SceneConfig &scene_config) {
    build(frames, scene_config.fps, scene_config.translation, scene_config.rotation);
  }
  void build(const std::vector<FrameData<Vec3, Mat3>> &frames, float fps, std::vector<float> &translation,
             std::vector<float> &rotation) {
    if (translation.size() != 3 || rotation.size() != 3)
      throw std::invalid_argument("trajectory requires positive fps and three-component mount transforms");
    std::vector<dvs_frame> c(frames.size());
    for (size_t i = 0; i < frames.size(); ++i) {
      const auto &f = frames[i];
      c[i].frame_index = f.frame_index;
      c[i].timestamp_seconds = f.timestamp_seconds;
      std::memcpy(c[i].camera_origin, f.camera_origin.data(), sizeof(c[i].camera_origin));
      std::memcpy(c[i].camera_world, f.camera_world.data(), sizeof(c[i].camera_world));
      std::memcpy(c[i].light_dir, f.light_dir.data(), sizeof(c[i].light_dir));
      c[i].light_intensity = f.light_intensity;
    }
    dvs_trajectory h = nullptr;
    detail::check(dvs_trajectory_create(c.data(), (int32_t)c.size(), fps, translation.data(), rotation.data(), &h));
    handle_.reset(h, [](dvs_trajectory p) { dvs_trajectory_destroy(p); });
  }
  bool valid() const { return size() > 0; }
  std::size_t size() const { return range().size; }
  float start_time() const { return range().start; }
  float end_time() const { return range().end; }
  CameraTrajectoryState evaluate(float t) const {
    CameraTrajectoryState out;
    out.timestamp = t;
    if (!handle_) return out;
    dvs_trajectory_state s;
    detail::check(dvs_trajectory_evaluate(handle_.get(), &t, 1, &s));
    out.timestamp = s.timestamp;
    out.position = Vec3(s.position[0], s.position[1], s.position[2]);
    std::memcpy(out.orientation.m, s.orientation, sizeof(s.orientation));
    std::memcpy(out.quaternion.q, s.quaternion, sizeof(s.quaternion));
    out.linear_velocity = Vec3(s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]);
    out.angular_velocity = Vec3(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2]);
    out.light_direction = Vec3(s.light_direction[0], s.light_direction[1], s.light_direction[2]);
    out.light_direction_rate =
        Vec3(s.light_direction_rate[0], s.light_direction_rate[1], s.light_direction_rate[2]);
    out.light_intensity = s.light_intensity;
    out.light_intensity_rate = s.light_intensity_rate;
    return out;
  }
  Vec3 position(float t) const { return evaluate(t).position; }
  Mat3 orientation(float t) const { return evaluate(t).orientation; }
  Quaternion quaternion(float t) const { return evaluate(t).quaternion; }
  Vec3 linear_velocity(float t) const { return evaluate(t).linear_velocity; }
  Vec3 angular_velocity(float t) const { return evaluate(t).angular_velocity; }
  Vec3 light_direction(float t) const { return evaluate(t).light_direction; }
  Vec3 light_direction_rate(float t) const { return evaluate(t).light_direction_rate; }
  float light_intensity(float t) const { return evaluate(t).light_intensity; }
  float light_intensity_rate(float t) const { return evaluate(t).light_intensity_rate; }
  // The intervals bin/dvs builds between consecutive host horizons
  // (apps/main_dvs.cpp:103-107), each endpoint evaluated once, on the device.
  std::vector<OpticalInterval> intervals(const std::vector<double> &times) const {
    if (!handle_) throw std::invalid_argument("trajectory has not been built");
    std::vector<OpticalInterval> out(times.size() > 1 ? times.size() - 1 : 0);
    detail::check(dvs_trajectory_intervals(handle_.get(), times.data(), (int64_t)times.size(),
                                           reinterpret_cast<dvs_optical_interval *>(out.data())));
    return out;
  }

  // The C-ABI trajectory (for the windows the sensor evaluates on the device).
  dvs_trajectory native_handle() const { return handle_.get(); }

 private:
  struct Range {
    float start = 0, end = 0;
    int32_t size = 0;
  };
  Range range() const {
    Range r;
    if (handle_) detail::check(dvs_trajectory_range(handle_.get(), &r.start, &r.end, &r.size));
    return r;
  }
  std::shared_ptr<dvs_trajectory_s> handle_;
};

// Scene bundle and trajectory files (src/scene_data.cpp:93-283) over
// libdvs_io.so: false with the reference's message in *err on failure.
template <class F3, class U3> struct MeshData {
  std::vector<F3> vertices, normals;
  std::vector<U3> indices;
};
namespace detail {
inline bool io_ok(int code, std::string *err) {
  if (code == DVS_OK) return true;
  if (err) *err = dvs_io_last_error();
  return false;
}
inline FrameData<Vec3, Mat3> from_c(const dvs_frame &c) {
  FrameData<Vec3, Mat3> f;
  f.frame_index = c.frame_index;
  f.timestamp_seconds = c.timestamp_seconds;
  f.camera_origin = Vec3(c.camera_origin[0], c.camera_origin[1], c.camera_origin[2]);
  std::memcpy(f.camera_world.m, c.camera_world, sizeof(c.camera_world));
  f.light_dir = Vec3(c.light_dir[0], c.light_dir[1], c.light_dir[2]);
  f.light_intensity = c.light_intensity;
  return f;
}
template <class Load>
bool load_frames(Load load, const std::string &path, std::vector<FrameData<Vec3, Mat3>> &out, std::string *err) {
  int64_t n = 0;
  if (!io_ok(load(path.c_str(), nullptr, 0, &n), err)) return false;
  std::vector<dvs_frame> c((size_t)n);
  if (!io_ok(load(path.c_str(), c.data(), n, &n), err)) return false;
  for (const dvs_frame &f : c) out.push_back(from_c(f));
  return true;
}
}  // namespace detail

inline bool load_mesh_data(const std::string &dir, MeshData<float3, uint3> &out, std::string *err) {
  int64_t nv = 0, nt = 0;
  if (!detail::io_ok(dvs_io_mesh_load(dir.c_str(), nullptr, nullptr, 0, nullptr, 0, &nv, &nt), err)) return false;
  out.vertices.resize((size_t)nv);
  out.normals.resize((size_t)nv);
  out.indices.resize((size_t)nt);
  return detail::io_ok(dvs_io_mesh_load(dir.c_str(), &out.vertices[0].x, &out.normals[0].x, nv, &out.indices[0].x,
                                        nt, &nv, &nt),
                       err);
}
inline bool load_frames_from_binary(const std::string &dir, std::vector<FrameData<Vec3, Mat3>> &out,
                                    std::string *err) {
  out.clear();
  return detail::load_frames(dvs_io_frames_binary, dir, out, err);
}
inline bool load_frames_from_text(const std::string &path, std::vector<FrameData<Vec3, Mat3>> &out,
                                  std::string *err) {
  return detail::load_frames(dvs_io_frames_text, path, out, err);
}
// load_scene_bundle(bundle_dir) (scene_data.cpp:200-236): mesh and frames.bin.
inline bool load_scene_bundle(const std::string &bundle_dir, SceneBundle<float3, uint3> &out_bundle,
                              std::vector<FrameData<Vec3, Mat3>> &frames, std::string *err) {
  MeshData<float3, uint3> mesh;
  if (!load_mesh_data(bundle_dir, mesh, err)) return false;
  std::cout << "Loaded mesh with " << mesh.vertices.size() << " vertices, " << mesh.normals.size()
            << " normals, and " << mesh.indices.size() << " triangles" << std::endl;
  if (!load_frames_from_binary(bundle_dir, frames, err)) return false;
  if (frames.empty()) {
    if (err) *err = "frames.bin is empty";
    return false;
  }
  out_bundle.bvh_triangles = std::move(mesh.indices);
  out_bundle.vertices = std::move(mesh.vertices);
  out_bundle.normals = std::move(mesh.normals);
  out_bundle.bvh_nodes.clear();
  return true;
}
// render_float_to_png_bytes (include/dvs/blenderize.hpp:50-69).
inline void render_float_to_png_bytes(const float *linear, int width, int height, float dither,
                                      std::vector<uint8_t> &out_bytes) {
  out_bytes.resize((size_t)width * height);
  detail::check_io(dvs_io_frame_bytes(linear, width, height, dither, out_bytes.data()));
}

// OpticalRefinementStats / ReceptorProbeResult of the adaptive solver.
using OpticalRefinementStats = dvs_refinement_stats;
using ReceptorProbeResult = dvs_receptor_probe;
using OpticalProbeResult = dvs_optical_probe;  // include/dvs/optical_sampler.hpp:15
constexpr double kDefaultOpticalLogTolerance = 1e-2;       // optical_sampler.hpp:6
constexpr double kOpticalVoltageThresholdFraction = 1e-3;  // optical_sampler.hpp:7

// DvsRayCaster (include/dvs/raycast.hpp:23-62).
class DvsRayCaster {
 public:
  DvsRayCaster(const SceneBundle<float3, uint3> &scene, const Config &config, int root_index)
      : DvsRayCaster(std::make_shared<DeviceScene>(scene, config, root_index), config) {}
  DvsRayCaster(std::shared_ptr<DeviceScene> scene, const Config &config) : scene_(std::move(scene)) {
    const dvs_physic p = detail::to_c(config.physic);
    detail::check(dvs_sensor_create(scene_->view(), &p, 1, 1, &sensor_));
  }
  ~DvsRayCaster() { dvs_sensor_destroy(sensor_); }
  DvsRayCaster(const DvsRayCaster &) = delete;
  DvsRayCaster &operator=(const DvsRayCaster &) = delete;

  template <class State>
  void initialize(const State &state, double timestamp) {
    const OpticalState o = to_optical_state(state);
    detail::check(dvs_sensor_initialize(sensor_, reinterpret_cast<const dvs_optical_state *>(&o), timestamp));
  }
  template <class State>
  void advance(const State &start, const State &end, double t0, double t1, std::vector<Event> &events) {
    advance_many({OpticalInterval{to_optical_state(start), to_optical_state(end), t0, t1}}, events);
  }
  // events receives the readout-delivered events of this call.
  void advance_many(const std::vector<OpticalInterval> &intervals, std::vector<Event> &events) {
    int64_t ready = 0;
    detail::check(dvs_sensor_process(sensor_, reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                     (int32_t)intervals.size(), &ready));
    take(events, ready);
  }
  int adaptive_cache_capacity() {
    int32_t c = 0;
    detail::check(dvs_sensor_adaptive_capacity(sensor_, &c));
    return c;
  }
  // false: the receptor check rejected the window; nothing changed and
  // events is empty.
  bool advance_adaptive(const std::vector<OpticalInterval> &intervals, double tolerance, std::vector<Event> &events,
                        bool check_convergence = false) {
    int32_t accepted = 0;
    int64_t ready = 0;
    detail::check(dvs_sensor_process_adaptive(sensor_,
                                              reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                              (int32_t)intervals.size(), tolerance, check_convergence ? 1 : 0,
                                              &accepted, &ready));
    take(events, ready);
    return accepted != 0;
  }
  // advance_adaptive over the intervals between consecutive `times` of
  // `trajectory`, their states evaluated on the device inside the window
  // (dvs_sensor_process_adaptive_times): the events advance_adaptive gives
  // for trajectory.intervals(times).  No reference counterpart.
  bool advance_adaptive(const CameraTrajectory &trajectory, const std::vector<double> &times, double tolerance,
                        std::vector<Event> &events, bool check_convergence = false) {
    int32_t accepted = 0;
    int64_t ready = 0;
    detail::check(dvs_sensor_process_adaptive_times(sensor_, trajectory.native_handle(), times.data(),
                                                    (int32_t)times.size(), tolerance, check_convergence ? 1 : 0,
                                                    &accepted, &ready));
    take(events, ready);
    return accepted != 0;
  }
  // The window the next advance_adaptive(trajectory, times, ...) call will
  // most likely run: when the current call's window is accepted, its setup
  // is queued on the device before that call returns
  // (dvs_sensor_expect_adaptive_times).  Events are the same either way; an
  // empty `times` forgets it.  No reference counterpart.
  void expect_adaptive(const CameraTrajectory &trajectory, const std::vector<double> &times, double tolerance) {
    detail::check(dvs_sensor_expect_adaptive_times(sensor_, times.empty() ? nullptr : trajectory.native_handle(),
                                                   times.data(), (int32_t)times.size(), tolerance));
  }
  ReceptorProbeResult last_convergence() const {
    ReceptorProbeResult r{};
    detail::check(dvs_sensor_last_convergence(sensor_, &r));
    return r;
  }
  OpticalRefinementStats refinement_statistics() const {
    OpticalRefinementStats r{};
    detail::check(dvs_sensor_refinement_stats(sensor_, &r));
    return r;
  }
  // raycast.hpp:45-53: every later adaptive window measures its receptor
  // check; verify_receptor returns that of the last window `intervals`,
  // verify_optics its knots against native traces (std::logic_error before
  // any adaptive window, or without the diagnostics for verify_receptor).
  void enable_numerical_diagnostics() { detail::check(dvs_sensor_enable_diagnostics(sensor_)); }
  ReceptorProbeResult verify_receptor(const std::vector<OpticalInterval> &intervals) {
    ReceptorProbeResult r{};
    detail::check_logic(dvs_sensor_verify_receptor(sensor_, reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                             (int32_t)intervals.size(), &r));
    return r;
  }
  OpticalProbeResult verify_optics(const std::vector<OpticalInterval> &intervals, double tolerance) {
    OpticalProbeResult r{};
    detail::check_logic(dvs_sensor_verify_optics(sensor_, reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                           (int32_t)intervals.size(), tolerance, &r));
    return r;
  }
  void drain(std::vector<Event> &events) {
    int64_t ready = 0;
    detail::check(dvs_sensor_drain(sensor_, &ready));
    take(events, ready);
  }
  std::size_t dropped_events() const {
    int64_t dropped = 0, peak = 0;
    detail::check(dvs_sensor_readout_stats(sensor_, &dropped, &peak));
    return (std::size_t)dropped;
  }
  std::size_t peak_queue() const {
    int64_t dropped = 0, peak = 0;
    detail::check(dvs_sensor_readout_stats(sensor_, &dropped, &peak));
    return (std::size_t)peak;
  }

 private:
  void take(std::vector<Event> &events, int64_t ready) {
    events.resize((size_t)ready);
    int64_t got = 0;
    detail::check(dvs_sensor_take(sensor_, events.data(), ready, &got));
    events.resize((size_t)got);
  }
  std::shared_ptr<DeviceScene> scene_;
  dvs_sensor sensor_ = nullptr;
};

// CamRayCaster (include/dvs/raycast.hpp:64-86).
class CamRayCaster {
 public:
  CamRayCaster(const SceneBundle<float3, uint3> &scene, const Config &config, const int root_index)
      : CamRayCaster(std::make_shared<DeviceScene>(scene, config, root_index)) {}
  explicit CamRayCaster(std::shared_ptr<DeviceScene> scene) : scene_(std::move(scene)) {}

  // camera_world: camera-to-world rotation, 3x3 column major.  The pose's
  // quaternion is the matrix's, by the trace / largest-diagonal method
  // (Shepperd 1978), normalised.
  void render_main(const float3 &camera_origin, const float *camera_world, const float3 &light_dir,
                   const float light_intensity, std::vector<float> &light_frame, std::vector<float> &depth_map) {
    dvs_optical_state pose{};
    pose.position[0] = camera_origin.x;
    pose.position[1] = camera_origin.y;
    pose.position[2] = camera_origin.z;
    pose.light[0] = light_dir.x;
    pose.light[1] = light_dir.y;
    pose.light[2] = light_dir.z;
    pose.intensity = light_intensity;
    quaternion_of(camera_world, pose.rotation);
    const size_t n = (size_t)scene_->width() * scene_->height();
    light_frame.resize(n);
    depth_map.resize(n);
    detail::check(dvs_frame_render(scene_->view(), &pose, 1, light_frame.data(), depth_map.data(), nullptr));
  }
  void render_main(const float3 &camera_origin, const float *camera_world, const float3 &light_dir,
                   const float light_intensity, std::vector<float> &light_frame) {
    std::vector<float> depth;
    render_main(camera_origin, camera_world, light_dir, light_intensity, light_frame, depth);
  }

  static void quaternion_of(const float *m, float q[4]) {  // m column major: m[c*3+r]
    auto at = [m](int r, int c) { return m[c * 3 + r]; };
    float x, y, z, w;
    const float trace = at(0, 0) + at(1, 1) + at(2, 2);
    if (trace > 0) {
      float t = std::sqrt(trace + 1.0f);
      w = 0.5f * t;
      t = 0.5f / t;
      x = (at(2, 1) - at(1, 2)) * t;
      y = (at(0, 2) - at(2, 0)) * t;
      z = (at(1, 0) - at(0, 1)) * t;
    } else {
      int i = 0;
      if (at(1, 1) > at(0, 0)) i = 1;
      if (at(2, 2) > at(i, i)) i = 2;
      const int j = (i + 1) % 3, k = (j + 1) % 3;
      float v[3];
      float t = std::sqrt(at(i, i) - at(j, j) - at(k, k) + 1.0f);
      v[i] = 0.5f * t;
      t = 0.5f / t;
      w = (at(k, j) - at(j, k)) * t;
      v[j] = (at(j, i) + at(i, j)) * t;
      v[k] = (at(k, i) + at(i, k)) * t;
      x = v[0];
      y = v[1];
      z = v[2];
    }
    const float norm = std::sqrt(x * x + y * y + z * z + w * w);
    q[0] = x / norm;
    q[1] = y / norm;
    q[2] = z / norm;
    q[3] = w / norm;
  }

  // One sub-exposure cell of a motion-blur exposure: the render_main
  // arguments of its pose and its weight in the exposure's sum.
  struct BlurCell {
    float3 camera_origin{};
    float camera_world[9]{};
    float3 light_dir{};
    float light_intensity = 0;
    double weight = 0;
  };
  // Motion-blur exposures: the render_main loop of apps/main_cam_blur.cpp:
  // 70-77 (and, with `wide`, apps/main_cam_blur_windows.cpp:70-81) over
  // whole groups of exposures per launch.  Exposure e sums cells
  // [offsets[e], offsets[e+1]) in cell order; `frames` receives
  // (offsets.size() - 1) frames of width * height floats.
  void render_blur(const std::vector<BlurCell> &cells, const std::vector<int32_t> &offsets, bool wide,
                   std::vector<float> &frames) {
    if (offsets.size() < 2 || offsets.back() != (int32_t)cells.size())
      throw std::invalid_argument("motion-blur offsets must end at the number of cells");
    std::vector<dvs_optical_state> poses(cells.size());
    std::vector<double> weights(cells.size());
    for (size_t c = 0; c < cells.size(); ++c) {
      dvs_optical_state &p = poses[c];
      p = dvs_optical_state{};
      p.position[0] = cells[c].camera_origin.x;
      p.position[1] = cells[c].camera_origin.y;
      p.position[2] = cells[c].camera_origin.z;
      p.light[0] = cells[c].light_dir.x;
      p.light[1] = cells[c].light_dir.y;
      p.light[2] = cells[c].light_dir.z;
      p.intensity = cells[c].light_intensity;
      quaternion_of(cells[c].camera_world, p.rotation);
      weights[c] = cells[c].weight;
    }
    const int exposures = (int)offsets.size() - 1;
    frames.resize((size_t)exposures * scene_->width() * scene_->height());
    detail::check(dvs_frame_blur(scene_->view(), poses.data(), weights.data(), offsets.data(), exposures,
                                 wide ? 1 : 0, frames.data()));
  }

 private:
  std::shared_ptr<DeviceScene> scene_;
};

// PixelParameters / PixelState / PixelEvent and the receptor model
// (include/dvs/pixel_model.hpp:7-102), host side.
struct PixelParameters {
  double threshold_pos = .3, threshold_neg = .3;
  double charge_scale = 200, refractory_time = .0002;
  double reference_leak = 0, reset_noise_std = 0;
};
struct PixelState {
  double receptor_log = 0, reference_log = 0, refractory_until = -1;
  double threshold_pos = .3, threshold_neg = .3;
  bool reset_pending = false;
};
struct PixelEvent { double time; int polarity; };
inline double receptor_log_at(double u, double rate, double dt, double charge) {
  return dvs_receptor_log_at(u, rate, dt, charge);
}
inline double contrast_at(double u, double rate, double dt, double ref, const PixelParameters &p) {
  return receptor_log_at(u, rate, dt, p.charge_scale) - ref + p.reference_leak * dt;
}
// advance_pixel<Rng> (pixel_model.hpp:35-102): rng.normal() supplies every
// reset-noise draw.  Returns the crossings written to out, -1 past capacity.
template <class Rng>
int advance_pixel(PixelState &s, const PixelParameters &p, double rate, double t0, double t1, PixelEvent *out,
                  int capacity, Rng &rng) {
  dvs_pixel_state c{};
  c.receptor_log = s.receptor_log;
  c.reference_log = s.reference_log;
  c.refractory_until = s.refractory_until;
  c.threshold_pos = s.threshold_pos;
  c.threshold_neg = s.threshold_neg;
  c.reset_pending = s.reset_pending ? 1 : 0;
  const dvs_pixel_params q{p.threshold_pos, p.threshold_neg, p.charge_scale, p.refractory_time, p.reference_leak,
                           p.reset_noise_std};
  std::vector<dvs_crossing> buf((size_t)std::max(capacity, 0));
  double (*normal)(void *) = [](void *context) -> double { return static_cast<Rng *>(context)->normal(); };
  const int n = dvs_advance_pixel(&c, &q, rate, t0, t1, buf.data(), capacity, normal, &rng);
  if (n == -2) throw std::invalid_argument("invalid pixel circuit arguments");
  s.receptor_log = c.receptor_log;
  s.reference_log = c.reference_log;
  s.refractory_until = c.refractory_until;
  s.reset_pending = c.reset_pending != 0;
  for (int i = 0; i < n; ++i) out[i] = PixelEvent{buf[i].time, buf[i].polarity};
  return n;
}

// Hapke BRDF (include/dvs/hapke.hpp:131-172), host side.
inline float hapke_reflectance(float cosi, float cose, float cosg, float albedo, const HapkeConfig &p) {
  const dvs_hapke h = detail::to_c(p);
  return dvs_hapke_reflectance(cosi, cose, cosg, albedo, &h);
}
inline float hapke_directional_reflectance(const float3 &normal, const float3 &light, const float3 &view,
                                           float albedo, const HapkeConfig &p) {
  const dvs_hapke h = detail::to_c(p);
  const float n[3] = {normal.x, normal.y, normal.z}, l[3] = {light.x, light.y, light.z},
              v[3] = {view.x, view.y, view.z};
  return dvs_hapke_directional(n, l, v, albedo, &h);
}
inline float invert_hapke_if(float inc, float emi, float phase, float target, const HapkeConfig &p) {
  const dvs_hapke h = detail::to_c(p);
  return dvs_invert_hapke_if(inc, emi, phase, target, &h);
}

// EventReadout (include/dvs/event_sensor.hpp:24-37): the bounded FIFO on
// its own.  submit sorts by (t, y, x), queues, and returns in `ready` the
// events delivered by `now`.
class EventReadout {
 public:
  EventReadout(double bandwidth, double latency, std::size_t capacity) {
    detail::check(dvs_readout_create(bandwidth, latency, (int64_t)capacity, &r_));
  }
  ~EventReadout() { dvs_readout_destroy(r_); }
  EventReadout(const EventReadout &) = delete;
  EventReadout &operator=(const EventReadout &) = delete;
  void submit(std::vector<Event> events, double now, std::vector<Event> &ready) {
    int64_t n = 0;
    detail::check(dvs_readout_submit(r_, events.data(), (int64_t)events.size(), now, &n));
    take(ready, n);
  }
  void drain(std::vector<Event> &ready) {
    int64_t n = 0;
    detail::check(dvs_readout_drain(r_, &n));
    take(ready, n);
  }
  std::size_t dropped() const {
    int64_t d = 0, p = 0;
    detail::check(dvs_readout_stats(r_, &d, &p));
    return (std::size_t)d;
  }
  std::size_t peak_queue() const {
    int64_t d = 0, p = 0;
    detail::check(dvs_readout_stats(r_, &d, &p));
    return (std::size_t)p;
  }

 private:
  void take(std::vector<Event> &ready, int64_t n) {
    ready.resize((size_t)n);
    int64_t got = 0;
    detail::check(dvs_readout_take(r_, ready.data(), n, &got));
    ready.resize((size_t)got);
  }
  dvs_readout r_ = nullptr;
};

// OpticalScene: the scene view the sensor classes take, DeviceScene::view().
using OpticalScene = dvs_scene;
// OpticalValue (optical_sampler.hpp:8): mean radiance, mean hit depth (-1
// without a hit) and the sub-ray grid size of one pixel.
struct OpticalValue {
  double radiance = 0;
  float depth = -1;
  int samples = 0;
};

// OpticalSampler (include/dvs/optical_sampler.hpp:17-48).  evaluate renders
// poses into host memory (the values stay readable through the returned
// pointer and download).  The adaptive refinement runs inside
// EventSensor::process_adaptive; statistics() and knot_capacity() report
// the sensor the sampler last refined with (zero before).
class OpticalSampler {
 public:
  OpticalSampler(int width, int height, int maximum_samples)
      : width_(width), height_(height), maximum_(maximum_samples) {}
  OpticalSampler(const OpticalSampler &) = delete;
  OpticalSampler &operator=(const OpticalSampler &) = delete;
  const OpticalValue *evaluate(const OpticalScene &scene, const std::vector<OpticalState> &states) {
    dvs_camera cam{};
    detail::check(dvs_scene_camera(scene, &cam));
    const size_t n = (size_t)width_ * height_;
    if (states.empty() || cam.width * cam.height != width_ * height_ || cam.samples != maximum_)
      throw std::invalid_argument("invalid parallel optical request");
    std::vector<float> image(n * states.size()), depth(n * states.size());
    std::vector<double> radiance(n * states.size());
    detail::check(dvs_frame_render(scene, reinterpret_cast<const dvs_optical_state *>(states.data()),
                                   (int32_t)states.size(), image.data(), depth.data(), radiance.data()));
    values_.resize(n * states.size());
    for (size_t i = 0; i < values_.size(); ++i) values_[i] = OpticalValue{radiance[i], depth[i], maximum_};
    return values_.data();
  }
  void download(std::vector<OpticalValue> &values) const { values = values_; }
  OpticalRefinementStats statistics() const {
    OpticalRefinementStats r{};
    if (sensor_) detail::check(dvs_sensor_refinement_stats(sensor_, &r));
    return r;
  }
  int knot_capacity() const {
    int32_t c = 0;
    if (sensor_) detail::check(dvs_sensor_adaptive_capacity(sensor_, &c));
    return c;
  }

 private:
  friend class EventSensor;
  int width_, height_, maximum_;
  std::vector<OpticalValue> values_;
  dvs_sensor sensor_ = nullptr;  // the sensor of the last refinement
};

// EventSensor (include/dvs/event_sensor.hpp:40-79).  The device state is
// created for the scene of the first initialize; later calls take the same
// scene.
class EventSensor {
 public:
  EventSensor(int width, int height, const PhysicConfig &config)
      : width_(width), height_(height), physic_(detail::to_c(config)) {
    if (width <= 0 || height <= 0 || config.max_events_per_step < 1)
      throw std::invalid_argument("invalid event sensor dimensions or capacity");
  }
  ~EventSensor() { dvs_sensor_destroy(sensor_); }
  EventSensor(const EventSensor &) = delete;
  EventSensor &operator=(const EventSensor &) = delete;
  void initialize(const OpticalScene &scene, const OpticalState &state, double timestamp) {
    bind(scene);
    detail::check(dvs_sensor_initialize(sensor_, reinterpret_cast<const dvs_optical_state *>(&state), timestamp));
  }
  void process(const OpticalScene &scene, const OpticalInterval &interval, std::vector<Event> &events) {
    process_many(scene, {interval}, events);
  }
  void process_many(const OpticalScene &scene, const std::vector<OpticalInterval> &intervals,
                    std::vector<Event> &events) {
    bind(scene);
    int64_t ready = 0;
    detail::check(dvs_sensor_process(sensor_, reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                     (int32_t)intervals.size(), &ready));
    take(events, ready);
  }
  bool process_adaptive(const OpticalScene &scene, const std::vector<OpticalInterval> &intervals,
                        OpticalSampler &sampler, double log_tolerance, std::vector<Event> &events,
                        bool check_convergence = false) {
    bind(scene);
    int32_t accepted = 0;
    int64_t ready = 0;
    detail::check(dvs_sensor_process_adaptive(sensor_,
                                              reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                              (int32_t)intervals.size(), log_tolerance, check_convergence ? 1 : 0,
                                              &accepted, &ready));
    sampler.sensor_ = sensor_;
    take(events, ready);
    return accepted != 0;
  }
  OpticalProbeResult verify_optics(const OpticalScene &scene, OpticalSampler &sampler,
                                   const std::vector<OpticalInterval> &ivs, double tolerance) {
    bind(scene);
    if (sampler.sensor_ != sensor_) throw std::logic_error("refine before validating optical knots");
    OpticalProbeResult r{};
    detail::check_logic(dvs_sensor_verify_optics(sensor_, reinterpret_cast<const dvs_optical_interval *>(ivs.data()),
                                           (int32_t)ivs.size(), tolerance, &r));
    return r;
  }
  const ReceptorProbeResult &last_convergence() const {
    if (sensor_) detail::check(dvs_sensor_last_convergence(sensor_, &convergence_));
    return convergence_;
  }
  void enable_numerical_diagnostics() {
    diagnostics_ = true;
    if (sensor_) detail::check(dvs_sensor_enable_diagnostics(sensor_));
  }
  ReceptorProbeResult verify_receptor(const OpticalScene &scene, OpticalSampler &,
                                      const std::vector<OpticalInterval> &intervals) {
    bind(scene);
    ReceptorProbeResult r{};
    detail::check_logic(dvs_sensor_verify_receptor(sensor_, reinterpret_cast<const dvs_optical_interval *>(intervals.data()),
                                             (int32_t)intervals.size(), &r));
    return r;
  }
  void drain(std::vector<Event> &events) {
    if (!sensor_) { events.clear(); return; }
    int64_t ready = 0;
    detail::check(dvs_sensor_drain(sensor_, &ready));
    take(events, ready);
  }
  std::size_t dropped() const { return readout_stat(false); }
  std::size_t peak_queue() const { return readout_stat(true); }

 private:
  void bind(const OpticalScene &scene) {
    if (sensor_) {
      if (scene != scene_) throw std::invalid_argument("EventSensor is bound to the scene it was initialized with");
      return;
    }
    dvs_camera cam{};
    detail::check(dvs_scene_camera(scene, &cam));
    if (cam.width != width_ || cam.height != height_)
      throw std::invalid_argument("event sensor dimensions differ from the scene's camera");
    detail::check(dvs_sensor_create(scene, &physic_, 1, 1, &sensor_));
    scene_ = scene;
    if (diagnostics_) detail::check(dvs_sensor_enable_diagnostics(sensor_));
  }
  std::size_t readout_stat(bool peak) const {
    if (!sensor_) return 0;
    int64_t d = 0, p = 0;
    detail::check(dvs_sensor_readout_stats(sensor_, &d, &p));
    return (std::size_t)(peak ? p : d);
  }
  void take(std::vector<Event> &events, int64_t ready) {
    events.resize((size_t)ready);
    int64_t got = 0;
    detail::check(dvs_sensor_take(sensor_, events.data(), ready, &got));
    events.resize((size_t)got);
  }
  int width_, height_;
  dvs_physic physic_;
  dvs_sensor sensor_ = nullptr;
  OpticalScene scene_ = nullptr;
  bool diagnostics_ = false;
  mutable ReceptorProbeResult convergence_{};
};

// EventStreamWriter / BinaryEventStreamWriter / Hdf5EventStreamWriter
// (include/dvs/dvs_writer.hpp) and make_event_writer (apps/main_dvs.cpp:23-34).
// DatasetSizeLimit / DatasetBudget (include/dvs/dataset_budget.hpp:18-114):
// every write to a guarded dataset root is checked, under a cross-process
// lock, against a stop level and a hard ceiling on the apparent bytes below
// the root, with the caller's bound on the growth.
class DatasetSizeLimit : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class DatasetBudget {
 public:
  explicit DatasetBudget(const std::filesystem::path &root, uint64_t hard = 500000000000ULL,
                         uint64_t stop = 480000000000ULL)
      : root_(std::filesystem::canonical(root)), hard_(hard), stop_(stop) {
    if (stop_ == 0 || stop_ > hard_ || hard_ > 500000000000ULL)
      throw std::invalid_argument("invalid dataset capacity limits");
    uint64_t key = 14695981039346656037ULL;  // FNV-1a of the root, names the lock outside it
    for (unsigned char c : root_.string()) {
      key ^= c;
      key *= 1099511628211ULL;
    }
    std::ostringstream name;
    name << "dvs-dataset-" << std::hex << key << ".lock";
    lock_path_ = std::filesystem::temp_directory_path() / name.str();
  }
  void require_path(const std::filesystem::path &p) const {
    const auto relative = std::filesystem::weakly_canonical(p).lexically_relative(root_);
    if (relative.empty() || relative.is_absolute() || *relative.begin() == "..")
      throw std::runtime_error("output is outside the guarded dataset: " + p.string());
  }
  uint64_t logical_bytes() const {
    uint64_t total = 0;
    for (const auto &e : std::filesystem::recursive_directory_iterator(root_)) {
      if (e.is_symlink()) throw std::runtime_error("dataset symlinks are not supported by capacity accounting");
      if (e.is_regular_file()) {
        const uint64_t n = e.file_size();
        if (n > std::numeric_limits<uint64_t>::max() - total) throw DatasetSizeLimit("dataset size overflow");
        total += n;
      }
    }
    return total;
  }
  const std::filesystem::path &root() const { return root_; }
  uint64_t hard_limit() const { return hard_; }
  uint64_t stop_limit() const { return stop_; }
  // The operation must flush what it writes; max_growth bounds the growth.
  template <class F> void write(uint64_t max_growth, F operation) const {
    Lock lock(lock_path_);
    const uint64_t before = logical_bytes();
    if (before > hard_ || before >= stop_ || max_growth > stop_ - before)
      throw DatasetSizeLimit("dataset capacity stop: current=" + std::to_string(before) +
                             ", reserved_write=" + std::to_string(max_growth) + ", stop=" + std::to_string(stop_) +
                             ", hard=" + std::to_string(hard_) + " bytes");
    const uint64_t available = std::filesystem::space(root_).available;
    if (max_growth > available || available - max_growth < 128ULL * 1024 * 1024)
      throw DatasetSizeLimit("insufficient filesystem space for reserved dataset write");
    operation();
    const uint64_t after = logical_bytes();
    if (after > before + max_growth)
      throw DatasetSizeLimit("writer exceeded its declared storage reservation; batch must stop");
    if (after > hard_) throw DatasetSizeLimit("dataset hard size limit reached; batch must stop");
  }
  // Metadata through a temporary sibling, renamed over the old file.
  void write_text(const std::filesystem::path &p, const std::string &text) const {
    require_path(p);
    write(text.size(), [&] {
      auto partial = p;
      partial += ".partial";
      std::ofstream out(partial, std::ios::binary | std::ios::trunc);
      if (!out) throw std::runtime_error("cannot open dataset metadata: " + partial.string());
      out.write(text.data(), (std::streamsize)text.size());
      out.flush();
      if (!out) throw std::runtime_error("cannot write dataset metadata: " + partial.string());
      out.close();
      std::filesystem::rename(partial, p);
    });
  }

 private:
  class Lock {
   public:
    explicit Lock(const std::filesystem::path &path) {
      fd_ = ::open(path.c_str(), O_CREAT | O_RDWR | O_CLOEXEC, 0600);
      if (fd_ < 0) throw std::runtime_error("cannot open dataset capacity lock: " + std::string(std::strerror(errno)));
      int result;
      do {
        result = flock(fd_, LOCK_EX);
      } while (result < 0 && errno == EINTR);
      if (result < 0) {
        ::close(fd_);
        fd_ = -1;
        throw std::runtime_error("cannot lock dataset capacity budget");
      }
    }
    ~Lock() {
      if (fd_ >= 0) {
        flock(fd_, LOCK_UN);
        ::close(fd_);
      }
    }
    Lock(const Lock &) = delete;
    Lock &operator=(const Lock &) = delete;

   private:
    int fd_ = -1;
  };
  std::filesystem::path root_, lock_path_;
  uint64_t hard_, stop_;
};

// EventStreamWriter and its events.bin / events.h5 forms
// (include/dvs/dvs_writer.hpp:13-264); with a budget every file operation
// is a budgeted write with the reference's reservations.
class EventStreamWriter {
 public:
  EventStreamWriter(std::string output_path, int32_t format, const DatasetBudget *budget = nullptr)
      : path_(std::move(output_path)), format_(format), budget_(budget) {
    if (budget_) budget_->require_path(path_);
    auto open = [&] { detail::check_io(dvs_writer_open(path_.c_str(), format, &writer_)); };
    if (budget_) budget_->write(format == DVS_EVENTS_HDF5 ? 4 * 1024 * 1024 : 0, open);
    else open();
  }
  virtual ~EventStreamWriter() { dvs_writer_destroy(writer_); }
  EventStreamWriter(const EventStreamWriter &) = delete;
  EventStreamWriter &operator=(const EventStreamWriter &) = delete;
  void append(std::vector<Event> events) {
    if (events.empty()) return;
    auto write = [&] {
      detail::check_io(dvs_writer_append(writer_, events.data(), (int64_t)events.size()));
      if (budget_) detail::check_io(dvs_writer_flush(writer_));
    };
    uint64_t reservation = events.size() * 13;  // BinaryRecord
    if (format_ == DVS_EVENTS_HDF5) {
      const uint64_t chunk = 1 << 16;
      reservation = ((events.size() + chunk - 1) / chunk + 1) * chunk * 13 + 4 * 1024 * 1024;
    }
    if (budget_) budget_->write(reservation, write);
    else write();
    for (const Event &e : events) last_time_ = std::max(last_time_, e.timestamp);
  }
  void close(double end_time_seconds) {
    auto finish = [&] { detail::check_io(dvs_writer_close(writer_, end_time_seconds)); };
    uint64_t reservation = 0;
    if (format_ == DVS_EVENTS_HDF5) {
      const double end_ms = std::max(end_time_seconds, last_time_) * 1000.0;
      const uint64_t index_count = (uint64_t)std::max(0.0, std::ceil(end_ms) + 2.0);
      reservation = index_count * sizeof(int64_t) + 4 * 1024 * 1024;
    }
    if (budget_) budget_->write(reservation, finish);
    else finish();
  }
  const std::string &path() const { return path_; }
  std::size_t event_count() const {
    int64_t n = 0;
    detail::check_io(dvs_writer_count(writer_, &n));
    return (std::size_t)n;
  }

 private:
  std::string path_;
  int32_t format_;
  const DatasetBudget *budget_;
  dvs_writer writer_ = nullptr;
  double last_time_ = 0;
};
class BinaryEventStreamWriter final : public EventStreamWriter {
 public:
  explicit BinaryEventStreamWriter(std::string output_path, const DatasetBudget *budget = nullptr)
      : EventStreamWriter(std::move(output_path), DVS_EVENTS_BIN, budget) {}
};
class Hdf5EventStreamWriter final : public EventStreamWriter {
 public:
  explicit Hdf5EventStreamWriter(std::string output_path, const DatasetBudget *budget = nullptr)
      : EventStreamWriter(std::move(output_path), DVS_EVENTS_HDF5, budget) {}
};
inline std::unique_ptr<EventStreamWriter> make_event_writer(const std::string &event_format,
                                                            const std::string &output_dir) {
  if (event_format == "hdf5") return std::make_unique<Hdf5EventStreamWriter>(output_dir + "/events.h5");
  if (event_format == "bin") return std::make_unique<BinaryEventStreamWriter>(output_dir + "/events.bin");
  throw std::runtime_error("unsupported event format: " + event_format);
}

// Hdf5FrameWriter (include/dvs/frame_writer.hpp:12-77): frames.h5 with t,
// images (the sRGB preview bytes) and depth, every write budgeted.
class Hdf5FrameWriter {
 public:
  Hdf5FrameWriter(const std::filesystem::path &path, int width, int height, const DatasetBudget &budget)
      : width_(width), height_(height), budget_(budget) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("invalid frame dimensions");
    budget.require_path(path);
    budget.write(4 * 1024 * 1024, [&] { detail::check_io(dvs_frames_open(path.c_str(), width, height, &frames_)); });
  }
  ~Hdf5FrameWriter() { dvs_frames_destroy(frames_); }
  Hdf5FrameWriter(const Hdf5FrameWriter &) = delete;
  Hdf5FrameWriter &operator=(const Hdf5FrameWriter &) = delete;
  void append(double time, const std::vector<float> &radiance, const std::vector<float> &depth) {
    const size_t pixels = (size_t)width_ * height_;
    if (radiance.size() != pixels || depth.size() != pixels)
      throw std::invalid_argument("invalid frame timestamp or array size");
    budget_.write(pixels * 5 + 4 * 1024 * 1024,
                  [&] { detail::check_io(dvs_frames_append(frames_, time, radiance.data(), depth.data())); });
  }
  void close() {
    budget_.write(4 * 1024 * 1024, [&] { detail::check_io(dvs_frames_close(frames_)); });
  }
  std::size_t count() const {
    int64_t n = 0;
    detail::check_io(dvs_frames_count(frames_, &n));
    return (std::size_t)n;
  }

 private:
  int width_, height_;
  const DatasetBudget &budget_;
  dvs_frames frames_ = nullptr;
};

// RepeatedTrajectory (include/dvs/repeated_trajectory.hpp:8-35): the source
// path at its own time scale, back and forth, with C2 reversals of
// turn_seconds at each end (phase = turn * (3x^2 - 3x^3 + x^4), rate 6x -
// 9x^2 + 4x^3); the state's rates are scaled by the phase rate.
class RepeatedTrajectory {
 public:
  explicit RepeatedTrajectory(const CameraTrajectory &source, double turn_seconds = .5)
      : source_(source), duration_(source.end_time() - source.start_time()),
        turn_(std::min(turn_seconds, duration_ / 4)) {
    if (!std::isfinite(duration_) || !std::isfinite(turn_seconds) || duration_ <= 0 || turn_seconds <= 0)
      throw std::invalid_argument("finite positive trajectory and turn duration required");
  }
  CameraTrajectoryState evaluate(double time) const {
    if (!std::isfinite(time)) throw std::invalid_argument("finite trajectory query time required");
    double r = std::fmod(time, 2 * duration_);
    if (r < 0) r += 2 * duration_;
    double phase = r <= duration_ ? r : 2 * duration_ - r;
    double rate = r <= duration_ ? 1 : -1;
    if (phase < turn_) {
      const double x = phase / turn_;
      phase = turn_ * (3 * x * x - 3 * x * x * x + x * x * x * x);
      rate *= 6 * x - 9 * x * x + 4 * x * x * x;
    } else if (duration_ - phase < turn_) {
      const double x = (duration_ - phase) / turn_;
      phase = duration_ - turn_ * (3 * x * x - 3 * x * x * x + x * x * x * x);
      rate *= 6 * x - 9 * x * x + 4 * x * x * x;
    }
    CameraTrajectoryState state = source_.evaluate((float)(source_.start_time() + phase));
    state.timestamp = (float)time;
    const float f = (float)rate;  // a float vector scaled by a double: the scale is a float
    for (int i = 0; i < 3; ++i) {
      state.linear_velocity.v[i] *= f;
      state.angular_velocity.v[i] *= f;
      state.light_direction_rate.v[i] *= f;
    }
    state.light_intensity_rate = (float)(state.light_intensity_rate * rate);
    return state;
  }
  double source_duration() const { return duration_; }

 private:
  const CameraTrajectory &source_;
  double duration_, turn_;
};

}  // namespace dvs

// ---- Terrain from rasters (include/tools/loader.hpp,
// src/tools/build_from_raster.cu, src/scene_data.cpp:284-395), without GDAL:
// GeoTIFF rasters through libdvs_io, the mesh built on the device.
namespace tools {

// Imagery (loader.hpp:17-22); projection holds the GeoKey identity of the
// raster's CRS rather than GDAL's WKT.
struct Imagery {
  std::vector<float> data, lats, lons, x_coords, y_coords;
  int width = 0, height = 0;
  double transform[6]{};
  std::string projection;
};

// loadImageryData (loader.hpp:31-100): messages on std::cerr and false on
// failure, the NoData count on std::cout.
inline bool loadImageryData(const std::string &path, Imagery &imagery) {
  dvs_raster r = nullptr;
  if (dvs_io_raster_load(path.c_str(), &r) != DVS_OK) {
    std::cerr << dvs_io_last_error() << std::endl;
    return false;
  }
  dvs_raster_info info{};
  dvs_io_raster_info(r, &info);
  imagery.width = info.width;
  imagery.height = info.height;
  imagery.data.resize((size_t)info.width * info.height);
  imagery.x_coords.resize((size_t)info.width);
  imagery.y_coords.resize((size_t)info.height);
  dvs_io_raster_data(r, imagery.data.data(), imagery.x_coords.data(), imagery.y_coords.data());
  std::copy(info.transform, info.transform + 6, imagery.transform);
  int64_t length = 0;
  dvs_io_raster_projection(r, nullptr, 0, &length);
  std::vector<char> proj((size_t)length + 1);
  dvs_io_raster_projection(r, proj.data(), length + 1, &length);
  imagery.projection.assign(proj.data(), (size_t)length);
  dvs_io_raster_free(r);
  if (info.missing)
    std::cout << path << ": " << info.missing << " NoData samples (NaN), scale=" << info.scale
              << ", offset=" << info.offset << "\n";
  return true;
}

// reference_height (loader.hpp:102-106).
inline float reference_height(const Imagery &dem) {
  const float center = dem.data[(size_t)dem.width * (dem.height / 2) + dem.width / 2];
  if (std::isfinite(center)) return center;
  throw std::runtime_error("DEM center is NoData; choose a valid crop/reference origin before rendering");
}

// require_same_projection (loader.hpp:107-114).
inline void require_same_projection(const Imagery &reference, const Imagery &source) {
  if (reference.projection == source.projection) return;
  throw std::runtime_error("Raster CRS differ; reproject to a common CRS before alignment");
}

namespace detail {
inline void raster_mesh(const Imagery &dem, const float *albedo, std::vector<dvs::float3> &vertices,
                        std::vector<dvs::float3> &normals, std::vector<dvs::uint3> &indices,
                        std::vector<float> *albedos) {
  const size_t npix = (size_t)dem.width * dem.height;
  const size_t ncell = (size_t)std::max(dem.width - 1, 0) * std::max(dem.height - 1, 0);
  vertices.resize(std::max<size_t>(npix, 1));
  normals.resize(std::max<size_t>(npix, 1));
  indices.resize(std::max<size_t>(2 * ncell, 1));
  if (albedos) albedos->resize(std::max<size_t>(npix, 1));
  int64_t nv = 0, nt = 0;
  const int code = dvs_raster_mesh(dem.data.data(), albedo, dem.x_coords.data(), dem.y_coords.data(), dem.width,
                                   dem.height, &vertices[0].x, &normals[0].x, albedos ? albedos->data() : nullptr,
                                   &nv, &indices[0].x, &nt);
  if (code != DVS_OK) throw std::runtime_error(dvs_last_error());
  vertices.resize((size_t)nv);
  normals.resize((size_t)nv);
  indices.resize((size_t)nt);
  if (albedos) albedos->resize((size_t)nv);
  std::cout << "Generated " << nv << " vertices, " << nv << " normals, and " << nt << " triangles." << std::endl;
  std::cout << "DEM center x: " << dem.x_coords[dem.width / 2] << ", y: " << dem.y_coords[dem.height / 2]
            << ", depth: " << reference_height(dem) << std::endl;
  std::cout << "Converted triangles to DVS format." << std::endl;
}
}  // namespace detail

// load_from_raster(dem, albedo, ...) (build_from_raster.cu:683-823): the
// albedo raster, in the DEM's CRS, resampled onto the DEM grid
// (batch_align_kernel), then the DEM's mesh with a per-vertex albedo.
inline bool load_from_raster(const std::string &dem_path, const std::string &albedo_path,
                             std::vector<dvs::float3> &vertices, std::vector<dvs::float3> &normals,
                             std::vector<dvs::uint3> &indices, std::vector<float> &albedos) {
  Imagery dem, albedo;
  if (!loadImageryData(dem_path, dem)) throw std::runtime_error("failed to load DEM data from: " + dem_path);
  reference_height(dem);
  std::cout << "DEM dimensions: " << dem.width << "x" << dem.height << std::endl;
  if (!loadImageryData(albedo_path, albedo)) throw std::runtime_error("failed to load albedo data from: " + albedo_path);
  require_same_projection(dem, albedo);
  std::cout << "Albedo dimensions: " << albedo.width << "x" << albedo.height << std::endl;
  std::vector<float> aligned((size_t)dem.width * dem.height);
  if (dvs_raster_align(dem.x_coords.data(), dem.y_coords.data(), dem.width, dem.height, albedo.data.data(),
                       albedo.x_coords.data(), albedo.y_coords.data(), albedo.width, albedo.height,
                       aligned.data()) != DVS_OK)
    throw std::runtime_error(dvs_last_error());
  detail::raster_mesh(dem, aligned.data(), vertices, normals, indices, &albedos);
  return true;
}

// load_from_raster(dem, ...) (build_from_raster.cu:825-920).
inline bool load_from_raster(const std::string &dem_path, std::vector<dvs::float3> &vertices,
                             std::vector<dvs::float3> &normals, std::vector<dvs::uint3> &indices) {
  Imagery dem;
  if (!loadImageryData(dem_path, dem)) throw std::runtime_error("failed to load DEM data from: " + dem_path);
  reference_height(dem);
  std::cout << "DEM dimensions: " << dem.width << "x" << dem.height << std::endl;
  detail::raster_mesh(dem, nullptr, vertices, normals, indices, nullptr);
  return true;
}

}  // namespace tools

namespace dvs {
namespace detail {
inline bool finish_raster_bundle(SceneBundle<float3, uint3> &out, std::string *err) {
  if (!out.albedos.empty()) out.use_albedo = true;
  out.bvh_nodes.clear();
  if (out.bvh_triangles.empty()) {
    if (err) *err = "mesh is empty after triangulation";
    return false;
  }
  return true;
}
}  // namespace detail

// load_scene_bundle over rasters (scene_data.cpp:284-395).
inline bool load_scene_bundle(const std::string &dem_path, const std::string &albedo_path,
                              SceneBundle<float3, uint3> &out_bundle, std::string *err) {
  if (!tools::load_from_raster(dem_path, albedo_path, out_bundle.vertices, out_bundle.normals,
                               out_bundle.bvh_triangles, out_bundle.albedos))
    return false;
  return detail::finish_raster_bundle(out_bundle, err);
}
inline bool load_scene_bundle(const std::string &dem_path, SceneBundle<float3, uint3> &out_bundle,
                              std::string *err) {
  if (!tools::load_from_raster(dem_path, out_bundle.vertices, out_bundle.normals, out_bundle.bvh_triangles))
    return false;
  return detail::finish_raster_bundle(out_bundle, err);
}
inline bool load_scene_bundle(const std::string &dem_path, const std::string &albedo_path,
                              const std::string &traj_path, SceneBundle<float3, uint3> &out_bundle,
                              std::vector<FrameData<Vec3, Mat3>> &frames, std::string *err) {
  if (!tools::load_from_raster(dem_path, albedo_path, out_bundle.vertices, out_bundle.normals,
                               out_bundle.bvh_triangles, out_bundle.albedos))
    return false;
  if (!load_frames_from_text(traj_path, frames, err)) return false;
  if (frames.empty()) {
    if (err) *err = "trajectory file is empty";
    return false;
  }
  return detail::finish_raster_bundle(out_bundle, err);
}
inline bool load_scene_bundle(const std::string &dem_path, const std::string &traj_path,
                              SceneBundle<float3, uint3> &out_bundle, std::vector<FrameData<Vec3, Mat3>> &frames,
                              std::string *err) {
  if (!tools::load_from_raster(dem_path, out_bundle.vertices, out_bundle.normals, out_bundle.bvh_triangles))
    return false;
  if (!load_frames_from_text(traj_path, frames, err)) return false;
  if (frames.empty()) {
    if (err) *err = "trajectory file is empty";
    return false;
  }
  return detail::finish_raster_bundle(out_bundle, err);
}
}  // namespace dvs
