// Trajectory of the DVS path: the host build of CameraTrajectory's curves
// (src/trajectory.cpp:19-82,125-193,389-436,604-716), single evaluations on
// the host and the intervals of a window evaluated on the device, over the
// shared curve math of trajectory.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"
#include "dvs_hip.h"
#include "trajectory.hpp"
#include "trajectory_device.h"

namespace dvstraj {

// The n - 1 intervals between consecutive times: each time's optical state
// (to_optical_state of evaluate((float)t)) is evaluated once and written as
// the end of the interval before it and the start of the one after it.  The
// times of one launch travel in its arguments (times base .. base + count - 1
// of n), so a window needs no copy before its launch.
constexpr int kArgTimes = 240;
struct TimesArg {
  double t[kArgTimes];
  int base, count, n;
};
__global__ void intervals_kernel(View v, TimesArg a, dvs_optical_interval *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.count) return;
  const int i = a.base + j, n = a.n;
  const double t = a.t[j];
  const Sample s = evaluate(v, (float)t, false);
  dvs_optical_state o;
  for (int c = 0; c < 3; ++c) {
    o.position[c] = s.position[c];
    o.velocity[c] = s.velocity[c];
    o.light[c] = s.light[c];
    o.light_rate[c] = s.light_rate[c];
  }
  o.rotation[0] = s.rotation.x;
  o.rotation[1] = s.rotation.y;
  o.rotation[2] = s.rotation.z;
  o.rotation[3] = s.rotation.w;
  o.intensity = s.intensity;
  o.intensity_rate = s.intensity_rate;
  if (i + 1 < n) {
    out[i].start = o;
    out[i].t0 = t;
  }
  if (i > 0) {
    out[i - 1].end = o;
    out[i - 1].t1 = t;
  }
}

}  // namespace dvstraj

namespace dvsamd {
namespace {
using dvstraj::Key;
using dvstraj::Quat;

// Monotone (Fritsch-Carlson) key slopes: the harmonic-type mean of the
// neighbouring secants, zero where they change sign
// (compute_auto_clamped_tangents, trajectory.cpp:19-82).
std::vector<float> auto_clamped_slopes(const std::vector<float> &t, const std::vector<float> &y) {
  const size_t n = t.size();
  std::vector<float> m(n, 0.0f);
  if (n <= 1) return m;
  if (n == 2) {
    const float h = t[1] - t[0];
    m[0] = m[1] = std::fabs(h) > 1e-12f ? (y[1] - y[0]) / h : 0.0f;
    return m;
  }
  std::vector<float> h(n - 1), d(n - 1);
  for (size_t i = 0; i + 1 < n; ++i) {
    h[i] = t[i + 1] - t[i];
    d[i] = std::fabs(h[i]) > 1e-12f ? (y[i + 1] - y[i]) / h[i] : 0.0f;
  }
  m[0] = d[0];
  m[n - 1] = d[n - 2];
  for (size_t i = 1; i + 1 < n; ++i) {
    const float d0 = d[i - 1], d1 = d[i];
    if (d0 == 0.0f || d1 == 0.0f || d0 * d1 <= 0.0f) continue;
    const float w1 = 2.0f * h[i] + h[i - 1], w2 = h[i] + 2.0f * h[i - 1];
    const float denom = w1 / d0 + w2 / d1;
    m[i] = std::fabs(denom) > 1e-12f ? (w1 + w2) / denom : 0.0f;
  }
  return m;
}

// Keys with handles a third of the neighbouring span along the slope
// (build_auto_clamped, trajectory.cpp:125-193).
void build_curve(const std::vector<float> &t, const std::vector<float> &y, Key *keys) {
  const size_t n = t.size();
  for (size_t i = 0; i < n; ++i) keys[i] = Key{t[i], y[i], t[i], y[i], t[i], y[i]};
  if (n < 2) return;
  const std::vector<float> m = auto_clamped_slopes(t, y);
  for (size_t i = 0; i + 1 < n; ++i) {
    const float h = t[i + 1] - t[i];
    keys[i].right_time = t[i] + h / 3.0f;
    keys[i].right_value = y[i] + m[i] * h / 3.0f;
    keys[i + 1].left_time = t[i + 1] - h / 3.0f;
    keys[i + 1].left_value = y[i + 1] - m[i + 1] * h / 3.0f;
  }
}

// Column-major 3x3 (m[3 * col + row]).
struct M3 {
  float m[9];
  float operator()(int r, int c) const { return m[3 * c + r]; }
  float &at(int r, int c) { return m[3 * c + r]; }
};
M3 product(const M3 &a, const M3 &b) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.at(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
  return r;
}
// Rotation by `angle` about coordinate axis `axis` (Rodrigues' formula).
M3 axis_rotation(float angle, int axis) {
  float e[3] = {0, 0, 0};
  e[axis] = 1;
  const float s = std::sin(angle), c = std::cos(angle);
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.at(i, j) = (1 - c) * e[i] * e[j] + (i == j ? c : 0.0f);
  r.at(1, 2) -= s * e[0]; r.at(2, 1) += s * e[0];
  r.at(2, 0) -= s * e[1]; r.at(0, 2) += s * e[1];
  r.at(0, 1) -= s * e[2]; r.at(1, 0) += s * e[2];
  return r;
}
// Unit quaternion of a rotation matrix (Shepperd's method: from the trace
// when it is positive, else from the largest diagonal entry).
Quat quaternion_of(const M3 &R) {
  float q[4];  // x, y, z, w
  const float tr = R(0, 0) + R(1, 1) + R(2, 2);
  if (tr > 0) {
    float t = std::sqrt(tr + 1.0f);
    q[3] = 0.5f * t;
    t = 0.5f / t;
    q[0] = (R(2, 1) - R(1, 2)) * t;
    q[1] = (R(0, 2) - R(2, 0)) * t;
    q[2] = (R(1, 0) - R(0, 1)) * t;
  } else {
    int i = 0;
    if (R(1, 1) > R(0, 0)) i = 1;
    if (R(2, 2) > R(i, i)) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    float t = std::sqrt(R(i, i) - R(j, j) - R(k, k) + 1.0f);
    q[i] = 0.5f * t;
    t = 0.5f / t;
    q[3] = (R(k, j) - R(j, k)) * t;
    q[j] = (R(j, i) + R(i, j)) * t;
    q[k] = (R(k, i) + R(i, k)) * t;
  }
  return dvstraj::qnormalized(Quat{q[0], q[1], q[2], q[3]});
}
float qdot(const Quat &a, const Quat &b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
bool finite(const float *v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

}  // namespace

struct Trajectory {
  int count = 0;
  std::vector<Key> keys;     // [kCurves][count]
  std::vector<float> quats;  // [count][4]
  DeviceArray<Key> device_keys;
  DeviceArray<float> device_quats;
  DeviceArray<dvs_optical_interval> device_intervals;
  PinnedArray<dvs_optical_interval> host_intervals;
  hipStream_t stream = nullptr;
  ~Trajectory() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  dvstraj::View host_view() const { return dvstraj::View{keys.data(), quats.data(), count}; }
};

namespace {
// CameraTrajectory::build (trajectory.cpp:604-716).
void build(Trajectory &tr, const dvs_frame *frames, int n, float fps, const float translation[3],
           const float rotation[3]) {
  if (!std::isfinite(fps) || fps <= 0)
    throw std::invalid_argument("trajectory requires positive fps and three-component mount transforms");
  const M3 mount = product(product(axis_rotation(rotation[2], 2), axis_rotation(rotation[1], 1)),
                           axis_rotation(rotation[0], 0));
  tr.count = n;
  if (n == 0) return;
  std::vector<float> times(n);
  for (int i = 0; i < n; ++i) {
    const dvs_frame &f = frames[i];
    times[i] = (float)(std::isnan(f.timestamp_seconds) ? f.frame_index / (double)fps : f.timestamp_seconds);
    if (!std::isfinite(times[i])) throw std::invalid_argument("non-finite trajectory timestamp");
  }
  for (int i = 1; i < n; ++i)
    if (!(times[i] > times[i - 1])) throw std::invalid_argument("trajectory timestamps must be strictly increasing");
  std::vector<std::vector<float>> values(dvstraj::kCurves, std::vector<float>(n));
  std::vector<Quat> q(n);
  for (int i = 0; i < n; ++i) {
    const dvs_frame &f = frames[i];
    M3 W;
    for (int k = 0; k < 9; ++k) W.m[k] = f.camera_world[k];
    const float light_norm = std::sqrt(f.light_dir[0] * f.light_dir[0] + f.light_dir[1] * f.light_dir[1] +
                                       f.light_dir[2] * f.light_dir[2]);
    // rigidity: |W^T W - I| (Frobenius) and det W
    float off = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const float g = W(0, r) * W(0, c) + W(1, r) * W(1, c) + W(2, r) * W(2, c) - (r == c ? 1.0f : 0.0f);
        off += g * g;
      }
    const float det = W(0, 0) * (W(1, 1) * W(2, 2) - W(1, 2) * W(2, 1)) -
                      W(0, 1) * (W(1, 0) * W(2, 2) - W(1, 2) * W(2, 0)) +
                      W(0, 2) * (W(1, 0) * W(2, 1) - W(1, 1) * W(2, 0));
    if (!finite(f.camera_origin, 3) || !finite(f.camera_world, 9) || !finite(f.light_dir, 3) ||
        light_norm < 1e-12f || !std::isfinite(f.light_intensity) || f.light_intensity < 0 ||
        std::sqrt(off) > 2e-3f || det <= 0)
      throw std::invalid_argument("trajectory requires finite rigid poses and valid world lighting");
    // mounted camera: origin + W t, orientation W M
    for (int r = 0; r < 3; ++r)
      values[dvstraj::kPx + r][i] =
          f.camera_origin[r] + (W(r, 0) * translation[0] + W(r, 1) * translation[1] + W(r, 2) * translation[2]);
    Quat qi = quaternion_of(product(W, mount));
    if (i > 0 && qdot(qi, q[i - 1]) < 0.0f) qi = Quat{-qi.x, -qi.y, -qi.z, -qi.w};
    q[i] = dvstraj::qnormalized(qi);
    for (int c = 0; c < 3; ++c) values[dvstraj::kLx + c][i] = f.light_dir[c] / light_norm;
    values[dvstraj::kIntensity][i] = f.light_intensity;
    values[dvstraj::kPhase][i] = (float)i;
  }
  // RotationBezierCurve::build (trajectory.cpp:389-436): unit keys, kept on
  // one hemisphere
  for (int i = 0; i < n; ++i) {
    q[i] = dvstraj::qnormalized(q[i]);
    if (i > 0 && qdot(q[i], q[i - 1]) < 0.0f) q[i] = Quat{-q[i].x, -q[i].y, -q[i].z, -q[i].w};
  }
  tr.keys.resize((size_t)dvstraj::kCurves * n);
  for (int c = 0; c < dvstraj::kCurves; ++c) build_curve(times, values[c], tr.keys.data() + (size_t)c * n);
  tr.quats.resize((size_t)4 * n);
  for (int i = 0; i < n; ++i) {
    tr.quats[4 * i] = q[i].x;
    tr.quats[4 * i + 1] = q[i].y;
    tr.quats[4 * i + 2] = q[i].z;
    tr.quats[4 * i + 3] = q[i].w;
  }
}

// The state of an empty trajectory (CameraTrajectoryState's defaults).
dvs_trajectory_state default_state(float t) {
  dvs_trajectory_state s{};
  s.timestamp = t;
  s.orientation[0] = s.orientation[4] = s.orientation[8] = 1;
  s.quaternion[3] = 1;
  s.light_direction[2] = 1;
  return s;
}
dvs_optical_state optical_of(const dvs_trajectory_state &s) {
  dvs_optical_state o{};
  for (int c = 0; c < 3; ++c) {
    o.position[c] = s.position[c];
    o.velocity[c] = s.linear_velocity[c];
    o.light[c] = s.light_direction[c];
    o.light_rate[c] = s.light_direction_rate[c];
  }
  for (int c = 0; c < 4; ++c) o.rotation[c] = s.quaternion[c];
  o.intensity = s.light_intensity;
  o.intensity_rate = s.light_intensity_rate;
  return o;
}
Trajectory &get(dvs_trajectory t) {
  if (!t) throw std::invalid_argument("null trajectory handle");
  return *reinterpret_cast<Trajectory *>(t);
}
// The intervals of n times (2 <= n <= INT32_MAX, tr.count > 0) into device
// memory `out`, by launches on st; the keys are uploaded on st at first use.
void launch_intervals(Trajectory &tr, const double *times, int64_t n, dvs_optical_interval *out, hipStream_t st) {
  if (!tr.device_keys.ptr) {
    tr.device_keys.upload(tr.keys.data(), tr.keys.size(), "trajectory keys", st);
    tr.device_quats.upload(tr.quats.data(), tr.quats.size(), "trajectory keys", st);
    // other streams may use the keys next: the copies are done before return
    hip_check(hipStreamSynchronize(st), "trajectory keys");
  }
  const dvstraj::View v{tr.device_keys.ptr, tr.device_quats.ptr, tr.count};
  for (int64_t base = 0; base < n; base += dvstraj::kArgTimes) {
    dvstraj::TimesArg a;
    a.base = (int)base;
    a.count = (int)std::min<int64_t>(dvstraj::kArgTimes, n - base);
    a.n = (int)n;
    std::memcpy(a.t, times + base, sizeof(double) * (size_t)a.count);
    dvstraj::intervals_kernel<<<(unsigned)((a.count + 63) / 64), 64, 0, st>>>(v, a, out);
    hip_check(hipGetLastError(), "trajectory launch");
  }
}
}  // namespace

void trajectory_intervals_on(dvs_trajectory trajectory, const double *times, int64_t n, dvs_optical_interval *out,
                             hipStream_t stream) {
  Trajectory &tr = get(trajectory);
  if (n < 2 || !times || !out) throw std::invalid_argument("invalid trajectory arguments");
  if (n > (int64_t)INT32_MAX) throw std::invalid_argument("too many trajectory times");
  if (tr.count == 0) {
    std::vector<dvs_optical_interval> host((size_t)n - 1);
    for (int64_t i = 0; i + 1 < n; ++i)
      host[(size_t)i] = dvs_optical_interval{optical_of(default_state((float)times[i])),
                                             optical_of(default_state((float)times[i + 1])), times[i], times[i + 1]};
    hip_check(hipMemcpyAsync(out, host.data(), sizeof(dvs_optical_interval) * host.size(), hipMemcpyHostToDevice,
                             stream),
              "trajectory intervals");
    hip_check(hipStreamSynchronize(stream), "trajectory intervals");  // `host` is released on return
    return;
  }
  launch_intervals(tr, times, n, out, stream);
}
}  // namespace dvsamd

using namespace dvsamd;

extern "C" {

int dvs_trajectory_create(const dvs_frame *frames, int32_t count, float fps, const float translation[3],
                          const float rotation[3], dvs_trajectory *out) {
  return guarded([&] {
    if (!out || count < 0 || (count > 0 && !frames) || !translation || !rotation)
      throw std::invalid_argument("invalid trajectory arguments");
    auto tr = std::make_unique<Trajectory>();
    build(*tr, frames, count, fps, translation, rotation);
    *out = reinterpret_cast<dvs_trajectory>(tr.release());
  });
}

int dvs_trajectory_destroy(dvs_trajectory trajectory) {
  return guarded([&] { delete reinterpret_cast<Trajectory *>(trajectory); });
}

int dvs_trajectory_range(dvs_trajectory trajectory, float *start_time, float *end_time, int32_t *size) {
  return guarded([&] {
    const Trajectory &tr = get(trajectory);
    const bool any = tr.count > 0;
    if (start_time) *start_time = any ? tr.keys.front().time : 0.0f;
    if (end_time) *end_time = any ? tr.keys[tr.count - 1].time : 0.0f;
    if (size) *size = tr.count;
  });
}

int dvs_trajectory_evaluate(dvs_trajectory trajectory, const float *times, int64_t n, dvs_trajectory_state *out) {
  return guarded([&] {
    const Trajectory &tr = get(trajectory);
    if (n < 0 || (n > 0 && (!times || !out))) throw std::invalid_argument("invalid trajectory arguments");
    for (int64_t i = 0; i < n; ++i) {
      if (tr.count == 0) {
        out[i] = default_state(times[i]);
        continue;
      }
      const dvstraj::Sample s = dvstraj::evaluate(tr.host_view(), times[i], true);
      dvs_trajectory_state &o = out[i];
      o.timestamp = s.timestamp;
      for (int c = 0; c < 3; ++c) {
        o.position[c] = s.position[c];
        o.linear_velocity[c] = s.velocity[c];
        o.angular_velocity[c] = s.angular_velocity[c];
        o.light_direction[c] = s.light[c];
        o.light_direction_rate[c] = s.light_rate[c];
      }
      for (int k = 0; k < 9; ++k) o.orientation[k] = s.orientation[k];
      o.quaternion[0] = s.rotation.x;
      o.quaternion[1] = s.rotation.y;
      o.quaternion[2] = s.rotation.z;
      o.quaternion[3] = s.rotation.w;
      o.light_intensity = s.intensity;
      o.light_intensity_rate = s.intensity_rate;
    }
  });
}

int dvs_trajectory_intervals(dvs_trajectory trajectory, const double *times, int64_t n,
                             dvs_optical_interval *out) {
  return guarded([&] {
    Trajectory &tr = get(trajectory);
    if (n < 2 || !times || !out) throw std::invalid_argument("invalid trajectory arguments");
    if (n > (int64_t)INT32_MAX) throw std::invalid_argument("too many trajectory times");
    if (tr.count == 0) {
      std::vector<dvs_optical_state> states((size_t)n);
      for (int64_t i = 0; i < n; ++i) states[i] = optical_of(default_state((float)times[i]));
      for (int64_t i = 0; i + 1 < n; ++i)
        out[i] = dvs_optical_interval{states[i], states[i + 1], times[i], times[i + 1]};
      return;
    }
    if (!tr.stream) hip_check(hipStreamCreateWithFlags(&tr.stream, hipStreamNonBlocking), "trajectory stream");
    hipStream_t st = tr.stream;
    // times in through the launches' arguments, intervals out through a
    // page-locked buffer
    stream_wait(st, "trajectory times");  // the previous call's copies are done
    tr.host_intervals.reserve((size_t)n - 1, "trajectory intervals");
    tr.device_intervals.reserve((size_t)n - 1, "trajectory intervals");
    launch_intervals(tr, times, n, tr.device_intervals.ptr, st);
    hip_check(hipMemcpyAsync(tr.host_intervals.ptr, tr.device_intervals.ptr,
                             sizeof(dvs_optical_interval) * (size_t)(n - 1), hipMemcpyDeviceToHost, st),
              "trajectory intervals");
    stream_wait(st, "trajectory intervals");
    std::memcpy(out, tr.host_intervals.ptr, sizeof(dvs_optical_interval) * (size_t)(n - 1));
  });
}

}  // extern "C"
