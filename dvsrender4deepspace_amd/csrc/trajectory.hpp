// Camera and light trajectory of the DVS path (CameraTrajectory,
// include/dvs/trajectory.hpp and src/trajectory.cpp): Blender-style cubic
// Bezier curves with auto-clamped (Fritsch-Carlson) handles for position,
// light direction and intensity, and an SO(3) curve for the orientation
// (q0 Exp(alpha Log(q0^-1 q1)) with alpha from a Bezier phase curve).  The
// evaluation is written once for host and device: the host answers single
// queries with the full state, a kernel builds the optical intervals of a
// whole window.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace dvstraj {

// One key of a scalar curve with its left and right Bezier handles.
struct Key {
  float time, value, left_time, left_value, right_time, right_value;
};

// Curves of one trajectory: kCurves scalar curves of `count` keys each
// (keys[curve * count + i]) and the continuous unit quaternions (x, y, z, w)
// of the orientation keys.
enum { kPx = 0, kPy, kPz, kLx, kLy, kLz, kIntensity, kPhase, kCurves };
struct View {
  const Key *keys;
  const float *quats;  // [count][4]
  int count;
};

__host__ __device__ inline float clampf(float x, float lo, float hi) { return fmaxf(lo, fminf(x, hi)); }

// De Casteljau in difference form (a constant curve stays exactly constant).
__host__ __device__ inline float bezier(float p0, float p1, float p2, float p3, float u) {
  const float a = p0 + u * (p1 - p0), b = p1 + u * (p2 - p1), c = p2 + u * (p3 - p2);
  const float d = a + u * (b - a), e = b + u * (c - b);
  return d + u * (e - d);
}
__host__ __device__ inline float bezier_derivative(float p0, float p1, float p2, float p3, float u) {
  const float v = 1.0f - u;
  return 3.0f * v * v * (p1 - p0) + 6.0f * v * u * (p2 - p1) + 3.0f * u * u * (p3 - p2);
}

// Segment [i, i + 1] holding t: the last key at or before t, within
// [0, count - 2].
__host__ __device__ inline int find_segment(const Key *k, int count, float t) {
  if (t <= k[0].time) return 0;
  if (t >= k[count - 1].time) return count - 2;
  int lo = 0, hi = count;  // first key with time > t
  while (lo < hi) {
    const int m = (lo + hi) / 2;
    if (t < k[m].time) hi = m; else lo = m + 1;
  }
  const int seg = lo - 1;
  return seg < 0 ? 0 : seg > count - 2 ? count - 2 : seg;
}

// The curve parameter u with x(u) = t: Newton from the linear guess, eight
// steps at most, then bisection (solve_bezier_time, trajectory.cpp:218-263).
__host__ __device__ inline float solve_time(const Key &k0, const Key &k1, float t) {
  const float x0 = k0.time, x1 = k0.right_time, x2 = k1.left_time, x3 = k1.time;
  const float span = x3 - x0;
  if (span <= 1e-12f) return 0.0f;
  float u = clampf((t - x0) / span, 0.0f, 1.0f);
  for (int i = 0; i < 8; ++i) {
    const float x = bezier(x0, x1, x2, x3, u), dx = bezier_derivative(x0, x1, x2, x3, u);
    if (fabsf(dx) < 1e-10f) break;
    const float next = clampf(u - (x - t) / dx, 0.0f, 1.0f);
    if (fabsf(next - u) < 1e-7f) return next;
    u = next;
  }
  float lo = 0.0f, hi = 1.0f;
  for (int i = 0; i < 24; ++i) {
    const float mid = 0.5f * (lo + hi);
    if (bezier(x0, x1, x2, x3, mid) < t) lo = mid; else hi = mid;
  }
  return 0.5f * (lo + hi);
}

// Every curve of a trajectory has its keys at the same times with the same
// time handles (t_i + h/3, t_{i+1} - h/3), so the segment and the curve
// parameter at a time are those of every curve: they are found once per
// time, on the key times of any one curve.
//
// Where a value comes from (curve value, trajectory.cpp:266-285): the first
// or last key outside the key range, else segment `seg` at parameter u.
struct ValueAt {
  int where;  // -1 first key, 1 last key, 0 segment seg at u
  int seg;
  float u;
};
__host__ __device__ inline ValueAt value_at(const Key *k, int count, float t) {
  if (count == 1 || t <= k[0].time) return ValueAt{-1, 0, 0.0f};
  if (t >= k[count - 1].time) return ValueAt{1, 0, 0.0f};
  const int s = find_segment(k, count, t);
  return ValueAt{0, s, solve_time(k[s], k[s + 1], t)};
}
__host__ __device__ inline float curve_value(const Key *k, int count, const ValueAt &at) {
  if (at.where < 0) return k[0].value;
  if (at.where > 0) return k[count - 1].value;
  const int s = at.seg;
  return bezier(k[s].value, k[s].right_value, k[s + 1].left_value, k[s + 1].value, at.u);
}

// Where a rate comes from (derivative, trajectory.cpp:287-314): the segment
// of t kept 1e-7 inside the key range, the curve parameter of t itself, and
// dx/du there (zero rate when |dx/du| < 1e-12 or with fewer than two keys).
struct RateAt {
  int valid;
  int seg;
  float u, dx;
};
__host__ __device__ inline RateAt rate_at(const Key *k, int count, float t) {
  if (count < 2) return RateAt{0, 0, 0.0f, 0.0f};
  const float lo = k[0].time, hi = k[count - 1].time, eps = 1e-7f;
  float tc = t;
  if (tc < lo + eps) tc = lo + eps;
  if (tc > hi - eps) tc = hi - eps;
  const int s = find_segment(k, count, tc);
  const float u = solve_time(k[s], k[s + 1], t);
  const float dx = bezier_derivative(k[s].time, k[s].right_time, k[s + 1].left_time, k[s + 1].time, u);
  return RateAt{fabsf(dx) < 1e-12f ? 0 : 1, s, u, dx};
}
__host__ __device__ inline float curve_rate(const Key *k, const RateAt &at) {
  if (!at.valid) return 0.0f;
  const int s = at.seg;
  const float dy = bezier_derivative(k[s].value, k[s].right_value, k[s + 1].left_value, k[s + 1].value, at.u);
  return dy / at.dx;
}

// Quaternions as (x, y, z, w).
struct Quat {
  float x, y, z, w;
};
__host__ __device__ inline Quat qmul(const Quat &a, const Quat &b) {  // Hamilton product a b
  return Quat{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__host__ __device__ inline Quat qnormalized(const Quat &q) {
  const float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return n > 0 ? Quat{q.x / n, q.y / n, q.z / n, q.w / n} : q;
}
// Exp of a rotation vector (quat_exp, trajectory.cpp:346-364).
__host__ __device__ inline Quat qexp(float rx, float ry, float rz) {
  const float theta = sqrtf(rx * rx + ry * ry + rz * rz);
  if (theta < 1e-8f) return qnormalized(Quat{0.5f * rx, 0.5f * ry, 0.5f * rz, 1.0f});
  const float h = 0.5f * theta, s = sinf(h);
  return qnormalized(Quat{s * (rx / theta), s * (ry / theta), s * (rz / theta), cosf(h)});
}
// Log of a unit quaternion on the short arc (quat_log, trajectory.cpp:366-387).
__host__ __device__ inline void qlog(Quat q, float r[3]) {
  q = qnormalized(q);
  if (q.w < 0.0f) q = Quat{-q.x, -q.y, -q.z, -q.w};
  const float s = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
  if (s < 1e-8f) {
    r[0] = 2.0f * q.x; r[1] = 2.0f * q.y; r[2] = 2.0f * q.z;
    return;
  }
  const float f = 2.0f * atan2f(s, q.w) / s;
  r[0] = q.x * f; r[1] = q.y * f; r[2] = q.z * f;
}
// Column-major rotation matrix of a unit quaternion (m[3 * c + r]).
__host__ __device__ inline void qmatrix(const Quat &q, float m[9]) {
  const float tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
  const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w, txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  m[0] = 1 - (tyy + tzz); m[3] = txy - twz;       m[6] = txz + twy;
  m[1] = txy + twz;       m[4] = 1 - (txx + tzz); m[7] = tyz - twx;
  m[2] = txz - twy;       m[5] = tyz + twx;       m[8] = 1 - (txx + tyy);
}

// Orientation at t with its phase rate (RotationBezierCurve::evaluate,
// trajectory.cpp:474-560): alpha from the phase curve, q = q0 Exp(alpha rho)
// with rho = Log(q0^-1 q1) on the short arc; the body angular velocity is
// phase_rate * rho.
__host__ __device__ inline Quat orientation(const View &v, const ValueAt &vat, const RateAt &rat,
                                            float rho_rate[3]) {
  const Key *ph = v.keys + kPhase * v.count;
  rho_rate[0] = rho_rate[1] = rho_rate[2] = 0;
  if (v.count < 2) return Quat{0, 0, 0, 1};  // one key: no segment, the identity (as the reference)
  const float phase = curve_value(ph, v.count, vat), phase_rate = curve_rate(ph, rat);
  int seg = (int)floorf(phase);
  seg = seg < 0 ? 0 : seg > v.count - 2 ? v.count - 2 : seg;
  const float alpha = clampf(phase - (float)seg, 0.0f, 1.0f);
  const float *a = v.quats + 4 * seg, *b = a + 4;
  const Quat q0{a[0], a[1], a[2], a[3]}, q1{b[0], b[1], b[2], b[3]};
  Quat rel = qnormalized(qmul(Quat{-q0.x, -q0.y, -q0.z, q0.w}, q1));
  if (rel.w < 0.0f) rel = Quat{-rel.x, -rel.y, -rel.z, -rel.w};
  float rho[3];
  qlog(rel, rho);
  const Quat q = qnormalized(qmul(q0, qexp(alpha * rho[0], alpha * rho[1], alpha * rho[2])));
  rho_rate[0] = phase_rate * rho[0];
  rho_rate[1] = phase_rate * rho[1];
  rho_rate[2] = phase_rate * rho[2];
  return q;
}

// The optical part of a state (to_optical_state, raycast.hpp:17-22) at t
// clamped to the key range; with the orientation matrix and the world
// angular velocity when asked (CameraTrajectory::evaluate,
// trajectory.cpp:719-787).
struct Sample {
  float timestamp;
  float position[3], velocity[3], light[3], light_rate[3];
  Quat rotation;
  float intensity, intensity_rate;
  float orientation[9], angular_velocity[3];
};
__host__ __device__ inline Sample evaluate(const View &v, float t, bool full) {
  Sample s{};
  const float lo = v.keys[0].time, hi = v.keys[v.count - 1].time;
  const float tc = clampf(t, lo, hi);
  s.timestamp = tc;
  const Key *k = v.keys;
  const int n = v.count;
  const ValueAt vat = value_at(k, n, tc);
  const RateAt rat = rate_at(k, n, tc);
  for (int c = 0; c < 3; ++c) {
    s.position[c] = curve_value(k + (kPx + c) * n, n, vat);
    s.velocity[c] = curve_rate(k + (kPx + c) * n, rat);
  }
  float body[3];
  s.rotation = orientation(v, vat, rat, body);
  if (full) {
    qmatrix(s.rotation, s.orientation);
    for (int r = 0; r < 3; ++r)
      s.angular_velocity[r] =
          s.orientation[r] * body[0] + s.orientation[3 + r] * body[1] + s.orientation[6 + r] * body[2];
  }
  float l[3], ld[3];
  for (int c = 0; c < 3; ++c) {
    l[c] = curve_value(k + (kLx + c) * n, n, vat);
    ld[c] = curve_rate(k + (kLx + c) * n, rat);
  }
  const float norm = sqrtf(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
  if (norm > 1e-12f) {
    // d(v / |v|)/dt = v' / |v| - v (v . v') / |v|^3 (normalized_vector_derivative)
    const float dot = l[0] * ld[0] + l[1] * ld[1] + l[2] * ld[2];
    for (int c = 0; c < 3; ++c) {
      s.light[c] = l[c] / norm;
      s.light_rate[c] = ld[c] / norm - l[c] * (dot / (norm * norm * norm));
    }
  } else {
    s.light[0] = s.light[1] = 0; s.light[2] = 1;
    s.light_rate[0] = s.light_rate[1] = s.light_rate[2] = 0;
  }
  s.intensity = curve_value(k + kIntensity * n, n, vat);
  s.intensity_rate = curve_rate(k + kIntensity * n, rat);
  return s;
}

}  // namespace dvstraj
