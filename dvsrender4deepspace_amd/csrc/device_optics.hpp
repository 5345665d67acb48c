// gfx950 device optics: SoA BVH traversal, Moller-Trumbore, shadow origin,
// Hapke shading and the per-subray sample.  Every function keeps the
// reference's IEEE operation sequence (the library is built with
// -ffp-contract=off) so that ray geometry is bit-identical to a host
// evaluation of the same source expressions; only the float transcendentals
// of the Hapke BRDF (ocml vs. a host libm) differ in the last bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dvs_hip.h"

namespace dvsdev {

struct f3 { float x, y, z; };
__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ f3 scale(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
__host__ __device__ __forceinline__ f3 divide(f3 a, float b) { return mk(a.x / b, a.y / b, a.z / b); }
__host__ __device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ f3 normalize(f3 v) {
  const float len = sqrtf(dot(v, v));
  return len > 0.0f ? divide(v, len) : mk(0.f, 0.f, 0.f);
}
__host__ __device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

constexpr float kRayEps = 1e-6f;                  // include/helper_math.hpp:11
constexpr float kPi = 3.14159265358979323846f;    // include/dvs/math_types.hpp

// The Hapke parameters' per-scene constants (the terms of
// hapke_reflectance that depend on the parameters alone), computed once on
// the host by hapke_constants() with the same float expressions.
struct HapkeConsts {
  float tan_tp, cot_tp, ch;  // tan, cot and chi of theta-bar (chi: 1/sqrt(1 + pi tan^2))
  float K;                   // porosity factor
  int rough;                 // theta-bar > 1e-8
};

// Device view of one scene (SoA BVH + indexed mesh), passed by value.
struct DevScene {
  const float4 *pairs;       // 4 x float4 per internal node
  const int2 *leaves;        // (start, count)
  const float4 *triangles;   // 3 x float4 per BVH-ordered triangle: v0 | e1 | e2
  const uint32_t *indices;   // 3 per BVH-ordered triangle
  const float *normals;      // 3 per vertex
  const float *albedos;      // per vertex or nullptr
  int root_ref, stack_depth;
  int width, height, samples;
  float fx, fy, cx, cy;
  dvs_hapke hapke;
  HapkeConsts hk;  // hapke_constants(hapke)
  // 1: the wave-packet traversal may serve this tree (every reference
  // subtree covers one triangle index range, left before right, and word 14
  // of each pair holds the right subtree's first index)
  int packet;
};

// Pose used by the ray kernels: the interpolated OpticalState fields rays read.
struct Pose {
  float px, py, pz;
  float lx, ly, lz;
  float qx, qy, qz, qw;
  float intensity;
  float pad;
};

// ---------------------------------------------------------------------------
// Pose interpolation: optical_state_at (include/dvs/device_optics.cuh:6-26)
// with the quaternion renormalised by 1/sqrtf (the reference's rsqrtf is an
// approximate intrinsic).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float hermite(float a, float va, float b, float vb, float u, float dt) {
  const float u2 = u * u, u3 = u2 * u;
  return a + (-2 * u3 + 3 * u2) * (b - a) + (u3 - 2 * u2 + u) * dt * va + (u3 - u2) * dt * vb;
}
__device__ inline Pose pose_at(const dvs_optical_interval &iv, double t) {
  const float dt = (float)(iv.t1 - iv.t0);
  const float u = dt > 0 ? (float)fmin(1.0, fmax(0.0, (t - iv.t0) / dt)) : 0.f;
  const dvs_optical_state &a = iv.start, &b = iv.end;
  Pose p;
  p.px = hermite(a.position[0], a.velocity[0], b.position[0], b.velocity[0], u, dt);
  p.py = hermite(a.position[1], a.velocity[1], b.position[1], b.velocity[1], u, dt);
  p.pz = hermite(a.position[2], a.velocity[2], b.position[2], b.velocity[2], u, dt);
  p.lx = hermite(a.light[0], a.light_rate[0], b.light[0], b.light_rate[0], u, dt);
  p.ly = hermite(a.light[1], a.light_rate[1], b.light[1], b.light_rate[1], u, dt);
  p.lz = hermite(a.light[2], a.light_rate[2], b.light[2], b.light_rate[2], u, dt);
  p.intensity = fmaxf(0.f, hermite(a.intensity, a.intensity_rate, b.intensity, b.intensity_rate, u, dt));
  float q0 = b.rotation[0], q1 = b.rotation[1], q2 = b.rotation[2], q3 = b.rotation[3];
  float d = a.rotation[0] * q0 + a.rotation[1] * q1 + a.rotation[2] * q2 + a.rotation[3] * q3;
  if (d < 0) { q0 = -q0; q1 = -q1; q2 = -q2; q3 = -q3; d = -d; }
  float wa = 1 - u, wb = u;
  if (d < .9995f) {
    const float angle = acosf(fminf(1.f, d)), inv = 1 / sinf(angle);
    wa = sinf((1 - u) * angle) * inv;
    wb = sinf(u * angle) * inv;
  }
  const float x = wa * a.rotation[0] + wb * q0, y = wa * a.rotation[1] + wb * q1;
  const float z = wa * a.rotation[2] + wb * q2, w = wa * a.rotation[3] + wb * q3;
  const float norm = 1.0f / sqrtf(x * x + y * y + z * z + w * w);
  p.qx = x * norm; p.qy = y * norm; p.qz = z * norm; p.qw = w * norm;
  p.pad = 0;
  return p;
}
__device__ __forceinline__ Pose pose_of(const dvs_optical_state &s) {
  Pose p;
  p.px = s.position[0]; p.py = s.position[1]; p.pz = s.position[2];
  p.lx = s.light[0]; p.ly = s.light[1]; p.lz = s.light[2];
  p.qx = s.rotation[0]; p.qy = s.rotation[1]; p.qz = s.rotation[2]; p.qw = s.rotation[3];
  p.intensity = s.intensity; p.pad = 0;
  return p;
}
// rotate_ray (device_optics.cuh:27-29)
__device__ __forceinline__ f3 rotate_ray(const Pose &q, f3 v) {
  const f3 qv = mk(q.qx, q.qy, q.qz);
  const f3 c = scale(cross(qv, v), 2.0f);
  return add(add(v, scale(c, q.qw)), cross(qv, c));
}

// ---------------------------------------------------------------------------
// Traversal over the SoA layout.  Order of node visits is the reference's
// near-first order (device_ray.hpp:120-127,153): the nearer child is taken at
// once, the farther one goes on the per-lane LDS stack, so the first triangle
// found at a tied distance is the same one the reference finds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 inverse_direction(f3 d) {  // device_ray.hpp:6-12
  f3 inv;
  inv.x = 1.0f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x));
  inv.y = 1.0f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y));
  inv.z = 1.0f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z));
  return inv;
}
// Scene arrays in the global address space and the traversal stack in LDS:
// generic (flat) accesses would wait on both memory counters.
typedef const __attribute__((address_space(1))) float *GlobalFloats;
typedef const __attribute__((address_space(1))) int *GlobalInts;
typedef __attribute__((address_space(3))) int *LdsInts;
// Wave-uniform reads of the scene: through the constant address space the
// compiler issues scalar loads (one per wave) for a uniform address.
typedef const __attribute__((address_space(4))) float *ConstFloats;
typedef const __attribute__((address_space(4))) int *ConstInts;
// float4 i of a record (16-byte aligned), generic, global or constant.
__device__ __forceinline__ float4 record4(const float4 *rec, int i) { return rec[i]; }
__device__ __forceinline__ float4 record4(GlobalFloats rec, int i) {
  return make_float4(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2], rec[4 * i + 3]);
}
__device__ __forceinline__ float4 record4(ConstFloats rec, int i) {
  return make_float4(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2], rec[4 * i + 3]);
}


// Slab test of both children of one pair record (device_ray.hpp:13-53):
// whether each child's box is entered (hit0, hit1), its entry distances and
// the child references.
template <class Rec>
__device__ __forceinline__ void slab_pair(Rec rec, f3 P, f3 idir, float t_min, float t_max, float &d0, float &d1,
                                          int &left, int &right, bool &hit0, bool &hit1) {
  const float4 a = record4(rec, 0), b = record4(rec, 1), c = record4(rec, 2), r = record4(rec, 3);
  const float t1x0 = (a.x - P.x) * idir.x, t1x1 = (a.w - P.x) * idir.x;
  const float t1y0 = (a.y - P.y) * idir.y, t1y1 = (b.x - P.y) * idir.y;
  const float t1z0 = (a.z - P.z) * idir.z, t1z1 = (b.y - P.z) * idir.z;
  float t1min = fmaxf(fmaxf(fminf(t1x0, t1x1), fminf(t1y0, t1y1)), fminf(t1z0, t1z1));
  float t1max = fminf(fminf(fmaxf(t1x0, t1x1), fmaxf(t1y0, t1y1)), fmaxf(t1z0, t1z1));
  t1min = fmaxf(t_min, t1min); t1max = fminf(t_max, t1max);
  const float t2x0 = (b.z - P.x) * idir.x, t2x1 = (c.y - P.x) * idir.x;
  const float t2y0 = (b.w - P.y) * idir.y, t2y1 = (c.z - P.y) * idir.y;
  const float t2z0 = (c.x - P.z) * idir.z, t2z1 = (c.w - P.z) * idir.z;
  float t2min = fmaxf(fmaxf(fminf(t2x0, t2x1), fminf(t2y0, t2y1)), fminf(t2z0, t2z1));
  float t2max = fminf(fminf(fmaxf(t2x0, t2x1), fmaxf(t2y0, t2y1)), fmaxf(t2z0, t2z1));
  t2min = fmaxf(t_min, t2min); t2max = fminf(t_max, t2max);
  d0 = t1min; d1 = t2min;
  left = __float_as_int(r.x); right = __float_as_int(r.y);
  hit0 = t1max >= t1min;
  hit1 = t2max >= t2min;
}
// The same as a mask: bit 0 the left child, bit 1 the right one.
template <class Rec>
__device__ __forceinline__ int test_pair(Rec rec, f3 P, f3 idir, float t_min, float t_max,
                                         float &d0, float &d1, int &left, int &right) {
  bool hit0, hit1;
  slab_pair(rec, P, idir, t_min, t_max, d0, d1, left, right, hit0, hit1);
  return (hit0 ? 1 : 0) | (hit1 ? 2 : 0);
}
// Moller-Trumbore on a pre-gathered record (device_ray.hpp:54-80).  `meta`
// receives the record's last float4: (e2.z, reference index, leaf lo, leaf hi).
template <class Rec>
__device__ __forceinline__ bool intersect(Rec rec, f3 ray, f3 o, float &ot, float &ou, float &ov, float4 &meta) {
  const float4 a = record4(rec, 0), b = record4(rec, 1), c = record4(rec, 2);
  meta = c;
  const f3 v0 = mk(a.x, a.y, a.z), e1 = mk(a.w, b.x, b.y), e2 = mk(b.z, b.w, c.x);
  const f3 pvec = cross(ray, e2);
  const float det = dot(e1, pvec);
  if (fabsf(det) < kRayEps) return false;
  const float inv_det = 1.0f / det;
  const f3 tvec = sub(o, v0);
  const float u = dot(tvec, pvec) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  const f3 qvec = cross(tvec, e1);
  const float v = dot(ray, qvec) * inv_det;
  if (v < 0.0f || (u + v) > 1.0f) return false;
  const float t = dot(e2, qvec) * inv_det;
  if (t <= kRayEps) return false;
  ot = t; ou = u; ov = v;
  return true;
}

// tri: the reference's BVH-order triangle index; rec: the device record.
struct Hit { int hit, tri, rec; float t, u, v; };

// Traversal work counters: the hot path uses NoCount (compiled away); the
// diagnostic traversal-statistics kernel uses WorkCount.
struct NoCount {
  __device__ __forceinline__ void pair() {}
  __device__ __forceinline__ void leaf() {}
  __device__ __forceinline__ void triangle() {}
};
struct WorkCount {
  unsigned pairs = 0, leaves = 0, triangles = 0;
  __device__ __forceinline__ void pair() { ++pairs; }
  __device__ __forceinline__ void leaf() { ++leaves; }
  __device__ __forceinline__ void triangle() { ++triangles; }
};

// Nearest hit (trace_nearest, device_ray.hpp:136-158).  `stack` points at
// this lane's first LDS slot, consecutive slots `stride` ints apart.
template <class Count>
__device__ inline Hit trace_nearest(const DevScene &s, f3 origin, f3 ray, int *stack, int stride, Count &cnt) {
  Hit h{0, -1, -1, 0.f, 0.f, 0.f};
  float closest = 1e30f;
  const f3 idir = inverse_direction(ray);
  const GlobalFloats pairs = (GlobalFloats)s.pairs, tris = (GlobalFloats)s.triangles;
  const GlobalInts leaves = (GlobalInts)s.leaves;
  const LdsInts lds = (LdsInts)stack;
  // copied once: `s` may be read through a generic pointer, which stores to
  // the LDS stack could alias, so its fields would be reloaded every push
  const int depth = s.stack_depth;
  int ref = s.root_ref, sp = 0;
  for (;;) {
    if (ref < 0) {
      const int2 leaf = make_int2(leaves[2 * ~ref], leaves[2 * ~ref + 1]);
      cnt.leaf();
      for (int i = 0; i < leaf.y; ++i) {
        const int pos = leaf.x + i;
        float t, u, v;
        float4 meta;
        cnt.triangle();
        if (!intersect(tris + 12 * (size_t)pos, ray, origin, t, u, v, meta)) continue;
        // Strictly nearer wins; an equal distance wins only against a hit of
        // the same reference leaf with a higher index (its sequential scan).
        const int id = __float_as_int(meta.y);
        if (t < closest || (t == closest && id < h.tri && h.tri >= __float_as_int(meta.z) &&
                            h.tri < __float_as_int(meta.w))) {
          closest = t; h.hit = 1; h.tri = id; h.rec = pos; h.t = t; h.u = u; h.v = v;
        }
      }
    } else {
      float d0, d1; int l, r;
      cnt.pair();
      const int mask = test_pair(pairs + 16 * (size_t)ref, origin, idir, kRayEps, closest, d0, d1, l, r);
      if (mask == 3) {
        if (sp >= depth) { h.hit = -1; return h; }
        const bool left_first = d0 < d1;
        lds[sp * stride] = left_first ? r : l;
        ++sp;
        ref = left_first ? l : r;
        continue;
      }
      if (mask) { ref = mask == 1 ? l : r; continue; }
    }
    if (sp == 0) break;
    --sp;
    ref = lds[sp * stride];
  }
  return h;
}
// Any hit toward the light skipping the origin's own triangle
// (trace_shadow_ray_any_hit, device_ray.hpp:81-135).  Overflow counts as
// occluded, as in the reference.
template <class Count>
__device__ inline bool trace_any(const DevScene &s, f3 origin, f3 dir, int self, float t_min,
                                 float t_max, int *stack, int stride, Count &cnt) {
  const f3 idir = inverse_direction(dir);
  const GlobalFloats pairs = (GlobalFloats)s.pairs, tris = (GlobalFloats)s.triangles;
  const GlobalInts leaves = (GlobalInts)s.leaves;
  const LdsInts lds = (LdsInts)stack;
  const int depth = s.stack_depth;
  int ref = s.root_ref, sp = 0;
  for (;;) {
    if (ref < 0) {
      const int2 leaf = make_int2(leaves[2 * ~ref], leaves[2 * ~ref + 1]);
      cnt.leaf();
      for (int i = 0; i < leaf.y; ++i) {
        const GlobalFloats rec = tris + 12 * (size_t)(leaf.x + i);
        if (__float_as_int(rec[9]) == self) continue;  // record word 9: the reference triangle index
        float t, u, v;
        float4 meta;
        cnt.triangle();
        if (intersect(rec, dir, origin, t, u, v, meta) && t > t_min && t < t_max) return true;
      }
    } else {
      float d0, d1; int l, r;
      cnt.pair();
      const int mask = test_pair(pairs + 16 * (size_t)ref, origin, idir, t_min, t_max, d0, d1, l, r);
      if (mask == 3) {
        if (sp >= depth) return true;
        const bool left_first = d0 < d1;
        lds[sp * stride] = left_first ? r : l;
        ++sp;
        ref = left_first ? l : r;
        continue;
      }
      if (mask) { ref = mask == 1 ? l : r; continue; }
    }
    if (sp == 0) break;
    --sp;
    ref = lds[sp * stride];
  }
  return false;
}

// ---------------------------------------------------------------------------
// Wave-packet traversal.  The 64 lanes of a wave walk the tree together: a
// node's record is read once per wave by scalar loads (its address is
// wave-uniform), every lane runs the slab and Moller-Trumbore arithmetic of
// the per-lane traversal above on its own ray, and a child is entered by the
// lanes whose own test passed.  The stack is wave-uniform: entry k (a node
// reference and the 64-bit mask of the lanes that entered it) sits in lane
// k of three registers, written by v_writelane and read by v_readlane, so
// it costs no LDS.  A lane tests a box only where its own tests let it in,
// with its own current nearest distance, and a popped entry is not tested
// again, as in the per-lane traversal.
//
// Why the results are those of the per-lane (reference) order:
// * any hit: whether some triangle other than `self` in a leaf under boxes
//   the ray passes (t_min and t_max fixed) hits inside (t_min, t_max) does
//   not depend on the order of the visits;
// * nearest hit: a box holding a hit at distance t is entered whenever the
//   lane's nearest distance is at least t, so every order finds the same
//   nearest distance.  At equal distances the reference keeps the first
//   triangle of its near-first order: within one reference leaf the lower
//   index (its sequential scan), and across reference leaves the one in the
//   subtree entered first at the two leaves' lowest common ancestor, the
//   child with the smaller entry distance, the right one at a tie
//   (device_ray.hpp:120-127).  earlier_leaf() finds that ancestor by the
//   first reference index of each pair's right subtree (pair word 14) and
//   tests its two boxes for the lane's ray.
// The stack holds at most one entry per level of the current path: a push
// at a node of depth k finds at most k entries (one sibling per level above
// it), and the node has children, so the tree is deeper than k.  Trees are
// at most 64 deep (dvs_scene_create), so entry k always has its lane k and
// the stack never overflows.
// ---------------------------------------------------------------------------
constexpr int kWaveStackDepth = 64;  // dvs_scene_create rejects deeper trees
__device__ __forceinline__ int uniform_int(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ unsigned long long lane_bit() { return 1ull << __lane_id(); }

// The wave-uniform stack in lanes of three registers (lane sp takes entry
// sp by a compare and select; v_readlane reads it back).
struct WaveStack {
  int ref = 0, lo = 0, hi = 0;
  // push when `when` (wave-uniform) holds, without a branch
  __device__ __forceinline__ void push_if(bool when, int sp, int node, unsigned long long mask) {
    const bool here = when && (int)__lane_id() == sp;
    ref = here ? node : ref;
    lo = here ? (int)(unsigned)mask : lo;
    hi = here ? (int)(unsigned)(mask >> 32) : hi;
  }
  __device__ __forceinline__ int node(int sp) const { return __builtin_amdgcn_readlane(ref, sp); }
  __device__ __forceinline__ unsigned long long mask(int sp) const {
    return ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, sp) << 32) |
           (unsigned)__builtin_amdgcn_readlane(lo, sp);
  }
};

// One geometry, wave-uniform.
struct WaveGeom {
  ConstFloats pairs, triangles;
  ConstInts leaves;
  int root_ref, depth;
};
__device__ __forceinline__ unsigned long long lane_u64(unsigned long long x, int lane) {
  return ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), lane) << 32) |
         (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, lane);
}
// Every lane's geometry is the scene `s` it traces (per lane); the lanes of
// `todo` are served geometry by geometry, each group by one walk.
template <class F>
__device__ __forceinline__ void for_each_geometry(const DevScene &s, unsigned long long todo, F &&walk) {
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const unsigned long long p = (unsigned long long)(uintptr_t)s.pairs;
    const unsigned long long leader_pairs = lane_u64(p, leader);
    const unsigned long long group = __ballot(p == leader_pairs) & todo;
    WaveGeom g;
    g.pairs = (ConstFloats)(uintptr_t)leader_pairs;
    g.triangles = (ConstFloats)(uintptr_t)lane_u64((unsigned long long)(uintptr_t)s.triangles, leader);
    g.leaves = (ConstInts)(uintptr_t)lane_u64((unsigned long long)(uintptr_t)s.leaves, leader);
    g.root_ref = __builtin_amdgcn_readlane(s.root_ref, leader);
    g.depth = __builtin_amdgcn_readlane(s.stack_depth, leader);
    walk(g, group);
    todo &= ~group;
  }
}

// Whether reference leaf lo_a comes before reference leaf lo_b (distinct
// leaves, given by their first reference triangle index) in the near-first
// order of the ray (origin, idir): the side entered first at their lowest
// common ancestor.  Per lane, only at exact distance ties.
__device__ inline bool earlier_leaf(const WaveGeom &g, int lo_a, int lo_b, f3 origin, f3 idir) {
  const GlobalFloats pairs = (GlobalFloats)(const float *)g.pairs;
  int ref = g.root_ref;
  for (int k = 0; k < kWaveStackDepth && ref >= 0; ++k) {
    const GlobalFloats rec = pairs + 16 * (size_t)ref;
    const int split = __float_as_int(rec[14]);
    const bool a_left = lo_a < split, b_left = lo_b < split;
    if (a_left == b_left) {
      ref = __float_as_int(rec[a_left ? 12 : 13]);
      continue;
    }
    float d0, d1;
    int l, r;
    test_pair(rec, origin, idir, kRayEps, 1e30f, d0, d1, l, r);
    return a_left == (d0 < d1);
  }
  return false;
}

// The side of a pair most of the lanes entering both sides (m0 & m1) would
// take first.  Lanes outside the walk's mask test the pair too (their results
// are masked off: a lane's own state is all its test reads), which keeps the
// lane mask out of the vector registers.
__device__ __forceinline__ bool wave_left_first(float d0, float d1, unsigned long long m0, unsigned long long m1) {
  const unsigned long long both = m0 & m1;
  return 2 * __popcll(both & __ballot(d0 < d1)) >= __popcll(both);
}

// Nearest hit of the lanes of `group` (trace_nearest's result per lane).
__device__ inline void wave_nearest(const WaveGeom &g, unsigned long long group, f3 origin, f3 ray, f3 idir, Hit &h) {
  WaveStack st;
  float closest = 1e30f;
  int best_lo = -1;
  int ref = g.root_ref, sp = 0;
  unsigned long long m = group;
  for (;;) {
    if (ref < 0) {
      const bool mine = (m & lane_bit()) != 0;
      const int2 leaf = make_int2(g.leaves[2 * ~ref], g.leaves[2 * ~ref + 1]);
      for (int i = 0; i < leaf.y; ++i) {
        const ConstFloats rec = g.triangles + 12 * (leaf.x + i);
        float t, u, v;
        float4 meta;
        if (mine && intersect(rec, ray, origin, t, u, v, meta)) {
          const int id = __float_as_int(meta.y), lo = __float_as_int(meta.z);
          bool take = t < closest;
          if (!take && t == closest && best_lo >= 0)
            take = lo == best_lo ? id < h.tri : earlier_leaf(g, lo, best_lo, origin, idir);
          if (take) {
            closest = t; best_lo = lo;
            h.hit = 1; h.tri = id; h.rec = leaf.x + i; h.t = t; h.u = u; h.v = v;
          }
        }
      }
    } else {
      float d0, d1;
      int l, r;
      bool hit0, hit1;
      slab_pair(g.pairs + 16 * ref, origin, idir, kRayEps, closest, d0, d1, l, r, hit0, hit1);
      const unsigned long long m0 = m & __ballot(hit0), m1 = m & __ballot(hit1);
      if (m0 | m1) {
        // both sides: the far one waits on the stack; one side: go there
        const bool both = m0 && m1, left_first = wave_left_first(d0, d1, m0, m1);
        const int near = left_first ? l : r, far = left_first ? r : l;
        const unsigned long long near_m = left_first ? m0 : m1, far_m = left_first ? m1 : m0;
        st.push_if(both, sp, far, far_m);
        sp += both ? 1 : 0;
        ref = near_m ? near : far;
        m = near_m ? near_m : far_m;
        continue;
      }
    }
    if (sp == 0) break;
    --sp;
    ref = st.node(sp);
    m = st.mask(sp);
  }
}

// Any hit toward `dir` of the lanes of `group`, skipping each lane's `self`
// triangle (trace_any's result per lane); returns the occluded lanes.
__device__ inline unsigned long long wave_any(const WaveGeom &g, unsigned long long group, f3 origin, f3 dir, f3 idir,
                                              int self, float t_min, float t_max) {
  WaveStack st;
  unsigned long long done = 0;
  int ref = g.root_ref, sp = 0;
  unsigned long long m = group;
  for (;;) {
    if (ref < 0) {
      const bool mine = (m & lane_bit()) != 0;
      const int2 leaf = make_int2(g.leaves[2 * ~ref], g.leaves[2 * ~ref + 1]);
      bool hit = false;
      for (int i = 0; i < leaf.y; ++i) {
        const ConstFloats rec = g.triangles + 12 * (leaf.x + i);
        float t, u, v;
        float4 meta;
        // record word 9: the reference triangle index
        if (mine && !hit && __float_as_int(rec[9]) != self && intersect(rec, dir, origin, t, u, v, meta) &&
            t > t_min && t < t_max)
          hit = true;
      }
      done |= __ballot(hit);
    } else {
      float d0, d1;
      int l, r;
      bool hit0, hit1;
      slab_pair(g.pairs + 16 * ref, origin, idir, t_min, t_max, d0, d1, l, r, hit0, hit1);
      const unsigned long long m0 = m & __ballot(hit0), m1 = m & __ballot(hit1);
      if (m0 | m1) {
        const bool both = m0 && m1, left_first = wave_left_first(d0, d1, m0, m1);
        const int near = left_first ? l : r, far = left_first ? r : l;
        const unsigned long long near_m = left_first ? m0 : m1, far_m = left_first ? m1 : m0;
        st.push_if(both, sp, far, far_m);
        sp += both ? 1 : 0;
        ref = near_m ? near : far;
        m = near_m ? near_m : far_m;
        continue;
      }
    }
    // the next entry with a lane still looking
    m = 0;
    while (sp > 0 && !m) {
      --sp;
      ref = st.node(sp);
      m = st.mask(sp) & ~done;
    }
    if (!m) break;
  }
  return done;
}

// shadow_origin (device_optics.cuh:37-51)
template <class Rec>
__device__ inline f3 shadow_origin(Rec rec, float u, float v, f3 light) {
  const float4 ra = record4(rec, 0), rb = record4(rec, 1), rc = record4(rec, 2);
  const f3 a = mk(ra.x, ra.y, ra.z), e1 = mk(ra.w, rb.x, rb.y), e2 = mk(rb.z, rb.w, rc.x);
  f3 point = add(add(a, scale(e1, u)), scale(e2, v));
  f3 normal = normalize(cross(e1, e2));
  if (dot(normal, light) < 0) normal = neg(normal);
  constexpr float epsilon = 5.960464477539063e-8f, gamma = 8 * epsilon / (1 - 8 * epsilon);
  const f3 err = mk(gamma * (fabsf(a.x) + fabsf(u * e1.x) + fabsf(v * e2.x)),
                    gamma * (fabsf(a.y) + fabsf(u * e1.y) + fabsf(v * e2.y)),
                    gamma * (fabsf(a.z) + fabsf(u * e1.z) + fabsf(v * e2.z)));
  const float distance = fabsf(normal.x) * err.x + fabsf(normal.y) * err.y + fabsf(normal.z) * err.z + 1e-6f;
  const f3 offset = scale(normal, distance);
  point = add(point, offset);
  if (offset.x != 0) point.x = nextafterf(point.x, copysignf(INFINITY, offset.x));
  if (offset.y != 0) point.y = nextafterf(point.y, copysignf(INFINITY, offset.y));
  if (offset.z != 0) point.z = nextafterf(point.z, copysignf(INFINITY, offset.z));
  return point;
}

// ---------------------------------------------------------------------------
// Hapke BRDF x cos(i) (include/dvs/hapke.hpp:14-157), float.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ float hk_safe_tan(float x) { return tanf(clampf(x, -kPi / 2.0f + 1e-6f, kPi / 2.0f - 1e-6f)); }
__host__ __device__ __forceinline__ float hk_safe_acos(float x) { return acosf(clampf(x, -1.0f, 1.0f)); }
__host__ __device__ __forceinline__ float hk_phase(float cosg, float b, float c) {
  const float omb = 1.0f - b * b;
  const float t1 = ((1.0f + c) * 0.5f) * omb / powf(1.0f - 2.0f * b * cosg + b * b, 1.5f);
  const float t2 = ((1.0f - c) * 0.5f) * omb / powf(1.0f + 2.0f * b * cosg + b * b, 1.5f);
  return t1 + t2;
}
__host__ __device__ __forceinline__ float hk_shoe(float tg, float h) {
  if (h <= 0.0f) return tg <= 0.0f ? 1.0f : 0.0f;
  return 1.0f / (1.0f + tg / h);
}
__host__ __device__ __forceinline__ float hk_cboe(float tg, float h) {
  if (h <= 0.0f) return tg <= 0.0f ? 1.0f : 0.0f;
  const float T = tg / h;
  if (T < 1e-8f) return 1.0f;
  return (1.0f + (1.0f - expf(-T)) / T) / (2.0f * (1.0f + T) * (1.0f + T));
}
__host__ __device__ __forceinline__ float hk_H(float x, float w) {
  if (x <= 1e-8f) return 1.0f;
  const float root = sqrtf(fmaxf(0.0f, 1.0f - w));
  const float r0 = (1.0f - root) / (1.0f + root);
  const float term = r0 + (1.0f - 2.0f * r0 * x) * 0.5f * logf(1.0f + 1.0f / x);
  return 1.0f / (1.0f - w * x * term);
}
__host__ __device__ __forceinline__ float hk_porosity(float phi) {
  if (phi <= 1e-8f) return 1.0f;
  const float p = 1.209f * powf(phi, 2.0f / 3.0f);
  return -logf(1.0f - p) / p;
}
// The roughness helpers take tan(thetap) and chi(thetap) precomputed; the
// reference recomputes the same float expressions (tanf of the same
// argument), so the values are identical.
__host__ __device__ __forceinline__ float hk_E1(float cot_tp, float cy, float sy) {
  return expf((-2.0f / kPi) * cot_tp * cy / fmaxf(sy, 1e-6f));
}
__host__ __device__ __forceinline__ float hk_E2(float cot_tp, float cy, float sy) {
  const float v = cot_tp * cy / fmaxf(sy, 1e-6f);
  return expf((-1.0f / kPi) * v * v);
}
__host__ __device__ __forceinline__ float hk_eta(float ch, float tan_tp, float cot_tp, float cy, float sy) {
  const float e1 = hk_E1(cot_tp, cy, sy), e2 = hk_E2(cot_tp, cy, sy);
  return ch * (cy + sy * tan_tp * e2 / (2.0f - e1));
}
__host__ __device__ inline float hapke_reflectance(float cosi, float cose, float cosg, float albedo, const dvs_hapke &p) {
  if (cosi <= 0 || cose <= 0) return 0;
  cosi = clampf(cosi, 0, 1); cose = clampf(cose, 0, 1); cosg = clampf(cosg, -1, 1);
  const float sini = sqrtf(fmaxf(0, 1 - cosi * cosi)), sine = sqrtf(fmaxf(0, 1 - cose * cose));
  const float tg = hk_safe_tan(hk_safe_acos(cosg) * 0.5f);
  const float psi = hk_safe_acos((cosg - cosi * cose) / (sini * sine + 1e-6f));
  const float pg = hk_phase(cosg, p.b, p.c);
  const float BS = hk_shoe(tg, p.hS), BC = hk_cboe(tg, p.hC), K = hk_porosity(p.phi);
  float mu0, mue, S;
  const float tp = p.thetap;
  if (tp <= 1e-8f) {
    mu0 = cosi; mue = cose; S = 1.0f;
  } else {
    const float tan_tp = tanf(tp);
    const float cot_tp = 1.0f / tan_tp;
    const float ch = 1.0f / sqrtf(1.0f + kPi * tan_tp * tan_tp);
    const float sh = sinf(psi * 0.5f), shv = sh * sh;
    const float cpsi = cosf(psi);
    const float e1i = hk_E1(cot_tp, cosi, sini), e1e = hk_E1(cot_tp, cose, sine);
    const float e2i = hk_E2(cot_tp, cosi, sini), e2e = hk_E2(cot_tp, cose, sine);
    const float eta_i = hk_eta(ch, tan_tp, cot_tp, cosi, sini);
    const float eta_e = hk_eta(ch, tan_tp, cot_tp, cose, sine);
    const float fpsi = expf(-2.0f * tanf(psi * 0.5f));
    if (cosi >= cose) {
      const float den = 2.0f - e1e - (psi / kPi) * e1i;
      mu0 = ch * (cosi + sini * tan_tp * (cpsi * e2e + shv * e2i) / den);
      mue = ch * (cose + sine * tan_tp * (e2e - shv * e2i) / den);
      S = mue / eta_e * cosi / eta_i * ch / (1.0f - fpsi + fpsi * ch * (cosi / eta_i));
    } else {
      const float den = 2.0f - e1i - (psi / kPi) * e1e;
      mu0 = ch * (cosi + sini * tan_tp * (e2i - shv * e2e) / den);
      mue = ch * (cose + sine * tan_tp * (cpsi * e2i + shv * e2e) / den);
      S = mue / eta_e * cosi / eta_i * ch / (1.0f - fpsi + fpsi * ch * (cose / eta_e));
    }
  }
  const float M = hk_H(mu0 / K, albedo) * hk_H(mue / K, albedo) - 1.0f;
  const float LS = mu0 / (mu0 + mue + 1e-6f);
  return (albedo / (4.0f * kPi)) * LS * K * (pg * (1.0f + BS * p.B_S0) + M) * (1.0f + BC * p.B_C0) * S;
}
__host__ __device__ __forceinline__ float hapke_directional(f3 normal, f3 light, f3 view, float albedo, const dvs_hapke &p) {
  const f3 n = normalize(normal), l = normalize(light), v = normalize(view);
  return hapke_reflectance(dot(n, l), dot(n, v), dot(l, v), albedo, p);
}
// Per-scene constants of hapke_reflectance (host).
__host__ __device__ inline HapkeConsts hapke_constants(const dvs_hapke &p) {
  HapkeConsts c{};
  c.rough = p.thetap > 1e-8f;
  c.tan_tp = c.rough ? tanf(p.thetap) : 0.f;
  c.cot_tp = c.rough ? 1.0f / c.tan_tp : 0.f;
  c.ch = c.rough ? 1.0f / sqrtf(1.0f + kPi * c.tan_tp * c.tan_tp) : 1.f;
  c.K = hk_porosity(p.phi);
  return c;
}
// hapke_reflectance for the ray kernels: the same expressions in the same
// order, with the terms that depend on the parameters alone (tan, cot and
// chi of theta-bar, the porosity factor) taken from HapkeConsts, computed
// once per scene on the host by the same float expressions.  The refinement
// decisions sit on radiance last bits (a segment whose residual is at the
// tolerance), so everything else keeps the exact operations.
__device__ inline float hapke_fast(float cosi, float cose, float cosg, float albedo, const dvs_hapke &p,
                                   const HapkeConsts &c) {
  if (cosi <= 0 || cose <= 0) return 0;
  cosi = clampf(cosi, 0, 1); cose = clampf(cose, 0, 1); cosg = clampf(cosg, -1, 1);
  const float sini = sqrtf(fmaxf(0, 1 - cosi * cosi)), sine = sqrtf(fmaxf(0, 1 - cose * cose));
  const float tg = hk_safe_tan(hk_safe_acos(cosg) * 0.5f);
  const float psi = hk_safe_acos((cosg - cosi * cose) / (sini * sine + 1e-6f));
  const float pg = hk_phase(cosg, p.b, p.c);
  const float BS = hk_shoe(tg, p.hS), BC = hk_cboe(tg, p.hC), K = c.K;
  float mu0, mue, S;
  if (!c.rough) {
    mu0 = cosi; mue = cose; S = 1.0f;
  } else {
    const float tan_tp = c.tan_tp, cot_tp = c.cot_tp, ch = c.ch;
    const float sh = sinf(psi * 0.5f), shv = sh * sh;
    const float cpsi = cosf(psi);
    const float e1i = hk_E1(cot_tp, cosi, sini), e1e = hk_E1(cot_tp, cose, sine);
    const float e2i = hk_E2(cot_tp, cosi, sini), e2e = hk_E2(cot_tp, cose, sine);
    const float eta_i = ch * (cosi + sini * tan_tp * e2i / (2.0f - e1i));
    const float eta_e = ch * (cose + sine * tan_tp * e2e / (2.0f - e1e));
    const float fpsi = expf(-2.0f * tanf(psi * 0.5f));
    if (cosi >= cose) {
      const float den = 2.0f - e1e - (psi / kPi) * e1i;
      mu0 = ch * (cosi + sini * tan_tp * (cpsi * e2e + shv * e2i) / den);
      mue = ch * (cose + sine * tan_tp * (e2e - shv * e2i) / den);
      S = mue / eta_e * cosi / eta_i * ch / (1.0f - fpsi + fpsi * ch * (cosi / eta_i));
    } else {
      const float den = 2.0f - e1i - (psi / kPi) * e1e;
      mu0 = ch * (cosi + sini * tan_tp * (e2i - shv * e2e) / den);
      mue = ch * (cose + sine * tan_tp * (cpsi * e2i + shv * e2e) / den);
      S = mue / eta_e * cosi / eta_i * ch / (1.0f - fpsi + fpsi * ch * (cose / eta_e));
    }
  }
  const float M = hk_H(mu0 / K, albedo) * hk_H(mue / K, albedo) - 1.0f;
  const float LS = mu0 / (mu0 + mue + 1e-6f);
  return (albedo / (4.0f * kPi)) * LS * K * (pg * (1.0f + BS * p.B_S0) + M) * (1.0f + BC * p.B_C0) * S;
}
__device__ __forceinline__ float hapke_directional_fast(f3 normal, f3 light, f3 view, float albedo,
                                                        const dvs_hapke &p, const HapkeConsts &c) {
  const f3 n = normalize(normal), l = normalize(light), v = normalize(view);
  return hapke_fast(dot(n, l), dot(n, v), dot(l, v), albedo, p, c);
}
__host__ __device__ __forceinline__ float material_albedo(float value, float fallback) {
  return __builtin_isfinite(value) && value >= 0 && value <= 1 ? value : fallback;
}
// I/F = pi * hapke_reflectance for the angles in degrees, inverted for the
// single-scattering albedo by 32 bisection steps on [0, 1]
// (invert_hapke_if, include/dvs/hapke.hpp:160-172); NaN for invalid
// geometry, missing values or a target outside [I/F(w=0), I/F(w=1)].
__host__ __device__ inline float invert_hapke_if(float inc, float emi, float phase, float target, const dvs_hapke &p) {
  if (!(inc >= 0 && inc < 90 && emi >= 0 && emi < 90 && phase >= 0 && phase <= 180 && target >= 0) ||
      !__builtin_isfinite(target))
    return NAN;
  if (phase > inc + emi + 1e-4f || phase < fabsf(inc - emi) - 1e-4f) return NAN;
  const float ci = cosf(inc * kPi / 180), ce = cosf(emi * kPi / 180), cg = cosf(phase * kPi / 180);
  const float top = kPi * hapke_reflectance(ci, ce, cg, 1, p);
  if (!__builtin_isfinite(top) || target > top) return NAN;
  float lo = 0, hi = 1;
  for (int k = 0; k < 32; ++k) {
    const float w = (lo + hi) * .5f;
    if (kPi * hapke_reflectance(ci, ce, cg, w, p) < target) lo = w; else hi = w;
  }
  return (lo + hi) * .5f;
}

// ---------------------------------------------------------------------------
// One sub-ray of one pixel sample (sample_subray / the body of sample_pixel,
// device_optics.cuh:54-94).  SHADE=false traces the primary ray only (depth).
// ---------------------------------------------------------------------------
struct SubSample {
  double radiance;
  float depth;   // hit distance, -1 on a miss
  int invalid;   // traversal overflow
  int shadow;    // a shadow ray was traced
};
// The sub-ray through sub-sample (sx, sy) of pixel `pixel` (sample_subray,
// device_optics.cuh:58-63).
struct SubRay {
  float px, py;
  f3 ray, origin;
};
__device__ __forceinline__ SubRay subray_of(const DevScene &s, int pixel, const Pose &pose, int sx, int sy) {
  SubRay r;
  const int ss = s.samples;
  r.px = (pixel % s.width + (sx + .5f) / ss - .5f - s.cx) / s.fx;
  r.py = (pixel / s.width + (sy + .5f) / ss - .5f - s.cy) / s.fy;
  r.ray = normalize(rotate_ray(pose, mk(r.px, -r.py, -1.f)));
  r.origin = mk(pose.px, pose.py, pose.pz);
  return r;
}
// The hit's interpolated normal and the light direction; true when the
// surface faces both the light and the camera, i.e. a shadow ray is traced
// from `so` (device_optics.cuh:64-70).
struct Surface {
  f3 normal, light, so;
  uint32_t i0, i1, i2;
  float w;
};
__device__ inline bool lit_surface(const DevScene &s, const Hit &hit, const Pose &pose, f3 ray, Surface &f) {
  const GlobalInts indices = (GlobalInts)s.indices;
  const GlobalFloats normals = (GlobalFloats)s.normals;
  f.i0 = (uint32_t)indices[3 * (size_t)hit.tri];
  f.i1 = (uint32_t)indices[3 * (size_t)hit.tri + 1];
  f.i2 = (uint32_t)indices[3 * (size_t)hit.tri + 2];
  f.w = 1 - hit.u - hit.v;
  const f3 n0 = mk(normals[3 * (size_t)f.i0], normals[3 * (size_t)f.i0 + 1], normals[3 * (size_t)f.i0 + 2]);
  const f3 n1 = mk(normals[3 * (size_t)f.i1], normals[3 * (size_t)f.i1 + 1], normals[3 * (size_t)f.i1 + 2]);
  const f3 n2 = mk(normals[3 * (size_t)f.i2], normals[3 * (size_t)f.i2 + 1], normals[3 * (size_t)f.i2 + 2]);
  f.normal = normalize(add(add(scale(n0, f.w), scale(n1, hit.u)), scale(n2, hit.v)));
  f.light = normalize(mk(pose.lx, pose.ly, pose.lz));
  if (dot(f.normal, f.light) <= 0 || dot(f.normal, neg(ray)) <= 0) return false;
  f.so = shadow_origin((GlobalFloats)s.triangles + 12 * (size_t)hit.rec, hit.u, hit.v, f.light);
  return true;
}
// Radiance of an unshadowed lit hit: albedo, Hapke, intensity and cos^4
// (device_optics.cuh:71-74, material_albedo :31-33).
__device__ inline double lit_radiance(const DevScene &s, const Hit &hit, const Surface &f, const SubRay &r,
                                      const Pose &pose) {
  const float albedo = s.albedos ? material_albedo(s.albedos[f.i0], s.hapke.w) * f.w +
                                       material_albedo(s.albedos[f.i1], s.hapke.w) * hit.u +
                                       material_albedo(s.albedos[f.i2], s.hapke.w) * hit.v
                                 : s.hapke.w;
  const double a = 1.0 + (double)(r.px * r.px) + (double)(r.py * r.py);
  const double cos4 = 1.0 / (a * a);
  return (double)fmaxf(0.f, hapke_directional_fast(f.normal, f.light, neg(r.ray), albedo, s.hapke, s.hk) *
                                pose.intensity) * cos4;
}

template <bool SHADE, class Count = NoCount>
__device__ inline SubSample sample_subray(const DevScene &s, int pixel, const Pose &pose, int sx, int sy,
                                          int *stack, int stride, Count *primary = nullptr,
                                          Count *shadow = nullptr) {
  Count unused_primary, unused_shadow;
  Count &cp = primary ? *primary : unused_primary;
  Count &cs = shadow ? *shadow : unused_shadow;
  SubSample out{0.0, -1.f, 0, 0};
  const SubRay r = subray_of(s, pixel, pose, sx, sy);
  const Hit hit = trace_nearest(s, r.origin, r.ray, stack, stride, cp);
  if (hit.hit < 0) { out.invalid = 1; return out; }
  if (!hit.hit) return out;
  out.depth = hit.t;
  if (!SHADE) return out;
  Surface f;
  if (!lit_surface(s, hit, pose, r.ray, f)) return out;
  out.shadow = 1;
  if (trace_any(s, f.so, f.light, hit.tri, kRayEps, 1e30f, stack, stride, cs)) return out;
  out.radiance = lit_radiance(s, hit, f, r, pose);
  return out;
}

// The same sub-ray sample by the wave-packet traversal.  Every lane of the
// wave calls it (`active` marks the lanes with a sub-ray); the lanes' scenes
// may differ.  The wave traversal cannot overflow (see above).
template <bool SHADE>
__device__ inline SubSample sample_subray_wave(const DevScene &s, int pixel, const Pose &pose, int sx, int sy,
                                               bool active) {
  SubSample out{0.0, -1.f, 0, 0};
  const SubRay r = subray_of(s, pixel, pose, sx, sy);
  const f3 idir = inverse_direction(r.ray);
  Hit hit{0, -1, -1, 0.f, 0.f, 0.f};
  for_each_geometry(s, __ballot(active), [&](const WaveGeom &g, unsigned long long group) {
    wave_nearest(g, group, r.origin, r.ray, idir, hit);
  });
  const bool surface = active && hit.hit > 0;
  if (surface) out.depth = hit.t;
  if (!SHADE) return out;
  Surface f{};
  const bool need = surface && lit_surface(s, hit, pose, r.ray, f);
  unsigned long long occluded = 0;
  const f3 light_idir = inverse_direction(f.light);
  for_each_geometry(s, __ballot(need), [&](const WaveGeom &g, unsigned long long group) {
    occluded |= wave_any(g, group, f.so, f.light, light_idir, hit.tri, kRayEps, 1e30f);
  });
  if (!need) return out;
  out.shadow = 1;
  if (!(occluded & lane_bit())) out.radiance = lit_radiance(s, hit, f, r, pose);
  return out;
}

}  // namespace dvsdev
