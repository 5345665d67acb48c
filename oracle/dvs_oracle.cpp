// CPU oracle for the DVS rendering hot path — TEST INFRASTRUCTURE ONLY.
// Restates the reference's device and host code operation for operation; the
// reference file:line each function follows is cited beside it.  Loaded only
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).
#include "dvs_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// float3 helpers: include/helper_math.hpp:14-77 (same operand order).
// ---------------------------------------------------------------------------
struct f3 { float x, y, z; };
static inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
static inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline f3 operator*(float b, f3 a) { return mk(a.x * b, a.y * b, a.z * b); }
static inline f3 operator*(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
static inline f3 operator/(f3 a, float b) { return mk(a.x / b, a.y / b, a.z / b); }
static inline f3 operator/(float a, f3 b) { return mk(a / b.x, a / b.y, a / b.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 normalize(f3 v) {
  float len = sqrtf(dot(v, v));
  if (len > 0.0f) return v / len;
  return mk(0.f, 0.f, 0.f);
}
static inline f3 fmax3(f3 a, f3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline f3 fmin3(f3 a, f3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline f3 fmax3(f3 a, float b) { return mk(fmaxf(a.x, b), fmaxf(a.y, b), fmaxf(a.z, b)); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline f3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

constexpr float kRayEps = 1e-6f;       // helper_math.hpp:11
constexpr int kMaxStackSize = 128;     // helper_math.hpp:12
constexpr float kPi = 3.14159265358979323846f;  // math_types.hpp

// ---------------------------------------------------------------------------
// BVH build: src/bvh/build.cpp:27-76,126-187 and src/bvh/bining.cpp:9-182.
// ---------------------------------------------------------------------------
struct Box { f3 min, max; };
static inline Box box_empty() {  // common.h:11-16
  return Box{mk(FLT_MAX, FLT_MAX, FLT_MAX), mk(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
}
static inline bool box_is_empty(const Box &b) {  // common.h:18-21
  return (b.min.x > b.max.x) | (b.min.y > b.max.y) | (b.min.z > b.max.z);
}
static inline void box_grow(Box &b, f3 p) {  // common.h:23-31
  b.min.x = (p.x < b.min.x ? p.x : b.min.x);
  b.min.y = (p.y < b.min.y ? p.y : b.min.y);
  b.min.z = (p.z < b.min.z ? p.z : b.min.z);
  b.max.x = (p.x > b.max.x ? p.x : b.max.x);
  b.max.y = (p.y > b.max.y ? p.y : b.max.y);
  b.max.z = (p.z > b.max.z ? p.z : b.max.z);
}
static inline void box_grow(Box &b, const Box &o) {  // common.h:33-53
  if (box_is_empty(o)) return;
  float xmin = std::min(o.min.x, o.max.x), ymin = std::min(o.min.y, o.max.y),
        zmin = std::min(o.min.z, o.max.z);
  float xmax = std::max(o.min.x, o.max.x), ymax = std::max(o.min.y, o.max.y),
        zmax = std::max(o.min.z, o.max.z);
  b.min.x = std::min(b.min.x, xmin); b.min.y = std::min(b.min.y, ymin);
  b.min.z = std::min(b.min.z, zmin);
  b.max.x = std::max(b.max.x, xmax); b.max.y = std::max(b.max.y, ymax);
  b.max.z = std::max(b.max.z, zmax);
}
static inline f3 box_center(const Box &b) { return (b.min + b.max) * 0.5f; }  // common.h:55-57
static inline float box_area(const Box &b) {  // common.h:59-62
  f3 d = b.max - b.min;
  return std::max(d.x * d.y + d.y * d.z + d.z * d.x, 1e-12f);
}
struct Ref { Box bounds; f3 center; int tri; };  // common.h:65-69
static inline float comp(f3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

struct Binning {  // bining.cpp:9-108
  float leafSAH = 0, splitSAH = 0;
  int dim = 0, pos = 0, num_bins = 0;
  f3 scale{};
  Box cent{};
  enum { NUM_BINS = 32 };
  Binning(const Box &range_bounds, int start, int size, const Ref *refs) {
    Box cb = box_empty();
    for (int i = 0; i < size; ++i) box_grow(cb, box_center(refs[start + i].bounds));
    cent = cb;
    num_bins = (int)std::min((size_t)NUM_BINS, (size_t)(4.0f + 0.05f * size));
    f3 extent = fmax3(cb.max - cb.min, 1e-12f);
    scale = (float)num_bins / extent;
    Box bb[3][NUM_BINS];
    int bc[3][NUM_BINS];
    for (int d = 0; d < 3; ++d)
      for (int b = 0; b < NUM_BINS; ++b) { bb[d][b] = box_empty(); bc[d][b] = 0; }
    for (int i = 0; i < size; ++i) {
      const Ref &r = refs[start + i];
      int bin[3];
      for (int d = 0; d < 3; ++d) {
        int idx = (int)((comp(r.center, d) - comp(cb.min, d)) * comp(scale, d));
        bin[d] = std::clamp(idx, 0, num_bins - 1);
      }
      for (int d = 0; d < 3; ++d) { box_grow(bb[d][bin[d]], r.bounds); bc[d][bin[d]]++; }
    }
    leafSAH = box_area(range_bounds) * size;
    float best = FLT_MAX; int bd = 0, bp = 0;
    for (int d = 0; d < 3; ++d) {
      Box rb[NUM_BINS]; int rc[NUM_BINS];
      Box cur = box_empty(); int cnt = 0;
      for (int b = num_bins - 1; b >= 0; --b) {
        box_grow(cur, bb[d][b]); cnt += bc[d][b]; rb[b] = cur; rc[b] = cnt;
      }
      Box lb = box_empty(); int lc = 0;
      for (int b = 1; b < num_bins; ++b) {
        box_grow(lb, bb[d][b - 1]); lc += bc[d][b - 1];
        float cost = box_area(lb) * lc + box_area(rb[b]) * rc[b];
        if (cost < best) { best = cost; bd = d; bp = b; }
      }
    }
    dim = bd; pos = bp; splitSAH = best;
  }
  // bining.cpp:110-182
  void split(int start, int size, Ref *refs, int &lstart, int &lsize, int &rstart,
             int &rsize) const {
    int end = start + size, l = start, r = end - 1;
    float split_val = comp(cent.min, dim) +
                      (float(pos) / num_bins) * (comp(cent.max, dim) - comp(cent.min, dim));
    while (l <= r) {
      if (comp(refs[l].center, dim) < split_val) l++;
      else { std::swap(refs[l], refs[r]); r--; }
    }
    int mid = l;
    lstart = start; lsize = mid - start; rstart = mid; rsize = end - mid;
    if (lsize == 0 || rsize == 0) {
      int mid2 = start + (end - start) / 2;
      lstart = start; lsize = mid2 - start; rstart = mid2; rsize = end - mid2;
    }
  }
};

struct BuildCtx {
  int max_leaf; float node_cost, prim_cost;
  std::vector<Ref> refs;
  std::vector<orc_node> nodes;
};
static int make_leaf(BuildCtx &c, int start, int end, const Box &b) {  // build.cpp:176-187
  orc_node n{};
  n.min[0] = b.min.x; n.min[1] = b.min.y; n.min[2] = b.min.z;
  n.max[0] = b.max.x; n.max[1] = b.max.y; n.max[2] = b.max.z;
  n.left = -1; n.right = -1; n.start = start; n.count = end - start;
  c.nodes.push_back(n);
  return (int)c.nodes.size() - 1;
}
static int build_node(BuildCtx &c, int start, int end) {  // build.cpp:126-174
  int size = end - start;
  if (size <= 0) throw std::runtime_error("BVH build_node empty range");
  Box rb = box_empty();
  for (int i = start; i < end; ++i) box_grow(rb, c.refs[i].bounds);
  if (size <= c.max_leaf) return make_leaf(c, start, end, rb);
  Binning bin(rb, start, size, c.refs.data());
  float leaf_cost = c.prim_cost * bin.leafSAH;
  float split_cost = c.node_cost * box_area(rb) + c.prim_cost * bin.splitSAH;
  if (leaf_cost <= split_cost) return make_leaf(c, start, end, rb);
  int ls, lz, rs, rz;
  bin.split(start, size, c.refs.data(), ls, lz, rs, rz);
  if (lz <= 0 || rz <= 0) return make_leaf(c, start, end, rb);
  int li = build_node(c, ls, ls + lz);
  int ri = build_node(c, rs, rs + rz);
  orc_node n{};
  n.min[0] = rb.min.x; n.min[1] = rb.min.y; n.min[2] = rb.min.z;
  n.max[0] = rb.max.x; n.max[1] = rb.max.y; n.max[2] = rb.max.z;
  n.left = li; n.right = ri; n.start = 0; n.count = 0;
  c.nodes.push_back(n);
  return (int)c.nodes.size() - 1;
}

// ---------------------------------------------------------------------------
// Scene + traversal: include/dvs/device_ray.hpp:6-158.
// ---------------------------------------------------------------------------
struct Scene {
  std::vector<orc_node> nodes;
  int root = 0;
  std::vector<f3> vertices, normals;
  std::vector<uint32_t> indices;  // 3 per triangle
  std::vector<float> albedos;
  bool has_albedo = false;
  orc_camera cam{};
  orc_hapke hapke{};
};

static inline f3 inverse_direction(f3 d) {  // device_ray.hpp:6-12
  f3 inv;
  inv.x = 1.0f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x));
  inv.y = 1.0f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y));
  inv.z = 1.0f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z));
  return inv;
}
static inline int two_children(f3 P, f3 idir, float t_min, float t_max, const orc_node &L,
                               const orc_node &R, float *dist) {  // device_ray.hpp:13-53
  float t1x0 = (L.min[0] - P.x) * idir.x, t1x1 = (L.max[0] - P.x) * idir.x;
  float t1y0 = (L.min[1] - P.y) * idir.y, t1y1 = (L.max[1] - P.y) * idir.y;
  float t1z0 = (L.min[2] - P.z) * idir.z, t1z1 = (L.max[2] - P.z) * idir.z;
  float t1min = fmaxf(fmaxf(fminf(t1x0, t1x1), fminf(t1y0, t1y1)), fminf(t1z0, t1z1));
  float t1max = fminf(fminf(fmaxf(t1x0, t1x1), fmaxf(t1y0, t1y1)), fmaxf(t1z0, t1z1));
  t1min = fmaxf(t_min, t1min); t1max = fminf(t_max, t1max);
  float t2x0 = (R.min[0] - P.x) * idir.x, t2x1 = (R.max[0] - P.x) * idir.x;
  float t2y0 = (R.min[1] - P.y) * idir.y, t2y1 = (R.max[1] - P.y) * idir.y;
  float t2z0 = (R.min[2] - P.z) * idir.z, t2z1 = (R.max[2] - P.z) * idir.z;
  float t2min = fmaxf(fmaxf(fminf(t2x0, t2x1), fminf(t2y0, t2y1)), fminf(t2z0, t2z1));
  float t2max = fminf(fminf(fmaxf(t2x0, t2x1), fmaxf(t2y0, t2y1)), fmaxf(t2z0, t2z1));
  t2min = fmaxf(t_min, t2min); t2max = fminf(t_max, t2max);
  dist[0] = t1min; dist[1] = t2min;
  int mask = 0;
  if (t1max >= t1min) mask |= 1;
  if (t2max >= t2min) mask |= 2;
  return mask;
}
static inline bool intersect_triangle(f3 ray, f3 o, f3 v0, f3 v1, f3 v2, float *ot, float *ou,
                                      float *ov) {  // device_ray.hpp:54-80
  const f3 e1 = v1 - v0, e2 = v2 - v0;
  const f3 pvec = cross(ray, e2);
  const float det = dot(e1, pvec);
  if (fabsf(det) < kRayEps) return false;
  const float inv_det = 1.0f / det;
  const f3 tvec = o - v0;
  const float u = dot(tvec, pvec) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  const f3 qvec = cross(tvec, e1);
  const float v = dot(ray, qvec) * inv_det;
  if (v < 0.0f || (u + v) > 1.0f) return false;
  const float t = dot(e2, qvec) * inv_det;
  if (t <= kRayEps) return false;
  *ot = t; *ou = u; *ov = v;
  return true;
}
static inline void tri_vertices(const Scene &s, int id, f3 &a, f3 &b, f3 &c) {
  a = s.vertices[s.indices[3 * id]]; b = s.vertices[s.indices[3 * id + 1]];
  c = s.vertices[s.indices[3 * id + 2]];
}
static bool shadow_any_hit(const Scene &s, f3 o, f3 dir, int self, float t_min,
                           float t_max) {  // device_ray.hpp:81-135
  f3 idir = inverse_direction(dir);
  int stack[kMaxStackSize]; int sp = 0;
  stack[sp++] = s.root;
  while (sp > 0) {
    const orc_node &node = s.nodes[stack[--sp]];
    if (node.count > 0) {
      for (int i = 0; i < node.count; ++i) {
        int id = node.start + i;
        if (id == self) continue;
        float t, u, v; f3 a, b, c; tri_vertices(s, id, a, b, c);
        if (intersect_triangle(dir, o, a, b, c, &t, &u, &v))
          if (t > t_min && t < t_max) return true;
      }
      continue;
    }
    float dist[2];
    int mask = two_children(o, idir, t_min, t_max, s.nodes[node.left], s.nodes[node.right], dist);
    if (mask == 0) continue;
    if (sp + (mask == 3 ? 2 : 1) > kMaxStackSize) return true;
    if (mask == 3) {
      if (dist[0] < dist[1]) { stack[sp++] = node.right; stack[sp++] = node.left; }
      else { stack[sp++] = node.left; stack[sp++] = node.right; }
    } else if (mask == 1) stack[sp++] = node.left;
    else stack[sp++] = node.right;
  }
  return false;
}
struct Hit { int hit = 0, tri = -1; float t = 0, u = 0, v = 0; };
static Hit trace_nearest(const Scene &s, f3 origin, f3 ray, int64_t *counts = nullptr) {
  // device_ray.hpp:136-158
  Hit hit; float closest = 1e30f;
  f3 idir = inverse_direction(ray);
  int stack[kMaxStackSize], sp = 0; stack[sp++] = s.root;
  while (sp > 0) {
    const orc_node &node = s.nodes[stack[--sp]];
    if (node.count > 0) {
      if (counts) { counts[1]++; counts[2] += node.count; }
      for (int i = 0; i < node.count; ++i) {
        int id = node.start + i; float t, u, v; f3 a, b, c; tri_vertices(s, id, a, b, c);
        if (intersect_triangle(ray, origin, a, b, c, &t, &u, &v) && t < closest) {
          closest = t; hit.hit = 1; hit.tri = id; hit.t = t; hit.u = u; hit.v = v;
        }
      }
    } else {
      if (counts) counts[0]++;
      float dist[2];
      int mask = two_children(origin, idir, kRayEps, closest, s.nodes[node.left],
                              s.nodes[node.right], dist);
      if (!mask) continue;
      if (sp + (mask == 3 ? 2 : 1) > kMaxStackSize) { hit.hit = -1; return hit; }
      if (mask == 3) {
        bool left = dist[0] < dist[1];
        stack[sp++] = left ? node.right : node.left;
        stack[sp++] = left ? node.left : node.right;
      } else stack[sp++] = mask == 1 ? node.left : node.right;
    }
  }
  return hit;
}

// ---------------------------------------------------------------------------
// Hapke: include/dvs/hapke.hpp:14-172.
// ---------------------------------------------------------------------------
static inline float safe_tan(float x) { return tanf(clampf(x, -kPi / 2.0f + 1e-6f, kPi / 2.0f - 1e-6f)); }
static inline float safe_acos(float x) { return acosf(clampf(x, -1.0f, 1.0f)); }
static inline float HG2(float cosg, float b, float c) {  // :22-29
  const float omb = 1.0f - b * b;
  const float t1 = ((1.0f + c) * 0.5f) * omb / powf(1.0f - 2.0f * b * cosg + b * b, 1.5f);
  const float t2 = ((1.0f - c) * 0.5f) * omb / powf(1.0f + 2.0f * b * cosg + b * b, 1.5f);
  return t1 + t2;
}
static inline float B_S(float tg, float hS) {  // :31-34
  if (hS <= 0.0f) return tg <= 0.0f ? 1.0f : 0.0f;
  return 1.0f / (1.0f + tg / hS);
}
static inline float B_C(float tg, float hC) {  // :36-43
  if (hC <= 0.0f) return tg <= 0.0f ? 1.0f : 0.0f;
  const float T = tg / hC;
  if (T < 1e-8f) return 1.0f;
  return (1.0f + (1.0f - expf(-T)) / T) / (2.0f * (1.0f + T) * (1.0f + T));
}
static inline float Hfunc(float x, float w) {  // :45-51
  if (x <= 1e-8f) return 1.0f;
  const float root = sqrtf(fmaxf(0.0f, 1.0f - w));
  const float r0 = (1.0f - root) / (1.0f + root);
  const float term = r0 + (1.0f - 2.0f * r0 * x) * 0.5f * logf(1.0f + 1.0f / x);
  return 1.0f / (1.0f - w * x * term);
}
static inline float porosity_K(float phi) {  // :53-57
  if (phi <= 1e-8f) return 1.0f;
  const float p = 1.209f * powf(phi, 2.0f / 3.0f);
  return -logf(1.0f - p) / p;
}
static inline float chi(float tp) { const float t = tanf(tp); return 1.0f / sqrtf(1.0f + kPi * t * t); }
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
static inline float E1(float tp, float cy, float sy) {
  const float ss = fmaxf(sy, 1e-6f);
#if ORC_PLANT == 8  // E1's factor 2/pi taken as 1/pi
  return expf((-1.0f / kPi) * (1.0f / tanf(tp)) * cy / ss);
#else
  return expf((-2.0f / kPi) * (1.0f / tanf(tp)) * cy / ss);
#endif
}
static inline float E2(float tp, float cy, float sy) {
  const float ss = fmaxf(sy, 1e-6f);
  const float v = (1.0f / tanf(tp)) * cy / ss;
  return expf((-1.0f / kPi) * v * v);
}
static inline float eta(float tp, float cy, float sy) {
  const float ch = chi(tp), e1 = E1(tp, cy, sy), e2 = E2(tp, cy, sy);
#if ORC_PLANT == 7  // eta's denominator with E2 in place of E1
  (void)e1;
  return ch * (cy + sy * tanf(tp) * e2 / (2.0f - e2));
#else
  return ch * (cy + sy * tanf(tp) * e2 / (2.0f - e1));
#endif
}
static inline float f_psi(float psi) { return expf(-2.0f * tanf(psi * 0.5f)); }
struct Rough { float mu0 = 0, mue = 0, S = 0; };
// ORC_PLANT (defined above E1) selects a deliberately wrong variant of one
// term of the rough branch (0 = the reference's formula).  The planted
// builds exist only so the tests can show that their checks of this branch
// (reciprocity, the smooth limit, the published form) fail on each
// plausible mistake.
static inline Rough rough_S(float cosi, float cose, float sini, float sine, float psi,
                            float tp) {  // :86-129
  if (tp <= 1e-8f) return Rough{cosi, cose, 1.0f};
#if ORC_PLANT == 6
  const float ch = 1.0f / sqrtf(1.0f + tanf(tp) * tanf(tp)), ttp = tanf(tp);  // chi without pi
#else
  const float ch = chi(tp), ttp = tanf(tp);
#endif
  const float sh = sinf(psi * 0.5f), shv = sh * sh;
  Rough o;
  if (cosi >= cose) {
#if ORC_PLANT == 1  // E2 operands swapped in mu0
    o.mu0 = ch * (cosi + sini * ttp * (cosf(psi) * E2(tp, cosi, sini) + shv * E2(tp, cose, sine)) /
                             (2.0f - E1(tp, cose, sine) - (psi / kPi) * E1(tp, cosi, sini)));
#elif ORC_PLANT == 2  // E1 operands swapped in mu0's denominator
    o.mu0 = ch * (cosi + sini * ttp * (cosf(psi) * E2(tp, cose, sine) + shv * E2(tp, cosi, sini)) /
                             (2.0f - E1(tp, cosi, sini) - (psi / kPi) * E1(tp, cose, sine)));
#else
    o.mu0 = ch * (cosi + sini * ttp * (cosf(psi) * E2(tp, cose, sine) + shv * E2(tp, cosi, sini)) /
                             (2.0f - E1(tp, cose, sine) - (psi / kPi) * E1(tp, cosi, sini)));
#endif
#if ORC_PLANT == 5  // sign of the shadowed term flipped in mue
    o.mue = ch * (cose + sine * ttp * (E2(tp, cose, sine) + shv * E2(tp, cosi, sini)) /
                             (2.0f - E1(tp, cose, sine) - (psi / kPi) * E1(tp, cosi, sini)));
#else
    o.mue = ch * (cose + sine * ttp * (E2(tp, cose, sine) - shv * E2(tp, cosi, sini)) /
                             (2.0f - E1(tp, cose, sine) - (psi / kPi) * E1(tp, cosi, sini)));
#endif
#if ORC_PLANT == 4  // the other branch's f(psi) denominator
    o.S = o.mue / eta(tp, cose, sine) * cosi / eta(tp, cosi, sini) * ch /
          (1.0f - f_psi(psi) + f_psi(psi) * ch * (cose / eta(tp, cose, sine)));
#else
    o.S = o.mue / eta(tp, cose, sine) * cosi / eta(tp, cosi, sini) * ch /
          (1.0f - f_psi(psi) + f_psi(psi) * ch * (cosi / eta(tp, cosi, sini)));
#endif
  } else {
    o.mu0 = ch * (cosi + sini * ttp * (E2(tp, cosi, sini) - shv * E2(tp, cose, sine)) /
                             (2.0f - E1(tp, cosi, sini) - (psi / kPi) * E1(tp, cose, sine)));
#if ORC_PLANT == 3  // cos(psi) dropped from mue
    o.mue = ch * (cose + sine * ttp * (E2(tp, cosi, sini) + shv * E2(tp, cose, sine)) /
                             (2.0f - E1(tp, cosi, sini) - (psi / kPi) * E1(tp, cose, sine)));
#else
    o.mue = ch * (cose + sine * ttp * (cosf(psi) * E2(tp, cosi, sini) + shv * E2(tp, cose, sine)) /
                             (2.0f - E1(tp, cosi, sini) - (psi / kPi) * E1(tp, cose, sine)));
#endif
    o.S = o.mue / eta(tp, cose, sine) * cosi / eta(tp, cosi, sini) * ch /
          (1.0f - f_psi(psi) + f_psi(psi) * ch * (cose / eta(tp, cose, sine)));
  }
  return o;
}
static float hapke_reflectance(float cosi, float cose, float cosg, float albedo,
                               const orc_hapke &p) {  // :131-151
  if (cosi <= 0 || cose <= 0) return 0;
  cosi = clampf(cosi, 0, 1); cose = clampf(cose, 0, 1); cosg = clampf(cosg, -1, 1);
  const float sini = sqrtf(fmaxf(0, 1 - cosi * cosi)), sine = sqrtf(fmaxf(0, 1 - cose * cose));
  const float tg = safe_tan(safe_acos(cosg) * 0.5f);
  const float psi = safe_acos((cosg - cosi * cose) / (sini * sine + 1e-6f));
  const float pg = HG2(cosg, p.b, p.c);
  const float BS = B_S(tg, p.hS), BC = B_C(tg, p.hC), K = porosity_K(p.phi);
  const Rough r = rough_S(cosi, cose, sini, sine, psi, p.thetap);
  const float M = Hfunc(r.mu0 / K, albedo) * Hfunc(r.mue / K, albedo) - 1.0f;
  const float LS = r.mu0 / (r.mu0 + r.mue + 1e-6f);
  return (albedo / (4.0f * kPi)) * LS * K * (pg * (1.0f + BS * p.B_S0) + M) *
         (1.0f + BC * p.B_C0) * r.S;
}
static float hapke_directional(f3 normal, f3 light, f3 view, float albedo,
                               const orc_hapke &p) {  // :153-157
  f3 n = normalize(normal), l = normalize(light), v = normalize(view);
  return hapke_reflectance(dot(n, l), dot(n, v), dot(l, v), albedo, p);
}

// ---------------------------------------------------------------------------
// Optics: include/dvs/device_optics.cuh:6-94.
// ---------------------------------------------------------------------------
static inline float hermite(float a, float va, float b, float vb, float u, float dt) {  // :6-9
  float u2 = u * u, u3 = u2 * u;
  return a + (-2 * u3 + 3 * u2) * (b - a) + (u3 - 2 * u2 + u) * dt * va + (u3 - u2) * dt * vb;
}
static void optical_state_at(const orc_interval &iv, double t, orc_state &s) {  // :13-26
  float dt = (float)(iv.t1 - iv.t0);
  float u = dt > 0 ? (float)fmin(1.0, fmax(0.0, (t - iv.t0) / dt)) : 0;
  s = iv.start;
  for (int k = 0; k < 3; ++k) {
    s.position[k] = hermite(iv.start.position[k], iv.start.velocity[k], iv.end.position[k],
                            iv.end.velocity[k], u, dt);
    s.light[k] = hermite(iv.start.light[k], iv.start.light_rate[k], iv.end.light[k],
                         iv.end.light_rate[k], u, dt);
  }
  s.intensity = fmaxf(0, hermite(iv.start.intensity, iv.start.intensity_rate, iv.end.intensity,
                                 iv.end.intensity_rate, u, dt));
  const float *a = iv.start.rotation;
  float b[4] = {iv.end.rotation[0], iv.end.rotation[1], iv.end.rotation[2], iv.end.rotation[3]};
  float d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  if (d < 0) { for (float &x : b) x = -x; d = -d; }
  float wa = 1 - u, wb = u;
  if (d < .9995f) {
    float angle = acosf(fminf(1, d)), inv = 1 / sinf(angle);
    wa = sinf((1 - u) * angle) * inv; wb = sinf(u * angle) * inv;
  }
  float q[4];
  for (int k = 0; k < 4; ++k) q[k] = wa * a[k] + wb * b[k];
  // rsqrtf in the reference; restated with IEEE sqrt and division.
  float norm = 1.0f / sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) s.rotation[k] = q[k] * norm;
}
static inline f3 rotate_ray(const float *q, f3 v) {  // :27-29
  f3 qv = mk(q[0], q[1], q[2]);
  f3 c = 2 * cross(qv, v);
  return v + q[3] * c + cross(qv, c);
}
static inline float material_albedo(float value, float fallback) {  // :31-33
  return std::isfinite(value) && value >= 0 && value <= 1 ? value : fallback;
}
static f3 shadow_origin(const Scene &s, int tri, float u, float v, f3 light) {  // :37-51
  f3 a = s.vertices[s.indices[3 * tri]];
  f3 e1 = s.vertices[s.indices[3 * tri + 1]] - a, e2 = s.vertices[s.indices[3 * tri + 2]] - a;
  f3 point = a + u * e1 + v * e2, normal = normalize(cross(e1, e2));
  if (dot(normal, light) < 0) normal = -normal;
  constexpr float epsilon = 5.960464477539063e-8f, gamma = 8 * epsilon / (1 - 8 * epsilon);
  f3 err = mk(gamma * (fabsf(a.x) + fabsf(u * e1.x) + fabsf(v * e2.x)),
              gamma * (fabsf(a.y) + fabsf(u * e1.y) + fabsf(v * e2.y)),
              gamma * (fabsf(a.z) + fabsf(u * e1.z) + fabsf(v * e2.z)));
  float distance = fabsf(normal.x) * err.x + fabsf(normal.y) * err.y + fabsf(normal.z) * err.z + 1e-6f;
  f3 offset = normal * distance;
  point = point + offset;
  if (offset.x != 0) point.x = nextafterf(point.x, copysignf(INFINITY, offset.x));
  if (offset.y != 0) point.y = nextafterf(point.y, copysignf(INFINITY, offset.y));
  if (offset.z != 0) point.z = nextafterf(point.z, copysignf(INFINITY, offset.z));
  return point;
}
struct SubRay { double radiance = 0; float depth = -1; bool valid = true; bool shadow_ray = false; };
// sample_subray (device_optics.cuh:54-72)
static SubRay sample_subray(const Scene &s, int pixel, const orc_state &pose, int sx, int sy,
                            int ss) {
  SubRay out;
  float px = (pixel % s.cam.width + (sx + .5f) / ss - .5f - s.cam.cx) / s.cam.fx;
  float py = (pixel / s.cam.width + (sy + .5f) / ss - .5f - s.cam.cy) / s.cam.fy;
  f3 ray = normalize(rotate_ray(pose.rotation, mk(px, -py, -1)));
  Hit hit = trace_nearest(s, ld3(pose.position), ray);
  if (hit.hit < 0) { out.valid = false; return out; }
  if (!hit.hit) return out;
  out.depth = hit.t;
  const uint32_t *tri = &s.indices[3 * hit.tri];
  float w = 1 - hit.u - hit.v;
  f3 normal = normalize(s.normals[tri[0]] * w + s.normals[tri[1]] * hit.u + s.normals[tri[2]] * hit.v);
  f3 light = normalize(ld3(pose.light));
  if (dot(normal, light) <= 0 || dot(normal, -ray) <= 0) return out;
  f3 origin = shadow_origin(s, hit.tri, hit.u, hit.v, light);
  out.shadow_ray = true;
  if (shadow_any_hit(s, origin, light, hit.tri, kRayEps, 1e30f)) return out;
  float albedo = s.has_albedo
                     ? material_albedo(s.albedos[tri[0]], s.hapke.w) * w +
                           material_albedo(s.albedos[tri[1]], s.hapke.w) * hit.u +
                           material_albedo(s.albedos[tri[2]], s.hapke.w) * hit.v
                     : s.hapke.w;
  double cos4 = 1.0 / ((1.0 + px * px + py * py) * (1.0 + px * px + py * py));
  out.radiance = fmaxf(0, hapke_directional(normal, light, -ray, albedo, s.hapke) * pose.intensity) * cos4;
  return out;
}
struct PixelSample { double radiance = 0; float depth = -1; bool valid = true; int shadow_rays = 0; };
// sample_pixel (device_optics.cuh:73-94).  The per-subray work is the same
// sequence of operations as sample_subray; accumulation is sequential sy,sx.
static PixelSample sample_pixel(const Scene &s, int pixel, const orc_state &pose) {
  PixelSample out; double depth = 0; int valid_depth = 0;
  int ss = s.cam.samples;
  for (int sy = 0; sy < ss; ++sy)
    for (int sx = 0; sx < ss; ++sx) {
      SubRay r = sample_subray(s, pixel, pose, sx, sy, ss);
      if (!r.valid) { out.valid = false; return out; }
      out.shadow_rays += r.shadow_ray;
      if (r.depth < 0) continue;  // miss
      depth += r.depth; ++valid_depth;
      out.radiance += r.radiance;
    }
  out.radiance /= ss * ss;
  out.depth = valid_depth ? (float)(depth / valid_depth) : -1;
  return out;
}

// ---------------------------------------------------------------------------
// Circuit: include/dvs/pixel_model.hpp:18-102.
// ---------------------------------------------------------------------------
static inline double receptor_log_at(double u, double rate, double dt, double charge) {
  if (dt <= 0) return u;
  rate = fmax(0.0, rate);
  double x = exp(u);
  if (rate == 0) return u - log1p(x * dt / charge);
  double a = exp(-rate * dt / charge), b = -expm1(-rate * dt / charge);
  return -log(a / x + b / rate);
}
static inline double contrast_at(double u, double rate, double dt, double ref,
                                 const orc_pixel_params &p) {
  return receptor_log_at(u, rate, dt, p.charge_scale) - ref + p.reference_leak * dt;
}
template <class Normal>
static int advance_pixel(orc_pixel_state &s, const orc_pixel_params &p, double rate, double t0,
                         double t1, orc_pixel_event *out, int capacity, Normal &&normal) {
  if (t1 <= t0) return 0;
  int count = 0; double cursor = t0, u = s.receptor_log;
  while (cursor < t1) {
    if (cursor < s.refractory_until) {
      double end = fmin(t1, s.refractory_until);
      u = receptor_log_at(u, rate, end - cursor, p.charge_scale);
      cursor = end; s.reference_log = u;
      if (cursor < t1 || cursor == s.refractory_until) {
        if (cursor >= s.refractory_until && s.reset_pending) {
          if (p.reset_noise_std > 0) s.reference_log += p.reset_noise_std * normal();
          s.reset_pending = 0;
        }
      }
      if (cursor >= t1) break;
    } else if (s.reset_pending) {
      if (p.reset_noise_std > 0) s.reference_log += p.reset_noise_std * normal();
      s.reset_pending = 0;
    }
    double remaining = t1 - cursor, c0 = u - s.reference_log, boundary = remaining;
    double x = exp(u), target = rate + p.charge_scale * p.reference_leak;
    if (p.reference_leak > 0 && x > target && target > rate) {
      double extremum;
      if (rate > 0) extremum = -p.charge_scale / rate * log((1.0 / target - 1.0 / rate) / (1.0 / x - 1.0 / rate));
      else extremum = p.charge_scale * (1.0 / target - 1.0 / x);
      if (extremum > 1e-15 && extremum < boundary) boundary = extremum;
    }
    double ub = receptor_log_at(u, rate, boundary, p.charge_scale);
    double cb = ub - s.reference_log + p.reference_leak * boundary;
    int polarity = -1; double level = 0;
    if (c0 >= s.threshold_pos) { polarity = 1; level = s.threshold_pos; }
    else if (c0 <= -s.threshold_neg) { polarity = 0; level = -s.threshold_neg; }
    else if (cb >= s.threshold_pos) { polarity = 1; level = s.threshold_pos; }
    else if (cb <= -s.threshold_neg) { polarity = 0; level = -s.threshold_neg; }
    if (polarity < 0) {
      u = ub; s.reference_log -= p.reference_leak * boundary; cursor += boundary;
      continue;
    }
    if (count == capacity) return -1;
    double lo = 0, hi = boundary;
    if (c0 >= s.threshold_pos || c0 <= -s.threshold_neg) hi = 0;
    else if (p.reference_leak == 0) {
      double target_inverse = exp(-(s.reference_log + level));
      double initial_inverse = exp(-u);
      if (rate > 0) {
        double stationary_inverse = 1.0 / rate;
        hi = -p.charge_scale / rate * log1p((target_inverse - initial_inverse) / (initial_inverse - stationary_inverse));
      } else hi = p.charge_scale * (target_inverse - initial_inverse);
      hi = fmin(boundary, fmax(0.0, hi));
    } else for (int j = 0; j < 48; ++j) {
      double mid = (lo + hi) * .5, c = contrast_at(u, rate, mid, s.reference_log, p);
      if (polarity ? c >= level : c <= level) hi = mid; else lo = mid;
    }
    u = receptor_log_at(u, rate, hi, p.charge_scale); cursor += hi;
    out[count].time = cursor; out[count].polarity = polarity; out[count].pad = 0; ++count;
    s.reference_log = u; s.refractory_until = cursor + p.refractory_time;
    s.reset_pending = 1;
  }
  s.receptor_log = u;
  return count;
}

// Model parameters: src/cuda/event_sensor.cu:13,151-176.
constexpr double kQ = 1.602176634e-19, kK = 1.380649e-23, kH = 6.62607015e-34, kC = 299792458;
static void parameters(const orc_physic &c, bool effects, orc_pixel_params &p, orc_photo_params &ph) {
  p.threshold_pos = c.threshold_pos; p.threshold_neg = c.threshold_neg;
  double voltage = kK * c.temperature_k / (kQ * c.subthreshold_kappa);
  p.charge_scale = c.receptor_capacitance_f * voltage / kQ;
  if (effects) {
    p.refractory_time = c.refractory_time;
    p.reference_leak = c.reset_leakage_a / (c.reset_capacitance_f * voltage);
    p.reset_noise_std = c.reset_thermal_noise ? sqrt(kK * c.temperature_k / c.reset_capacitance_f) / voltage : 0;
  } else {
    p.refractory_time = 0; p.reference_leak = 0; p.reset_noise_std = 0;
  }
  double pitch = c.pixel_pitch_um * 1e-6;
  ph.electrons_per_radiance = c.radiance_scale * 3.141592653589793 / (4 * c.f_number * c.f_number) *
                              c.lens_transmission * pitch * pitch * c.fill_factor *
                              c.quantum_efficiency * (c.wavelength_nm * 1e-9) / (kH * kC);
  ph.integration_step = c.circuit_step;
  if (effects) { ph.dark_electrons = c.dark_current_a / kQ; ph.shot_noise = c.photon_shot_noise; }
  else { ph.dark_electrons = 0; ph.shot_noise = 0; }
  ph.pad = 0;
}

// ---------------------------------------------------------------------------
// Noise.  The reference draws cuRAND XORWOW (event_sensor.cu:21,26-27,
// photon_noise.cuh:4-9), a closed library; this build replaces it with a
// counter-based Philox4x64-10 (Salmon et al., SC'11) keyed by (seed, stream)
// with counter (block, pixel).  Samplers: Poisson by sequential-search
// inversion for lambda < 10 (Devroye 1986, X.3), by the transformed rejection
// PTRS method (Hoermann 1993) up to 1e6, and the reference's rounded normal
// approximation above 1e6; normals by Box-Muller.
// ---------------------------------------------------------------------------
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64); lo = (uint64_t)p;
}
static void philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(0xD2E7470EE14C6C93ULL, c0, hi0, lo0);
    mulhilo64(0xCA5A826395121157ULL, c2, hi1, lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
enum { kStreamShot = 0, kStreamThreshold = 1, kStreamReset = 2, kStreamShotShared = 3 };
constexpr uint64_t kStreamTag = 0x4456535F4E4F4953ULL;  // key word 1 = tag ^ stream
// Counter (block, pixel, word, 0).  Shot noise: one stream per circuit cell,
// word = the pixel's cell serial n since initialisation, blocks from 0,
// except that the inversion branch (lambda < 10) takes one 32-bit uniform:
// half n mod 2 of word (n / 2) mod 4 of block n / 8 of the shared shot stream
// (word 0); reset and threshold noise: word 0 and a per-pixel block counter.
struct Stream {
  uint64_t seed, pixel, stream;
  uint64_t *block;  // next unused block index for this stream
  uint64_t word = 0;
  uint64_t buf[4]; int used = 4;
  double uniform() {
    if (used == 4) {
      uint64_t ctr[4] = {(*block)++, pixel, word, 0}, key[2] = {seed, kStreamTag ^ stream};
      philox4x64(ctr, key, buf); used = 0;
    }
    return ((double)(buf[used++] >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
};
static double normal_draw(Stream &st) {
  st.used = 4;  // every normal starts a fresh block
  double u1 = st.uniform(), u2 = st.uniform();
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static double poisson_draw(Stream &st, double lambda, double first = -1) {
  if (lambda != lambda) return lambda;  // NaN in, NaN out (rejection would never accept)
  if (lambda <= 0) return 0;
  if (lambda > 1e6) return fmax(0.0, floor(lambda + sqrt(lambda) * normal_draw(st) + .5));
  st.used = 4;  // every draw starts a fresh block
  if (lambda < 10) {
    double u = first > 0 ? first : st.uniform(), p = exp(-lambda), F = p; int k = 0;
    while (u > F && k < 1000) { ++k; p *= lambda / k; F += p; }
    return k;
  }
  double slam = sqrt(lambda), loglam = log(lambda);
  double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
  for (;;) {
    double U = st.uniform() - 0.5, V = st.uniform();
    double us = 0.5 - fabs(U);
    double k = floor((2 * a / us + b) * U + lambda + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lambda + k * loglam - lgamma(k + 1)) return k;
  }
}
// Photocurrent of one circuit cell with its shot-noise draw (cell serial n).
static double cell_rate(double light, const orc_photo_params &ph, double dt, uint64_t seed, uint64_t pixel,
                        uint64_t serial) {
  double rate = light * ph.electrons_per_radiance + ph.dark_electrons;
  if (ph.shot_noise) {
    uint64_t block = 0;
    Stream shot{seed, pixel, kStreamShot, &block, serial};
    uint64_t ctr[4] = {serial >> 3, pixel, 0, 0}, key[2] = {seed, kStreamTag ^ kStreamShotShared}, words[4];
    philox4x64(ctr, key, words);
    const uint64_t w = words[(serial >> 1) & 3];
    const uint32_t half = (serial & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
    const double first = ((double)half + 0.5) * (1.0 / 4294967296.0);
    rate = poisson_draw(shot, rate * dt, first) / dt;
  }
  return rate;
}

// ---------------------------------------------------------------------------
// Readout FIFO: src/event_readout.cpp:6-34.
// ---------------------------------------------------------------------------
struct Readout {
  double bandwidth, latency, available = 0;
  size_t capacity, dropped = 0, peak = 0;
  struct Pending { double ready; orc_event e; };
  std::deque<Pending> pending;
  Readout(double b, double l, size_t c) : bandwidth(b), latency(l), capacity(c) {}
  void submit(std::vector<orc_event> events, double now, std::vector<orc_event> &ready) {
    ready.clear();
    std::stable_sort(events.begin(), events.end(), [](const orc_event &a, const orc_event &b) {
      if (a.timestamp != b.timestamp) return a.timestamp < b.timestamp;
      if (a.y != b.y) return a.y < b.y;
      return a.x < b.x;
    });
    for (const auto &e : events) {
      while (!pending.empty() && pending.front().ready <= e.timestamp) {
        ready.push_back(pending.front().e); pending.pop_front();
      }
      if (bandwidth <= 0) { ready.push_back(e); continue; }
      if (pending.size() >= capacity) { ++dropped; continue; }
      double completed = std::max(available, (double)e.timestamp) + 1.0 / bandwidth;
      available = completed;
      orc_event d = e; d.timestamp = completed + latency;
      pending.push_back({completed + latency, d});
      peak = std::max(peak, pending.size());
    }
    while (!pending.empty() && pending.front().ready <= now) {
      ready.push_back(pending.front().e); pending.pop_front();
    }
  }
  void drain(std::vector<orc_event> &ready) {
    ready.clear();
    while (!pending.empty()) { ready.push_back(pending.front().e); pending.pop_front(); }
  }
};

// ---------------------------------------------------------------------------
// Sensor: src/cuda/event_sensor.cu:18-60,244-305.
// ---------------------------------------------------------------------------
struct Sensor {
  int width, height, count, max_events;
  uint64_t seed;
  orc_pixel_params p{};
  orc_photo_params photo{};
  double threshold_std, threshold_min, last_time = 0;
  bool initialized = false;
  std::vector<orc_pixel_state> states;
  std::vector<uint64_t> shot_block, reset_block;
  Readout readout;
  int64_t shadow_rays = 0;
  // the last adaptive window's knots ({fraction, radiance} pairs, `capacity`
  // per pixel; count -1 for a native fallback pixel) for verify_optics
  std::vector<double> win_knots;
  std::vector<int> win_kcount;
  std::vector<orc_interval> win_ivs;
  int win_capacity = 0;
  Sensor(int w, int h, const orc_physic &c, bool effects, bool aer)
      : width(w), height(h), count(w * h), max_events(c.max_events_per_step),
        seed((uint64_t)(uint32_t)c.random_seed), threshold_std(c.threshold_std),
        threshold_min(c.threshold_min),
        readout(aer ? c.readout_bandwidth : 0, aer ? c.readout_latency : 0,
                (size_t)c.readout_queue_capacity) {
    if (!effects) threshold_std = 0;
    parameters(c, effects, p, photo);
    states.assign(count, orc_pixel_state{});
    shot_block.assign(count, 0); reset_block.assign(count, 0);
  }
};

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_bvh_build(int max_leaf, float node_cost, float prim_cost, const float *vertices, int nv,
                  uint32_t *indices, int nt, orc_node *nodes, int *node_count) {
  if (nt <= 0) { *node_count = 0; return -1; }
  BuildCtx c{max_leaf, node_cost, prim_cost, {}, {}};
  c.refs.reserve(nt);
  for (int i = 0; i < nt; ++i) {
    f3 a = ld3(vertices + 3 * (size_t)indices[3 * i]), b = ld3(vertices + 3 * (size_t)indices[3 * i + 1]),
       d = ld3(vertices + 3 * (size_t)indices[3 * i + 2]);
    Box box{fmin3(fmin3(a, b), d), fmax3(fmax3(a, b), d)};
    c.refs.push_back(Ref{box, box_center(box), i});
  }
  (void)nv;
  c.nodes.reserve(2 * (size_t)nt);
  int root = build_node(c, 0, nt);
  std::vector<uint32_t> out(3 * (size_t)nt);
  for (int i = 0; i < nt; ++i)
    for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = indices[3 * (size_t)c.refs[i].tri + k];
  std::memcpy(indices, out.data(), out.size() * sizeof(uint32_t));
  std::memcpy(nodes, c.nodes.data(), c.nodes.size() * sizeof(orc_node));
  *node_count = (int)c.nodes.size();
  return root;
}

int orc_bvh_validate(const orc_node *nodes, int n, int root) {  // build.cpp:188-242
  if (root < 0 || root >= n) return -1;
  std::vector<char> seen(n, 0);
  std::vector<int> st{root};
  while (!st.empty()) {
    int i = st.back(); st.pop_back();
    if (i < 0 || i >= n) return -2;
    if (seen[i]) return -3;
    seen[i] = 1;
    if (nodes[i].count > 0) {
      if (nodes[i].left != -1 || nodes[i].right != -1) return -4;
      continue;
    }
    if (nodes[i].left < 0 || nodes[i].left >= n || nodes[i].right < 0 || nodes[i].right >= n) return -5;
    st.push_back(nodes[i].right); st.push_back(nodes[i].left);
  }
  return 0;
}

void *orc_scene_create(const orc_node *nodes, int node_count, int root, const float *vertices,
                       const float *normals, int nv, const uint32_t *indices, int nt,
                       const float *albedos, orc_camera camera, orc_hapke hapke) {
  auto *s = new Scene();
  s->nodes.assign(nodes, nodes + node_count);
  s->root = root;
  s->vertices.resize(nv); s->normals.resize(nv);
  for (int i = 0; i < nv; ++i) { s->vertices[i] = ld3(vertices + 3 * (size_t)i); s->normals[i] = ld3(normals + 3 * (size_t)i); }
  s->indices.assign(indices, indices + 3 * (size_t)nt);
  if (albedos) { s->albedos.assign(albedos, albedos + nv); s->has_albedo = true; }
  s->cam = camera; s->hapke = hapke;
  return s;
}
void orc_scene_destroy(void *s) { delete static_cast<Scene *>(s); }

int orc_trace_nearest(const void *sp, const float o[3], const float d[3], int *tri, float *t,
                      float *u, float *v) {
  Hit h = trace_nearest(*static_cast<const Scene *>(sp), ld3(o), ld3(d));
  *tri = h.tri; *t = h.t; *u = h.u; *v = h.v;
  return h.hit;
}
int orc_trace_nearest_counted(const void *sp, const float o[3], const float d[3], int64_t counts[3]) {
  return trace_nearest(*static_cast<const Scene *>(sp), ld3(o), ld3(d), counts).hit;
}
int orc_trace_shadow(const void *sp, const float o[3], const float d[3], int self, float t_min,
                     float t_max) {
  return shadow_any_hit(*static_cast<const Scene *>(sp), ld3(o), ld3(d), self, t_min, t_max);
}
void orc_shadow_origin(const void *sp, int tri, float u, float v, const float light[3], float out[3]) {
  f3 p = shadow_origin(*static_cast<const Scene *>(sp), tri, u, v, ld3(light));
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
}
int orc_sample_pixel(const void *sp, int pixel, const orc_state *pose, double *rad, float *depth) {
  PixelSample r = sample_pixel(*static_cast<const Scene *>(sp), pixel, *pose);
  *rad = r.radiance; *depth = r.depth;
  return r.valid ? 1 : 0;
}
void orc_optical_state_at(const orc_interval *iv, double t, orc_state *out) { optical_state_at(*iv, t, *out); }

int orc_render(const void *sp, const orc_state *pose, float *image, float *depth, double *radiance,
               int threads) {
  const Scene &s = *static_cast<const Scene *>(sp);
  set_threads(threads);
  int n = s.cam.width * s.cam.height, bad = 0, R = s.cam.samples * s.cam.samples;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    // sample_warps (optical_sampler.cu:8-23): ordered sum over j = sy*ss+sx.
    double total = 0, ranges = 0; int hits = 0;
    for (int j = 0; j < R; ++j) {
      SubRay r = sample_subray(s, i, *pose, j % s.cam.samples, j / s.cam.samples, s.cam.samples);
      if (!r.valid) bad = 1;
      total += r.radiance;
      if (r.depth >= 0) { ranges += r.depth; ++hits; }
    }
    double rad = total / R;
    if (radiance) radiance[i] = rad;
    if (image) image[i] = (float)rad;
    if (depth) depth[i] = hits ? (float)(ranges / hits) : -1;
  }
  return bad ? 2 : 0;
}

float orc_hapke_reflectance(float ci, float ce, float cg, float a, orc_hapke p) { return hapke_reflectance(ci, ce, cg, a, p); }
float orc_hapke_directional(const float n[3], const float l[3], const float v[3], float a, orc_hapke p) {
  return hapke_directional(ld3(n), ld3(l), ld3(v), a, p);
}
float orc_invert_hapke_if(float inc, float emi, float phase, float target, orc_hapke p) {  // hapke.hpp:160-172
  if (!(inc >= 0 && inc < 90 && emi >= 0 && emi < 90 && phase >= 0 && phase <= 180 && target >= 0) || !std::isfinite(target)) return NAN;
  if (phase > inc + emi + 1e-4f || phase < fabsf(inc - emi) - 1e-4f) return NAN;
  float ci = cosf(inc * kPi / 180), ce = cosf(emi * kPi / 180), cg = cosf(phase * kPi / 180);
  float top = kPi * hapke_reflectance(ci, ce, cg, 1, p);
  if (!std::isfinite(top) || target > top) return NAN;
  float lo = 0, hi = 1;
  for (int k = 0; k < 32; ++k) {
    float w = (lo + hi) * .5f;
    if (kPi * hapke_reflectance(ci, ce, cg, w, p) < target) lo = w; else hi = w;
  }
  return (lo + hi) * .5f;
}

double orc_receptor_log_at(double u, double rate, double dt, double charge) { return receptor_log_at(u, rate, dt, charge); }
int orc_advance_pixel(orc_pixel_state *s, const orc_pixel_params *p, double rate, double t0,
                      double t1, orc_pixel_event *out, int capacity, double fixed_normal) {
  return advance_pixel(*s, *p, rate, t0, t1, out, capacity, [&] { return fixed_normal; });
}
void orc_parameters(const orc_physic *c, int effects, orc_pixel_params *p, orc_photo_params *photo) {
  parameters(*c, effects != 0, *p, *photo);
}

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) { philox4x64(ctr, key, out); }
// Sequential-search inversion (the lambda < 10 branch of poisson_draw).
double orc_poisson_inversion(double lambda, double u) {
  double p = exp(-lambda), F = p;
  int k = 0;
  while (u > F && k < 1000) { ++k; p *= lambda / k; F += p; }
  return k;
}
void orc_poisson_series(uint64_t seed, uint64_t pixel, double lambda, int n, double *out) {
  uint64_t block = 0; Stream st{seed, pixel, kStreamShot, &block};
  for (int i = 0; i < n; ++i) out[i] = poisson_draw(st, lambda);
}
void orc_shot_counts(uint64_t seed, uint64_t pixel, double lambda, uint64_t serial0, int n, double *out) {
  const orc_photo_params ph{1.0, 0.0, 1.0, 1, 0};  // rate = light, dt = 1: the draw is the electron count
  for (int i = 0; i < n; ++i) out[i] = cell_rate(lambda, ph, 1.0, seed, pixel, serial0 + (uint64_t)i);
}
void orc_normal_series(uint64_t seed, uint64_t pixel, int n, double *out) {
  uint64_t block = 0; Stream st{seed, pixel, kStreamReset, &block};
  for (int i = 0; i < n; ++i) out[i] = normal_draw(st);
}

void *orc_sensor_create(int w, int h, const orc_physic *c, int effects, int aer) {
  if (w <= 0 || h <= 0 || w > 32767 || h > 32767 || c->max_events_per_step < 1 || c->max_events_per_step > 64)
    return nullptr;
  return new Sensor(w, h, *c, effects != 0, aer != 0);
}
void orc_sensor_destroy(void *s) { delete static_cast<Sensor *>(s); }

// initialize_pixels (event_sensor.cu:18-28) + EventSensor::initialize (:262-270).
int orc_sensor_initialize(void *sp, const void *scp, const orc_state *pose, double timestamp, int threads) {
  Sensor &s = *static_cast<Sensor *>(sp);
  const Scene &sc = *static_cast<const Scene *>(scp);
  if (!std::isfinite(timestamp) || s.initialized) return 3;
  set_threads(threads);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int i = 0; i < s.count; ++i) {
    PixelSample o = sample_pixel(sc, i, *pose);
    if (!o.valid || !std::isfinite(o.radiance)) { bad = 1; continue; }
    orc_pixel_state st{};
    st.refractory_until = -1; st.reset_pending = 0;
    st.receptor_log = st.reference_log = log(fmax(1e-30, o.radiance * s.photo.electrons_per_radiance + s.photo.dark_electrons));
    double zp = 0, zn = 0;
    if (s.threshold_std != 0) {
      uint64_t block = 0; Stream rs{s.seed, (uint64_t)i, kStreamThreshold, &block};
      zp = normal_draw(rs); zn = normal_draw(rs);
    }
    st.threshold_pos = fmax(s.threshold_min, s.p.threshold_pos + s.threshold_std * zp);
    st.threshold_neg = fmax(s.threshold_min, s.p.threshold_neg + s.threshold_std * zn);
    s.states[i] = st;
  }
  if (bad) return 2;
  s.last_time = timestamp; s.initialized = true;
  return 0;
}

// process_many (event_sensor.cu:274-305) over generate_pixels (:29-60).
int orc_sensor_process(void *sp, const void *scp, const orc_interval *ivs, int n, orc_event *raw,
                       int64_t raw_cap, int64_t *raw_count, orc_event *ready, int64_t ready_cap,
                       int64_t *ready_count, int threads) {
  Sensor &s = *static_cast<Sensor *>(sp);
  const Scene &sc = *static_cast<const Scene *>(scp);
  if (!s.initialized || n <= 0) return 3;
  double previous = s.last_time, cells = 0;
  for (int k = 0; k < n; ++k) {
    const orc_interval &iv = ivs[k];
    if (!std::isfinite(iv.t0) || !std::isfinite(iv.t1) || iv.t0 != previous || iv.t1 <= iv.t0) return 3;
    cells += (iv.t1 - iv.t0) / s.photo.integration_step; previous = iv.t1;
  }
  if (cells > 1000000) return 3;
  set_threads(threads);
  std::vector<std::vector<orc_event>> per_pixel(s.count);
  int err_capacity = 0, err_invalid = 0; int64_t shadow = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err_capacity, err_invalid) reduction(+ : shadow)
  for (int i = 0; i < s.count; ++i) {
    std::vector<orc_pixel_event> local(s.max_events);
    int count = 0; bool failed = false;
    Stream reset{s.seed, (uint64_t)i, kStreamReset, &s.reset_block[i]};
    auto normal = [&] { return normal_draw(reset); };
    for (int k = 0; k < n && !failed; ++k) {
      const orc_interval &iv = ivs[k];
      int steps = (int)ceil((iv.t1 - iv.t0) / s.photo.integration_step);
      double dt = (iv.t1 - iv.t0) / steps;
      for (int step = 0; step < steps; ++step) {
        double t0 = iv.t0 + step * dt, t1 = step + 1 == steps ? iv.t1 : iv.t0 + (step + 1) * dt;
        orc_state pose; optical_state_at(iv, (t0 + t1) * .5, pose);
        PixelSample o = sample_pixel(sc, i, pose);
        shadow += o.shadow_rays;
        if (!o.valid || !std::isfinite(o.radiance)) { err_invalid = 1; failed = true; break; }
        const double rate = cell_rate(o.radiance, s.photo, t1 - t0, s.seed, (uint64_t)i, s.shot_block[i]++);
        int emitted = advance_pixel(s.states[i], s.p, rate, t0, t1, local.data() + count, s.max_events - count, normal);
        if (emitted < 0) { err_capacity = 1; failed = true; break; }
        count += emitted;
      }
    }
    if (failed) continue;
    int interval = 0;
    per_pixel[i].reserve(count);
    for (int j = 0; j < count; ++j) {
      while (interval + 1 < n && local[j].time > ivs[interval].t1) ++interval;
      orc_state pose; optical_state_at(ivs[interval], local[j].time, pose);
      PixelSample o = sample_pixel(sc, i, pose);
      if (!o.valid) { err_invalid = 1; break; }
      orc_event e; e.timestamp = local[j].time; e.x = (int16_t)(i % s.width);
      e.y = (int16_t)(i / s.width); e.polarity = (int8_t)local[j].polarity; e.depth = o.depth;
      per_pixel[i].push_back(e);
    }
  }
  s.shadow_rays += shadow;
  if (err_capacity) return 1;
  if (err_invalid) return 2;
  std::vector<orc_event> generated;
  for (auto &v : per_pixel) generated.insert(generated.end(), v.begin(), v.end());
  if (raw_count) *raw_count = (int64_t)generated.size();
  if (raw) std::memcpy(raw, generated.data(), sizeof(orc_event) * std::min<int64_t>(raw_cap, generated.size()));
  std::vector<orc_event> out;
  s.readout.submit(std::move(generated), ivs[n - 1].t1, out);
  s.last_time = ivs[n - 1].t1;
  if (ready_count) *ready_count = (int64_t)out.size();
  if (ready) std::memcpy(ready, out.data(), sizeof(orc_event) * std::min<int64_t>(ready_cap, out.size()));
  return 0;
}
}  // extern "C"

// ---------------------------------------------------------------------------
// Adaptive optics: OpticalSampler::refine / prepare_exact_cells /
// sample_cells (src/cuda/optical_sampler.cu:25-84,114-138,176-264) and
// EventSensor::process_adaptive / verify_receptor
// (src/cuda/event_sensor.cu:88-150,306-374), one pixel at a time.
// ---------------------------------------------------------------------------
namespace orc {
struct Knot { double fraction, radiance; };
struct ExactCell { double start, end; int interval, index; };

// native_pose (optical_sampler.cu:26-30).
static void native_pose(const orc_interval *ivs, int count, double t, orc_state &out) {
  int lo = 0, hi = count - 1;
  while (lo < hi) { int m = (lo + hi) / 2; if (t > ivs[m].t1) lo = m + 1; else hi = m; }
  optical_state_at(ivs[lo], t, out);
}

// refine_pixels (optical_sampler.cu:39-84) for one pixel: returns the knot
// count or -1 (fallback), adding to the statistics.
static int refine_pixel(const Scene &sc, int pixel, const orc_interval *ivs, int n, double start, double end,
                        const orc_photo_params &photo, double tolerance, const orc_pixel_state &state,
                        double charge, int capacity, Knot *knots, uint64_t &queries, int &exact_segments,
                        bool &invalid) {
  auto radiance = [&](double t) {
    orc_state pose; native_pose(ivs, n, t, pose);
    PixelSample o = sample_pixel(sc, pixel, pose);
    if (!o.valid || !std::isfinite(o.radiance)) invalid = true;
    ++queries;
    return o.radiance;
  };
  double initial[5];
  for (int j = 0; j < 5; ++j) initial[j] = radiance(start + (end - start) * j * .25);
  struct Entry { double a, b, left, middle, right; };
  Entry stack[16];
  stack[0] = {.5, 1, initial[2], initial[3], initial[4]};
  stack[1] = {0, .5, initial[0], initial[1], initial[2]};
  knots[0] = {0, initial[0]};
  int top = 2, written = 1;
  while (top) {
    const Entry e = stack[--top];
    const double actual = e.middle * photo.electrons_per_radiance + photo.dark_electrons;
    const double estimate = (e.left + e.right) * .5 * photo.electrons_per_radiance + photo.dark_electrons;
    const double residual = fabs(log(fmax(1e-30, actual) / fmax(1e-30, estimate)));
    const double leaf = (e.b - e.a) * (end - start);
    const double verr = fabs(receptor_log_at(state.receptor_log, actual, leaf, charge) -
                             receptor_log_at(state.receptor_log, estimate, leaf, charge));
    const double budget = 1e-3 * fmin(state.threshold_pos, state.threshold_neg);
    if (verr <= budget && (residual <= tolerance || fmax(actual, estimate) * leaf < 1.0)) {
      if (written + 2 > capacity) return -1;
      knots[written] = {(e.a + e.b) * .5, e.middle};
      knots[written + 1] = {e.b, e.right};
      written += 2;
    } else if (leaf <= photo.integration_step || top + 2 > 16) {
      if (written >= capacity) return -1;
      knots[written++] = {-e.b, e.right};
      ++exact_segments;
    } else {
      const double mid = (e.a + e.b) * .5;
      const double left_mid = radiance(start + (end - start) * (e.a + mid) * .5);
      const double right_mid = radiance(start + (end - start) * (mid + e.b) * .5);
      stack[top] = {mid, e.b, e.middle, right_mid, e.right};
      stack[top + 1] = {e.a, mid, e.left, left_mid, e.middle};
      top += 2;
    }
  }
  return written;
}

// Knot bracket of a cell midpoint fraction (the shared walk of
// integrate_optical_knots / identify_exact_cells / advance_mean_probes).
static inline int knot_bracket(const Knot *v, int size, double fraction, int a) {
  while (a + 2 < size && fraction > fabs(v[a + 1].fraction)) ++a;
  return a;
}

template <class F>
static void for_each_cell(const orc_interval *ivs, int n, double step_len, F &&f) {
  int cell = 0;
  for (int k = 0; k < n; ++k) {
    const orc_interval &iv = ivs[k];
    const int steps = (int)ceil((iv.t1 - iv.t0) / step_len);
    const double dt = (iv.t1 - iv.t0) / steps;
    for (int step = 0; step < steps; ++step, ++cell)
      f(k, cell, iv.t0 + step * dt, step + 1 == steps ? iv.t1 : iv.t0 + (step + 1) * dt);
  }
}
}  // namespace orc

extern "C" {

int orc_sensor_process_adaptive(void *sp, const void *scp, const orc_interval *ivs, int n, double tolerance,
                                int check, int capacity, int *accepted, orc_refinement_stats *stats_out,
                                orc_receptor_probe *probe_out, orc_event *raw, int64_t raw_cap,
                                int64_t *raw_count, orc_event *ready, int64_t ready_cap, int64_t *ready_count,
                                int threads) {
  Sensor &s = *static_cast<Sensor *>(sp);
  const Scene &sc = *static_cast<const Scene *>(scp);
  if (accepted) *accepted = 0;
  if (n <= 0 || capacity < 3) return 3;
  const double start = ivs[0].t0, end = ivs[n - 1].t1;
  if (!std::isfinite(tolerance) || tolerance <= 0 || s.p.refractory_time <= 0 ||
      (end - start) / s.p.refractory_time + 2 > s.max_events)
    return 3;
  if (!s.initialized) return 3;
  double previous = s.last_time, cells_total = 0;
  for (int k = 0; k < n; ++k) {
    const orc_interval &iv = ivs[k];
    if (!std::isfinite(iv.t0) || !std::isfinite(iv.t1) || iv.t0 != previous || iv.t1 <= iv.t0) return 3;
    cells_total += (iv.t1 - iv.t0) / s.photo.integration_step; previous = iv.t1;
  }
  if (cells_total > 1000000) return 3;
  set_threads(threads);
  const double step_len = s.photo.integration_step, charge = s.p.charge_scale;
  std::vector<Knot> knots((size_t)s.count * capacity);
  std::vector<int> kcount(s.count);
  std::vector<std::vector<ExactCell>> exact(s.count);
  std::vector<std::vector<double>> exact_rad(s.count);
  uint64_t queries = 0; int exact_segments = 0, fallback = 0, max_knots = 0, native_queries = 0, invalid = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : queries, exact_segments, fallback, native_queries) \
    reduction(max : max_knots) reduction(| : invalid)
  for (int i = 0; i < s.count; ++i) {
    bool bad = false;
    Knot *v = knots.data() + (size_t)i * capacity;
    int size = refine_pixel(sc, i, ivs, n, start, end, s.photo, tolerance, s.states[i], charge, capacity, v,
                            queries, exact_segments, bad);
    if (bad) invalid = 1;
    if (size < 0) { kcount[i] = -1; ++fallback; continue; }
    max_knots = std::max(max_knots, size);
    // identify_exact_cells (optical_sampler.cu:114-131).
    int a = 0; bool overflow = false;
    for_each_cell(ivs, n, step_len, [&](int k, int cell, double left, double right) {
      if (overflow) return;
      const double fraction = ((left + right) * .5 - start) / (end - start);
      a = knot_bracket(v, size, fraction, a);
      if (v[a + 1].fraction < 0) {
        if ((int)exact[i].size() == capacity) { overflow = true; return; }
        exact[i].push_back({left, right, k, cell});
      }
    });
    if (overflow) { kcount[i] = -1; ++fallback; exact[i].clear(); continue; }
    kcount[i] = size;
    native_queries += (int)exact[i].size();
    // sample_exact_queries (optical_sampler.cu:132-138).
    for (const ExactCell &c : exact[i]) {
      orc_state pose; optical_state_at(ivs[c.interval], (c.start + c.end) * .5, pose);
      PixelSample o = sample_pixel(sc, i, pose);
      if (!o.valid || !std::isfinite(o.radiance)) invalid = 1;
      exact_rad[i].push_back(o.radiance);
    }
  }
  if (invalid) return 2;
  if (stats_out) {
    stats_out->traced_pixel_samples = queries; stats_out->fallback_pixels = fallback;
    stats_out->maximum_knots = max_knots; stats_out->exact_segments = exact_segments;
    stats_out->native_cell_queries = native_queries;
  }
  s.win_knots.resize(knots.size() * 2);
  std::memcpy(s.win_knots.data(), knots.data(), sizeof(Knot) * knots.size());
  s.win_kcount = kcount;
  s.win_ivs.assign(ivs, ivs + n);
  s.win_capacity = capacity;
  // Convergence gate: capture_mean_probes / advance_mean_probes /
  // verify_receptor (event_sensor.cu:134-150,318-326,357-374).
  if (check) {
    const int np = std::min(1024, s.count);
    std::vector<double> errors(np);
    int cells_seen = 0, bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int i = 0; i < np; ++i) {
      const int pixel = (int)(((uint64_t)i * 2654435761ULL) % (uint64_t)s.count);
      double ex = s.states[pixel].receptor_log, ca = ex;
      const Knot *v = knots.data() + (size_t)pixel * capacity;
      const int size = kcount[pixel];
      for_each_cell(ivs, n, step_len, [&](int k, int, double left, double right) {
        orc_state pose; optical_state_at(ivs[k], (left + right) * .5, pose);
        PixelSample o = sample_pixel(sc, pixel, pose);
        if (!o.valid || !std::isfinite(o.radiance)) bad = 1;
        double light = o.radiance, cached = light;
        if (size >= 0) {
          const double fraction = ((left + right) * .5 - start) / (end - start);
          const int a = knot_bracket(v, size, fraction, 0);
          const double w = (fraction - fabs(v[a].fraction)) / (fabs(v[a + 1].fraction) - fabs(v[a].fraction));
          if (v[a + 1].fraction >= 0) cached = v[a].radiance + (v[a + 1].radiance - v[a].radiance) * w;
        }
        ex = receptor_log_at(ex, light * s.photo.electrons_per_radiance + s.photo.dark_electrons, right - left, charge);
        ca = receptor_log_at(ca, fmax(0.0, cached) * s.photo.electrons_per_radiance + s.photo.dark_electrons,
                             right - left, charge);
      });
      errors[i] = fabs(ex - ca);
    }
    if (bad) return 2;
    for_each_cell(ivs, n, step_len, [&](int, int, double, double) { ++cells_seen; });
    orc_receptor_probe r{};
    r.pixels = np; r.circuit_cells = cells_seen;
    for (double e : errors) r.mean_log_voltage_error += e;
    std::sort(errors.begin(), errors.end());
    r.maximum_log_voltage_error = errors.back();
    r.p99_log_voltage_error = errors[(size_t)(np * .99)];
    r.mean_log_voltage_error /= np;
    if (probe_out) *probe_out = r;
    const double threshold = std::min(s.p.threshold_pos, s.p.threshold_neg);
    if (r.maximum_log_voltage_error > .02 * threshold || r.p99_log_voltage_error > .005 * threshold) {
      if (raw_count) *raw_count = 0;
      if (ready_count) *ready_count = 0;
      return 0;
    }
  }
  // integrate_optical_knots and the exact fallback tiles
  // (event_sensor.cu:88-132,331-353).
  std::vector<std::vector<orc_event>> per_pixel(s.count);
  int err_capacity = 0, err_invalid = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err_capacity, err_invalid)
  for (int i = 0; i < s.count; ++i) {
    std::vector<orc_pixel_event> local(s.max_events);
    int count = 0, tile_start = 0; bool failed = false;
    Stream reset{s.seed, (uint64_t)i, kStreamReset, &s.reset_block[i]};
    auto normal = [&] { return normal_draw(reset); };
    const Knot *v = knots.data() + (size_t)i * capacity;
    const int size = kcount[i];
    int a = 0; size_t e = 0;
    for_each_cell(ivs, n, step_len, [&](int k, int cell, double t0, double t1) {
      if (failed) return;
      double light;
      if (size < 0) {
        // fallback tiles of 64 cells, each with its own event capacity
        if (cell % 64 == 0) tile_start = count;
        orc_state pose; optical_state_at(ivs[k], (t0 + t1) * .5, pose);
        PixelSample o = sample_pixel(sc, i, pose);
        if (!o.valid || !std::isfinite(o.radiance)) { err_invalid = 1; failed = true; return; }
        light = o.radiance;
      } else {
        const double fraction = ((t0 + t1) * .5 - start) / (end - start);
        a = knot_bracket(v, size, fraction, a);
        const double w = (fraction - fabs(v[a].fraction)) / (fabs(v[a + 1].fraction) - fabs(v[a].fraction));
        light = v[a].radiance + (v[a + 1].radiance - v[a].radiance) * w;
        if (e < exact[i].size() && exact[i][e].index == cell) light = exact_rad[i][e++];
        light = fmax(0.0, light);
      }
      const double rate = cell_rate(light, s.photo, t1 - t0, s.seed, (uint64_t)i, s.shot_block[i]++);
      const int cap = size < 0 ? s.max_events - (count - tile_start) : s.max_events - count;
      if ((int)local.size() < count + cap) local.resize(count + cap);
      const int emitted = advance_pixel(s.states[i], s.p, rate, t0, t1, local.data() + count, cap, normal);
      if (emitted < 0) { err_capacity = 1; failed = true; return; }
      count += emitted;
    });
    if (failed) continue;
    for (int j = 0; j < count; ++j) {
      orc_event ev; ev.timestamp = local[j].time; ev.x = (int16_t)(i % s.width); ev.y = (int16_t)(i / s.width);
      ev.polarity = (int8_t)local[j].polarity; ev.depth = -1;
      per_pixel[i].push_back(ev);
    }
  }
  if (err_capacity) return 1;
  if (err_invalid) return 2;
  std::vector<orc_event> generated;
  for (auto &pv : per_pixel) generated.insert(generated.end(), pv.begin(), pv.end());
  if (raw_count) *raw_count = (int64_t)generated.size();
  if (raw) std::memcpy(raw, generated.data(), sizeof(orc_event) * std::min<int64_t>(raw_cap, generated.size()));
  std::vector<orc_event> out;
  s.readout.submit(std::move(generated), end, out);
  s.last_time = end;
  if (accepted) *accepted = 1;
  if (ready_count) *ready_count = (int64_t)out.size();
  if (ready) std::memcpy(ready, out.data(), sizeof(orc_event) * std::min<int64_t>(ready_cap, out.size()));
  return 0;
}
// OpticalSampler::verify / verify_knots (optical_sampler.cu:86-103,195-215) on
// the last adaptive window: 1024 hashed pixels x 17 times, the knot
// interpolation of the photocurrent against the native trace.  out: max,
// p99, mean log error; iout: probes, over tolerance, fallback probes.
int orc_sensor_verify_optics(void *sp, const void *scp, double tolerance, double out[3], int iout[3], int threads) {
  Sensor &s = *static_cast<Sensor *>(sp);
  const Scene &sc = *static_cast<const Scene *>(scp);
  if (s.win_ivs.empty()) return 5;
  set_threads(threads);
  constexpr int positions = 1024, times = 17, total = positions * times;
  const int n = (int)s.win_ivs.size();
  const double start = s.win_ivs[0].t0, end = s.win_ivs[n - 1].t1;
  std::vector<double> values(total);
  int invalid = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : invalid)
  for (int work = 0; work < total; ++work) {
    const int pixel = (int)((uint64_t)(work % positions) * 2654435761ULL % (uint64_t)s.count), probe = work / positions;
    if (s.win_kcount[pixel] < 0) { values[work] = -1; continue; }
    const double fraction = (probe + .5) / times, t = start + (end - start) * fraction;
    orc_state pose; native_pose(s.win_ivs.data(), n, t, pose);
    const PixelSample o = sample_pixel(sc, pixel, pose);
    if (!o.valid) invalid = 1;
    const Knot *v = reinterpret_cast<const Knot *>(s.win_knots.data()) + (size_t)pixel * s.win_capacity;
    int a = 0;
    while (a + 2 < s.win_kcount[pixel] && fraction > fabs(v[a + 1].fraction)) ++a;
    if (v[a + 1].fraction < 0) { values[work] = -1; continue; }
    const double weight = (fraction - fabs(v[a].fraction)) / (fabs(v[a + 1].fraction) - fabs(v[a].fraction));
    const double estimate = v[a].radiance + (v[a + 1].radiance - v[a].radiance) * weight;
    const double actual = o.radiance * s.photo.electrons_per_radiance + s.photo.dark_electrons;
    const double predicted = estimate * s.photo.electrons_per_radiance + s.photo.dark_electrons;
    values[work] = fabs(log(fmax(1e-30, actual) / fmax(1e-30, predicted)));
  }
  if (invalid) return 2;
  double mx = 0, mean = 0; int over = 0, fb = 0;
  for (double value : values) {
    if (value < 0) { ++fb; continue; }
    mx = std::max(mx, value); mean += value;
    if (value > tolerance) ++over;
  }
  mean /= total - fb > 0 ? total - fb : 1;
  std::sort(values.begin(), values.end());
  out[0] = mx; out[1] = values[(size_t)(total * .99)]; out[2] = mean;
  iout[0] = total; iout[1] = over; iout[2] = fb;
  return 0;
}
int orc_sensor_drain(void *sp, orc_event *ready, int64_t cap, int64_t *count) {
  Sensor &s = *static_cast<Sensor *>(sp);
  std::vector<orc_event> out; s.readout.drain(out);
  *count = (int64_t)out.size();
  if (ready) std::memcpy(ready, out.data(), sizeof(orc_event) * std::min<int64_t>(cap, out.size()));
  return 0;
}
int64_t orc_sensor_dropped(const void *sp) { return (int64_t)static_cast<const Sensor *>(sp)->readout.dropped; }
int64_t orc_sensor_peak_queue(const void *sp) { return (int64_t)static_cast<const Sensor *>(sp)->readout.peak; }
int orc_sensor_states(const void *sp, orc_pixel_state *out) {
  const Sensor &s = *static_cast<const Sensor *>(sp);
  std::memcpy(out, s.states.data(), sizeof(orc_pixel_state) * s.states.size());
  return 0;
}
int64_t orc_sensor_shadow_rays(const void *sp) { return static_cast<const Sensor *>(sp)->shadow_rays; }

void *orc_readout_create(double b, double l, int64_t c) { return new Readout(b, l, (size_t)c); }
void orc_readout_destroy(void *r) { delete static_cast<Readout *>(r); }
int64_t orc_readout_submit(void *rp, const orc_event *ev, int64_t n, double now, orc_event *ready, int64_t cap) {
  std::vector<orc_event> in(ev, ev + n), out;
  static_cast<Readout *>(rp)->submit(std::move(in), now, out);
  if (ready) std::memcpy(ready, out.data(), sizeof(orc_event) * std::min<int64_t>(cap, out.size()));
  return (int64_t)out.size();
}
int64_t orc_readout_drain(void *rp, orc_event *ready, int64_t cap) {
  std::vector<orc_event> out; static_cast<Readout *>(rp)->drain(out);
  if (ready) std::memcpy(ready, out.data(), sizeof(orc_event) * std::min<int64_t>(cap, out.size()));
  return (int64_t)out.size();
}
int64_t orc_readout_dropped(const void *r) { return (int64_t)static_cast<const Readout *>(r)->dropped; }
int64_t orc_readout_peak(const void *r) { return (int64_t)static_cast<const Readout *>(r)->peak; }

// render_float_to_png_bytes (blenderize.hpp:12-69).
void orc_render_to_bytes(const float *linear, int width, int height, float dither, uint8_t *out) {
  auto to_byte = [](float v) -> uint8_t {
    if (!std::isfinite(v) || v <= 0.0f) return 0;
    if (v >= 1.0f) return 255;
    return (uint8_t)(v * 255.0f + 0.5f);
  };
  auto srgb = [](float c) { return c < 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f; };
  auto hash = [](uint32_t x, uint32_t y) {
    const uint32_t qx = 1103515245u * ((x >> 1u) ^ (y));
    const uint32_t qy = 1103515245u * ((y >> 1u) ^ (x));
    const uint32_t n = 1103515245u * ((qx) ^ (qy >> 3u));
    return float(n) * (1.0f / float(0xffffffffu));
  };
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      float r = linear[y * width + x];
      r = std::clamp(r / 100, 0.0f, 1.0f);
      r = srgb(r);
      if (dither != 0) {
        float v = hash(x, y) * 2.0f - 1.0f;
        v = std::copysign(1.0f, v) * (1.0f - sqrtf(1.0f - fabsf(v)));
        r += dither * v;
      }
      out[y * width + x] = to_byte(r);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Camera trajectory: src/trajectory.cpp (CameraTrajectory), float arithmetic.
// Eigen 3.4 (quaternion <-> matrix, AngleAxis, norms) is not in the
// reference tree; its published formulas are restated: Shepperd's method for
// Quaternion(Matrix3), the Rodrigues matrix for AngleAxis, the Hamilton
// product and the quaternion's rotation matrix.
// ---------------------------------------------------------------------------
namespace orc {
namespace traj {

struct Key { float time, value, lt, lv, rt, rv; };

struct Scalar {
  std::vector<Key> k;
  // bezier / bezier_derivative (trajectory.cpp:114-128)
  static float bz(float p0, float p1, float p2, float p3, float u) {
    const float a = p0 + u * (p1 - p0), b = p1 + u * (p2 - p1), c = p2 + u * (p3 - p2);
    const float d = a + u * (b - a), e = b + u * (c - b);
    return d + u * (e - d);
  }
  static float dbz(float p0, float p1, float p2, float p3, float u) {
    const float v = 1.0f - u;
    return 3.0f * v * v * (p1 - p0) + 6.0f * v * u * (p2 - p1) + 3.0f * u * u * (p3 - p2);
  }
  // compute_auto_clamped_tangents (trajectory.cpp:26-82) and
  // build_auto_clamped (:130-193)
  void build(const std::vector<float> &t, const std::vector<float> &y) {
    const size_t n = t.size();
    k.assign(n, Key{});
    for (size_t i = 0; i < n; ++i) k[i] = Key{t[i], y[i], t[i], y[i], t[i], y[i]};
    if (n < 2) return;
    std::vector<float> m(n, 0.0f), h(n - 1), d(n - 1);
    for (size_t i = 0; i + 1 < n; ++i) {
      h[i] = t[i + 1] - t[i];
      d[i] = std::fabs(h[i]) > 1e-12f ? (y[i + 1] - y[i]) / h[i] : 0.0f;
    }
    if (n == 2) {
      m[0] = m[1] = d[0];
    } else {
      m[0] = d[0];
      m[n - 1] = d[n - 2];
      for (size_t i = 1; i + 1 < n; ++i) {
        if (d[i - 1] == 0.0f || d[i] == 0.0f || d[i - 1] * d[i] <= 0.0f) continue;
        const float w1 = 2.0f * h[i] + h[i - 1], w2 = h[i] + 2.0f * h[i - 1];
        const float den = w1 / d[i - 1] + w2 / d[i];
        m[i] = std::fabs(den) > 1e-12f ? (w1 + w2) / den : 0.0f;
      }
    }
    for (size_t i = 0; i + 1 < n; ++i) {
      const float span = t[i + 1] - t[i];
      k[i].rt = t[i] + span / 3.0f;
      k[i].rv = y[i] + m[i] * span / 3.0f;
      k[i + 1].lt = t[i + 1] - span / 3.0f;
      k[i + 1].lv = y[i + 1] - m[i + 1] * span / 3.0f;
    }
  }
  // find_segment (trajectory.cpp:195-215): upper_bound of t, minus one
  int segment(float t) const {
    const int n = (int)k.size();
    if (t <= k.front().time) return 0;
    if (t >= k.back().time) return n - 2;
    int idx = (int)(std::upper_bound(k.begin(), k.end(), t, [](float x, const Key &q) { return x < q.time; }) -
                    k.begin()) - 1;
    return std::max(0, std::min(idx, n - 2));
  }
  // solve_bezier_time (trajectory.cpp:217-267)
  static float solve(const Key &a, const Key &b, float t) {
    const float x0 = a.time, x1 = a.rt, x2 = b.lt, x3 = b.time;
    if (x3 - x0 <= 1e-12f) return 0.0f;
    float u = std::min(std::max((t - x0) / (x3 - x0), 0.0f), 1.0f);
    for (int it = 0; it < 8; ++it) {
      const float x = bz(x0, x1, x2, x3, u), dx = dbz(x0, x1, x2, x3, u);
      if (std::fabs(dx) < 1e-10f) break;
      const float nu = std::min(std::max(u - (x - t) / dx, 0.0f), 1.0f);
      if (std::fabs(nu - u) < 1e-7f) return nu;
      u = nu;
    }
    float lo = 0.0f, hi = 1.0f;
    for (int it = 0; it < 24; ++it) {
      const float mid = 0.5f * (lo + hi);
      (bz(x0, x1, x2, x3, mid) < t ? lo : hi) = mid;
    }
    return 0.5f * (lo + hi);
  }
  // evaluate (trajectory.cpp:269-295)
  float value(float t) const {
    if (k.empty()) return 0.0f;
    if (k.size() == 1 || t <= k.front().time) return k.front().value;
    if (t >= k.back().time) return k.back().value;
    const int s = segment(t);
    const float u = solve(k[s], k[s + 1], t);
    return bz(k[s].value, k[s].rv, k[s + 1].lv, k[s + 1].value, u);
  }
  // derivative (trajectory.cpp:297-328): segment of the clamped time, curve
  // parameter of the unclamped one
  float rate(float t) const {
    if (k.size() < 2) return 0.0f;
    float tc = t;
    if (tc < k.front().time + 1e-7f) tc = k.front().time + 1e-7f;
    if (tc > k.back().time - 1e-7f) tc = k.back().time - 1e-7f;
    const int s = segment(tc);
    const float u = solve(k[s], k[s + 1], t);
    const float dx = dbz(k[s].time, k[s].rt, k[s + 1].lt, k[s + 1].time, u);
    const float dy = dbz(k[s].value, k[s].rv, k[s + 1].lv, k[s + 1].value, u);
    return std::fabs(dx) < 1e-12f ? 0.0f : dy / dx;
  }
};

// Quaternion stored (w, x, y, z) here.
struct Q { float w, x, y, z; };
static Q qmul(const Q &a, const Q &b) {
  return Q{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
           a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
static float qnorm(const Q &q) { return std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z); }
static Q qunit(Q q) {
  const float n = qnorm(q);
  if (n > 0) { q.w /= n; q.x /= n; q.y /= n; q.z /= n; }
  return q;
}
static float qdot(const Q &a, const Q &b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
static Q qneg(const Q &q) { return Q{-q.w, -q.x, -q.y, -q.z}; }

// Row-major 3x3 here: R[r][c].
struct M { float a[3][3]; };
static M mmul(const M &p, const M &q) {
  M r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = p.a[i][0] * q.a[0][j] + p.a[i][1] * q.a[1][j] + p.a[i][2] * q.a[2][j];
  return r;
}
// AngleAxis(angle, unit e).toRotationMatrix(): (1-c) e e^T + c I + s [e]x
static M angle_axis(float angle, int axis) {
  const float e[3] = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
  const float s = std::sin(angle), c = std::cos(angle);
  const float cross[3][3] = {{0, -e[2], e[1]}, {e[2], 0, -e[0]}, {-e[1], e[0], 0}};
  M r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = (1.0f - c) * e[i] * e[j] + (i == j ? c : 0.0f) + s * cross[i][j];
  return r;
}
// Quaternion(R) by Shepperd's method (Eigen quaternion_assign_impl<3x3>).
static Q from_matrix(const M &R) {
  const float tr = R.a[0][0] + R.a[1][1] + R.a[2][2];
  Q q;
  if (tr > 0) {
    const float s = std::sqrt(tr + 1.0f);
    const float f = 0.5f / s;
    q = Q{0.5f * s, (R.a[2][1] - R.a[1][2]) * f, (R.a[0][2] - R.a[2][0]) * f, (R.a[1][0] - R.a[0][1]) * f};
  } else {
    int i = R.a[1][1] > R.a[0][0] ? 1 : 0;
    if (R.a[2][2] > R.a[i][i]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    const float s = std::sqrt(R.a[i][i] - R.a[j][j] - R.a[k][k] + 1.0f);
    const float f = 0.5f / s;
    float v[3];
    v[i] = 0.5f * s;
    v[j] = (R.a[j][i] + R.a[i][j]) * f;
    v[k] = (R.a[k][i] + R.a[i][k]) * f;
    q = Q{(R.a[k][j] - R.a[j][k]) * f, v[0], v[1], v[2]};
  }
  return q;
}
// Quaternion::toRotationMatrix
static M to_matrix(const Q &q) {
  M r{};
  const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z, wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  r.a[0][0] = 1 - (2 * yy + 2 * zz); r.a[0][1] = 2 * xy - 2 * wz;       r.a[0][2] = 2 * xz + 2 * wy;
  r.a[1][0] = 2 * xy + 2 * wz;       r.a[1][1] = 1 - (2 * xx + 2 * zz); r.a[1][2] = 2 * yz - 2 * wx;
  r.a[2][0] = 2 * xz - 2 * wy;       r.a[2][1] = 2 * yz + 2 * wx;       r.a[2][2] = 1 - (2 * xx + 2 * yy);
  return r;
}
// quat_exp / quat_log (trajectory.cpp:346-387)
static Q qexp(const float v[3]) {
  const float th = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (th < 1e-8f) return qunit(Q{1.0f, 0.5f * v[0], 0.5f * v[1], 0.5f * v[2]});
  const float s = std::sin(0.5f * th);
  return qunit(Q{std::cos(0.5f * th), s * (v[0] / th), s * (v[1] / th), s * (v[2] / th)});
}
static void qlog(Q q, float out[3]) {
  q = qunit(q);
  if (q.w < 0.0f) q = qneg(q);
  const float sh = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (sh < 1e-8f) { out[0] = 2.0f * q.x; out[1] = 2.0f * q.y; out[2] = 2.0f * q.z; return; }
  const float f = 2.0f * std::atan2(sh, q.w) / sh;
  out[0] = q.x * f; out[1] = q.y * f; out[2] = q.z * f;
}

struct Trajectory {
  Scalar px, py, pz, lx, ly, lz, intensity, phase;
  std::vector<float> times;
  std::vector<Q> rot;
  size_t size() const { return px.k.size(); }
};

// CameraTrajectory::build (trajectory.cpp:609-731), messages included.
static void build(Trajectory &T, const orc_frame *f, int n, float fps, const float tr[3], const float ro[3]) {
  if (!std::isfinite(fps) || fps <= 0)
    throw std::invalid_argument("trajectory requires positive fps and three-component mount transforms");
  const M mount = mmul(mmul(angle_axis(ro[2], 2), angle_axis(ro[1], 1)), angle_axis(ro[0], 0));
  if (n == 0) return;
  std::vector<float> t(n), x(n), y(n), z(n), lx(n), ly(n), lz(n), li(n), ph(n);
  std::vector<Q> q(n);
  for (int i = 0; i < n; ++i) {
    const double secs = std::isnan(f[i].timestamp_seconds) ? f[i].frame_index / (double)fps : f[i].timestamp_seconds;
    t[i] = (float)secs;
    if (!std::isfinite(t[i])) throw std::invalid_argument("non-finite trajectory timestamp");
  }
  for (int i = 1; i < n; ++i)
    if (!(t[i] > t[i - 1])) throw std::invalid_argument("trajectory timestamps must be strictly increasing");
  for (int i = 0; i < n; ++i) {
    M W;  // camera_world is column major
    bool finite = true;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        W.a[r][c] = f[i].camera_world[3 * c + r];
        finite = finite && std::isfinite(W.a[r][c]);
      }
    for (int c = 0; c < 3; ++c)
      finite = finite && std::isfinite(f[i].camera_origin[c]) && std::isfinite(f[i].light_dir[c]);
    const float *L = f[i].light_dir;
    const float ln = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    float frob = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const float g = W.a[0][r] * W.a[0][c] + W.a[1][r] * W.a[1][c] + W.a[2][r] * W.a[2][c] - (r == c);
        frob += g * g;
      }
    const float det = W.a[0][0] * (W.a[1][1] * W.a[2][2] - W.a[2][1] * W.a[1][2]) -
                      W.a[1][0] * (W.a[0][1] * W.a[2][2] - W.a[2][1] * W.a[0][2]) +
                      W.a[2][0] * (W.a[0][1] * W.a[1][2] - W.a[1][1] * W.a[0][2]);
    if (!finite || ln < 1e-12f || !std::isfinite(f[i].light_intensity) || f[i].light_intensity < 0 ||
        std::sqrt(frob) > 2e-3f || det <= 0)
      throw std::invalid_argument("trajectory requires finite rigid poses and valid world lighting");
    const float *o = f[i].camera_origin;
    x[i] = o[0] + (W.a[0][0] * tr[0] + W.a[0][1] * tr[1] + W.a[0][2] * tr[2]);
    y[i] = o[1] + (W.a[1][0] * tr[0] + W.a[1][1] * tr[1] + W.a[1][2] * tr[2]);
    z[i] = o[2] + (W.a[2][0] * tr[0] + W.a[2][1] * tr[1] + W.a[2][2] * tr[2]);
    Q qi = qunit(from_matrix(mmul(W, mount)));  // matrix_to_quaternion_continuous (:580-592)
    if (i > 0 && qdot(qi, q[i - 1]) < 0.0f) qi = qneg(qi);
    q[i] = qunit(qi);
    lx[i] = L[0] / ln; ly[i] = L[1] / ln; lz[i] = L[2] / ln;
    li[i] = f[i].light_intensity;
    ph[i] = (float)i;
  }
  T.px.build(t, x); T.py.build(t, y); T.pz.build(t, z);
  // RotationBezierCurve::build (:389-436)
  T.times = t;
  T.rot.resize(n);
  for (int i = 0; i < n; ++i) {
    T.rot[i] = qunit(q[i]);
    if (i > 0 && qdot(T.rot[i], T.rot[i - 1]) < 0.0f) T.rot[i] = qneg(T.rot[i]);
  }
  T.phase.build(t, ph);
  T.lx.build(t, lx); T.ly.build(t, ly); T.lz.build(t, lz);
  T.intensity.build(t, li);
}

// CameraTrajectory::evaluate (trajectory.cpp:733-790) with
// RotationBezierCurve::evaluate (:477-564).
static void evaluate(const Trajectory &T, float t, orc_trajectory_state *s) {
  std::memset(s, 0, sizeof(*s));
  s->timestamp = t;
  s->orientation[0] = s->orientation[4] = s->orientation[8] = 1;
  s->quaternion[3] = 1;
  s->light_direction[2] = 1;
  if (T.size() == 0) return;
  const float tc = std::min(std::max(t, T.px.k.front().time), T.px.k.back().time);
  s->timestamp = tc;
  const Scalar *P[3] = {&T.px, &T.py, &T.pz};
  for (int c = 0; c < 3; ++c) {
    s->position[c] = P[c]->value(tc);
    s->linear_velocity[c] = P[c]->rate(tc);
  }
  const int n = (int)T.rot.size();
  if (n >= 2) {
    const float phase = T.phase.value(tc), pdot = T.phase.rate(tc);
    const int seg = std::max(0, std::min((int)std::floor(phase), n - 2));
    const float alpha = std::min(std::max(phase - (float)seg, 0.0f), 1.0f);
    const Q &q0 = T.rot[seg], &q1 = T.rot[seg + 1];
    Q rel = qunit(qmul(Q{q0.w, -q0.x, -q0.y, -q0.z}, q1));
    if (rel.w < 0.0f) rel = qneg(rel);
    float rho[3];
    qlog(rel, rho);
    const float arho[3] = {alpha * rho[0], alpha * rho[1], alpha * rho[2]};
    const Q q = qunit(qmul(q0, qexp(arho)));
    const M R = to_matrix(q);
    const float wb[3] = {pdot * rho[0], pdot * rho[1], pdot * rho[2]};
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) s->orientation[3 * c + r] = R.a[r][c];
      s->angular_velocity[r] = R.a[r][0] * wb[0] + R.a[r][1] * wb[1] + R.a[r][2] * wb[2];
    }
    s->quaternion[0] = q.x; s->quaternion[1] = q.y; s->quaternion[2] = q.z; s->quaternion[3] = q.w;
  }
  const float L[3] = {T.lx.value(tc), T.ly.value(tc), T.lz.value(tc)};
  const float Ld[3] = {T.lx.rate(tc), T.ly.rate(tc), T.lz.rate(tc)};
  const float ln = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  if (ln > 1e-12f) {
    // normalized_vector_derivative (:594-607)
    const float vd = L[0] * Ld[0] + L[1] * Ld[1] + L[2] * Ld[2];
    for (int c = 0; c < 3; ++c) {
      s->light_direction[c] = L[c] / ln;
      s->light_direction_rate[c] = Ld[c] / ln - L[c] * (vd / (ln * ln * ln));
    }
  }
  s->light_intensity = T.intensity.value(tc);
  s->light_intensity_rate = T.intensity.rate(tc);
}

}  // namespace traj
}  // namespace orc

extern "C" {

void *orc_trajectory_create(const orc_frame *frames, int count, float fps, const float translation[3],
                            const float rotation[3], char *error, int error_len) {
  try {
    auto T = std::make_unique<orc::traj::Trajectory>();
    orc::traj::build(*T, frames, count, fps, translation, rotation);
    if (error && error_len > 0) error[0] = 0;
    return T.release();
  } catch (const std::exception &e) {
    if (error && error_len > 0) {
      std::strncpy(error, e.what(), error_len - 1);
      error[error_len - 1] = 0;
    }
    return nullptr;
  }
}
void orc_trajectory_destroy(void *t) { delete static_cast<orc::traj::Trajectory *>(t); }
void orc_trajectory_range(const void *tp, float *start, float *end, int *size) {
  const auto &T = *static_cast<const orc::traj::Trajectory *>(tp);
  const bool any = T.size() > 0;
  *start = any ? T.px.k.front().time : 0.0f;
  *end = any ? T.px.k.back().time : 0.0f;
  *size = (int)T.size();
}
void orc_trajectory_evaluate(const void *tp, const float *times, int64_t n, orc_trajectory_state *out) {
  const auto &T = *static_cast<const orc::traj::Trajectory *>(tp);
  for (int64_t i = 0; i < n; ++i) orc::traj::evaluate(T, times[i], out + i);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Terrain from rasters (src/tools/build_from_raster.cu), planar path
// ---------------------------------------------------------------------------
namespace orc {
namespace raster {

struct P3 {
  float x, y, z;
};
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(P3 a) { return std::sqrt(dot(a, a)); }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline P3 scaled_div(P3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

// accumulate_triangle_normal_for_vertex (build_from_raster.cu:37-127) with
// NORMAL_ORIENT_PLANAR_Z_UP.
void accumulate(int target, int ta, int tb, int tc, const std::vector<P3> &v, P3 &acc) {
  if (target != ta && target != tb && target != tc) return;
  const P3 pa = v[ta], pb = v[tb], pc = v[tc];
  P3 face = cross(pb - pa, pc - pa);
  const float face_len = length(face);
  if (!std::isfinite(face_len) || face_len < 1e-12f) return;
  face = scaled_div(face, face_len);
  P3 p, o1, o2;
  if (target == ta) {
    p = pa; o1 = pb; o2 = pc;
  } else if (target == tb) {
    p = pb; o1 = pc; o2 = pa;
  } else {
    p = pc; o1 = pa; o2 = pb;
  }
  P3 e1 = o1 - p, e2 = o2 - p;
  const float l1 = length(e1), l2 = length(e2);
  if (!std::isfinite(l1) || !std::isfinite(l2) || l1 < 1e-12f || l2 < 1e-12f) return;
  e1 = scaled_div(e1, l1);
  e2 = scaled_div(e2, l2);
  const float w = std::acos(std::fmin(1.0f, std::fmax(-1.0f, dot(e1, e2))));
  if (!std::isfinite(w)) return;
  if (face.z < 0.0f) face = {face.x * -1.0f, face.y * -1.0f, face.z * -1.0f};
  acc = {acc.x + face.x * w, acc.y + face.y * w, acc.z + face.z * w};
}

}  // namespace raster
}  // namespace orc

extern "C" {

// load_from_raster (build_from_raster.cu:825-920) after loadImageryData:
// build_from_raster's vertices (raster order here), triangles and normals
// (:213-245, 333-472), then the recentring on the DEM's centre pixel
// (:898-914, reference_height, loader.hpp:102-106).  Returns 0, or -1 for a
// NoData centre.
int orc_raster_mesh(const float *dem, const float *albedo, const float *xs, const float *ys, int width, int height,
                    float *vertices, float *normals, float *albedos, int64_t *vertex_count, uint32_t *indices,
                    int64_t *triangle_count) {
  using orc::raster::P3;
  const float centre = dem[(size_t)width * (height / 2) + width / 2];
  if (!std::isfinite(centre)) return -1;
  std::vector<int> hash((size_t)width * height, -1);
  std::vector<P3> v;
  for (int y = 0; y < height; ++y)  // batch_build_vertices_kernel (:213-245)
    for (int x = 0; x < width; ++x) {
      const float z = dem[(size_t)y * width + x];
      if (!std::isfinite(z) || std::fabs(z) > 1e8f) continue;
      hash[(size_t)y * width + x] = (int)v.size();
      if (albedo) albedos[v.size()] = albedo[(size_t)y * width + x];
      v.push_back({xs[x], ys[y], z});
    }
  auto cell = [&](int cx, int cy, int &v1, int &v2, int &v3, int &v4) {  // fetch_cell_vertices (:129-159)
    if (cx < 0 || cy < 0 || cx >= width - 1 || cy >= height - 1) return false;
    v1 = hash[(size_t)cy * width + cx];
    v2 = hash[(size_t)cy * width + cx + 1];
    v3 = hash[(size_t)(cy + 1) * width + cx];
    v4 = hash[(size_t)(cy + 1) * width + cx + 1];
    return v1 >= 0 && v2 >= 0 && v3 >= 0 && v4 >= 0;
  };
  int64_t nt = 0;
  for (int cy = 0; cy + 1 < height; ++cy)  // batch_build_triangles_kernel (:333-386)
    for (int cx = 0; cx + 1 < width; ++cx) {
      int v1, v2, v3, v4;
      if (!cell(cx, cy, v1, v2, v3, v4)) continue;
      const uint32_t t[6] = {(uint32_t)v1, (uint32_t)v3, (uint32_t)v2, (uint32_t)v2, (uint32_t)v3, (uint32_t)v4};
      std::memcpy(indices + 3 * nt, t, sizeof(t));
      nt += 2;
    }
  for (int y = 0; y < height; ++y)  // batch_gather_vertex_normals_kernel (:388-472)
    for (int x = 0; x < width; ++x) {
      const int me = hash[(size_t)y * width + x];
      if (me < 0) continue;
      P3 acc = {0, 0, 0};
      for (int oy = -1; oy <= 0; ++oy)
        for (int ox = -1; ox <= 0; ++ox) {
          int v1, v2, v3, v4;
          if (!cell(x + ox, y + oy, v1, v2, v3, v4)) continue;
          orc::raster::accumulate(me, v1, v3, v2, v, acc);
          orc::raster::accumulate(me, v2, v3, v4, v, acc);
        }
      const float len = orc::raster::length(acc);
      P3 n = {0, 0, 0};
      if (std::isfinite(len) && len >= 1e-12f) n = orc::raster::scaled_div(acc, len);
      // apply_rigid_transform with the identity quaternion (:612-622): the
      // rotation's zero terms and the zero translation leave no -0
      normals[3 * me] = n.x + 0.0f;
      normals[3 * me + 1] = n.y + 0.0f;
      normals[3 * me + 2] = n.z + 0.0f;
    }
  const float tx = -xs[width / 2], ty = -ys[height / 2], tz = -centre;
  for (size_t i = 0; i < v.size(); ++i) {
    vertices[3 * i] = v[i].x + tx;
    vertices[3 * i + 1] = v[i].y + ty;
    vertices[3 * i + 2] = v[i].z + tz;
  }
  *vertex_count = (int64_t)v.size();
  *triangle_count = nt;
  return 0;
}

// batch_align_kernel (build_from_raster.cu:653-682).
void orc_raster_align(const float *ref_x, const float *ref_y, int width, int height, const float *source,
                      const float *source_x, const float *source_y, int source_width, int source_height, float *out) {
  auto bracket = [](const float *axis, int n, float value) {
    if (n < 2 || !std::isfinite(value)) return -1;
    const bool increasing = axis[n - 1] > axis[0];
    if (value < std::fmin(axis[0], axis[n - 1]) || value > std::fmax(axis[0], axis[n - 1])) return -1;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if ((axis[mid] <= value) == increasing) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const int row = bracket(source_y, source_height, ref_y[y]), col = bracket(source_x, source_width, ref_x[x]);
      float value = std::numeric_limits<float>::quiet_NaN();
      if (row >= 0 && col >= 0) {
        const float u = (ref_x[x] - source_x[col]) / (source_x[col + 1] - source_x[col]);
        const float v = (ref_y[y] - source_y[row]) / (source_y[row + 1] - source_y[row]);
        double total = 0, weight = 0;
        bool valid = true;
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            const float w = (dx ? u : 1 - u) * (dy ? v : 1 - v);
            const float s = source[(size_t)(row + dy) * source_width + col + dx];
            if (w > 0) {
              if (!std::isfinite(s)) {
                valid = false;
                continue;
              }
              const float product = w * s;
              total += product;
              weight += w;
            }
          }
        if (valid && weight > 0) value = (float)(total / weight);
      }
      out[(size_t)y * width + x] = value;
    }
}

}  // extern "C"
