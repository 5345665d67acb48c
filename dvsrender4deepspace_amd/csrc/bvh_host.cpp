// Host-side BVH: surface-area-heuristic build, validation and the re-layout
// into the SoA node-pair / leaf / triangle-record arrays the gfx950 kernels
// traverse.
//
// The build reproduces the reference's tree exactly (same topology, node
// numbering and leaf triangle order: BVHBuilder::build / build_node /
// create_leaf_node, src/bvh/build.cpp:27-76,126-187, and the 32-bin centroid
// binning with its in-place partition, src/bvh/bining.cpp:9-182), because
// traversal order decides ties between triangles at equal distance.  It is
// organised differently: triangle bounds and centroids live in SoA arrays and
// the partition permutes a 32-bit index array instead of 40-byte records.
#include "bvh_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

namespace dvsamd {
namespace {

struct Aabb {
  float lo[3], hi[3];
  static Aabb empty() {
    Aabb b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; }
    return b;
  }
  // Surface measure used by the SAH; an empty box evaluates to +inf exactly
  // as the reference's (min = FLT_MAX, max = -FLT_MAX) representation does.
  float half_area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return std::max(dx * dy + dy * dz + dz * dx, 1e-12f);
  }
  void take(const Aabb &o) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], o.lo[k]);
      hi[k] = std::max(hi[k], o.hi[k]);
    }
  }
};

struct Primitives {
  // Triangle bounds and centroids, SoA.
  std::vector<float> lo[3], hi[3], centre[3];
  Aabb bounds(int i) const {
    Aabb b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = lo[k][i]; b.hi[k] = hi[k][i]; }
    return b;
  }
};

class Builder {
 public:
  Builder(const BuildOptions &opt, const Primitives &prims, std::vector<int> &order,
          std::vector<dvs_bvh_node> &nodes)
      : opt_(opt), prims_(prims), order_(order), nodes_(nodes) {}

  int subtree(int first, int last) {
    const int n = last - first;
    if (n <= 0) throw std::runtime_error("BVH build_node empty range");
    Aabb box = Aabb::empty();
    for (int i = first; i < last; ++i) box.take(prims_.bounds(order_[i]));
    if (n <= opt_.max_leaf_size) return emit_leaf(first, last, box);

    Split s = choose_split(box, first, n);
    const float leaf_cost = opt_.primitive_cost * s.leaf_sah;
    const float split_cost = opt_.node_cost * box.half_area() + opt_.primitive_cost * s.split_sah;
    if (leaf_cost <= split_cost) return emit_leaf(first, last, box);

    const int mid = partition(s, first, last);
    if (mid - first <= 0 || last - mid <= 0) return emit_leaf(first, last, box);
    const int left = subtree(first, mid);
    const int right = subtree(mid, last);
    dvs_bvh_node node{};
    store_box(node, box);
    node.left = left; node.right = right; node.start = 0; node.count = 0;
    nodes_.push_back(node);
    return static_cast<int>(nodes_.size()) - 1;
  }

 private:
  static constexpr int kMaxBins = 32;
  struct Split {
    int axis = 0, bin = 0, bins = 0;
    float leaf_sah = 0, split_sah = 0;
    float cmin[3], cmax[3], scale[3];
  };

  static void store_box(dvs_bvh_node &node, const Aabb &b) {
    for (int k = 0; k < 3; ++k) { node.min[k] = b.lo[k]; node.max[k] = b.hi[k]; }
  }
  int emit_leaf(int first, int last, const Aabb &box) {
    dvs_bvh_node node{};
    store_box(node, box);
    node.left = -1; node.right = -1; node.start = first; node.count = last - first;
    nodes_.push_back(node);
    return static_cast<int>(nodes_.size()) - 1;
  }

  Split choose_split(const Aabb &box, int first, int n) const {
    Split s;
    for (int k = 0; k < 3; ++k) { s.cmin[k] = FLT_MAX; s.cmax[k] = -FLT_MAX; }
    for (int i = first; i < first + n; ++i) {
      const int t = order_[i];
      for (int k = 0; k < 3; ++k) {
        // centre of the triangle box, (lo + hi) * 0.5f
        const float c = (prims_.lo[k][t] + prims_.hi[k][t]) * 0.5f;
        s.cmin[k] = c < s.cmin[k] ? c : s.cmin[k];
        s.cmax[k] = c > s.cmax[k] ? c : s.cmax[k];
      }
    }
    s.bins = static_cast<int>(std::min<size_t>(kMaxBins, static_cast<size_t>(4.0f + 0.05f * n)));
    for (int k = 0; k < 3; ++k) s.scale[k] = static_cast<float>(s.bins) / std::max(s.cmax[k] - s.cmin[k], 1e-12f);

    Aabb bin_box[3][kMaxBins];
    int bin_count[3][kMaxBins];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < kMaxBins; ++b) { bin_box[a][b] = Aabb::empty(); bin_count[a][b] = 0; }
    for (int i = first; i < first + n; ++i) {
      const int t = order_[i];
      const Aabb tb = prims_.bounds(t);
      for (int a = 0; a < 3; ++a) {
        int b = static_cast<int>((prims_.centre[a][t] - s.cmin[a]) * s.scale[a]);
        b = std::clamp(b, 0, s.bins - 1);
        bin_box[a][b].take(tb);
        ++bin_count[a][b];
      }
    }
    s.leaf_sah = box.half_area() * n;
    float best = FLT_MAX;
    for (int a = 0; a < 3; ++a) {
      Aabb suffix_box[kMaxBins];
      int suffix_count[kMaxBins];
      Aabb acc = Aabb::empty();
      int cnt = 0;
      for (int b = s.bins - 1; b >= 0; --b) {
        if (bin_count[a][b]) acc.take(bin_box[a][b]);
        cnt += bin_count[a][b];
        suffix_box[b] = acc; suffix_count[b] = cnt;
      }
      Aabb prefix = Aabb::empty();
      int pcnt = 0;
      for (int b = 1; b < s.bins; ++b) {
        if (bin_count[a][b - 1]) prefix.take(bin_box[a][b - 1]);
        pcnt += bin_count[a][b - 1];
        const float cost = prefix.half_area() * pcnt + suffix_box[b].half_area() * suffix_count[b];
        if (cost < best) { best = cost; s.axis = a; s.bin = b; }
      }
    }
    s.split_sah = best;
    return s;
  }

  // Two-index sweep of the reference split: entries below the plane stay,
  // others swap to the back; an empty side falls back to the index median.
  int partition(const Split &s, int first, int last) {
    const int a = s.axis;
    const float plane = s.cmin[a] + (static_cast<float>(s.bin) / s.bins) * (s.cmax[a] - s.cmin[a]);
    int l = first, r = last - 1;
    const std::vector<float> &c = prims_.centre[a];
    while (l <= r) {
      if (c[order_[l]] < plane) ++l;
      else { std::swap(order_[l], order_[r]); --r; }
    }
    if (l == first || l == last) return first + (last - first) / 2;
    return l;
  }

  const BuildOptions &opt_;
  const Primitives &prims_;
  std::vector<int> &order_;
  std::vector<dvs_bvh_node> &nodes_;
};

}  // namespace

int build_bvh(const BuildOptions &opt, const float *vertices, int64_t vertex_count,
              uint32_t *indices, int64_t triangle_count, std::vector<dvs_bvh_node> &nodes) {
  nodes.clear();
  if (triangle_count <= 0) return -1;
  if (triangle_count > INT32_MAX / 2) throw std::invalid_argument("too many triangles for 32-bit BVH indices");
  const int nt = static_cast<int>(triangle_count);
  Primitives prims;
  for (int k = 0; k < 3; ++k) { prims.lo[k].resize(nt); prims.hi[k].resize(nt); prims.centre[k].resize(nt); }
  for (int i = 0; i < nt; ++i) {
    const uint32_t *tri = indices + 3 * static_cast<size_t>(i);
    for (int c = 0; c < 3; ++c)
      if (tri[c] >= static_cast<uint64_t>(vertex_count))
        throw std::invalid_argument("mesh triangle index outside vertex array");
    const float *a = vertices + 3 * static_cast<size_t>(tri[0]);
    const float *b = vertices + 3 * static_cast<size_t>(tri[1]);
    const float *d = vertices + 3 * static_cast<size_t>(tri[2]);
    for (int k = 0; k < 3; ++k) {
      const float lo = std::fmin(std::fmin(a[k], b[k]), d[k]);
      const float hi = std::fmax(std::fmax(a[k], b[k]), d[k]);
      prims.lo[k][i] = lo; prims.hi[k][i] = hi;
      prims.centre[k][i] = (lo + hi) * 0.5f;
    }
  }
  std::vector<int> order(nt);
  for (int i = 0; i < nt; ++i) order[i] = i;
  nodes.reserve(2 * static_cast<size_t>(nt));
  Builder builder(opt, prims, order, nodes);
  const int root = builder.subtree(0, nt);
  std::vector<uint32_t> reordered(3 * static_cast<size_t>(nt));
  for (int i = 0; i < nt; ++i)
    std::memcpy(&reordered[3 * static_cast<size_t>(i)], indices + 3 * static_cast<size_t>(order[i]), 3 * sizeof(uint32_t));
  std::memcpy(indices, reordered.data(), reordered.size() * sizeof(uint32_t));
  return root;
}

std::string validate_bvh(const dvs_bvh_node *nodes, int64_t count, int32_t root) {
  if (root < 0 || root >= count) return "BVH root index out of range";
  std::vector<char> seen(static_cast<size_t>(count), 0);
  std::vector<int32_t> pending{root};
  while (!pending.empty()) {
    const int32_t i = pending.back();
    pending.pop_back();
    if (i < 0 || i >= count) return "BVH invalid node index";
    if (seen[i]) return "BVH cycle detected";
    seen[i] = 1;
    const dvs_bvh_node &n = nodes[i];
    if (n.count > 0) {
      if (n.left != -1 || n.right != -1) return "BVH leaf has non-negative children";
      continue;
    }
    if (n.left < 0 || n.left >= count || n.right < 0 || n.right >= count) return "BVH child index out of range";
    pending.push_back(n.right);
    pending.push_back(n.left);
  }
  return {};
}

SoaBvh relayout(const dvs_bvh_node *nodes, int64_t count, int32_t root, const float *vertices,
                const uint32_t *indices, int64_t triangle_count, int device_leaf_size) {
  SoaBvh out;
  // Box inflation of the device-only sub-boxes: far above the float rounding
  // of the slab test and of the Moller-Trumbore distance (a few ulps of the
  // coordinates and of t), so a sub-box never hides a triangle the
  // reference's full leaf scan would accept.
  float extent = 1.0f;
  for (int64_t t = 0; t < 3 * triangle_count; ++t)
    for (int k = 0; k < 3; ++k) extent = std::max(extent, std::fabs(vertices[3 * static_cast<size_t>(indices[t]) + k]));
  const float margin = 8e-6f * (extent + 1.0f);

  auto new_pair = [&]() {
    out.pairs.resize(out.pairs.size() + 16, 0.0f);
    return static_cast<int32_t>(out.pairs.size() / 16 - 1);
  };
  auto set_child = [&](int32_t pair, int side, const float lo[3], const float hi[3], int32_t ref) {
    float *rec = &out.pairs[16 * static_cast<size_t>(pair)];
    float *box = rec + 6 * side;
    box[0] = lo[0]; box[1] = lo[1]; box[2] = lo[2]; box[3] = hi[0]; box[4] = hi[1]; box[5] = hi[2];
    std::memcpy(&rec[12 + side], &ref, 4);
  };
  auto corner = [&](int32_t tri, int c) { return vertices + 3 * static_cast<size_t>(indices[3 * static_cast<size_t>(tri) + c]); };
  auto make_leaf = [&](const std::vector<int32_t> &ids, int32_t lo, int32_t hi, int depth) {
    const int32_t first = static_cast<int32_t>(out.triangles.size() / 12);
    for (int32_t id : ids) {
      const float *a = corner(id, 0), *b = corner(id, 1), *c = corner(id, 2);
      float rec[12] = {a[0], a[1], a[2], b[0] - a[0], b[1] - a[1], b[2] - a[2], c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      std::memcpy(&rec[9], &id, 4);
      std::memcpy(&rec[10], &lo, 4);
      std::memcpy(&rec[11], &hi, 4);
      out.triangles.insert(out.triangles.end(), rec, rec + 12);
    }
    out.leaf_ranges.push_back(first);
    out.leaf_ranges.push_back(static_cast<int32_t>(ids.size()));
    out.max_depth = std::max(out.max_depth, depth);
    return ~static_cast<int32_t>(out.leaf_ranges.size() / 2 - 1);
  };
  auto bounds = [&](const int32_t *ids, size_t n, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const float *p = corner(ids[i], c);
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
      }
    for (int k = 0; k < 3; ++k) { lo[k] -= margin; hi[k] += margin; }
  };
  // Device-only subtree over one reference leaf: median splits on the longest
  // centroid axis down to device_leaf_size triangles.
  std::function<int32_t(std::vector<int32_t> &, int32_t, int32_t, int)> split =
      [&](std::vector<int32_t> &ids, int32_t lo, int32_t hi, int depth) -> int32_t {
    if (static_cast<int>(ids.size()) <= device_leaf_size) return make_leaf(ids, lo, hi, depth);
    float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    auto centre = [&](int32_t id, int k) { return corner(id, 0)[k] + corner(id, 1)[k] + corner(id, 2)[k]; };
    for (int32_t id : ids)
      for (int k = 0; k < 3; ++k) { cmin[k] = std::min(cmin[k], centre(id, k)); cmax[k] = std::max(cmax[k], centre(id, k)); }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (cmax[k] - cmin[k] > cmax[axis] - cmin[axis]) axis = k;
    std::sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) {
      const float ca = centre(a, axis), cb = centre(b, axis);
      return ca < cb || (ca == cb && a < b);
    });
    std::vector<int32_t> left(ids.begin(), ids.begin() + ids.size() / 2), right(ids.begin() + ids.size() / 2, ids.end());
    const int32_t pair = new_pair();
    float llo[3], lhi[3], rlo[3], rhi[3];
    bounds(left.data(), left.size(), llo, lhi);
    bounds(right.data(), right.size(), rlo, rhi);
    const int32_t l = split(left, lo, hi, depth + 1);
    const int32_t r = split(right, lo, hi, depth + 1);
    set_child(pair, 0, llo, lhi, l);
    set_child(pair, 1, rlo, rhi, r);
    return pair;
  };
  auto leaf_ref = [&](const dvs_bvh_node &n, int depth) -> int32_t {
    if (n.start < 0 || static_cast<int64_t>(n.start) + n.count > triangle_count)
      throw std::invalid_argument("BVH leaf range outside triangle array");
    std::vector<int32_t> ids(n.count);
    for (int32_t i = 0; i < n.count; ++i) ids[i] = n.start + i;
    if (device_leaf_size <= 0 || n.count <= device_leaf_size) return make_leaf(ids, n.start, n.start + n.count, depth);
    return split(ids, n.start, n.start + n.count, depth);
  };
  // Reference tree in pre-order (left first); its internal nodes become pair
  // records holding both children's reference boxes.  Word 14 of a pair
  // holds the first reference triangle index of its right subtree: where
  // every subtree covers one index range and a left subtree's range precedes
  // its right one's (the SAH build partitions its index range in place, so
  // it always does), that index tells on which side of the pair a reference
  // leaf lies (`ordered_ranges`).
  out.ordered_ranges = true;
  std::function<int32_t(int32_t, int, int32_t &, int32_t &)> emit = [&](int32_t node, int depth, int32_t &lo,
                                                                         int32_t &hi) -> int32_t {
    const dvs_bvh_node &n = nodes[node];
    if (n.count > 0) {
      lo = n.start;
      hi = n.start + n.count;
      return leaf_ref(n, depth);
    }
    const int32_t pair = new_pair();
    int32_t llo, lhi, rlo, rhi;
    const int32_t l = emit(n.left, depth + 1, llo, lhi);
    const int32_t r = emit(n.right, depth + 1, rlo, rhi);
    set_child(pair, 0, nodes[n.left].min, nodes[n.left].max, l);
    set_child(pair, 1, nodes[n.right].min, nodes[n.right].max, r);
    if (lhi != rlo) out.ordered_ranges = false;
    std::memcpy(&out.pairs[16 * static_cast<size_t>(pair) + 14], &rlo, 4);
    lo = llo;
    hi = rhi;
    return pair;
  };
  out.triangles.reserve(12 * static_cast<size_t>(triangle_count));
  int32_t all_lo = 0, all_hi = 0;
  out.root_ref = emit(root, 0, all_lo, all_hi);
  out.internal_nodes = static_cast<int64_t>(out.pairs.size() / 16);
  out.leaf_count = static_cast<int64_t>(out.leaf_ranges.size() / 2);
  return out;
}

}  // namespace dvsamd
