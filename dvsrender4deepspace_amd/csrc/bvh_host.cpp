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
                const uint32_t *indices, int64_t triangle_count) {
  SoaBvh out;
  std::vector<int32_t> ref(static_cast<size_t>(count), 0);
  // Pre-order numbering, left child first: internal nodes become pair records,
  // leaves become (start, count) ranges.  Child refs >= 0 name pair records,
  // refs < 0 are ~leaf_index.
  struct Item { int32_t node, depth; };
  std::vector<Item> pending{{root, 0}};
  std::vector<int32_t> internal_order;
  int32_t pairs = 0, leaves = 0;
  while (!pending.empty()) {
    Item it = pending.back();
    pending.pop_back();
    const dvs_bvh_node &n = nodes[it.node];
    if (n.count > 0) {
      ref[it.node] = ~leaves++;
      out.leaf_ranges.push_back(n.start);
      out.leaf_ranges.push_back(n.count);
      if (n.start < 0 || static_cast<int64_t>(n.start) + n.count > triangle_count)
        throw std::invalid_argument("BVH leaf range outside triangle array");
      out.max_depth = std::max(out.max_depth, it.depth);
      continue;
    }
    ref[it.node] = pairs++;
    internal_order.push_back(it.node);
    pending.push_back({n.right, it.depth + 1});
    pending.push_back({n.left, it.depth + 1});
  }
  out.root_ref = ref[root];
  out.internal_nodes = pairs;
  out.leaf_count = leaves;
  out.pairs.resize(16 * static_cast<size_t>(pairs));
  for (int32_t p = 0; p < pairs; ++p) {
    const dvs_bvh_node &n = nodes[internal_order[p]];
    const dvs_bvh_node &l = nodes[n.left];
    const dvs_bvh_node &r = nodes[n.right];
    float *rec = &out.pairs[16 * static_cast<size_t>(p)];
    // 64-byte record: left box, right box, left ref, right ref, 2 spare words.
    rec[0] = l.min[0]; rec[1] = l.min[1]; rec[2] = l.min[2]; rec[3] = l.max[0];
    rec[4] = l.max[1]; rec[5] = l.max[2]; rec[6] = r.min[0]; rec[7] = r.min[1];
    rec[8] = r.min[2]; rec[9] = r.max[0]; rec[10] = r.max[1]; rec[11] = r.max[2];
    std::memcpy(&rec[12], &ref[n.left], 4);
    std::memcpy(&rec[13], &ref[n.right], 4);
    rec[14] = 0; rec[15] = 0;
  }
  // 48-byte triangle records in BVH order: v0, e1 = v1 - v0, e2 = v2 - v0.
  out.triangles.resize(12 * static_cast<size_t>(triangle_count));
  for (int64_t t = 0; t < triangle_count; ++t) {
    const uint32_t *tri = indices + 3 * t;
    const float *a = vertices + 3 * static_cast<size_t>(tri[0]);
    const float *b = vertices + 3 * static_cast<size_t>(tri[1]);
    const float *c = vertices + 3 * static_cast<size_t>(tri[2]);
    float *rec = &out.triangles[12 * t];
    rec[0] = a[0]; rec[1] = a[1]; rec[2] = a[2];
    rec[3] = b[0] - a[0]; rec[4] = b[1] - a[1]; rec[5] = b[2] - a[2];
    rec[6] = c[0] - a[0]; rec[7] = c[1] - a[1]; rec[8] = c[2] - a[2];
    rec[9] = 0; rec[10] = 0; rec[11] = 0;
  }
  return out;
}

}  // namespace dvsamd
