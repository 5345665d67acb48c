// Scene handle: resident BVH + indexed mesh (shared, reference counted) and
// the camera of one view of it.
#pragma once
#include <atomic>
#include <memory>

#include "common.h"
#include "kernels.h"

namespace dvsamd {

// Device arrays of one uploaded scene geometry; several camera views share it.
struct Geometry {
  DeviceArray<float4> pairs, triangles;
  DeviceArray<int2> leaves;
  DeviceArray<uint32_t> indices;
  DeviceArray<float> normals, albedos;
  bool has_albedo = false;
  bool ordered_ranges = false;  // the wave-packet traversal may serve the tree
  int32_t root_ref = 0, stack_depth = 1;
  int64_t internal_nodes = 0, leaf_count = 0, triangle_count = 0, vertex_count = 0;
  int64_t device_bytes() const {
    return (int64_t)(pairs.bytes() + triangles.bytes() + leaves.bytes() + indices.bytes() + normals.bytes() +
                     albedos.bytes());
  }
};

}  // namespace dvsamd

struct dvs_scene_s {
  std::atomic<int> refs{1};
  std::shared_ptr<dvsamd::Geometry> geometry;
  dvsamd::DevScene view{};
  hipStream_t stream = nullptr;
  dvsamd::DeviceArray<dvsamd::Pose> poses;
  dvsamd::DeviceArray<float> image, depth, blur;
  dvsamd::DeviceArray<double> radiance;
  dvsamd::DeviceArray<int> status;
  dvsamd::DeviceArray<unsigned long long> counter;
  dvs_stats stats{};
  ~dvs_scene_s() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace dvsamd {
inline void scene_retain(dvs_scene s) { s->refs.fetch_add(1); }
inline void scene_release(dvs_scene s) {
  if (s && s->refs.fetch_sub(1) == 1) delete s;
}
}  // namespace dvsamd
