// C ABI of the gfx950 DVS library: BVH build, scene upload, frame render and
// the unit entry points.  Host-side orchestration only; every per-ray
// computation runs in the kernels of dvs_kernels.hip.  The event sensor and
// the multi-segment engine are in engine.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bvh_host.h"
#include "common.h"
#include "dvs_hip.h"
#include "device_circuit.hpp"
#include "device_optics.hpp"
#include "kernels.h"
#include "raster.h"
#include "scan_sort.h"
#include "scene.h"

namespace dvsamd {
std::string &last_error() {
  thread_local std::string message;
  return message;
}
}  // namespace dvsamd

using namespace dvsamd;

namespace {
DevScene make_view(const Geometry &g, const dvs_camera &camera, const dvs_hapke &hapke) {
  DevScene v{};
  v.pairs = g.pairs.ptr;
  v.leaves = g.leaves.ptr;
  v.triangles = g.triangles.ptr;
  v.indices = g.indices.ptr;
  v.normals = g.normals.ptr;
  v.albedos = g.has_albedo ? g.albedos.ptr : nullptr;
  v.root_ref = g.root_ref;
  v.stack_depth = g.stack_depth;
  v.width = camera.width;
  v.height = camera.height;
  v.samples = camera.samples;
  v.fx = camera.fx; v.fy = camera.fy; v.cx = camera.cx; v.cy = camera.cy;
  v.hapke = hapke;
  v.hk = dvsdev::hapke_constants(hapke);
  v.packet = g.ordered_ranges ? 1 : 0;
  return v;
}
}  // namespace

namespace {
// Scratch of dvs_frame_render_device_noalloc, which may not allocate: the
// status word (bytes 0-3), the shadow ray counter (bytes 8-15) and the
// converted poses of up to kNoAllocMaxPoses frames from byte kNoAllocHeader,
// 196,864 bytes in all, held by every GPU the library runs on until the
// process ends.  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU race.  In host code `scratch` names a host-side
// stand-in; its device address on each GPU comes from hipGetSymbolAddress.
constexpr int kNoAllocMaxPoses = 4096;
constexpr size_t kNoAllocHeader = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocHeader + sizeof(Pose) * kNoAllocMaxPoses;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kNoAllocDevices = 64;
std::mutex noalloc_mutex;
void *noalloc_ptr[kNoAllocDevices] = {};

// The device address of `scratch` on the GPU that owns `st`, looked up once
// per device with that device current.
void *noalloc_scratch(hipStream_t st) {
  int device = 0;
  hip_check(hipStreamGetDevice(st, &device), "scratch device");
  if (device < 0 || device >= kNoAllocDevices) fail(DVS_E_ARGUMENT, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(noalloc_mutex);
  if (!noalloc_ptr[device]) {
    int current = 0;
    hip_check(hipGetDevice(&current), "scratch device");
    if (current != device) hip_check(hipSetDevice(device), "scratch device");
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) hip_check(hipSetDevice(current), "scratch device");
    hip_check(e, "scratch address");
    if (!ptr) fail(DVS_E_HIP, "scratch address");
    noalloc_ptr[device] = ptr;
  }
  return noalloc_ptr[device];
}

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

const char *dvs_last_error(void) { return last_error().c_str(); }

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
    auto g = std::make_shared<Geometry>();
    std::unique_ptr<dvs_scene_s> s(new dvs_scene_s());
    hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "scene stream");
    g->pairs.upload(reinterpret_cast<const float4 *>(soa.pairs.data()), soa.pairs.size() / 4, "BVH node pairs", s->stream);
    g->leaves.upload(reinterpret_cast<const int2 *>(soa.leaf_ranges.data()), soa.leaf_ranges.size() / 2, "BVH leaves", s->stream);
    g->triangles.upload(reinterpret_cast<const float4 *>(soa.triangles.data()), soa.triangles.size() / 4, "triangle records", s->stream);
    g->indices.upload(indices, 3 * (size_t)triangle_count, "triangle indices", s->stream);
    g->normals.upload(normals, 3 * (size_t)vertex_count, "vertex normals", s->stream);
    if (albedos) g->albedos.upload(albedos, (size_t)vertex_count, "vertex albedo", s->stream);
    hip_check(hipStreamSynchronize(s->stream), "scene upload");
    g->has_albedo = albedos != nullptr;
    g->ordered_ranges = soa.ordered_ranges;
    g->root_ref = soa.root_ref;
    g->stack_depth = std::max(1, soa.max_depth);
    g->internal_nodes = soa.internal_nodes;
    g->leaf_count = soa.leaf_count;
    g->triangle_count = triangle_count;
    g->vertex_count = vertex_count;
    s->geometry = g;
    s->view = make_view(*g, *camera, *hapke);
    *out = s.release();
  });
}

int dvs_scene_view(dvs_scene base, const dvs_camera *camera, const dvs_hapke *hapke, dvs_scene *out) {
  return guarded([&] {
    if (!base || !camera || !out) fail(DVS_E_ARGUMENT, "invalid scene view arguments");
    *out = nullptr;
    if (camera->width <= 0 || camera->height <= 0 || camera->samples < 1 || camera->samples > 16)
      fail(DVS_E_ARGUMENT, "camera grid must be positive with 1..16 sub-rays per axis");
    std::unique_ptr<dvs_scene_s> s(new dvs_scene_s());
    hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "scene stream");
    s->geometry = base->geometry;
    s->view = make_view(*s->geometry, *camera, hapke ? *hapke : base->view.hapke);
    *out = s.release();
  });
}

int dvs_scene_set_traversal(dvs_scene s, int32_t form, int32_t *effective) {
  return guarded([&] {
    if (!s || (form != 0 && form != 1)) fail(DVS_E_ARGUMENT, "invalid traversal form");
    s->view.packet = form == 1 && s->geometry->ordered_ranges ? 1 : 0;
    if (effective) *effective = s->view.packet;
  });
}

int dvs_scene_destroy(dvs_scene scene) {
  scene_release(scene);
  return DVS_OK;
}

int dvs_scene_camera(dvs_scene s, dvs_camera *out) {
  return guarded([&] {
    if (!s || !out) fail(DVS_E_ARGUMENT, "null scene");
    *out = dvs_camera{s->view.width, s->view.height, s->view.fx, s->view.fy, s->view.cx, s->view.cy,
                      s->view.samples};
  });
}

int dvs_scene_info(dvs_scene s, int32_t *max_depth, int64_t *internal_nodes, int64_t *leaves, int64_t *device_bytes) {
  return guarded([&] {
    if (!s) fail(DVS_E_ARGUMENT, "null scene");
    if (max_depth) *max_depth = s->view.stack_depth;
    if (internal_nodes) *internal_nodes = s->geometry->internal_nodes;
    if (leaves) *leaves = s->geometry->leaf_count;
    if (device_bytes) *device_bytes = s->geometry->device_bytes();
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

int dvs_frame_render_device_noalloc(dvs_scene s, const dvs_optical_state *device_poses, int32_t n,
                                    float *device_image, float *device_depth) {
  return guarded([&] {
    if (!s || !device_poses || !device_image || !device_depth || n <= 0)
      fail(DVS_E_ARGUMENT, "invalid parallel optical request");
    if (n > kNoAllocMaxPoses)
      fail(DVS_E_ARGUMENT, "dvs_frame_render_device_noalloc renders at most " + std::to_string(kNoAllocMaxPoses) +
                               " poses per call");
    hipStream_t st = s->stream;
    unsigned char *base = static_cast<unsigned char *>(noalloc_scratch(st));
    int *status = reinterpret_cast<int *>(base);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(base + 8);
    Pose *poses = reinterpret_cast<Pose *>(base + kNoAllocHeader);
    // status word and ray counter start at zero; a fill kernel clears them
    hip_check(launch_fill_int(status, 0, 4, st), "clear optical status");
    hip_check(launch_state_poses(n, device_poses, poses, st), "pose conversion");
    const int npix = s->view.width * s->view.height;
    hip_check(launch_trace(dvsdev::MODE_FRAME, s->view, poses, n, npix, (long long)npix * n, nullptr, nullptr,
                           device_image, device_depth, status, counter, st),
              "parallel optical launch");
    int words[4] = {0, 0, 0, 0};
    hip_check(hipMemcpyAsync(words, base, sizeof(words), hipMemcpyDeviceToHost, st), "parallel optical status");
    hip_check(hipStreamSynchronize(st), "parallel optical wait");
    unsigned long long shadows = 0;
    std::memcpy(&shadows, words + 2, sizeof(shadows));
    s->stats.trace_launches += 1;
    s->stats.primary_rays += (uint64_t)npix * n * s->view.samples * s->view.samples;
    s->stats.shadow_rays += shadows;
    if (words[0]) fail(DVS_E_OPTICAL, "invalid optical sample/BVH stack overflow");
  });
}

int dvs_frame_blur(dvs_scene s, const dvs_optical_state *poses, const double *weights, const int32_t *cell_offsets,
                   int32_t exposure_count, int32_t wide, float *image) {
  return guarded([&] {
    if (!s || !poses || !weights || !cell_offsets || !image || exposure_count <= 0)
      fail(DVS_E_ARGUMENT, "invalid motion-blur request");
    if (cell_offsets[0] != 0) fail(DVS_E_ARGUMENT, "motion-blur cell offsets must start at 0");
    for (int e = 0; e < exposure_count; ++e)
      if (cell_offsets[e + 1] < cell_offsets[e]) fail(DVS_E_ARGUMENT, "motion-blur cell offsets must not decrease");
    if (cell_offsets[exposure_count] <= 0) fail(DVS_E_ARGUMENT, "motion-blur request without cells");
    for (int c = 0; c < cell_offsets[exposure_count]; ++c)
      if (!std::isfinite(weights[c])) fail(DVS_E_ARGUMENT, "motion-blur weights must be finite");
    const size_t npix = (size_t)s->view.width * s->view.height;
    hipStream_t st = s->stream;
    // Exposures go in groups of whole exposures whose frames fit kBlurFrameBytes
    // (a single larger exposure forms a group of its own): one multi-pose trace
    // launch renders every cell of a group, one launch sums each exposure.
    constexpr size_t kBlurFrameBytes = size_t(1) << 30;
    const size_t max_cells = std::max<size_t>(1, kBlurFrameBytes / (npix * sizeof(float)));
    DeviceArray<dvs_optical_state> states;
    DeviceArray<double> wdev;
    DeviceArray<int> odev;
    std::vector<int> local;
    int e0 = 0;
    while (e0 < exposure_count) {
      int e1 = e0 + 1;
      while (e1 < exposure_count && e1 - e0 < 65535 &&
             (size_t)(cell_offsets[e1 + 1] - cell_offsets[e0]) <= max_cells)
        ++e1;
      const int c0 = cell_offsets[e0], n = cell_offsets[e1] - c0;
      float *out = image + (size_t)e0 * npix;
      if (n == 0) {
        std::fill(out, out + (size_t)(e1 - e0) * npix, 0.0f);
        e0 = e1;
        continue;
      }
      local.assign(cell_offsets + e0, cell_offsets + e1 + 1);
      for (int &o : local) o -= c0;
      states.upload(poses + c0, n, "upload motion-blur poses", st);
      wdev.upload(weights + c0, n, "upload motion-blur weights", st);
      odev.upload(local.data(), local.size(), "upload motion-blur offsets", st);
      s->poses.reserve(n, "optical poses");
      hip_check(launch_state_poses(n, states.ptr, s->poses.ptr, st), "pose conversion");
      s->image.reserve(npix * n, "optical image");
      frame_trace(s, s->poses.ptr, n, nullptr, s->image.ptr, nullptr, st);
      s->blur.reserve(npix * (e1 - e0), "motion-blur image");
      hip_check(launch_blur_accumulate((int)npix, e1 - e0, odev.ptr, s->image.ptr, wdev.ptr, wide ? 1 : 0,
                                       s->blur.ptr, st),
                "motion-blur accumulation");
      hip_check(hipMemcpyAsync(out, s->blur.ptr, npix * (e1 - e0) * sizeof(float), hipMemcpyDeviceToHost, st),
                "download motion-blur image");
      hip_check(hipStreamSynchronize(st), "motion-blur wait");
      e0 = e1;
    }
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

float dvs_hapke_reflectance(float cosi, float cose, float cosg, float albedo, const dvs_hapke *hapke) {
  return hapke ? dvsdev::hapke_reflectance(cosi, cose, cosg, albedo, *hapke) : NAN;
}

float dvs_hapke_directional(const float normal[3], const float light[3], const float view[3], float albedo,
                            const dvs_hapke *hapke) {
  if (!normal || !light || !view || !hapke) return NAN;
  using dvsdev::mk;
  return dvsdev::hapke_directional(mk(normal[0], normal[1], normal[2]), mk(light[0], light[1], light[2]),
                                   mk(view[0], view[1], view[2]), albedo, *hapke);
}

float dvs_invert_hapke_if(float incidence, float emission, float phase, float target, const dvs_hapke *hapke) {
  return hapke ? dvsdev::invert_hapke_if(incidence, emission, phase, target, *hapke) : NAN;
}

double dvs_receptor_log_at(double u, double rate, double dt, double charge) {
  return dvsdev::receptor_log_at(u, rate, dt, charge);
}

int32_t dvs_advance_pixel(dvs_pixel_state *state, const dvs_pixel_params *params, double rate, double t0, double t1,
                          dvs_crossing *out, int32_t capacity, double (*normal)(void *), void *context) {
  if (!state || !params || capacity < 0 || (capacity > 0 && !out)) return -2;
  dvsdev::PixelRegs r{state->receptor_log, state->reference_log, state->refractory_until, state->threshold_pos,
                      state->threshold_neg, state->reset_pending};
  const dvsdev::CircuitParams p{params->threshold_pos, params->threshold_neg, params->charge_scale,
                                params->refractory_time, params->reference_leak, params->reset_noise_std};
  auto draw = [&] { return normal ? normal(context) : 0.0; };
  const int count = dvsdev::advance_pixel(r, p, rate, t0, t1, reinterpret_cast<dvsdev::CrossingEvent *>(out),
                                          capacity, draw);
  state->receptor_log = r.receptor_log;
  state->reference_log = r.reference_log;
  state->refractory_until = r.refractory_until;
  state->reset_pending = r.reset_pending;
  state->receptor_inverse = std::exp(-r.receptor_log);
  return count;
}

int dvs_invert_hapke_eval(const float *inc, const float *emi, const float *phase, const float *target, int64_t n,
                          const dvs_hapke *h, float *out) {
  return guarded([&] {
    if (!inc || !emi || !phase || !target || !h || !out || n < 0) fail(DVS_E_ARGUMENT, "invalid Hapke arguments");
    DeviceArray<float> a, b, c, d, o;
    a.upload(inc, n, "hapke input"); b.upload(emi, n, "hapke input"); c.upload(phase, n, "hapke input");
    d.upload(target, n, "hapke input"); o.reserve(n, "hapke output");
    hip_check(launch_invert_hapke(n, a.ptr, b.ptr, c.ptr, d.ptr, *h, o.ptr, nullptr), "hapke kernel");
    hip_check(hipMemcpy(out, o.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "hapke output");
  });
}

int dvs_advance_pixel_eval(dvs_pixel_state *states, const dvs_pixel_params *params, const double *rates,
                           const double *t0, const double *t1, int64_t n, int32_t capacity, double normal,
                           int32_t *counts, dvs_crossing *events) {
  return guarded([&] {
    if (!states || !params || !rates || !t0 || !t1 || !counts || n < 0 || capacity < 0 || (capacity > 0 && !events))
      fail(DVS_E_ARGUMENT, "invalid advance_pixel arguments");
    DeviceArray<dvs_pixel_state> st;
    DeviceArray<dvs_pixel_params> pa;
    DeviceArray<double> r, a, b;
    DeviceArray<int> c;
    DeviceArray<dvs_crossing> ev;
    st.upload(states, n, "advance input"); pa.upload(params, n, "advance input"); r.upload(rates, n, "advance input");
    a.upload(t0, n, "advance input"); b.upload(t1, n, "advance input");
    c.reserve(n, "advance output");
    ev.reserve(std::max<size_t>(1, (size_t)n * capacity), "advance output");
    hip_check(launch_advance(n, st.ptr, pa.ptr, r.ptr, a.ptr, b.ptr, capacity, normal, c.ptr, ev.ptr, nullptr),
              "advance kernel");
    hip_check(hipMemcpy(states, st.ptr, n * sizeof(dvs_pixel_state), hipMemcpyDeviceToHost), "advance output");
    hip_check(hipMemcpy(counts, c.ptr, n * sizeof(int), hipMemcpyDeviceToHost), "advance output");
    if (capacity > 0)
      hip_check(hipMemcpy(events, ev.ptr, (size_t)n * capacity * sizeof(dvs_crossing), hipMemcpyDeviceToHost),
                "advance output");
  });
}

int dvs_poisson_eval(const double *lambda, const double *u, int64_t n, double *out) {
  return guarded([&] {
    if (!lambda || !u || !out || n < 0) fail(DVS_E_ARGUMENT, "invalid poisson arguments");
    DeviceArray<double> a, b, o;
    a.upload(lambda, n, "poisson input");
    b.upload(u, n, "poisson input");
    o.reserve(n, "poisson output");
    hip_check(launch_poisson(n, a.ptr, b.ptr, o.ptr, nullptr), "poisson kernel");
    hip_check(hipMemcpy(out, o.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "poisson output");
  });
}

int dvs_raster_mesh(const float *dem, const float *albedo, const float *x_coords, const float *y_coords,
                    int32_t width, int32_t height, float *vertices, float *normals, float *albedos,
                    int64_t *vertex_count, uint32_t *indices, int64_t *triangle_count) {
  return guarded([&] {
    if (!dem || !x_coords || !y_coords || width <= 0 || height <= 0 || !vertices || !normals || !indices ||
        !vertex_count || !triangle_count || (albedo && !albedos) || (int64_t)width * height > INT32_MAX)
      fail(DVS_E_ARGUMENT, "invalid raster mesh arguments");
    const size_t npix = (size_t)width * height, ncell = (size_t)std::max(width - 1, 0) * std::max(height - 1, 0);
    const float centre = dem[(size_t)width * (height / 2) + width / 2];
    if (!std::isfinite(centre))
      fail(DVS_E_ARGUMENT, "DEM center is NoData; choose a valid crop/reference origin before rendering");
    DeviceArray<float> d_dem, d_alb, d_x, d_y, d_v, d_n, d_a;
    DeviceArray<int> pflags, cflags, pscan, cscan;
    DeviceArray<unsigned> d_i;
    DeviceArray<unsigned char> temp;
    d_dem.upload(dem, npix, "upload DEM");
    if (albedo) d_alb.upload(albedo, npix, "upload albedo");
    d_x.upload(x_coords, width, "upload raster axes");
    d_y.upload(y_coords, height, "upload raster axes");
    pflags.reserve(npix, "raster flags");
    cflags.reserve(std::max<size_t>(ncell, 1), "raster flags");
    pscan.reserve(npix, "raster scan");
    cscan.reserve(std::max<size_t>(ncell, 1), "raster scan");
    hip_check(launch_raster_flags(d_dem.ptr, width, height, pflags.ptr, cflags.ptr, nullptr), "raster flags");
    size_t bytes = 0, bytes2 = 0;
    hip_check(inclusive_scan_int(nullptr, bytes, pflags.ptr, pscan.ptr, npix, nullptr), "scan size");
    hip_check(inclusive_scan_int(nullptr, bytes2, cflags.ptr, cscan.ptr, std::max<size_t>(ncell, 1), nullptr),
              "scan size");
    temp.reserve(std::max(bytes, bytes2), "scan scratch");
    bytes = std::max(bytes, bytes2);
    hip_check(inclusive_scan_int(temp.ptr, bytes, pflags.ptr, pscan.ptr, npix, nullptr), "raster vertex scan");
    int nv = 0, nc = 0;
    if (ncell) hip_check(inclusive_scan_int(temp.ptr, bytes, cflags.ptr, cscan.ptr, ncell, nullptr), "raster cell scan");
    hip_check(hipMemcpy(&nv, pscan.ptr + npix - 1, sizeof(int), hipMemcpyDeviceToHost), "raster vertex count");
    if (ncell) hip_check(hipMemcpy(&nc, cscan.ptr + ncell - 1, sizeof(int), hipMemcpyDeviceToHost), "raster cell count");
    d_v.reserve(std::max(nv, 1) * 3, "raster vertices");
    d_n.reserve(std::max(nv, 1) * 3, "raster normals");
    if (albedo) d_a.reserve(std::max(nv, 1), "raster albedos");
    d_i.reserve(std::max(nc, 1) * 6, "raster triangles");
    const float cx = -x_coords[width / 2], cy = -y_coords[height / 2], cz = -centre;
    hip_check(launch_raster_mesh(d_dem.ptr, albedo ? d_alb.ptr : nullptr, d_x.ptr, d_y.ptr, width, height, pscan.ptr,
                                 cscan.ptr, cflags.ptr, cx, cy, cz, d_v.ptr, d_n.ptr, albedo ? d_a.ptr : nullptr,
                                 d_i.ptr, nullptr),
              "raster mesh");
    hip_check(hipMemcpy(vertices, d_v.ptr, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost), "raster vertices");
    hip_check(hipMemcpy(normals, d_n.ptr, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost), "raster normals");
    if (albedo) hip_check(hipMemcpy(albedos, d_a.ptr, sizeof(float) * nv, hipMemcpyDeviceToHost), "raster albedos");
    hip_check(hipMemcpy(indices, d_i.ptr, sizeof(unsigned) * 6 * nc, hipMemcpyDeviceToHost), "raster triangles");
    *vertex_count = nv;
    *triangle_count = 2 * (int64_t)nc;
  });
}

int dvs_raster_align(const float *ref_x, const float *ref_y, int32_t width, int32_t height, const float *source,
                     const float *source_x, const float *source_y, int32_t source_width, int32_t source_height,
                     float *out) {
  return guarded([&] {
    if (!ref_x || !ref_y || !source || !source_x || !source_y || !out || width <= 0 || height <= 0 ||
        source_width <= 0 || source_height <= 0)
      fail(DVS_E_ARGUMENT, "invalid raster alignment arguments");
    DeviceArray<float> rx, ry, src, sx, sy, o;
    rx.upload(ref_x, width, "upload raster axes");
    ry.upload(ref_y, height, "upload raster axes");
    src.upload(source, (size_t)source_width * source_height, "upload raster");
    sx.upload(source_x, source_width, "upload raster axes");
    sy.upload(source_y, source_height, "upload raster axes");
    o.reserve((size_t)width * height, "aligned raster");
    hip_check(launch_raster_align(rx.ptr, ry.ptr, width, height, src.ptr, sx.ptr, sy.ptr, source_width, source_height,
                                  o.ptr, nullptr),
              "raster alignment");
    hip_check(hipMemcpy(out, o.ptr, sizeof(float) * width * height, hipMemcpyDeviceToHost), "aligned raster");
  });
}

int dvs_scan_eval(const int32_t *in, int64_t n, int32_t *out) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!in || !out)) || n > INT32_MAX) fail(DVS_E_ARGUMENT, "invalid scan arguments");
    if (n == 0) return;
    DeviceArray<int> a;
    DeviceArray<unsigned char> temp;
    a.upload(in, n, "scan input");
    size_t bytes = 0;
    hip_check(inclusive_scan_int(nullptr, bytes, a.ptr, a.ptr, n, nullptr), "scan size");
    temp.reserve(bytes, "scan scratch");
    hip_check(inclusive_scan_int(temp.ptr, bytes, a.ptr, a.ptr, n, nullptr), "scan");
    hip_check(hipMemcpy(out, a.ptr, n * sizeof(int), hipMemcpyDeviceToHost), "scan output");
  });
}

int dvs_sort_eval(const uint64_t *keys, const uint32_t *values, int64_t n, int32_t begin_bit, int32_t end_bit,
                  uint64_t *keys_out, uint32_t *values_out) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!keys || !values || !keys_out || !values_out)) || n > INT32_MAX || begin_bit < 0 ||
        end_bit > 64 || begin_bit > end_bit)
      fail(DVS_E_ARGUMENT, "invalid sort arguments");
    if (n == 0) return;
    DeviceArray<unsigned long long> k, k2;
    DeviceArray<unsigned> v, v2;
    DeviceArray<unsigned char> temp;
    k.upload(reinterpret_cast<const unsigned long long *>(keys), n, "sort keys");
    v.upload(values, n, "sort values");
    k2.reserve(n, "sort keys");
    v2.reserve(n, "sort values");
    size_t bytes = 0;
    hip_check(sort_pairs_u64(nullptr, bytes, k.ptr, k2.ptr, v.ptr, v2.ptr, n, begin_bit, end_bit, nullptr),
              "sort size");
    temp.reserve(bytes, "sort scratch");
    hip_check(sort_pairs_u64(temp.ptr, bytes, k.ptr, k2.ptr, v.ptr, v2.ptr, n, begin_bit, end_bit, nullptr), "sort");
    hip_check(hipMemcpy(keys_out, k2.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "sort output");
    hip_check(hipMemcpy(values_out, v2.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost), "sort output");
  });
}

int dvs_sort_bucketed_eval(const uint64_t *keys, const uint32_t *values, int64_t n, uint64_t key_lo,
                           int32_t key_bits, int32_t bucket_bits, uint64_t *keys_out, uint32_t *values_out,
                           int32_t *overflow) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!keys || !values || !keys_out || !values_out)) || !overflow || n > INT32_MAX ||
        key_bits < 0 || key_bits > 64 || bucket_bits <= 0 || bucket_bits > 64)
      fail(DVS_E_ARGUMENT, "invalid sort arguments");
    *overflow = 0;
    if (n == 0) return;
    DeviceArray<unsigned long long> k, k2;
    DeviceArray<unsigned> v, v2;
    DeviceArray<int> flag;
    DeviceArray<unsigned char> temp;
    k.upload(reinterpret_cast<const unsigned long long *>(keys), n, "sort keys");
    v.upload(values, n, "sort values");
    k2.reserve(n, "sort keys");
    v2.reserve(n, "sort values");
    const int zero = 0;
    flag.upload(&zero, 1, "sort flag");
    size_t bytes = 0;
    hip_check(sort_pairs_u64_bucketed(nullptr, bytes, k.ptr, k2.ptr, v.ptr, v2.ptr, n, key_lo, key_bits, bucket_bits,
                                      flag.ptr, nullptr),
              "sort size");
    temp.reserve(bytes, "sort scratch");
    hip_check(sort_pairs_u64_bucketed(temp.ptr, bytes, k.ptr, k2.ptr, v.ptr, v2.ptr, n, key_lo, key_bits, bucket_bits,
                                      flag.ptr, nullptr),
              "sort");
    hip_check(hipMemcpy(keys_out, k2.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "sort output");
    hip_check(hipMemcpy(values_out, v2.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost), "sort output");
    hip_check(hipMemcpy(overflow, flag.ptr, sizeof(int), hipMemcpyDeviceToHost), "sort output");
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
