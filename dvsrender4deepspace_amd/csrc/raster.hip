// DEM -> triangle mesh on the device (tools::build_from_raster, planar
// path, src/tools/build_from_raster.cu:213-245, 288-331, 333-472, 580-610, and
// the recentring of load_from_raster, :786-802, 898-914) and the bilinear
// alignment of an albedo raster onto the DEM grid (batch_align_kernel,
// :653-682).
//
// The reference numbers vertices and triangles with atomic counters, so
// their order changes from run to run.  Here the order is the raster's:
// valid pixels get their vertex index from a prefix sum over pixel flags,
// complete cells their triangle pair from a prefix sum over cell flags
// (row-major), so one DEM always gives the same mesh, index for index.
// Per pixel and per cell the arithmetic is the reference's: vertices at
// (x_coord, y_coord, z), triangles (v1, v3, v2) and (v2, v3, v4) of each
// cell, vertex normals gathered from the up-to-eight triangles around a
// vertex with angle weights and faces turned to +z, computed before the
// mesh is moved so that the DEM's centre pixel sits at the origin.
#include <hip/hip_runtime.h>

#include <cmath>

#include "raster.h"

namespace dvsamd {
namespace {

__device__ inline bool valid_height(float z) { return isfinite(z) && fabsf(z) <= 1e8f; }

// Flags of valid pixels (vertices) and of cells with four valid corners.
__global__ void raster_flags_kernel(const float *dem, int width, int height, int *pixel_flags, int *cell_flags) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= width) return;
  const size_t i = (size_t)y * width + x;
  pixel_flags[i] = valid_height(dem[i]) ? 1 : 0;
  if (x < width - 1 && y < height - 1) {
    const bool ok = valid_height(dem[i]) && valid_height(dem[i + 1]) && valid_height(dem[i + width]) &&
                    valid_height(dem[i + width + 1]);
    cell_flags[(size_t)y * (width - 1) + x] = ok ? 1 : 0;
  }
}

struct V3 {
  float x, y, z;
};
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float len3(V3 a) { return sqrtf(dot3(a, a)); }
__device__ inline V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

__device__ inline V3 grid_point(const float *dem, const float *xs, const float *ys, int width, int px, int py) {
  return {xs[px], ys[py], dem[(size_t)py * width + px]};
}

// accumulate_triangle_normal_for_vertex (build_from_raster.cu:37-127),
// planar orientation: the face normal of (a, b, c) turned to +z, weighted by
// the triangle's angle at corner `k` (0: a, 1: b, 2: c).
__device__ inline void add_corner(V3 pa, V3 pb, V3 pc, int k, V3 &acc) {
  V3 face = cross3(sub(pb, pa), sub(pc, pa));
  const float face_len = len3(face);
  if (!isfinite(face_len) || face_len < 1e-12f) return;
  face = divs(face, face_len);
  V3 t, o1, o2;
  if (k == 0) {
    t = pa; o1 = pb; o2 = pc;
  } else if (k == 1) {
    t = pb; o1 = pc; o2 = pa;
  } else {
    t = pc; o1 = pa; o2 = pb;
  }
  V3 e1 = sub(o1, t), e2 = sub(o2, t);
  const float l1 = len3(e1), l2 = len3(e2);
  if (!isfinite(l1) || !isfinite(l2) || l1 < 1e-12f || l2 < 1e-12f) return;
  e1 = divs(e1, l1);
  e2 = divs(e2, l2);
  const float c = fminf(1.0f, fmaxf(-1.0f, dot3(e1, e2)));
  const float angle = acosf(c);
  if (!isfinite(angle)) return;
  if (face.z < 0.0f) face = {face.x * -1.0f, face.y * -1.0f, face.z * -1.0f};
  acc = {acc.x + face.x * angle, acc.y + face.y * angle, acc.z + face.z * angle};
}

// Vertices (recentred), normals and albedos of the valid pixels, at the
// vertex index the prefix sum gives them.
__global__ void raster_vertices_kernel(const float *dem, const float *albedo, const float *xs, const float *ys,
                                       int width, int height, const int *pixel_scan, float cx, float cy, float cz,
                                       float *vertices, float *normals, float *albedos) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= width) return;
  const size_t i = (size_t)y * width + x;
  const float z = dem[i];
  if (!valid_height(z)) return;
  const int v = pixel_scan[i] - 1;
  const V3 me = grid_point(dem, xs, ys, width, x, y);
  V3 acc = {0.0f, 0.0f, 0.0f};
  // the four cells around the pixel, (-1,-1), (0,-1), (-1,0), (0,0), each
  // with T1 = (v1, v3, v2) and T2 = (v2, v3, v4) (build_from_raster.cu:388-472)
  for (int oy = -1; oy <= 0; ++oy)
    for (int ox = -1; ox <= 0; ++ox) {
      const int cx0 = x + ox, cy0 = y + oy;
      if (cx0 < 0 || cy0 < 0 || cx0 >= width - 1 || cy0 >= height - 1) continue;
      const size_t c = (size_t)cy0 * width + cx0;
      if (!valid_height(dem[c]) || !valid_height(dem[c + 1]) || !valid_height(dem[c + width]) ||
          !valid_height(dem[c + width + 1]))
        continue;
      const V3 p1 = grid_point(dem, xs, ys, width, cx0, cy0), p2 = grid_point(dem, xs, ys, width, cx0 + 1, cy0);
      const V3 p3 = grid_point(dem, xs, ys, width, cx0, cy0 + 1), p4 = grid_point(dem, xs, ys, width, cx0 + 1, cy0 + 1);
      // this pixel's corner in each triangle: v1 = (0,0), v2 = (1,0), v3 = (0,1), v4 = (1,1) of the cell
      const int corner = (-ox) + 2 * (-oy);  // 0: v1, 1: v2, 2: v3, 3: v4
      if (corner == 0) add_corner(p1, p3, p2, 0, acc);
      else if (corner == 2) add_corner(p1, p3, p2, 1, acc);
      else if (corner == 1) add_corner(p1, p3, p2, 2, acc);
      if (corner == 1) add_corner(p2, p3, p4, 0, acc);
      else if (corner == 2) add_corner(p2, p3, p4, 1, acc);
      else if (corner == 3) add_corner(p2, p3, p4, 2, acc);
    }
  const float l = len3(acc);
  V3 n = {0.0f, 0.0f, 0.0f};
  if (isfinite(l) && l >= 1e-12f) n = divs(acc, l);
  vertices[3 * (size_t)v + 0] = me.x + cx;
  vertices[3 * (size_t)v + 1] = me.y + cy;
  vertices[3 * (size_t)v + 2] = me.z + cz;
  // + 0: the reference passes the normals through its identity rotation,
  // which leaves no negative zero
  normals[3 * (size_t)v + 0] = n.x + 0.0f;
  normals[3 * (size_t)v + 1] = n.y + 0.0f;
  normals[3 * (size_t)v + 2] = n.z + 0.0f;
  if (albedos) albedos[v] = albedo[i];
}

__global__ void raster_triangles_kernel(int width, int height, const int *pixel_scan, const int *cell_scan,
                                        const int *cell_flags, unsigned *indices) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= width - 1) return;
  const size_t c = (size_t)y * (width - 1) + x;
  if (!cell_flags[c]) return;
  const size_t t = 2 * (size_t)(cell_scan[c] - 1);
  const size_t i = (size_t)y * width + x;
  const unsigned v1 = pixel_scan[i] - 1, v2 = pixel_scan[i + 1] - 1;
  const unsigned v3 = pixel_scan[i + width] - 1, v4 = pixel_scan[i + width + 1] - 1;
  unsigned *tri = indices + 3 * t;
  tri[0] = v1;
  tri[1] = v3;
  tri[2] = v2;
  tri[3] = v2;
  tri[4] = v3;
  tri[5] = v4;
}

// raster_bracket (build_from_raster.cu:653-660): the cell of a regular axis,
// increasing or decreasing, that holds `value`; -1 outside.
__device__ int raster_bracket(const float *axis, int n, float value) {
  if (n < 2 || !isfinite(value)) return -1;
  const bool increasing = axis[n - 1] > axis[0];
  if (value < fminf(axis[0], axis[n - 1]) || value > fmaxf(axis[0], axis[n - 1])) return -1;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if ((axis[mid] <= value) == increasing) lo = mid;
    else hi = mid;
  }
  return lo;
}

// batch_align_kernel (build_from_raster.cu:661-682): bilinear weights in
// float, their products with the samples in float, the sums in double;
// NaN outside the source or where a weighted sample is NoData.
__global__ void raster_align_kernel(const float *ref_x, const float *ref_y, int width, int height,
                                    const float *source, const float *source_x, const float *source_y,
                                    int source_width, int source_height, float *out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= width) return;
  const int row = raster_bracket(source_y, source_height, ref_y[y]);
  const int col = raster_bracket(source_x, source_width, ref_x[x]);
  float value = __builtin_nanf("");
  if (row >= 0 && col >= 0) {
    const float u = (ref_x[x] - source_x[col]) / (source_x[col + 1] - source_x[col]);
    const float v = (ref_y[y] - source_y[row]) / (source_y[row + 1] - source_y[row]);
    double total = 0, weight = 0;
    bool valid = true;
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const float w = (dx ? u : 1 - u) * (dy ? v : 1 - v);
        const float sample = source[(size_t)(row + dy) * source_width + col + dx];
        if (w > 0) {
          if (!isfinite(sample)) {
            valid = false;
            continue;
          }
          const float product = w * sample;
          total += (double)product;
          weight += (double)w;
        }
      }
    if (valid && weight > 0) value = (float)(total / weight);
  }
  out[(size_t)y * width + x] = value;
}

}  // namespace

hipError_t launch_raster_flags(const float *dem, int width, int height, int *pixel_flags, int *cell_flags,
                               hipStream_t st) {
  raster_flags_kernel<<<dim3((width + 255) / 256, height), 256, 0, st>>>(dem, width, height, pixel_flags, cell_flags);
  return hipGetLastError();
}

hipError_t launch_raster_mesh(const float *dem, const float *albedo, const float *xs, const float *ys, int width,
                              int height, const int *pixel_scan, const int *cell_scan, const int *cell_flags,
                              float cx, float cy, float cz, float *vertices, float *normals, float *albedos,
                              unsigned *indices, hipStream_t st) {
  raster_vertices_kernel<<<dim3((width + 255) / 256, height), 256, 0, st>>>(
      dem, albedo, xs, ys, width, height, pixel_scan, cx, cy, cz, vertices, normals, albedos);
  if (width > 1 && height > 1)
    raster_triangles_kernel<<<dim3((width - 1 + 255) / 256, height - 1), 256, 0, st>>>(width, height, pixel_scan,
                                                                                        cell_scan, cell_flags, indices);
  return hipGetLastError();
}

hipError_t launch_raster_align(const float *ref_x, const float *ref_y, int width, int height, const float *source,
                               const float *source_x, const float *source_y, int source_width, int source_height,
                               float *out, hipStream_t st) {
  raster_align_kernel<<<dim3((width + 255) / 256, height), 256, 0, st>>>(
      ref_x, ref_y, width, height, source, source_x, source_y, source_width, source_height, out);
  return hipGetLastError();
}

}  // namespace dvsamd
