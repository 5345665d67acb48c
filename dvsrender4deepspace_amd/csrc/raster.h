// DEM -> mesh and raster alignment kernels (raster.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace dvsamd {
// pixel_flags[width*height]: valid elevation; cell_flags[(width-1)*(height-1)]:
// four valid corners.
hipError_t launch_raster_flags(const float *dem, int width, int height, int *pixel_flags, int *cell_flags,
                               hipStream_t st);
// From the inclusive prefix sums of the flags: vertices (moved by (cx, cy,
// cz)), normals, albedos (if albedo is given) and triangle indices.
hipError_t launch_raster_mesh(const float *dem, const float *albedo, const float *xs, const float *ys, int width,
                              int height, const int *pixel_scan, const int *cell_scan, const int *cell_flags,
                              float cx, float cy, float cz, float *vertices, float *normals, float *albedos,
                              unsigned *indices, hipStream_t st);
hipError_t launch_raster_align(const float *ref_x, const float *ref_y, int width, int height, const float *source,
                               const float *source_x, const float *source_y, int source_width, int source_height,
                               float *out, hipStream_t st);
}  // namespace dvsamd
