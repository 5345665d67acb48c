/*
 * Host I/O of the DVS path: the event stream writers of bin/dvs, the YAML
 * configuration, the scene bundle and trajectory loaders, and the frame
 * images of bin/cam.  C ABI of libdvs_io.so (host code, HDF5 1.10 C library);
 * the status codes and thread-local message follow dvs_hip.h.
 */
#ifndef DVS_IO_H
#define DVS_IO_H
#include <stdint.h>

#include "dvs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the calling thread's last I/O failure ("" if none). */
const char *dvs_io_last_error(void);

/* ---- Event stream writers (include/dvs/dvs_writer.hpp) -----------------
 * DVS_EVENTS_BIN:  BinaryEventStreamWriter (:22-58), packed 13-byte records
 *                  {float t, int16 x, int16 y, int8 p, float depth}.
 * DVS_EVENTS_HDF5: Hdf5EventStreamWriter (:60-264), extendable datasets x, y
 *                  (int16), t (int64 microseconds, llround(t * 1e6)), p (int8)
 *                  in chunks of 65,536, and ms_to_idx (int64) written at close:
 *                  entry m is the index of the first event at or after m ms.
 *                  Appends must be globally time ordered (each batch is
 *                  stably sorted by time first). */
enum { DVS_EVENTS_BIN = 0, DVS_EVENTS_HDF5 = 1 };
typedef struct dvs_writer_s *dvs_writer;
int dvs_writer_open(const char *path, int32_t format, dvs_writer *out);
int dvs_writer_append(dvs_writer writer, const dvs_event *events, int64_t count);
/* close(end_time_seconds) (:127-143); closing twice is a no-op. */
int dvs_writer_close(dvs_writer writer, double end_time_seconds);
/* Flushes buffered records to the file (for size accounting). */
int dvs_writer_flush(dvs_writer writer);
int dvs_writer_count(dvs_writer writer, int64_t *count);
int dvs_writer_destroy(dvs_writer writer);

/* Reads an events.h5 file back: x, y, t (microseconds), p into caller
 * buffers of `capacity` entries (any may be NULL); *count = events. */
int dvs_events_h5_read(const char *path, int16_t *x, int16_t *y, int64_t *t, int8_t *p, int64_t capacity,
                       int64_t *count);
/* Reads the ms_to_idx dataset of an events.h5 file. */
int dvs_events_h5_ms_to_idx(const char *path, int64_t *out, int64_t capacity, int64_t *count);

/* ---- Configuration (Config, include/dvs/scene_types.hpp:64-109,
 * src/scene_types.cpp) -------------------------------------------------- */
/* SceneConfig (scene_types.hpp:80-86). */
typedef struct {
  int32_t frame_count;
  float fps, imu_fps, cam_fps, time_resolution, blur_resolution;
  int32_t supersample_factor;
  float translation[3], rotation[3];
} dvs_scene_config;

typedef struct {
  dvs_camera camera; /* samples = scene.supersample_factor */
  dvs_physic physic;
  dvs_scene_config scene;
  dvs_hapke hapke;
} dvs_config;

/* Config::Config(path, err): two-level "section: / key: value" YAML,
 * sections intrinsic, physic, scene (required) and hapke; required keys,
 * optional physic keys over the PhysicConfig defaults, and the reference's
 * range checks.  Failures return DVS_E_ARGUMENT with the reference's
 * message. */
int dvs_config_load(const char *path, dvs_config *out);

/* ---- Scene bundle and trajectory files (src/scene_data.cpp) ------------
 * Every loader fails with DVS_E_ARGUMENT and the reference's message
 * ("failed to open <path>", "invalid f32 binary size for <path>", ...). */
/* load_mesh_data (scene_data.cpp:93-163): <dir>/mesh_vertices.bin and
 * mesh_normals.bin (float32, N x 3) and mesh_indices.bin (uint32, M x 3).
 * With any buffer NULL only the counts are returned (from the file sizes);
 * otherwise the buffers hold vertex_capacity vertices and normals and
 * triangle_capacity triangles, and indices and finiteness are checked. */
int dvs_io_mesh_load(const char *dir, float *vertices, float *normals, int64_t vertex_capacity, uint32_t *indices,
                     int64_t triangle_capacity, int64_t *vertex_count, int64_t *triangle_count);
/* load_frames_from_binary (scene_data.cpp:165-198): <dir>/frames.bin, 24
 * floats per frame {t, origin[3], camera_world row-major 4x4, light_dir[3],
 * intensity}; timestamp_seconds = t.  out == NULL returns the count only. */
int dvs_io_frames_binary(const char *dir, dvs_frame *out, int64_t capacity, int64_t *count);
/* load_frames_from_text (scene_data.cpp:238-283): non-blank lines
 * "frame_idx x y z qx qy qz qw lx ly lz L"; camera_world is the rotation
 * matrix of the quaternion as given (not normalised); timestamps NaN. */
int dvs_io_frames_text(const char *path, dvs_frame *out, int64_t capacity, int64_t *count);

/* ---- frames.h5 (Hdf5FrameWriter, include/dvs/frame_writer.hpp:12-77) ----
 * Datasets t (int64 microseconds, llround(time * 1e6)), images (uint8
 * N x H x W, render_float_to_png_bytes of the radiance) and depth (float32
 * N x H x W), extendable, one frame per chunk, byte shuffle + Deflate 1 on
 * images and depth.  The file must not exist.  append checks strictly
 * increasing finite times and finite radiance and depth (depth >= 0 or -1):
 * DVS_E_ARGUMENT with the reference's messages. */
typedef struct dvs_frames_s *dvs_frames;
int dvs_frames_open(const char *path, int32_t width, int32_t height, dvs_frames *out);
int dvs_frames_append(dvs_frames frames, double time, const float *radiance, const float *depth);
int dvs_frames_close(dvs_frames frames);
int dvs_frames_count(dvs_frames frames, int64_t *count);
int dvs_frames_destroy(dvs_frames frames);
/* Reads a frames.h5 back: the frame count and size always; with
 * capacity >= count and all buffers given, the three datasets too. */
int dvs_frames_h5_read(const char *path, int64_t *times, uint8_t *images, float *depth, int64_t capacity,
                       int32_t *width, int32_t *height, int64_t *count);

/* ---- Raster input (include/tools/loader.hpp, without GDAL) -------------
 * tools::loadImageryData (loader.hpp:31-100) over a minimal GeoTIFF reader:
 * band 1 as float32, NoData and non-finite samples as NaN, the rest through
 * the band's scale and offset, pixel-centre x / y coordinates from the
 * geotransform; rotated, sheared or degenerate transforms are rejected with
 * the reference's message.  Classic and BigTIFF, strips or tiles, no / LZW /
 * Deflate compression, horizontal and floating-point predictors, 8-64 bit
 * samples.  The projection string identifies the GeoKey directory (two
 * rasters share a CRS when theirs are equal). */
typedef struct dvs_raster_s *dvs_raster;
typedef struct {
  int32_t width, height;
  double transform[6];
  int32_t has_nodata, has_scale, has_offset;
  double nodata, scale, offset;
  int64_t missing;  /* NoData samples (NaN) */
  float center_raw;  /* the stored sample at (width/2, height/2), before NoData / scale / offset */
} dvs_raster_info;
int dvs_io_raster_load(const char *path, dvs_raster *out);
int dvs_io_raster_info(dvs_raster raster, dvs_raster_info *out);
/* data: width*height floats (row major); x_coords: width; y_coords: height
 * (any may be NULL). */
int dvs_io_raster_data(dvs_raster raster, float *data, float *x_coords, float *y_coords);
/* Copies up to capacity bytes of the projection key (NUL terminated when
 * room allows); *length receives its length. */
int dvs_io_raster_projection(dvs_raster raster, char *out, int64_t capacity, int64_t *length);
/* The value of GeoKey `key` (its SHORT value or its GeoDoubleParams
 * entry); DVS_E_ARGUMENT when the raster has none. */
int dvs_io_raster_geokey(dvs_raster raster, int32_t key, double *value);
int dvs_io_raster_free(dvs_raster raster);

/* ---- Frame images of bin/cam (apps/main_cam.cpp:79-96) ----------------- */
/* render_float_to_png_bytes (include/dvs/blenderize.hpp:50-69): exposure
 * x/100 clamped to [0,1], sRGB, optional hash dither, to_byte. */
int dvs_io_frame_bytes(const float *linear, int32_t width, int32_t height, float dither, uint8_t *out);
/* 8-bit greyscale PNG (images/<t>.png). */
int dvs_io_write_png(const char *path, const uint8_t *gray, int32_t width, int32_t height);
/* Uncompressed float32 TIFF (depth/<t>.tif: mean hit range, -1 on a miss). */
int dvs_io_write_tiff(const char *path, const float *data, int32_t width, int32_t height);

#ifdef __cplusplus
}
#endif
#endif
