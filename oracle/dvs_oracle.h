/*
 * CPU oracle for the DVS rendering hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, operation for operation, the reference's device
 * code (BugBubbles/DvsRender4DeepSpace) on the host in C++17 so that the HIP
 * product can be checked against it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker or the
 * CPU baseline; the product never links or calls it.
 *
 * Pinning (what checks each function independently of this restatement):
 *   bvh build / traversal   brute-force nearest/any hit over all triangles;
 *                           structural invariants; plane fixtures.
 *   sample_pixel            closed-form plane depth 10*sqrt(1+px^2+py^2),
 *                           +Y-up orientation, miss -> 0/-1, shadow origin side
 *                           test at km coordinates (tests/test_render.cu:67-106).
 *   hapke                   limiting cases, opposition > 30 deg, I/F inversion
 *                           round trip, textbook H-function comparison.
 *   circuit                 closed-form crossing (tests/test_sensor.cu:40-55),
 *                           RK4 integration of C du/dt = r - e^u.
 *   readout                 EventReadout(100,1e-3,2) known answer
 *                           (tests/test_sensor.cu:176-177); serial FIFO.
 *   philox4x64-10           numpy.random.Philox outputs and the Random123
 *                           published answer for the zero counter/key.
 *   trajectory              constant-trajectory exactness
 *                           (tests/test_sensor.cu:32-36), local mount and
 *                           seconds (tests/test_render.cu:115-130), linear
 *                           and monotone data, Bezier/Fritsch-Carlson closed
 *                           forms on 2-3 keys, constant-rate rotation about
 *                           one axis, quaternion <-> matrix round trips.
 *   poisson / normal        moments (tests/test_sensor.cu:57-65).  The
 *                           reference draws from cuRAND (closed library), so
 *                           noise-on event streams are "parity unpinned" at
 *                           the stream level; only statistics are compared.
 *
 * Floating point: built with -ffp-contract=off and IEEE division/sqrt.  The
 * one deliberate departure from the reference's CUDA source is rsqrtf in the
 * quaternion renormalisation (device_optics.cuh:26), an approximate CUDA
 * intrinsic, restated as 1/sqrtf.
 */
#ifndef DVS_ORACLE_H
#define DVS_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Reference BVHNode<float3> (include/dvs/scene_types.hpp:29-35): 40 bytes. */
typedef struct {
  float min[3], max[3];
  int32_t left, right, start, count;
} orc_node;

/* HapkeConfig (include/dvs/scene_types.hpp:89-99), same field order. */
typedef struct {
  float b, c, hS, hC, thetap, w, B_S0, B_C0, phi;
} orc_hapke;

/* CameraConfig + supersample_factor (scene_types.hpp:59-62,85). */
typedef struct {
  int32_t width, height;
  float fx, fy, cx, cy;
  int32_t samples;
} orc_camera;

/* OpticalState (include/dvs/optical_sensor.hpp:15-19), packed floats.
 * rotation is (x,y,z,w). */
typedef struct {
  float position[3], velocity[3], light[3], light_rate[3];
  float rotation[4];
  float intensity, intensity_rate;
} orc_state;

/* OpticalInterval (optical_sensor.hpp:20). */
typedef struct {
  orc_state start, end;
  double t0, t1;
} orc_interval;

/* PhysicConfig (scene_types.hpp:64-79), same field order. */
typedef struct {
  float threshold_pos, threshold_neg, refractory_time;
  float readout_bandwidth, readout_latency;
  float threshold_std, threshold_min;
  float wavelength_nm, quantum_efficiency, pixel_pitch_um;
  float fill_factor, f_number, lens_transmission;
  float radiance_scale;
  float temperature_k, subthreshold_kappa;
  float receptor_capacitance_f;
  float reset_capacitance_f, dark_current_a;
  float reset_leakage_a, circuit_step;
  int32_t photon_shot_noise, reset_thermal_noise;
  int32_t random_seed, max_events_per_step;
  int32_t readout_queue_capacity;
} orc_physic;

/* PixelParameters / PixelState / PixelEvent (include/dvs/pixel_model.hpp:7-17). */
typedef struct {
  double threshold_pos, threshold_neg, charge_scale, refractory_time;
  double reference_leak, reset_noise_std;
} orc_pixel_params;
typedef struct {
  double receptor_log, reference_log, refractory_until;
  double threshold_pos, threshold_neg;
  int32_t reset_pending;
  uint32_t pad;
} orc_pixel_state;
typedef struct {
  double time;
  int32_t polarity;
  int32_t pad;
} orc_pixel_event;

/* Event (include/dvs/dvs_base.hpp:9-18): packed 17 bytes. */
#pragma pack(push, 1)
typedef struct {
  double timestamp;
  int16_t x, y;
  int8_t polarity;
  float depth;
} orc_event;
#pragma pack(pop)

/* Derived model parameters (src/cuda/event_sensor.cu:151-176). */
typedef struct {
  double electrons_per_radiance, dark_electrons, integration_step;
  int32_t shot_noise;
  int32_t pad;
} orc_photo_params;

/* ---- BVH (src/bvh/build.cpp, src/bvh/bining.cpp) ---- */
/* Builds into nodes[0..2*nt) and reorders indices in place; returns root or -1. */
int orc_bvh_build(int max_leaf, float sah_node_cost, float sah_primitive_cost,
                  const float *vertices, int nv, uint32_t *indices, int nt,
                  orc_node *nodes, int *node_count);
/* Returns 0 when valid, else a negative code (build.cpp:188-242). */
int orc_bvh_validate(const orc_node *nodes, int node_count, int root);

/* ---- Scene ---- */
void *orc_scene_create(const orc_node *nodes, int node_count, int root,
                       const float *vertices, const float *normals, int nv,
                       const uint32_t *indices, int nt, const float *albedos,
                       orc_camera camera, orc_hapke hapke);
void orc_scene_destroy(void *scene);

/* trace_nearest (device_ray.hpp:136-158); hit = 1, 0 or -1 (overflow). */
int orc_trace_nearest(const void *scene, const float origin[3], const float dir[3],
                      int *tri, float *t, float *u, float *v);
/* trace_shadow_ray_any_hit (device_ray.hpp:81-135). */
int orc_trace_shadow(const void *scene, const float origin[3], const float dir[3],
                     int self_triangle, float t_min, float t_max);
/* Counts node-pair visits, leaf visits and triangle tests of one nearest query. */
int orc_trace_nearest_counted(const void *scene, const float origin[3], const float dir[3],
                              int64_t counts[3]);
void orc_shadow_origin(const void *scene, int tri, float u, float v,
                       const float light[3], float out[3]);

/* sample_pixel (device_optics.cuh:73-94): radiance, depth, valid. */
int orc_sample_pixel(const void *scene, int pixel, const orc_state *pose,
                     double *radiance, float *depth);
/* optical_state_at (device_optics.cuh:13-26). */
void orc_optical_state_at(const orc_interval *iv, double t, orc_state *out);
/* Frame render = sample_warps + render_main (optical_sampler.cu:8-23,
 * cam_render.cu:16-28): image and depth per pixel, OpenMP over pixels.
 * Returns 0, or 2 on an invalid sample. */
int orc_render(const void *scene, const orc_state *pose, float *image,
               float *depth, double *radiance, int threads);

/* ---- Hapke (include/dvs/hapke.hpp) ---- */
float orc_hapke_reflectance(float cosi, float cose, float cosg, float albedo, orc_hapke p);
float orc_hapke_directional(const float n[3], const float l[3], const float v[3],
                            float albedo, orc_hapke p);
float orc_invert_hapke_if(float inc, float emi, float phase, float target, orc_hapke p);

/* ---- Circuit (include/dvs/pixel_model.hpp) ---- */
double orc_receptor_log_at(double u, double rate, double dt, double charge);
/* advance_pixel with a normal() that returns `fixed_normal`. */
int orc_advance_pixel(orc_pixel_state *s, const orc_pixel_params *p, double rate,
                      double t0, double t1, orc_pixel_event *out, int capacity,
                      double fixed_normal);
void orc_parameters(const orc_physic *c, int physic_effects, orc_pixel_params *p,
                    orc_photo_params *photo);

/* ---- Noise (this build's Philox design) ---- */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
/* Draw `n` Poisson(lambda) variates for (seed, pixel) from block counter 0. */
void orc_poisson_series(uint64_t seed, uint64_t pixel, double lambda, int n, double *out);
/* Shot-noise electron counts of cells serial0 .. serial0 + n - 1 of one
 * pixel at mean lambda, as the photocurrent draws them. */
void orc_shot_counts(uint64_t seed, uint64_t pixel, double lambda, uint64_t serial0, int n, double *out);
/* The Poisson inversion of one uniform (0 < lambda < 10). */
double orc_poisson_inversion(double lambda, double u);
void orc_normal_series(uint64_t seed, uint64_t pixel, int n, double *out);

/* ---- Sensor (src/cuda/event_sensor.cu) ---- */
void *orc_sensor_create(int width, int height, const orc_physic *config,
                        int physic_effects, int aer_delay);
void orc_sensor_destroy(void *sensor);
int orc_sensor_initialize(void *sensor, const void *scene, const orc_state *pose,
                          double timestamp, int threads);
/* process_many: raw comparator events (pre-readout) in pixel-major order are
 * written to raw (may be NULL); readout-delivered events to ready.  Counts are
 * returned through raw_count / ready_count; capacities bound the copies.
 * Returns 0 ok, 1 capacity, 2 invalid optical, 3 invalid argument. */
int orc_sensor_process(void *sensor, const void *scene, const orc_interval *intervals,
                       int n, orc_event *raw, int64_t raw_capacity, int64_t *raw_count,
                       orc_event *ready, int64_t ready_capacity, int64_t *ready_count,
                       int threads);
/* OpticalRefinementStats (include/dvs/optical_sampler.hpp:16). */
typedef struct {
  uint64_t traced_pixel_samples;
  int32_t fallback_pixels, maximum_knots, exact_segments, native_cell_queries;
} orc_refinement_stats;
/* ReceptorProbeResult (include/dvs/event_sensor.hpp:39). */
typedef struct {
  double maximum_log_voltage_error, p99_log_voltage_error, mean_log_voltage_error;
  int32_t pixels, circuit_cells;
} orc_receptor_probe;
/* process_adaptive (src/cuda/event_sensor.cu:306-355) with OpticalSampler
 * refine / prepare_exact_cells / sample_cells (src/cuda/optical_sampler.cu)
 * at knot and exact-cell capacity `capacity`.  *accepted is 0 when the
 * convergence gate rejected the window (no state changed).  Events carry
 * depth -1.  Returns as orc_sensor_process. */
int orc_sensor_process_adaptive(void *sensor, const void *scene, const orc_interval *intervals,
                                int n, double tolerance, int check_convergence, int capacity,
                                int *accepted, orc_refinement_stats *stats,
                                orc_receptor_probe *probe, orc_event *raw, int64_t raw_capacity,
                                int64_t *raw_count, orc_event *ready, int64_t ready_capacity,
                                int64_t *ready_count, int threads);
int orc_sensor_drain(void *sensor, orc_event *ready, int64_t capacity, int64_t *count);
/* OpticalSampler::verify (optical_sampler.cu:86-103,195-215) on the last
 * adaptive window's knots: out = {max, p99, mean} log error, iout = {probes,
 * over tolerance, fallback probes}. */
int orc_sensor_verify_optics(void *sensor, const void *scene, double tolerance, double out[3], int iout[3],
                             int threads);
int64_t orc_sensor_dropped(const void *sensor);
int64_t orc_sensor_peak_queue(const void *sensor);
int orc_sensor_states(const void *sensor, orc_pixel_state *out);
int64_t orc_sensor_shadow_rays(const void *sensor);

/* ---- Readout FIFO (src/event_readout.cpp) ---- */
void *orc_readout_create(double bandwidth, double latency, int64_t capacity);
void orc_readout_destroy(void *r);
int64_t orc_readout_submit(void *r, const orc_event *events, int64_t n, double now,
                           orc_event *ready, int64_t capacity);
int64_t orc_readout_drain(void *r, orc_event *ready, int64_t capacity);
int64_t orc_readout_dropped(const void *r);
int64_t orc_readout_peak(const void *r);

/* ---- Camera trajectory (src/trajectory.cpp) ---- */
/* FrameData<Vec3, Mat3> (include/dvs/scene_types.hpp:43-50); camera_world
 * column major; NaN timestamp -> frame_index / fps. */
typedef struct {
  int32_t frame_index;
  double timestamp_seconds;
  float camera_origin[3];
  float camera_world[9];
  float light_dir[3];
  float light_intensity;
} orc_frame;
/* CameraTrajectoryState (include/dvs/trajectory.hpp:44-60); orientation
 * column major, quaternion (x, y, z, w). */
typedef struct {
  float timestamp;
  float position[3];
  float orientation[9];
  float quaternion[4];
  float linear_velocity[3], angular_velocity[3];
  float light_direction[3], light_direction_rate[3];
  float light_intensity, light_intensity_rate;
} orc_trajectory_state;
/* NULL with the reference's std::invalid_argument message in `error`. */
void *orc_trajectory_create(const orc_frame *frames, int count, float fps, const float translation[3],
                            const float rotation[3], char *error, int error_len);
void orc_trajectory_destroy(void *t);
void orc_trajectory_range(const void *t, float *start, float *end, int *size);
void orc_trajectory_evaluate(const void *t, const float *times, int64_t n, orc_trajectory_state *out);

/* ---- Terrain from rasters (src/tools/build_from_raster.cu) ---- */
int orc_raster_mesh(const float *dem, const float *albedo, const float *xs, const float *ys, int width, int height,
                    float *vertices, float *normals, float *albedos, int64_t *vertex_count, uint32_t *indices,
                    int64_t *triangle_count);
void orc_raster_align(const float *ref_x, const float *ref_y, int width, int height, const float *source,
                      const float *source_x, const float *source_y, int source_width, int source_height, float *out);

/* ---- Image encode (include/dvs/blenderize.hpp) ---- */
void orc_render_to_bytes(const float *linear, int width, int height, float dither,
                         uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
