/*
 * dvs_hip.h — C ABI of the MI355X (gfx950) DVS rendering library.
 *
 * Plain C types, plain pointers and sizes; no HIP or C++ types cross this
 * boundary.  Every entry point returns an int status:
 *   DVS_OK 0, DVS_E_CAPACITY 1 (pixel event capacity exceeded),
 *   DVS_E_OPTICAL 2 (invalid optical sample / BVH traversal overflow),
 *   DVS_E_ARGUMENT 3, DVS_E_HIP 4, DVS_E_STATE 5,
 * and dvs_last_error() returns the message of the calling thread's last
 * failure, worded as the reference's exceptions.  The C++ layer
 * (include/dvs_render.hpp and include/dvs/raycast.hpp in this repository,
 * over dvs_render_amd.hpp) rethrows them as the reference does.
 *
 * Each declaration names the reference interface it replaces
 * (BugBubbles/DvsRender4DeepSpace, paths relative to its root).
 */
#ifndef DVS_HIP_H
#define DVS_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  DVS_OK = 0,
  DVS_E_CAPACITY = 1,
  DVS_E_OPTICAL = 2,
  DVS_E_ARGUMENT = 3,
  DVS_E_HIP = 4,
  DVS_E_STATE = 5
};

/* BVHNode<float3> (include/dvs/scene_types.hpp:29-35): 40 bytes, the node
 * format BVHBuilder::build emits and DeviceScene consumes. */
typedef struct {
  float min[3], max[3];
  int32_t left, right, start, count;
} dvs_bvh_node;

/* HapkeConfig (include/dvs/scene_types.hpp:89-99), same field order. */
typedef struct {
  float b, c, hS, hC, thetap, w, B_S0, B_C0, phi;
} dvs_hapke;

/* CameraConfig (scene_types.hpp:59-62) plus scene.supersample_factor. */
typedef struct {
  int32_t width, height;
  float fx, fy, cx, cy;
  int32_t samples;
} dvs_camera;

/* OpticalState (include/dvs/optical_sensor.hpp:15-19); rotation is the
 * camera-to-world quaternion (x,y,z,w). */
typedef struct {
  float position[3], velocity[3], light[3], light_rate[3];
  float rotation[4];
  float intensity, intensity_rate;
} dvs_optical_state;

/* OpticalInterval (optical_sensor.hpp:20). */
typedef struct {
  dvs_optical_state start, end;
  double t0, t1;
} dvs_optical_interval;

/* PhysicConfig (scene_types.hpp:64-79), same field order and defaults. */
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
} dvs_physic;

/* Event (include/dvs/dvs_base.hpp:9-18): packed, 17 bytes. */
#pragma pack(push, 1)
typedef struct {
  double timestamp;
  int16_t x, y;
  int8_t polarity;
  float depth;
} dvs_event;
#pragma pack(pop)

/* PixelState (include/dvs/pixel_model.hpp:11-15) as held on the device,
 * plus the noise counters: reset_block is the reset stream's next Philox
 * block, shot_block the number of circuit cells integrated so far (the
 * serial that keys each cell's shot-noise draw).  receptor_inverse is
 * exp(-receptor_log): without reference leakage the circuit integrates this
 * variable and carries it exactly from step to step (receptor_log is then
 * derived from it); with leakage it is derived from receptor_log. */
typedef struct {
  double receptor_log, reference_log, refractory_until;
  double threshold_pos, threshold_neg;
  int32_t reset_pending;
  uint32_t reset_block;
  uint64_t shot_block;
  double receptor_inverse;
} dvs_pixel_state;

/* PixelParameters (include/dvs/pixel_model.hpp:7-10) and PixelEvent
 * (:15-17): the circuit constants and one comparator crossing. */
typedef struct {
  double threshold_pos, threshold_neg, charge_scale, refractory_time;
  double reference_leak, reset_noise_std;
} dvs_pixel_params;
typedef struct {
  double time;
  int32_t polarity;
  int32_t pad;
} dvs_crossing;

/* OpticalProbeResult (include/dvs/optical_sampler.hpp:15): knot
 * interpolation against native traces over 1024 pixels x 17 times. */
typedef struct {
  double maximum_log_error, p99_log_error, mean_log_error;
  int32_t probes, over_tolerance, exact_fallback_probes;
} dvs_optical_probe;

/* Device timings and work counters of one sensor or scene handle. */
typedef struct {
  double trace_ms, circuit_ms, events_ms, total_ms;
  double trace_launches;
  uint64_t primary_rays, shadow_rays, event_rays;
  uint64_t pixel_cells, comparator_events, delivered_events;
  /* rays the native path would trace for the same accepted windows, and the
   * adaptive windows accepted / rejected by the convergence gate */
  uint64_t native_equivalent_rays, adaptive_windows, rejected_windows;
  /* adaptive steps: photocurrent kernels (knot interpolation, fallback
   * tiles, shot noise) apart from the serial circuit in circuit_ms */
  double rates_ms, circuit_launches;
  /* adaptive refinement's ray launches, timed on their stream: the five
   * initial probes per pixel (probe) and the bisections' midpoints (query);
   * pixel samples traced by each (ss*ss primary rays apiece) */
  double probe_ms, query_ms, probe_launches, query_launches;
  uint64_t probe_samples, query_samples;
  /* readout FIFO on the device: segment-steps run, and those the scan form
   * handed to the serial recurrence */
  uint64_t readout_tasks, readout_serial_tasks;
  /* adaptive windows whose setup was queued ahead, at the end of the window
   * before them (dvs_sensor_expect_adaptive_times) */
  uint64_t windows_set_up_ahead;
} dvs_stats;

/* OpticalRefinementStats (include/dvs/optical_sampler.hpp:16) of the last
 * adaptive window. */
typedef struct {
  uint64_t traced_pixel_samples;
  int32_t fallback_pixels, maximum_knots, exact_segments, native_cell_queries;
} dvs_refinement_stats;

/* ReceptorProbeResult (include/dvs/event_sensor.hpp:39) of the last checked
 * adaptive window. */
typedef struct {
  double maximum_log_voltage_error, p99_log_voltage_error, mean_log_voltage_error;
  int32_t pixels, circuit_cells;
} dvs_receptor_probe;

typedef struct dvs_scene_s *dvs_scene;
typedef struct dvs_sensor_s *dvs_sensor;

/* Message of the calling thread's last failure ("" if none). */
const char *dvs_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int dvs_device_count(void);

/* ---- BVH ------------------------------------------------------------
 * BVHBuilder::build(params, nodes, indices, vertices) (include/bvh/build.h:19-24,
 * src/bvh/build.cpp:27-76): host SAH build; `indices` (3*nt uint32) is
 * reordered into leaf order in place; nodes must hold 2*nt entries.
 * Writes the node count and the root index (last node). */
int dvs_bvh_build(int max_leaf_size, float sah_node_cost, float sah_primitive_cost,
                  const float *vertices, int64_t vertex_count, uint32_t *indices,
                  int64_t triangle_count, dvs_bvh_node *nodes, int64_t *node_count,
                  int32_t *root);
/* BVHBuilder::validate (src/bvh/build.cpp:188-242). */
int dvs_bvh_validate(const dvs_bvh_node *nodes, int64_t node_count, int32_t root);

/* ---- Scene -----------------------------------------------------------
 * DeviceScene(const SceneBundle&, const Config&, int root)
 * (include/dvs/device_scene.hpp:5-15, src/cuda/device_scene.cu:15-37):
 * validates the indexed mesh and BVH and uploads it as the aligned SoA
 * layout the gfx950 kernels traverse.  albedos may be NULL. */
int dvs_scene_create(const dvs_bvh_node *nodes, int64_t node_count, int32_t root,
                     const float *vertices, const float *normals, int64_t vertex_count,
                     const uint32_t *indices, int64_t triangle_count, const float *albedos,
                     const dvs_camera *camera, const dvs_hapke *hapke, dvs_scene *out);
/* Same with the device leaf size chosen: every reference leaf larger than
 * device_leaf_size triangles is subdivided into a device-only subtree
 * (0 keeps the reference leaves).  Results are those of the reference's leaf
 * scan: sub-boxes are inflated beyond rounding and equal-distance hits inside
 * one reference leaf resolve to the lowest reference index.
 * dvs_scene_create uses DVS_DEFAULT_DEVICE_LEAF_SIZE. */
#define DVS_DEFAULT_DEVICE_LEAF_SIZE 4
int dvs_scene_create_ex(const dvs_bvh_node *nodes, int64_t node_count, int32_t root,
                        const float *vertices, const float *normals, int64_t vertex_count,
                        const uint32_t *indices, int64_t triangle_count, const float *albedos,
                        const dvs_camera *camera, const dvs_hapke *hapke, int32_t device_leaf_size,
                        dvs_scene *out);
/* A second camera on the same resident geometry (no reference counterpart:
 * the reference builds one DeviceScene per camera configuration).  The view
 * shares the uploaded BVH and mesh; hapke may be NULL to keep the base's. */
int dvs_scene_view(dvs_scene base, const dvs_camera *camera, const dvs_hapke *hapke, dvs_scene *out);
/* Traversal form of the scene's frame, depth and ray launches (no reference
 * counterpart; results do not depend on it): 1 (the default) the
 * wave-packet traversal, 0 one traversal per lane with its stack in LDS.
 * A tree whose reference subtrees do not cover ordered triangle index
 * ranges (not built by dvs_bvh_build) always takes 0; *effective (may be
 * NULL) receives the form in use.  Sensors and batches choose theirs by
 * dvs_tuning::traversal. */
int dvs_scene_set_traversal(dvs_scene scene, int32_t form, int32_t *effective);
/* ~DeviceScene (device_scene.cu:38-43).  The scene is reference counted:
 * sensors created on it keep it alive. */
int dvs_scene_destroy(dvs_scene scene);
/* The camera of a scene view (OpticalScene::width, height, fx, fy, cx, cy,
 * samples, include/dvs/optical_sensor.hpp:5-13). */
int dvs_scene_camera(dvs_scene scene, dvs_camera *out);
/* Tree depth and SoA sizes of an uploaded scene (diagnostics). */
int dvs_scene_info(dvs_scene scene, int32_t *max_depth, int64_t *internal_nodes,
                   int64_t *leaves, int64_t *device_bytes);

/* ---- Frame / depth ---------------------------------------------------
 * OpticalSampler::evaluate + download (src/cuda/optical_sampler.cu:154-172) as
 * used by CamRayCaster::render_main (src/cuda/cam_render.cu:16-28): renders
 * `pose_count` poses; image/depth hold width*height*pose_count floats
 * (host memory).  radiance (double, may be NULL) receives the unrounded mean. */
int dvs_frame_render(dvs_scene scene, const dvs_optical_state *poses, int32_t pose_count,
                     float *image, float *depth, double *radiance);
/* Same on device buffers (no host copies); for timing. */
int dvs_frame_render_device(dvs_scene scene, const dvs_optical_state *device_poses,
                            int32_t pose_count, float *device_image, float *device_depth);
/* No reference counterpart.  dvs_frame_render_device for callers that cannot
 * allocate device memory (code that must not call hipMalloc, such as tight
 * serving loops and harnesses that forbid it): it allocates nothing.  Its
 * pose table, status word and ray counter live in a fixed device array of
 * the library (196,864 bytes per GPU, held until the process ends), so
 * pose_count is at most 4096 (larger requests fail with DVS_E_ARGUMENT).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Synchronous like dvs_frame_render_device; it adds rays and launches
 * to the scene's statistics but no trace time. */
int dvs_frame_render_device_noalloc(dvs_scene scene, const dvs_optical_state *device_poses,
                                    int32_t pose_count, float *device_image, float *device_depth);
/* Motion-blur exposures: the render_main loop of apps/main_cam_blur.cpp:64-83
 * and apps/main_cam_blur_windows.cpp:64-81 with every sub-exposure cell of
 * an exposure rendered in one multi-pose launch.  Exposure e sums the cells
 * c in [cell_offsets[e], cell_offsets[e+1]) in cell order:
 *   wide = 0: acc = float(acc + float(frame_c * float(weights[c])))
 *             (bin/cam_blur's cv::Mat sum, main_cam_blur.cpp:76);
 *   wide = 1: acc = acc + weights[c] * double(frame_c) in double, cast to
 *             float at the end (main_cam_blur_windows.cpp:77,81);
 * where frame_c is the render_main image of poses[c].  cell_offsets holds
 * exposure_count + 1 non-decreasing entries starting at 0; image receives
 * exposure_count * width * height floats (host memory). */
int dvs_frame_blur(dvs_scene scene, const dvs_optical_state *poses, const double *weights,
                   const int32_t *cell_offsets, int32_t exposure_count, int32_t wide, float *image);
int dvs_scene_stats(dvs_scene scene, dvs_stats *out);
/* Traversal work of every sub-ray of one frame, for the roofline's
 * algorithmic bytes per ray (no reference counterpart): totals[0..7] =
 * primary rays, their node-pair visits, leaf visits and triangle tests, then
 * shadow rays, their node-pair visits, leaf visits and triangle tests. */
int dvs_frame_traversal_stats(dvs_scene scene, const dvs_optical_state *pose, uint64_t totals[8]);

/* ---- Event sensor ----------------------------------------------------
 * EventSensor(width, height, PhysicConfig) (src/cuda/event_sensor.cu:244-260).
 * physic_effects / aer_delay are the reference's compile-time
 * ENABLE_PHYSIC_EFFECTS / ENABLE_AER_DELAY switches as runtime flags. */
int dvs_sensor_create(dvs_scene scene, const dvs_physic *physic, int32_t physic_effects,
                      int32_t aer_delay, dvs_sensor *out);
int dvs_sensor_destroy(dvs_sensor sensor);
/* EventSensor::initialize (event_sensor.cu:262-270). */
int dvs_sensor_initialize(dvs_sensor sensor, const dvs_optical_state *pose, double timestamp);
/* EventSensor::process_many (event_sensor.cu:274-305): integrates the
 * contiguous intervals, then runs the readout FIFO; the delivered events are
 * held until dvs_sensor_take.  *ready_count receives their number. */
int dvs_sensor_process(dvs_sensor sensor, const dvs_optical_interval *intervals,
                       int32_t interval_count, int64_t *ready_count);
/* EventSensor::process_adaptive (event_sensor.cu:306-355) through
 * DvsRayCaster::advance_adaptive (include/dvs/raycast.hpp:40-43): the
 * intervals form one window; per-pixel optical refinement
 * (OpticalSampler::refine, src/cuda/optical_sampler.cu:39-84) to the log
 * tolerance, local exact cells and whole-pixel fallback, then the circuit.
 * With check_convergence the receptor check (verify_receptor, :357-374)
 * runs first; *accepted = 0 means it rejected the window and nothing
 * changed.  Events of this path carry depth -1. */
int dvs_sensor_process_adaptive(dvs_sensor sensor, const dvs_optical_interval *intervals,
                                int32_t interval_count, double tolerance, int32_t check_convergence,
                                int32_t *accepted, int64_t *ready_count);
/* DvsRayCaster::adaptive_cache_capacity (raycast.hpp:36-39): knot and exact
 * cell capacity per pixel (optical_sampler.cu:141-150). */
/* EventSensor::enable_numerical_diagnostics (include/dvs/event_sensor.hpp:61):
 * every later adaptive window also measures its receptor check (1024
 * hashed pixels, native photocurrent against the knot interpolation), for
 * verify_receptor; only windows processed with check_convergence use it to
 * reject. */
int dvs_sensor_enable_diagnostics(dvs_sensor sensor);
/* EventSensor::verify_receptor (event_sensor.hpp:62, src/cuda/event_sensor.cu:
 * 357-374) of the last adaptive window `intervals`: DVS_E_STATE "enable
 * numerical diagnostics before processing" without diagnostics. */
int dvs_sensor_verify_receptor(dvs_sensor sensor, const dvs_optical_interval *intervals, int32_t count,
                               dvs_receptor_probe *out);
/* EventSensor::verify_optics (event_sensor.hpp:57-59) = OpticalSampler::verify
 * (src/cuda/optical_sampler.cu:86-103,195-215) on the last adaptive window
 * `intervals`, right after it (DVS_E_STATE "refine before validating optical
 * knots" otherwise): log errors of the knot-interpolated photocurrent against
 * native traces of 1024 hashed pixels at 17 times. */
int dvs_sensor_verify_optics(dvs_sensor sensor, const dvs_optical_interval *intervals, int32_t count,
                             double tolerance, dvs_optical_probe *out);
int dvs_sensor_adaptive_capacity(dvs_sensor sensor, int32_t *capacity);
/* DvsRayCaster::refinement_statistics (raycast.hpp:55). */
int dvs_sensor_refinement_stats(dvs_sensor sensor, dvs_refinement_stats *out);
/* EventSensor::last_convergence (include/dvs/event_sensor.hpp:54). */
int dvs_sensor_last_convergence(dvs_sensor sensor, dvs_receptor_probe *out);
/* Copies and clears the held delivered events (time-ordered). */
int dvs_sensor_take(dvs_sensor sensor, dvs_event *out, int64_t capacity, int64_t *count);
/* Comparator events of the last process call before the readout FIFO, in
 * (t,y,x) order (tests and diagnostics). */
int dvs_sensor_raw_events(dvs_sensor sensor, dvs_event *out, int64_t capacity, int64_t *count);
/* EventSensor::drain (event_sensor.hpp:71): moves the FIFO backlog to the
 * held delivered events. */
int dvs_sensor_drain(dvs_sensor sensor, int64_t *ready_count);
/* EventSensor::dropped / peak_queue (event_sensor.hpp:72-73). */
int dvs_sensor_readout_stats(dvs_sensor sensor, int64_t *dropped, int64_t *peak_queue);
/* Copies the device pixel states (width*height entries). */
int dvs_sensor_states(dvs_sensor sensor, dvs_pixel_state *out);
int dvs_sensor_stats(dvs_sensor sensor, dvs_stats *out);
int dvs_sensor_reset_stats(dvs_sensor sensor);

/* ---- EventReadout (include/dvs/event_sensor.hpp:24-37) on its own ------
 * The bounded FIFO of src/event_readout.cpp:6-34, host side: submit sorts
 * the events stably by (t, y, x) and queues them (completion = max(available,
 * t) + 1 / bandwidth, delivered at completion + latency, dropped while
 * `capacity` events are pending; bandwidth <= 0 delivers at once), then
 * delivers every event ready by `now`.  Delivered events accumulate until
 * dvs_readout_take (count: the number waiting). */
typedef struct dvs_readout_s *dvs_readout;
int dvs_readout_create(double bandwidth, double latency, int64_t capacity, dvs_readout *out);
int dvs_readout_destroy(dvs_readout readout);
int dvs_readout_submit(dvs_readout readout, const dvs_event *events, int64_t count, double now,
                       int64_t *ready_count);
int dvs_readout_drain(dvs_readout readout, int64_t *ready_count);
int dvs_readout_take(dvs_readout readout, dvs_event *out, int64_t capacity, int64_t *count);
int dvs_readout_stats(dvs_readout readout, int64_t *dropped, int64_t *peak_queue);

/* ---- Multi-segment engine (batched and served forms) -----------------
 * No reference counterpart: many EventSensors ("segments"), each with its own
 * scene view, circuit state, Philox key and readout FIFO, advance together.
 * dvs_batch_step takes up to horizons_per_step queued horizons of every
 * segment and runs them in shared launches (optionally replayed from a
 * captured HIP graph while the step shape repeats).  A segment's events are
 * exactly those of a dvs_sensor fed the same intervals in the same calls. */
typedef struct dvs_batch_s *dvs_batch;
typedef struct dvs_trajectory_s *dvs_trajectory;
int dvs_batch_create(int32_t horizons_per_step, int32_t use_graph, dvs_batch *out);
int dvs_batch_destroy(dvs_batch batch);
/* Adds and initialises a segment (EventSensor ctor + initialize). */
int dvs_batch_add(dvs_batch batch, dvs_scene scene, const dvs_physic *physic, int32_t physic_effects,
                  int32_t aer_delay, const dvs_optical_state *initial_pose, double t0, int32_t *segment);
/* Queues contiguous intervals of a segment. */
int dvs_batch_feed(dvs_batch batch, int32_t segment, const dvs_optical_interval *intervals, int64_t count);
/* Queues the n - 1 intervals between consecutive times (n >= 2, times[0]
 * the end of what the segment has queued so far, strictly increasing) whose
 * states the engine evaluates from `trajectory` on the device, in the step
 * that integrates them: the states dvs_trajectory_intervals gives for the
 * same times (apps/main_dvs.cpp:103-107 builds each window's intervals this
 * way on the host).  The trajectory must outlive the segment's queued
 * intervals.  No reference counterpart: the reference feeds host intervals. */
int dvs_batch_feed_trajectory(dvs_batch batch, int32_t segment, dvs_trajectory trajectory, const double *times,
                              int64_t n);
/* dvs_sensor_process_adaptive over the n - 1 intervals between consecutive
 * times (times[0] the end of what the sensor has integrated, strictly
 * increasing) whose states the sensor evaluates from `trajectory` on the
 * device inside the window's launch sequence: the window bin/dvs builds
 * (apps/main_dvs.cpp:103-116) without the host evaluating, downloading and
 * uploading its intervals.  The same events as dvs_sensor_process_adaptive
 * fed dvs_trajectory_intervals of the same times.  No reference
 * counterpart. */
int dvs_sensor_process_adaptive_times(dvs_sensor sensor, dvs_trajectory trajectory, const double *times, int32_t n,
                                      double tolerance, int32_t check_convergence, int32_t *accepted,
                                      int64_t *ready_count);
/* The window the next dvs_sensor_process_adaptive_times call on this sensor
 * will most likely run (its times, trajectory and tolerance; n = 0 forgets
 * it).  When the current call's window is accepted, the expected window's
 * setup (its horizon states, cell tables and the five initial probes of
 * every pixel, none of which reads pixel state) is queued on the device
 * before that call returns, so the device works on it while the host turns
 * to the next call.  The next call uses it when its window is the expected
 * one and the sensor has done nothing else in between; otherwise it is
 * dropped and the window is set up as usual.  Events and pixel states are
 * the same either way.  The expectation lasts until the end of the next
 * adaptive-times call.  With diagnostics enabled nothing is queued ahead;
 * after a setup was queued, dvs_sensor_verify_optics and
 * dvs_sensor_verify_receptor report no window to verify.  No reference
 * counterpart: it is for window loops such as bin/dvs's
 * (apps/main_dvs.cpp:96-116), whose next window is known before the current
 * one ends unless the current one is rejected. */
int dvs_sensor_expect_adaptive_times(dvs_sensor sensor, dvs_trajectory trajectory, const double *times, int32_t n,
                                     double tolerance);
/* Launch shapes of the adaptive solver's kernels, for autotuning; results
 * do not depend on them.  No reference counterpart.
 *   rate_cells     circuit cells each photocurrent thread draws in one run
 *                  (a multiple of 8 in [8, 1024]; 128 by default);
 *   trace_threads  lanes of a refinement trace block, one sub-ray each (64,
 *                  128, the default, or 256; 256 where a pixel's sub-rays
 *                  outnumber the block);
 *   traversal      1 (the default): the wave-packet traversal, the 64 lanes
 *                  of a wave walking the tree together; 0: one traversal
 *                  per lane with its stack in LDS.  Scenes whose reference
 *                  subtrees do not cover ordered index ranges (not built by
 *                  dvs_bvh_build) always take 0. */
typedef struct dvs_tuning {
  int32_t rate_cells;
  int32_t trace_threads;
  int32_t traversal;
} dvs_tuning;
int dvs_batch_set_tuning(dvs_batch batch, const dvs_tuning *tuning);
int dvs_batch_get_tuning(dvs_batch batch, dvs_tuning *tuning);
/* The same for a sensor (a one-segment engine). */
int dvs_sensor_set_tuning(dvs_sensor sensor, const dvs_tuning *tuning);
/* One step; *pending_segments = segments with queued horizons left. */
int dvs_batch_step(dvs_batch batch, int32_t *pending_segments);
int dvs_batch_take(dvs_batch batch, int32_t segment, dvs_event *out, int64_t capacity, int64_t *count);
int dvs_batch_drain(dvs_batch batch, int32_t segment, int64_t *ready_count);
int dvs_batch_segment_info(dvs_batch batch, int32_t segment, int64_t *queued_horizons, int64_t *done_horizons,
                           int64_t *events, int64_t *device_bytes);
int dvs_batch_remove(dvs_batch batch, int32_t segment);
/* Adaptive windows for a segment, with the host driver policy of
 * apps/main_dvs.cpp:96-116 applied per segment: windows of `horizons`
 * horizons (process_adaptive at `tolerance`), the receptor check at the
 * first window and then once every check_period seconds (never when
 * negative); a rejected window is retried at half the length, and a rejected
 * one-horizon window is integrated natively. */
int dvs_batch_set_adaptive(dvs_batch batch, int32_t segment, double tolerance, int32_t horizons,
                           double check_period);
int dvs_batch_refinement_stats(dvs_batch batch, int32_t segment, dvs_refinement_stats *out);
int dvs_batch_stats(dvs_batch batch, dvs_stats *out);
int dvs_batch_reset_stats(dvs_batch batch);

/* ---- Trajectory (CameraTrajectory, include/dvs/trajectory.hpp:81-213,
 * src/trajectory.cpp) ----------------------------------------------------
 * FrameData<Vec3, Mat3> (include/dvs/scene_types.hpp:43-50): camera_world
 * is the camera-to-world rotation, column major (m[3 * col + row]); a NaN
 * timestamp means frame_index / fps (frame_time, scene_types.hpp:53-56). */
typedef struct {
  int32_t frame_index;
  double timestamp_seconds;
  float camera_origin[3];
  float camera_world[9];
  float light_dir[3];
  float light_intensity;
} dvs_frame;

/* CameraTrajectoryState (trajectory.hpp:44-60): orientation column major,
 * quaternion (x, y, z, w). */
typedef struct {
  float timestamp;
  float position[3];
  float orientation[9];
  float quaternion[4];
  float linear_velocity[3], angular_velocity[3];
  float light_direction[3], light_direction_rate[3];
  float light_intensity, light_intensity_rate;
} dvs_trajectory_state;


/* CameraTrajectory(frames, scene_config) / build (trajectory.cpp:604-716):
 * auto-clamped Bezier keys of the mounted camera position, continuous
 * quaternions of its orientation (mount translation and ZYX Euler rotation
 * in the camera frame), the unit light direction and intensity.  Invalid
 * frames, fps or mount, and non-increasing times are DVS_E_ARGUMENT with the
 * reference's messages.  Host only: the keys go to the device on first use. */
int dvs_trajectory_create(const dvs_frame *frames, int32_t count, float fps, const float translation[3],
                          const float rotation[3], dvs_trajectory *out);
int dvs_trajectory_destroy(dvs_trajectory trajectory);
/* start_time(), end_time(), size(). */
int dvs_trajectory_range(dvs_trajectory trajectory, float *start_time, float *end_time, int32_t *size);
/* evaluate(t) (trajectory.cpp:719-787) for n times, on the host. */
int dvs_trajectory_evaluate(dvs_trajectory trajectory, const float *times, int64_t n, dvs_trajectory_state *out);
/* The n - 1 intervals between consecutive times (n >= 2) as bin/dvs builds
 * them (apps/main_dvs.cpp:103-107): {to_optical_state(evaluate((float)t_i)),
 * to_optical_state(evaluate((float)t_{i+1})), t_i, t_{i+1}}, every state
 * evaluated once, on the device. */
int dvs_trajectory_intervals(dvs_trajectory trajectory, const double *times, int64_t n,
                             dvs_optical_interval *out);

/* ---- Terrain from rasters (src/tools/build_from_raster.cu) -------------
 * build_from_raster, planar path (build_from_raster.cu:213-245, 333-472,
 * 580-610), and the recentring of load_from_raster (:786-802), on the device: a vertex per
 * pixel with a valid elevation (finite, |z| <= 1e8) at (x_coords[col],
 * y_coords[row], z), triangles (v1, v3, v2) and (v2, v3, v4) per cell with
 * four valid corners, angle-weighted vertex normals turned to +z; then every
 * vertex moved by -(x_coords[width/2], y_coords[height/2], dem[centre]).
 * Vertices and triangles are numbered in raster order (the reference's
 * atomic counters give a run-dependent order).  albedo (NULL for none) is
 * per pixel, already on the DEM grid; albedos receives one per vertex.
 * vertices and normals hold 3*width*height floats, albedos width*height,
 * indices 6*(width-1)*(height-1); the counts are returned.  A DEM whose
 * centre pixel is NoData fails (tools::reference_height, loader.hpp:102-106). */
int dvs_raster_mesh(const float *dem, const float *albedo, const float *x_coords, const float *y_coords,
                    int32_t width, int32_t height, float *vertices, float *normals, float *albedos,
                    int64_t *vertex_count, uint32_t *indices, int64_t *triangle_count);
/* batch_align_kernel (build_from_raster.cu:661-682): `source` (source_width x
 * source_height, pixel-centre axes source_x / source_y, either direction)
 * resampled bilinearly at the reference grid's pixel centres (ref_x, ref_y);
 * NaN outside the source or next to a NoData sample.  out: width*height. */
int dvs_raster_align(const float *ref_x, const float *ref_y, int32_t width, int32_t height, const float *source,
                     const float *source_x, const float *source_y, int32_t source_width, int32_t source_height,
                     float *out);

/* ---- The reference's __host__ __device__ physics, called on the host ---- */
/* hapke_reflectance (include/dvs/hapke.hpp:131-151): BRDF x cos(i). */
float dvs_hapke_reflectance(float cosi, float cose, float cosg, float albedo, const dvs_hapke *hapke);
/* hapke_directional_reflectance (include/dvs/hapke.hpp:153-157). */
float dvs_hapke_directional(const float normal[3], const float light[3], const float view[3], float albedo,
                            const dvs_hapke *hapke);
/* invert_hapke_if (include/dvs/hapke.hpp:160-172): angles in degrees,
 * target I/F; NaN for invalid geometry or an unreachable target. */
float dvs_invert_hapke_if(float incidence, float emission, float phase, float target, const dvs_hapke *hapke);
/* receptor_log_at (include/dvs/pixel_model.hpp:18-26). */
double dvs_receptor_log_at(double u, double rate, double dt, double charge);
/* advance_pixel (include/dvs/pixel_model.hpp:35-102): integrates one pixel
 * (the first six fields of `state` are the reference's PixelState; state
 * is updated in place) at the constant photo-electron rate over [t0, t1],
 * writing crossings to out[0 .. capacity).  normal(context) supplies every
 * reset-noise draw (the reference's Rng::normal; NULL draws 0).  Returns
 * the number of crossings, -1 past capacity, -2 for invalid arguments. */
int32_t dvs_advance_pixel(dvs_pixel_state *state, const dvs_pixel_params *params, double rate, double t0,
                          double t1, dvs_crossing *out, int32_t capacity, double (*normal)(void *context),
                          void *context);

/* ---- Device kernels exposed for unit parity (no reference counterpart;
 * they expose the reference's __host__ __device__ functions run on the GPU) */
/* hapke_reflectance (include/dvs/hapke.hpp:131-151) over n inputs, in the
 * form the ray kernels evaluate (the parameter-only terms computed once on
 * the host). */
int dvs_hapke_eval(const float *cosi, const float *cose, const float *cosg, const float *albedo,
                   int64_t n, const dvs_hapke *hapke, float *out);
/* invert_hapke_if (include/dvs/hapke.hpp:160-172) over n inputs (the
 * per-pixel inversion of tools/cpp_tools/main_albedo_map.cu:97-101). */
int dvs_invert_hapke_eval(const float *incidence, const float *emission, const float *phase,
                          const float *target, int64_t n, const dvs_hapke *hapke, float *out);
/* receptor_log_at (include/dvs/pixel_model.hpp:18-26) over n inputs. */
int dvs_receptor_eval(const double *u, const double *rate, const double *dt, double charge,
                      int64_t n, double *out);
/* advance_pixel (include/dvs/pixel_model.hpp:35-102) of n independent pixels
 * on the device: pixel i starts from states[i] (updated in place; the first
 * six fields are the reference's PixelState), integrates the constant
 * photo-electron rate rates[i] over [t0[i], t1[i]] under params[i], takes
 * `normal` as every reset-noise draw, and writes its crossings to
 * events[i * capacity ...]; counts[i] is their number, or -1 past
 * `capacity` (the reference's overflow). */
int dvs_advance_pixel_eval(dvs_pixel_state *states, const dvs_pixel_params *params, const double *rates,
                           const double *t0, const double *t1, int64_t n, int32_t capacity, double normal,
                           int32_t *counts, dvs_crossing *events);
/* The device's sequential-search Poisson inversion of uniform u[i] at
 * lambda[i] (0 < lambda < 10), the shot-noise draw of a circuit cell. */
int dvs_poisson_eval(const double *lambda, const double *u, int64_t n, double *out);
/* Poisson / normal draws of the device noise streams for (seed, pixel). */
int dvs_noise_eval(uint64_t seed, int64_t pixels, double lambda, int32_t draws, double *poisson,
                   double *normal);
/* The event compaction's device prefix sum and stable radix sort (no
 * reference counterpart; they replace the host's std::stable_sort of
 * EventReadout::submit, src/event_readout.cpp:8-12): the inclusive prefix
 * sum of n ints, and (key, value) pairs sorted stably by key bits
 * [begin_bit, end_bit). */
int dvs_scan_eval(const int32_t *in, int64_t n, int32_t *out);
int dvs_sort_eval(const uint64_t *keys, const uint32_t *values, int64_t n, int32_t begin_bit, int32_t end_bit,
                  uint64_t *keys_out, uint32_t *values_out);
/* The event sort's bucketed form over keys in [key_lo, key_lo + 2^key_bits):
 * stably by the top bucket_bits of (key - key_lo), then inside each bucket by
 * the whole key.  *overflow = 1 when a bucket held more than 64 pairs (that
 * bucket is then left in input order). */
int dvs_sort_bucketed_eval(const uint64_t *keys, const uint32_t *values, int64_t n, uint64_t key_lo,
                           int32_t key_bits, int32_t bucket_bits, uint64_t *keys_out, uint32_t *values_out,
                           int32_t *overflow);
/* Nearest / shadow queries of n rays against a scene (traversal parity). */
int dvs_trace_rays(dvs_scene scene, const float *origins, const float *directions, int64_t n,
                   int32_t *hit, int32_t *triangle, float *t, float *u, float *v);
int dvs_trace_shadow_rays(dvs_scene scene, const float *origins, const float *directions,
                          const int32_t *self_triangle, int64_t n, int32_t *occluded);

#ifdef __cplusplus
}
#endif
#endif
