// bin/sequence (apps/main_sequence.cpp): one dataset sequence -- 60 s of
// events (events.h5) and camera frames (frames.h5) over a source trajectory
// repeated back and forth, with the pose, motion and geographic tables, all
// writes checked against the dataset's size budget.
//
//   sequence [--prepare-poses|--optimized] dataset_root sequence_dir dem albedo_or_empty
//
// sequence_dir holds source_config.yaml and source_trajectory.txt.  The
// optimized form integrates adaptive-optics windows (200 horizons, halved
// on a rejected receptor check, doubled again after four good checks in a
// row); the default form native windows of 64.  DVS_SEQUENCE_SECONDS, if
// set, replaces the 60 s duration (for short runs).
//
// Longitude and latitude need the DEM's CRS inverted: geographic rasters
// and equirectangular projections (spherical body, GeoKeys) are handled
// here; any other projection fails as an unconvertible map does in the
// reference.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "dvs_render_amd.hpp"

namespace {

double sequence_duration() {
  const char *s = std::getenv("DVS_SEQUENCE_SECONDS");
  if (!s || !*s) return 60.0;
  const double d = std::strtod(s, nullptr);
  if (!std::isfinite(d) || d <= 0) throw std::runtime_error("DVS_SEQUENCE_SECONDS must be a positive number");
  return d;
}

// GeographicReference (main_sequence.cpp:18-47): the DEM's centre pixel in
// projected coordinates and height, and the inverse of its projection.
struct GeographicReference {
  double x = 0, y = 0, z = 0;
  std::function<void(double &, double &)> to_lonlat;
  explicit GeographicReference(const std::filesystem::path &path) {
    dvs_raster r = nullptr;
    if (dvs_io_raster_load(path.c_str(), &r) != DVS_OK) throw std::runtime_error("cannot open source DEM for geographic poses");
    std::unique_ptr<dvs_raster_s, int (*)(dvs_raster)> hold(r, dvs_io_raster_free);
    dvs_raster_info info{};
    dvs_io_raster_info(r, &info);
    const double *gt = info.transform;
    const int col = info.width / 2, row = info.height / 2;
    x = gt[0] + (col + .5) * gt[1] + (row + .5) * gt[2];
    y = gt[3] + (col + .5) * gt[4] + (row + .5) * gt[5];
    const float value = info.center_raw;
    if (!std::isfinite(value) || (info.has_nodata && value == float(info.nodata)))
      throw std::runtime_error("DEM center is NoData");
    z = value * (info.has_scale ? info.scale : 1) + (info.has_offset ? info.offset : 0);
    auto key = [&](int id, double &v) { return dvs_io_raster_geokey(r, id, &v) == DVS_OK; };
    double model = 0;
    if (!key(1024, model)) throw std::runtime_error("invalid DEM CRS");
    if (model == 2) {  // geographic: the raster axes are longitude and latitude
      to_lonlat = [](double &, double &) {};
      return;
    }
    double transform = 0;
    if (model != 1 || !key(3075, transform) || transform != 17)  // CT_Equirectangular
      throw std::runtime_error("cannot convert source map to lunar longitude/latitude");
    double radius = 0, parallel = 0, lon0 = 0, lat0 = 0, east = 0, north = 0;
    if (!key(2057, radius) || !(radius > 0)) throw std::runtime_error("missing source planetary geographic CRS");
    key(3078, parallel);
    if (!key(3080, lon0)) key(3088, lon0);
    if (!key(3081, lat0)) key(3089, lat0);
    key(3082, east);
    key(3083, north);
    const double deg = 180.0 / 3.14159265358979323846;
    const double scale = radius * std::cos(parallel / deg);
    to_lonlat = [=](double &px, double &py) {
      const double lon = lon0 + (px - east) / scale * deg, lat = lat0 + (py - north) / radius * deg;
      px = lon;
      py = lat;
    };
  }
};

// The quaternion of a column-major rotation matrix (normalised), as
// CamRayCaster derives it for render_main.
dvs::float4 quaternion_of(const float *m) {
  float q[4];
  dvs::CamRayCaster::quaternion_of(m, q);
  return dvs::make_float4(q[0], q[1], q[2], q[3]);
}

void prepare_poses(const dvs::RepeatedTrajectory &trajectory, const dvs::Config &c, double duration,
                   const std::filesystem::path &dem, const std::filesystem::path &out,
                   const dvs::DatasetBudget &budget) {
  GeographicReference map(dem);
  std::ostringstream poses, motion, geo, camera;
  poses << std::setprecision(12) << "t_s,x_m,y_m,z_m,qx,qy,qz,qw\n";
  motion << std::setprecision(12)
         << "t_s,vx_world_m_s,vy_world_m_s,vz_world_m_s,wx_world_rad_s,wy_world_rad_s,wz_world_rad_s,sun_x,sun_y,"
            "sun_z,sun_intensity\n";
  geo << std::setprecision(15) << "t_s,projected_x_m,projected_y_m,source_altitude_m,longitude_deg,latitude_deg\n";
  dvs::float4 previous = dvs::make_float4(0, 0, 0, 1);
  const std::size_t n = (std::size_t)std::ceil(duration * c.scene.imu_fps);
  for (std::size_t i = 0; i <= n; ++i) {
    const double t = std::min(duration, i / double(c.scene.imu_fps));
    const auto s = trajectory.evaluate(t);
    // orientation * diag(1, -1, -1): camera axes from Blender's to OpenCV's
    float m[9];
    const float *o = s.orientation.data();
    for (int r = 0; r < 3; ++r) {
      m[r] = o[r];
      m[3 + r] = -o[3 + r];
      m[6 + r] = -o[6 + r];
    }
    dvs::float4 q = quaternion_of(m);
    if (i && q.x * previous.x + q.y * previous.y + q.z * previous.z + q.w * previous.w < 0) {
      q.x = -q.x;
      q.y = -q.y;
      q.z = -q.z;
      q.w = -q.w;
    }
    previous = q;
    poses << t << ',' << s.position.x() << ',' << s.position.y() << ',' << s.position.z() << ',' << q.x << ','
          << q.y << ',' << q.z << ',' << q.w << '\n';
    motion << t << ',' << s.linear_velocity.x() << ',' << s.linear_velocity.y() << ',' << s.linear_velocity.z()
           << ',' << s.angular_velocity.x() << ',' << s.angular_velocity.y() << ',' << s.angular_velocity.z() << ','
           << s.light_direction.x() << ',' << s.light_direction.y() << ',' << s.light_direction.z() << ','
           << s.light_intensity << '\n';
    double px = map.x + s.position.x(), py = map.y + s.position.y(), longitude = px, latitude = py;
    map.to_lonlat(longitude, latitude);
    geo << t << ',' << px << ',' << py << ',' << map.z + s.position.z() << ',' << longitude << ',' << latitude
        << '\n';
  }
  camera << std::setprecision(9) << "model: pinhole\nwidth: " << c.camera.width << "\nheight: " << c.camera.height
         << "\nfx: " << c.camera.fx << "\nfy: " << c.camera.fy << "\ncx: " << c.camera.cx << "\ncy: " << c.camera.cy
         << "\nsupersample_factor: " << c.scene.supersample_factor << "\ntime_resolution: " << c.scene.time_resolution
         << "\ncircuit_step: " << c.physic.circuit_step << "\nthreshold_pos: " << c.physic.threshold_pos
         << "\nthreshold_neg: " << c.physic.threshold_neg << "\ncam_fps: " << c.scene.cam_fps
         << "\npose_fps: " << c.scene.imu_fps << "\nduration_seconds: " << duration
         << "\npixel_centers: integer\npose_axes: opencv_x_right_y_down_z_forward\nsource_model: source_config.yaml"
         << "\ndistortion: ideal_pinhole_zero_distortion_not_hardware_calibration"
         << "\nphysics_calibration: generic_uncalibrated_optical_and_circuit_parameters"
         << "\ngeographic_axes: local_projected_raster_x_y_and_source_height_relative_to_center"
         << "\nvertical_datum: as_provided_not_independently_verified"
         << "\nimage_model: instantaneous_fixed_exposure_sRGB_preview_not_calibrated_APS"
         << "\nimage_encoding: sRGB_of_linear_render_units_divided_by_100"
         << "\ndepth_model: mean_hit_ray_range_m_minus_one_on_miss"
         << "\nevent_timestamp: int64_microseconds_transport_completion"
         << "\nframe_timestamp: int64_microseconds_instantaneous_sample"
         << "\nmotion: original_speed_ping_pong_with_0.5_second_C2_reversals\n";
  budget.write_text(out / "poses.csv", poses.str());
  budget.write_text(out / "motion.csv", motion.str());
  budget.write_text(out / "pose_geographic.csv", geo.str());
  budget.write_text(out / "camera_used.yaml", camera.str());
}

int run(int argc, char **argv) {
  const bool poses_only = argc > 1 && std::string(argv[1]) == "--prepare-poses";
  const bool optimized = argc > 1 && std::string(argv[1]) == "--optimized";
  const int shift = (poses_only || optimized) ? 1 : 0;
  if (argc != 5 + shift)
    throw std::runtime_error("sequence [--prepare-poses|--optimized] dataset_root sequence_dir dem albedo_or_empty");
  const std::filesystem::path root = argv[1 + shift], out = argv[2 + shift], dem = argv[3 + shift];
  const std::string albedo = argv[4 + shift];
  const double duration = sequence_duration();
  dvs::DatasetBudget budget(root);
  budget.require_path(out / "sequence.json");
  std::string error;
  dvs::Config config((out / "source_config.yaml").string(), error);
  if (!error.empty()) throw std::runtime_error(error);
  if (config.scene.time_resolution <= 0 || config.scene.cam_fps <= 0 || config.scene.imu_fps <= 0)
    throw std::runtime_error("invalid source timing parameters");
  std::vector<dvs::FrameData<dvs::Vec3, dvs::Mat3>> frames;
  if (!dvs::load_frames_from_text((out / "source_trajectory.txt").string(), frames, &error) || frames.size() < 2)
    throw std::runtime_error("failed to load source trajectory: " + error);
  dvs::CameraTrajectory source(frames, config.scene);
  dvs::RepeatedTrajectory trajectory(source);
  if (std::filesystem::exists(out / "events.h5") || std::filesystem::exists(out / "frames.h5"))
    throw std::runtime_error("refusing to overwrite an existing rendered or partial sequence");
  prepare_poses(trajectory, config, duration, dem, out, budget);
  if (poses_only) {
    std::cout << "Prepared " << duration << " s native-camera poses and geographic coordinates: " << out << '\n';
    return 0;
  }
  budget.write(0, [] {});
  dvs::SceneBundle<dvs::float3, dvs::uint3> mesh;
  const bool loaded = albedo.empty() ? dvs::load_scene_bundle(dem.string(), mesh, &error)
                                     : dvs::load_scene_bundle(dem.string(), albedo, mesh, &error);
  if (!loaded) throw std::runtime_error("failed to load exact source DEM/material: " + error);
  dvs::BVHBuildParams params;
  params.max_leaf_size = 64;
  const int bvh_root = dvs::BVHBuilder::build(params, mesh.bvh_nodes, mesh.bvh_triangles, mesh.vertices);
  // one device copy of the terrain shared by image and event rays
  auto scene = std::make_shared<dvs::DeviceScene>(mesh, config, bvh_root);
  mesh = {};
  std::cout << "SCENE_GPU_READY" << std::endl;
  dvs::DvsRayCaster sensor(scene, config);
  dvs::CamRayCaster camera(scene);
  dvs::Hdf5EventStreamWriter events((out / "events.h5").string(), &budget);
  dvs::Hdf5FrameWriter images(out / "frames.h5", config.camera.width, config.camera.height, budget);
  sensor.initialize(trajectory.evaluate(0), 0);
  std::vector<float> radiance, depth;
  auto frame_at = [&](double t) {
    const auto p = trajectory.evaluate(t);
    camera.render_main(dvs::make_float3(p.position.x(), p.position.y(), p.position.z()), p.orientation.data(),
                       dvs::make_float3(p.light_direction.x(), p.light_direction.y(), p.light_direction.z()),
                       p.light_intensity, radiance, depth);
    images.append(t, radiance, depth);
  };
  frame_at(0);
  std::size_t frame_index = 1, step = 0;
  double now = 0, last_report = 0;
  const double dt = config.scene.time_resolution;
  const auto start = std::chrono::steady_clock::now();
  std::vector<dvs::Event> ready;
  std::vector<dvs::OpticalInterval> intervals;
  intervals.reserve(200);
  constexpr double optical_tolerance = dvs::kDefaultOpticalLogTolerance;
  unsigned long long optical_queries = 0, fallback_queries = 0, fallback_pixel_windows = 0, windows = 0;
  int maximum_knots = 0;
  unsigned convergence_checks = 0, convergence_rejections = 0, native_guard_windows = 0;
  double next_convergence_check = 0, maximum_receptor_error = 0, maximum_p99_receptor_error = 0;
  unsigned batch = optimized ? 200 : 64;
  if (config.physic.refractory_time <= 0) {
    batch = 1;
  } else {
    batch = std::max(1U, std::min(batch, unsigned(std::max(1.0, (config.physic.max_events_per_step - 2) *
                                                                     double(config.physic.refractory_time) / dt))));
  }
  const unsigned maximum_batch = batch;
  unsigned stable_quality_checks = 0;
  if (optimized) {
    std::ostringstream model;
    model << std::setprecision(12)
          << "{\"model\":\"ray_photocurrent_adaptive_optics_native_circuit\",\"maximum_window_seconds\":" << batch * dt
          << ",\"empirical_midpoint_log_photocurrent_tolerance\":" << optical_tolerance
          << ",\"dim_leaf_max_expected_electrons\":1.0"
          << ",\"leaf_voltage_error_threshold_fraction\":" << dvs::kOpticalVoltageThresholdFraction
          << ",\"maximum_knots_per_pixel\":" << sensor.adaptive_cache_capacity()
          << ",\"maximum_local_exact_cells_per_pixel\":" << sensor.adaptive_cache_capacity()
          << ",\"native_spatial_grid\":" << config.scene.supersample_factor
          << ",\"native_trajectory_interval_seconds\":" << dt
          << ",\"native_circuit_step_seconds\":" << config.physic.circuit_step
          << ",\"unresolved_pixel_action\":\"direct_rays_at_every_native_circuit_midpoint\""
          << ",\"optical_error_guarantee\":\"empirical_probe_convergence_not_global_visibility_bound\""
          << ",\"spatial_adaptation\":false,\"per_event_auxiliary_depth\":\"not_available_HDF5_uses_x_y_t_p_only\""
          << ",\"camera_parameters\":\"source_config.yaml_unchanged\",\"physical_parameters\":\"generic_uncalibrated\""
          << ",\"quality_check_interval_seconds\":0.25,\"mean_current_receptor_probe_pixels\":1024,"
             "\"quality_max_error_threshold_fraction\":0.02,\"quality_p99_error_threshold_fraction\":0.005"
          << ",\"shadow_origin\":\"barycentric_geometry_float_error_offset_and_outward_rounding\"}";
    budget.write_text(out / "numerical_model.json", model.str());
  }
  auto left = dvs::to_optical_state(trajectory.evaluate(0));
  while (now < duration) {
    const double window_start = now;
    const std::size_t window_step = step;
    const auto window_left = left;
    intervals.clear();
    for (unsigned i = 0; i < batch && now < duration; ++i) {
      const double next = std::min(duration, (++step) * dt);
      if (next <= now) throw std::runtime_error("non-increasing integration grid");
      const auto right = dvs::to_optical_state(trajectory.evaluate(next));
      intervals.push_back({left, right, now, next});
      left = right;
      now = next;
    }
    if (optimized) {
      const bool check = window_start >= next_convergence_check;
      const bool accepted = sensor.advance_adaptive(intervals, optical_tolerance, ready, check);
      if (check) {
        const auto result = sensor.last_convergence();
        ++convergence_checks;
        maximum_receptor_error = std::max(maximum_receptor_error, result.maximum_log_voltage_error);
        maximum_p99_receptor_error = std::max(maximum_p99_receptor_error, result.p99_log_voltage_error);
      }
      if (!accepted) {
        ++convergence_rejections;
        stable_quality_checks = 0;
        if (batch > 1) {
          now = window_start;
          step = window_step;
          left = window_left;
          batch = std::max(1U, batch / 2);
          continue;
        }
        sensor.advance_many(intervals, ready);
        ++native_guard_windows;
      }
      if (check) {
        next_convergence_check = now + .25;
        const auto quality = sensor.last_convergence();
        const double threshold = std::min(config.physic.threshold_pos, config.physic.threshold_neg);
        if (accepted && quality.maximum_log_voltage_error < .005 * threshold &&
            quality.p99_log_voltage_error < .001 * threshold)
          ++stable_quality_checks;
        else
          stable_quality_checks = 0;
        if (stable_quality_checks >= 4 && batch < maximum_batch) {
          batch = std::min(maximum_batch, batch * 2);
          stable_quality_checks = 0;
        }
      }
      const auto statistics = sensor.refinement_statistics();
      optical_queries += statistics.traced_pixel_samples + statistics.native_cell_queries;
      fallback_pixel_windows += statistics.fallback_pixels;
      unsigned long long cells = 0;
      for (const auto &iv : intervals) cells += (unsigned long long)std::ceil((iv.t1 - iv.t0) / config.physic.circuit_step);
      fallback_queries += cells * statistics.fallback_pixels;
      maximum_knots = std::max(maximum_knots, statistics.maximum_knots);
      ++windows;
    } else {
      sensor.advance_many(intervals, ready);
    }
    events.append(std::move(ready));
    ready.clear();
    while (frame_index / double(config.scene.cam_fps) <= now + 1e-12 &&
           frame_index / double(config.scene.cam_fps) <= duration) {
      frame_at(frame_index / double(config.scene.cam_fps));
      ++frame_index;
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    if (elapsed - last_report >= 30) {
      std::ostringstream progress;
      progress << std::setprecision(12) << "{\"status\":\"rendering\",\"observed_seconds\":" << now
               << ",\"wall_seconds\":" << elapsed << ",\"events\":" << events.event_count()
               << ",\"frames\":" << images.count() << ",\"adaptive_optics\":" << (optimized ? "true" : "false")
               << ",\"optical_pixel_queries\":" << optical_queries
               << ",\"exact_fallback_pixel_queries\":" << fallback_queries
               << ",\"fallback_pixel_windows\":" << fallback_pixel_windows
               << ",\"convergence_checks\":" << convergence_checks
               << ",\"convergence_rejections\":" << convergence_rejections << ",\"current_batch_intervals\":" << batch
               << "}";
      budget.write_text(out / "render_progress.json", progress.str());
      std::cout << progress.str() << std::endl;
      last_report = elapsed;
    }
  }
  if (std::abs((frame_index - 1) / double(config.scene.cam_fps) - duration) > 1e-9) frame_at(duration);
  sensor.drain(ready);
  const double transport_end = ready.empty() ? duration : std::max(duration, ready.back().timestamp);
  events.append(std::move(ready));
  events.close(transport_end);
  images.close();
  std::ostringstream result;
  result << std::setprecision(12) << "{\"status\":\"rendered_pending_validation\",\"duration_seconds\":" << duration
         << ",\"event_transport_end_seconds\":" << transport_end << ",\"events\":" << events.event_count()
         << ",\"frames\":" << images.count() << ",\"readout_dropped_events\":" << sensor.dropped_events()
         << ",\"peak_readout_queue\":" << sensor.peak_queue() << ",\"numerical_model\":\""
         << (optimized ? "adaptive_optics_native_circuit_cells" : "native_rays_and_circuit_cells")
         << "\",\"optical_pixel_queries\":" << optical_queries << ",\"exact_fallback_pixel_queries\":" << fallback_queries
         << ",\"fallback_pixel_windows\":" << fallback_pixel_windows << ",\"windows\":" << windows
         << ",\"maximum_knots\":" << maximum_knots << ",\"convergence_checks\":" << convergence_checks
         << ",\"convergence_rejections\":" << convergence_rejections
         << ",\"native_guard_windows\":" << native_guard_windows
         << ",\"maximum_probed_mean_current_receptor_log_error\":" << maximum_receptor_error
         << ",\"maximum_probed_p99_mean_current_receptor_log_error\":" << maximum_p99_receptor_error << "}";
  budget.write_text(out / "render_progress.json", result.str());
  std::cout << result.str() << '\n';
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const dvs::DatasetSizeLimit &e) {
    std::cerr << "STOP_SIZE_LIMIT: " << e.what() << '\n';
    return 42;
  } catch (const std::exception &e) {
    std::cerr << "FAILED: " << e.what() << '\n';
    return 1;
  }
}
