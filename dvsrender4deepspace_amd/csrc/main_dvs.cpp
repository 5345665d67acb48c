// bin/dvs (apps/main_dvs.cpp): events of a scene bundle flown along its
// trajectory, written to events.bin or events.h5 with the metadata files.
//
// The loop is the reference's: host horizons t_k = min(end, start + k *
// double(time_resolution)), windows of 64 (native) or 200 (adaptive)
// intervals capped by the event capacity, the receptor convergence check at
// the first window and every 0.25 s, a rejected window retried at half the
// size and a rejected one-horizon window integrated natively, writes every
// 1/imu_fps.  The trajectory states of a window's horizons are evaluated in
// one device launch instead of one host evaluation per horizon (inside the
// adaptive window itself, or by CameraTrajectory::intervals for native
// windows); each time is still evaluated once, at float(t), and consecutive
// intervals share their common state.
#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <memory>
#include <string>
#include <vector>

#include "app_common.hpp"

namespace {
constexpr double kDefaultOpticalLogTolerance = 1e-2;        // include/dvs/optical_sampler.hpp:6
constexpr double kOpticalVoltageThresholdFraction = 1e-3;   // include/dvs/optical_sampler.hpp:7

int run(int argc, char **argv) {
  if (argc <= 1) {
    dvsapp::Options::print_usage(argv[0]);
    return 1;
  }
  const dvsapp::Options opts(argc, argv);
  std::string error;
  dvs::Config config(opts.config_path, error);
  if (!error.empty()) throw std::runtime_error("failed to load config: " + error);

  const bool adaptive = opts.event_solver == "adaptive-optics" ||
                        (opts.event_solver == "auto" && opts.event_format == "hdf5" &&
                         config.physic.refractory_time > 0);
  if (adaptive && opts.event_format != "hdf5")
    throw std::runtime_error(
        "adaptive optics exports HDF5 x/y/t/p; per-event auxiliary binary depth requires --solver native");
  dvs::SceneBundle<dvs::float3, dvs::uint3> scene;
  dvsapp::Frames frames;
  dvsapp::load_inputs(opts, config, scene, frames);

  dvs::CameraTrajectory traj(frames, config.scene);
  dvs::BVHBuildParams params;
  params.max_leaf_size = 64;
  const int root = dvs::BVHBuilder::build(params, scene.bvh_nodes, scene.bvh_triangles, scene.vertices);
  dvs::DvsRayCaster raycaster(scene, config, root);
  std::filesystem::create_directories(opts.output_dir);
  auto events_out = dvs::make_event_writer(opts.event_format, opts.output_dir);

  const size_t frame_count = std::min(frames.size(), static_cast<size_t>(std::max(config.scene.frame_count, 0)));
  if (frame_count == 0) throw std::runtime_error("no trajectory frames selected");
  const float end_time = (float)dvsapp::frame_time(frames[frame_count - 1], config.scene.fps);
  const float start_time = (float)dvsapp::frame_time(frames.front(), config.scene.fps);
  const size_t total_steps =
      (size_t)std::ceil(((double)end_time - (double)start_time) / (double)config.scene.time_resolution);
  dvsapp::ProgressBar bar(total_steps, "Generating events");

  unsigned batch = adaptive ? 200 : 64;
  if (config.physic.refractory_time <= 0) {
    batch = 1;
  } else {
    const double cap = std::max(1.0, (config.physic.max_events_per_step - 2) * double(config.physic.refractory_time) /
                                         config.scene.time_resolution);
    batch = std::max(1U, std::min(batch, unsigned(cap)));
  }
  // Host horizons of the window starting after `step`, prefixed by its
  // start time; stops early where a horizon would not advance.
  auto window_times = [&](size_t step, double previous_time, std::vector<double> &times) {
    times.assign(1, previous_time);
    for (unsigned j = 0; j < batch && step < total_steps; ++j) {
      const double t = std::min(double(end_time), double(start_time) + (++step) * double(config.scene.time_resolution));
      if (t <= times.back()) break;
      times.push_back(t);
    }
  };

  std::vector<dvs::Event> events, events_buffer;
  double next_write_time = start_time;
  const float write_interval = 1.0f / config.scene.imu_fps;
  double previous_time = start_time, next_check = start_time;
  size_t step = 0;
  std::vector<double> times, next_times;
  std::vector<dvs::OpticalInterval> intervals;
  bool initialized = false;
  while (step < total_steps) {
    window_times(step, previous_time, times);
    if (times.size() < 2) break;
    if (!initialized) {
      // the state at start_time, as the first interval starts from it
      raycaster.initialize(traj.intervals({times[0], times[1]}).front().start, start_time);
      std::cout << "Initialized continuous photocurrent/circuit state from scene rays.\n";
      initialized = true;
    }
    const double first_time = previous_time;
    if (adaptive) {
      // the window's states are evaluated on the device inside the window;
      // the next window (if this one is accepted) is set up on the device
      // before this call returns
      const bool check = first_time >= next_check;
      window_times(step + times.size() - 1, times.back(), next_times);
      if (next_times.size() >= 2) raycaster.expect_adaptive(traj, next_times, kDefaultOpticalLogTolerance);
      if (!raycaster.advance_adaptive(traj, times, kDefaultOpticalLogTolerance, events, check)) {
        if (batch > 1) {
          batch = std::max(1U, batch / 2);
          continue;
        }
        raycaster.advance_many(traj.intervals(times), events);
      }
      if (check) next_check = times.back() + .25;
    } else {
      intervals = traj.intervals(times);
      raycaster.advance_many(intervals, events);
    }
    const size_t advanced = times.size() - 1;
    previous_time = times.back();
    std::move(events.begin(), events.end(), std::back_inserter(events_buffer));
    events.clear();
    if (previous_time >= next_write_time) {
      next_write_time = previous_time + write_interval;
      events_out->append(std::move(events_buffer));
      events_buffer.clear();
    }
    for (size_t j = 0; j < advanced; ++j) bar.update();
    step += advanced;
  }
  if (!initialized) raycaster.initialize(traj.evaluate(start_time), start_time);  // no horizon advanced
  if (!events_buffer.empty()) {
    events_out->append(std::move(events_buffer));
    events_buffer.clear();
  }
  // the last interval's events and the transport backlog
  raycaster.drain(events);
  const double output_end =
      events.empty() ? end_time : std::max(static_cast<double>(end_time), static_cast<double>(events.back().timestamp));
  events_out->append(std::move(events));
  events_out->close(output_end);

  std::ifstream source_config(opts.config_path);
  std::ofstream copied_config(opts.output_dir + "/sensor_config.yaml");
  copied_config << source_config.rdbuf();
  std::ofstream meta(opts.output_dir + "/sensor_model.txt");
  meta << std::setprecision(12) << "model=ray_photocurrent_circuit_v2\n"
       << "timestamp="
       << (opts.event_format == "hdf5" ? "int64_microseconds_transport_completion"
                                       : "float32_seconds_transport_completion")
       << "\n"
       << "config=" << opts.config_path << "\n"
       << "random_seed=" << config.physic.random_seed << "\n"
       << "dropped_readout_events=" << raycaster.dropped_events() << "\n"
       << "peak_readout_queue=" << raycaster.peak_queue() << "\n"
       << "physical_effects=on\n"
       << "photoreceptor=nonlinear_current_balance\n"
       << "optical_sampling="
       << (adaptive ? "adaptive_direct_ray_nodes_with_native_cell_fallback" : "midpoint_rays_within_circuit_cells")
       << "\n"
       << "optical_log_tolerance=" << (adaptive ? kDefaultOpticalLogTolerance : 0) << "\n"
       << "leaf_voltage_error_threshold_fraction=" << (adaptive ? kOpticalVoltageThresholdFraction : 0) << "\n"
       << "optical_cache_capacity=" << (adaptive ? raycaster.adaptive_cache_capacity() : 0) << "\n"
       << "camera_width=" << config.camera.width << "\ncamera_height=" << config.camera.height << "\n"
       << "fx=" << config.camera.fx << "\nfy=" << config.camera.fy << "\ncx=" << config.camera.cx
       << "\ncy=" << config.camera.cy << "\n"
       << "spatial_grid=" << config.scene.supersample_factor << "\n"
       << "distortion=ideal_pinhole_not_hardware_calibration\n"
       << "physics_calibration=generic_uncalibrated\n"
       << "photon_noise=charge_count_tau_leap\n"
       << "circuit_step_seconds=" << config.physic.circuit_step << "\n"
       << "radiance_scale=" << config.physic.radiance_scale << "\n"
       << "readout_queue=on\n";
  std::cout << "Readout dropped " << raycaster.dropped_events() << " events; peak queue " << raycaster.peak_queue()
            << "\n";
  std::cout << "Finished. Wrote " << events_out->event_count() << " events to " << events_out->path() << "\n";
  return 0;
}
}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &ex) {
    std::cerr << "Error: " << ex.what() << "\n";
    return 1;
  }
}
