// Shared parts of the bin/dvs and bin/cam front ends: the command line
// (tools::Options, include/tools/cli.hpp:9-100), input loading and the
// metadata files (load_app_inputs, apps/app_inputs.hpp:6-33) and the
// progress bar (include/dvs/progress.hpp).  Raster input (--dem, --albedo,
// --traj) is read as GeoTIFF and meshed on the device (dvs_render_amd.hpp's
// tools::load_from_raster).
#pragma once
#include <chrono>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dvs_render_amd.hpp"

namespace dvsapp {

// Same flags, defaults, aliases and messages as tools::Options.
struct Options {
  std::string scene_dir, dem_path, albedo_path, traj_path;
  std::string output_dir = "./output";
  std::string config_path = "./config/davis346.yaml";
  std::string event_format = "bin";
  std::string event_solver = "auto";

  static void print_usage(const char *prog) {
    std::cout << "Usage: " << prog << "[options]\n"
              << "Options:\n"
              << "  --scene_dir <path>         path to scene bundle directory\n"
              << "  --dem <path>               path to DEM file (optional)\n"
              << "  --albedo <path>            path to albedo file (optional)\n"
              << "  --traj <path>              path to camera trajectory file (optional)\n"
              << "  --output <path>            output images file (default: ./)\n"
              << "  --config <path>            path to YAML config file (default: ./config/davis346.yaml)\n"
              << "  --event_format <bin|hdf5>  event stream output format (default: bin)\n"
              << "  --solver <auto|native|adaptive-optics>  auto selects adaptive optics for physical HDF5 output"
              << std::endl;
  }

  Options(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
      const std::string arg = argv[i];
      auto value = [&]() -> std::string {
        if (i + 1 >= argc) throw std::runtime_error("missing value for " + arg);
        return argv[++i];
      };
      if (arg == "--scene_dir" || arg == "--scene-dir") {
        scene_dir = value();
      } else if (arg == "--dem") {
        dem_path = value();
      } else if (arg == "--albedo") {
        albedo_path = value();
      } else if (arg == "--traj") {
        traj_path = value();
      } else if (arg == "--output") {
        output_dir = value();
      } else if (arg == "--config") {
        config_path = value();
      } else if (arg == "--event_format" || arg == "--event-format" || arg == "--events_format" ||
                 arg == "--events-format") {
        event_format = value();
        if (event_format == "h5") event_format = "hdf5";
        if (event_format == "binary") event_format = "bin";
        if (event_format != "bin" && event_format != "hdf5")
          throw std::runtime_error("invalid event format for --event_format: " + event_format +
                                   " (expected bin or hdf5)");
      } else if (arg == "--solver") {
        event_solver = value();
        if (event_solver != "auto" && event_solver != "native" && event_solver != "adaptive-optics")
          throw std::runtime_error("solver must be auto, native, or adaptive-optics");
      } else if (arg == "--help" || arg == "-h") {
        print_usage(argv[0]);
        std::exit(0);
      } else {
        throw std::runtime_error("unknown argument: " + arg);
      }
    }
    if (scene_dir.empty() && (dem_path.empty() || traj_path.empty()))
      throw std::runtime_error("either --scene_dir or --dem and --traj are required");
    std::ifstream config_in(config_path);
    if (!config_in.good()) throw std::runtime_error("failed to open config file: " + config_path);
  }
};

using Frames = std::vector<dvs::FrameData<dvs::Vec3, dvs::Mat3>>;

// load_app_inputs: the scene bundle and its frames (cut to
// scene.frame_count), source_config.yaml and camera_used.yaml.
inline void load_inputs(const Options &opts, const dvs::Config &config, dvs::SceneBundle<dvs::float3, dvs::uint3> &scene,
                        Frames &frames) {
  std::string error;
  bool loaded = false;
  if (!opts.dem_path.empty() || !opts.traj_path.empty() || !opts.albedo_path.empty()) {
    if (opts.dem_path.empty() || opts.traj_path.empty())
      throw std::runtime_error("raster input requires both --dem and --traj");
    loaded = opts.albedo_path.empty()
                 ? dvs::load_scene_bundle(opts.dem_path, opts.traj_path, scene, frames, &error)
                 : dvs::load_scene_bundle(opts.dem_path, opts.albedo_path, opts.traj_path, scene, frames, &error);
  } else {
    loaded = dvs::load_scene_bundle(opts.scene_dir, scene, frames, &error);
  }
  if (!loaded) throw std::runtime_error("failed to load scene: " + error);
  if (frames.empty()) throw std::runtime_error("empty trajectory");
  frames.resize(std::min(frames.size(), static_cast<size_t>(std::max(config.scene.frame_count, 0))));
  namespace fs = std::filesystem;
  fs::create_directories(opts.output_dir);
  const fs::path source = fs::absolute(opts.config_path);
  const fs::path destination = fs::absolute(fs::path(opts.output_dir) / "source_config.yaml");
  if (source.lexically_normal() != destination.lexically_normal())
    fs::copy_file(source, destination, fs::copy_options::overwrite_existing);
  std::ofstream out(fs::path(opts.output_dir) / "camera_used.yaml");
  if (!out) throw std::runtime_error("cannot write camera metadata");
  out << std::setprecision(9) << "source: " << source << "\nwidth: " << config.camera.width
      << "\nheight: " << config.camera.height << "\nfx: " << config.camera.fx << "\nfy: " << config.camera.fy
      << "\ncx: " << config.camera.cx << "\ncy: " << config.camera.cy
      << "\nsupersample_factor: " << config.scene.supersample_factor
      << "\npixel_centers: integer\nimage_axes: x_right_y_down\npose_axes: Blender_x_right_y_up_z_backward"
      << "\ndepth: mean_hit_ray_range_m; miss=-1"
      << "\ndistortion: not_implemented"
      << "\nphysics_calibration: generic_uncalibrated_optical_and_circuit_parameters\n";
}

// frame_time (include/dvs/scene_types.hpp:53-56).
inline double frame_time(const dvs::FrameData<dvs::Vec3, dvs::Mat3> &f, double fps) {
  return std::isnan(f.timestamp_seconds) ? f.frame_index / fps : f.timestamp_seconds;
}

// "\r<prefix> |####----| 12.50% Elapsed: hh:mm:ss ETA: hh:mm:ss".
class ProgressBar {
 public:
  explicit ProgressBar(size_t total, std::string prefix = "", size_t width = 40)
      : total_(total), width_(width), start_(std::chrono::steady_clock::now()), prefix_(std::move(prefix)) {}
  void update() { update(++completed_); }
  void update(size_t completed) {
    if (total_ == 0) return;
    completed_ = std::min(completed, total_);
    const double frac = (double)completed_ / (double)total_;
    const size_t filled = (size_t)(frac * (double)width_);
    const long long elapsed =
        std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - start_).count();
    const double eta = frac > 1e-8 ? elapsed / frac - elapsed : 0.0;
    std::ostringstream line;
    line << "\r" << prefix_ << " |" << std::string(filled, '#') << std::string(width_ - filled, '-') << "| "
         << std::fixed << std::setprecision(2) << frac * 100.0 << "% Elapsed: " << hms(elapsed)
         << " ETA: " << hms((long long)eta);
    std::cout << line.str() << std::flush;
    if (completed_ >= total_) std::cout << '\n';
  }

 private:
  static std::string hms(long long s) {
    std::ostringstream o;
    o << std::setfill('0') << std::setw(2) << std::max(s / 3600, 0LL) << ":" << std::setw(2)
      << std::max((s % 3600) / 60, 0LL) << ":" << std::setw(2) << std::max(s % 60, 0LL);
    return o.str();
  }
  size_t total_, completed_ = 0, width_;
  std::chrono::steady_clock::time_point start_;
  std::string prefix_;
};

}  // namespace dvsapp
