// bin/cam (apps/main_cam.cpp): frames of a scene bundle along its
// trajectory at cam_fps, written as images/<t>.png (8-bit, the
// render_float_to_png_bytes encode) and depth/<t>.tif (float32 mean hit
// range, -1 on a miss).
#include <cmath>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "app_common.hpp"

namespace {
int run(int argc, char **argv) {
  if (argc <= 1) {
    dvsapp::Options::print_usage(argv[0]);
    return 1;
  }
  const dvsapp::Options opts(argc, argv);
  std::string error;
  dvs::Config config(opts.config_path, error);
  if (!error.empty()) throw std::runtime_error("failed to load config: " + error);
  dvs::SceneBundle<dvs::float3, dvs::uint3> scene;
  dvsapp::Frames frames;
  dvsapp::load_inputs(opts, config, scene, frames);

  dvs::CameraTrajectory traj(frames, config.scene);
  dvs::BVHBuildParams params;
  params.max_leaf_size = 64;
  const int root = dvs::BVHBuilder::build(params, scene.bvh_nodes, scene.bvh_triangles, scene.vertices);
  dvs::CamRayCaster raycaster(scene, config, root);
  const std::filesystem::path image_dir(opts.output_dir + "/images"), depth_dir(opts.output_dir + "/depth");
  std::filesystem::create_directories(image_dir);
  std::filesystem::create_directories(depth_dir);

  const size_t frame_count = std::min(frames.size(), static_cast<size_t>(std::max(config.scene.frame_count, 0)));
  if (frame_count == 0) throw std::runtime_error("no trajectory frames selected");
  const float end_time = (float)dvsapp::frame_time(frames[frame_count - 1], config.scene.fps);
  const float start_time = (float)dvsapp::frame_time(frames.front(), config.scene.fps);
  const size_t total_steps = (size_t)std::ceil((double(end_time) - start_time) * config.scene.cam_fps);
  dvsapp::ProgressBar bar(total_steps, "Rendering frames");
  const float write_interval = 1.0f / config.scene.cam_fps;
  const int w = config.camera.width, h = config.camera.height;
  std::vector<float> light(w * h), depth(w * h);
  std::vector<uint8_t> bytes(w * h);
  size_t total_images = 0;
  for (size_t step = 0; step < total_steps; ++step) {
    const float current_time = start_time + step * write_interval;
    const dvs::CameraTrajectoryState state = traj.evaluate(current_time);
    const dvs::float3 origin{state.position.x(), state.position.y(), state.position.z()};
    const dvs::float3 light_dir{state.light_direction.x(), state.light_direction.y(), state.light_direction.z()};
    raycaster.render_main(origin, state.orientation.data(), light_dir, state.light_intensity, light, depth);
    const std::string stem = std::to_string(current_time);
    dvs::render_float_to_png_bytes(light.data(), w, h, 0.0f, bytes);
    dvs::detail::check_io(dvs_io_write_png((image_dir / (stem + ".png")).c_str(), bytes.data(), w, h));
    dvs::detail::check_io(dvs_io_write_tiff((depth_dir / (stem + ".tif")).c_str(), depth.data(), w, h));
    ++total_images;
    bar.update();
  }
  std::cout << "Finished. Wrote " << total_images << " images to " << opts.output_dir << "\n";
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
