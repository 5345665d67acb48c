// bin/cam_blur (apps/main_cam_blur.cpp): motion-blur frames of a scene
// bundle along its trajectory, one per 1/cam_fps exposure, written as
// blur_images/<begin>.png.  Each exposure is cut into cells of
// blur_resolution seconds; the frame at each cell's midpoint is weighted by
// the cell's share of the exposure and summed in cell order.  The reference
// renders the cells one render_main call at a time; here a group of
// exposures is rendered by one multi-pose launch and summed on the device
// (CamRayCaster::render_blur), with the same float arithmetic.
#include <algorithm>
#include <cmath>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "app_common.hpp"

namespace {
// Exposures per render_blur call: enough cells to fill the device, few
// enough that the sub-exposure frames stay a modest buffer.
constexpr size_t kExposuresPerCall = 16;

struct Exposure {
  double begin, end;
};

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
  const std::filesystem::path image_dir(opts.output_dir + "/blur_images");
  std::filesystem::create_directories(image_dir);

  const size_t frame_count = std::min(frames.size(), static_cast<size_t>(std::max(config.scene.frame_count, 0)));
  if (frame_count == 0) throw std::runtime_error("no trajectory frames selected");
  const float end_time = (float)dvsapp::frame_time(frames[frame_count - 1], config.scene.fps);
  const float start_time = (float)dvsapp::frame_time(frames.front(), config.scene.fps);
  const double interval = 1.0 / config.scene.cam_fps;
  const size_t exposures = (size_t)std::ceil((double(end_time) - start_time) / interval);
  const double resolution = double(config.scene.blur_resolution);
  dvsapp::ProgressBar bar(exposures, "Rendering motion-blur frames");
  const int w = config.camera.width, h = config.camera.height;
  const size_t npix = (size_t)w * h;
  std::vector<uint8_t> bytes(npix);
  std::vector<float> images;
  std::vector<dvs::CamRayCaster::BlurCell> cells;
  std::vector<int32_t> offsets;
  std::vector<Exposure> batch;
  size_t total_images = 0;

  auto flush = [&]() {
    if (batch.empty()) return;
    raycaster.render_blur(cells, offsets, false, images);
    for (size_t e = 0; e < batch.size(); ++e) {
      dvs::render_float_to_png_bytes(images.data() + e * npix, w, h, 0.0f, bytes);
      const std::string name = std::to_string(batch[e].begin) + ".png";
      dvs::detail::check_io(dvs_io_write_png((image_dir / name).c_str(), bytes.data(), w, h));
      ++total_images;
      bar.update();
    }
    batch.clear();
    cells.clear();
    offsets.assign(1, 0);
  };

  offsets.assign(1, 0);
  for (size_t exposure = 0; exposure < exposures; ++exposure) {
    const double begin = start_time + exposure * interval;
    const double end = std::min(double(end_time), start_time + (exposure + 1) * interval);
    if (end <= begin) break;
    const size_t count = (size_t)std::ceil((end - begin) / resolution);
    for (size_t cell = 0; cell < count; ++cell) {
      const double left = begin + cell * resolution;
      const double right = std::min(end, begin + (cell + 1) * resolution);
      const dvs::CameraTrajectoryState state = traj.evaluate((float)((left + right) * .5));
      dvs::CamRayCaster::BlurCell c;
      c.camera_origin = {state.position.x(), state.position.y(), state.position.z()};
      std::copy(state.orientation.data(), state.orientation.data() + 9, c.camera_world);
      c.light_dir = {state.light_direction.x(), state.light_direction.y(), state.light_direction.z()};
      c.light_intensity = state.light_intensity;
      c.weight = (double)(float)((right - left) / (end - begin));
      cells.push_back(c);
    }
    offsets.push_back((int32_t)cells.size());
    batch.push_back({begin, end});
    if (batch.size() == kExposuresPerCall) flush();
  }
  flush();
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
