// bin/cam_blur_windows (apps/main_cam_blur_windows.cpp): three 50 ms
// motion-blur frames of a dataset sequence, centred on the given times of
// its repeated trajectory, over the sequence's GeoTIFF terrain.
//
//   cam_blur_windows <sequence_dir> <dem> <albedo_or_empty> <output_dir> <center_s> <center_s> <center_s>
//
// Each window is cut into cells of blur_resolution seconds; the frame at a
// cell's midpoint is weighted by the cell's share of the window and summed
// in double precision in cell order, then rounded to float and encoded
// (motion_blur_<begin>_<end>s.png).  All cells of the three windows go to
// the device in one multi-pose launch and are summed there
// (CamRayCaster::render_blur with the double accumulator).
#include <algorithm>
#include <cmath>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "dvs_render_amd.hpp"

namespace {

struct Window {
  double center, begin, end;
  size_t cells;
};

int run(int argc, char **argv) {
  if (argc != 8) {
    std::cerr << "Usage: cam_blur_windows <sequence_dir> <dem> <albedo_or_empty> "
                 "<output_dir> <center_s> <center_s> <center_s>\n";
    return 2;
  }
  const std::filesystem::path sequence = argv[1], dem = argv[2], output = argv[4];
  const std::string albedo = argv[3];
  std::string error;
  dvs::Config config((sequence / "source_config.yaml").string(), error);
  if (!error.empty()) throw std::runtime_error("cannot load source camera config: " + error);
  if (config.scene.blur_resolution <= 0 || config.scene.supersample_factor <= 0)
    throw std::runtime_error("invalid native blur sampling configuration");
  std::vector<dvs::FrameData<dvs::Vec3, dvs::Mat3>> frames;
  if (!dvs::load_frames_from_text((sequence / "source_trajectory.txt").string(), frames, &error) || frames.size() < 2)
    throw std::runtime_error("cannot load source trajectory: " + error);
  dvs::CameraTrajectory source(frames, config.scene);
  dvs::RepeatedTrajectory trajectory(source);
  dvs::SceneBundle<dvs::float3, dvs::uint3> mesh;
  const bool loaded = albedo.empty() ? dvs::load_scene_bundle(dem.string(), mesh, &error)
                                     : dvs::load_scene_bundle(dem.string(), albedo, mesh, &error);
  if (!loaded) throw std::runtime_error("cannot load source terrain/material: " + error);
  dvs::BVHBuildParams params;
  params.max_leaf_size = 64;
  const int root = dvs::BVHBuilder::build(params, mesh.bvh_nodes, mesh.bvh_triangles, mesh.vertices);
  dvs::CamRayCaster camera(mesh, config, root);
  mesh = {};
  std::cout << "SCENE_GPU_READY" << std::endl;
  std::filesystem::create_directories(output);

  const int width = config.camera.width, height = config.camera.height;
  const size_t pixels = (size_t)width * height;
  const double exposure = 0.050, step = config.scene.blur_resolution;
  std::vector<Window> windows;
  std::vector<dvs::CamRayCaster::BlurCell> cells;
  std::vector<int32_t> offsets(1, 0);
  for (int i = 5; i < 8; ++i) {
    Window win;
    win.center = std::stod(argv[i]);
    win.begin = win.center - exposure * .5;
    win.end = win.center + exposure * .5;
    if (!std::isfinite(win.center) || win.begin < 0 || win.end > 60.0)
      throw std::runtime_error("50 ms exposure lies outside the prepared 60 s trajectory");
    win.cells = (size_t)std::ceil(exposure / step - 1e-10);
    for (size_t c = 0; c < win.cells; ++c) {
      const double left = win.begin + c * step, right = std::min(win.end, win.begin + (c + 1) * step);
      if (right <= left) continue;
      const auto state = trajectory.evaluate((left + right) * .5);
      dvs::CamRayCaster::BlurCell cell;
      cell.camera_origin = {state.position.x(), state.position.y(), state.position.z()};
      std::copy(state.orientation.data(), state.orientation.data() + 9, cell.camera_world);
      cell.light_dir = {state.light_direction.x(), state.light_direction.y(), state.light_direction.z()};
      cell.light_intensity = state.light_intensity;
      cell.weight = (right - left) / exposure;
      cells.push_back(cell);
    }
    offsets.push_back((int32_t)cells.size());
    windows.push_back(win);
  }
  std::vector<float> images;
  camera.render_blur(cells, offsets, true, images);
  if (images.size() != windows.size() * pixels) throw std::runtime_error("camera returned an invalid radiance frame");
  std::vector<uint8_t> bytes;
  for (size_t k = 0; k < windows.size(); ++k) {
    const Window &win = windows[k];
    for (size_t c = 0; c < win.cells; ++c)
      if ((c + 1) % 10 == 0 || c + 1 == win.cells)
        std::cout << "SAMPLE center_s=" << win.center << " sample=" << (c + 1) << "/" << win.cells << std::endl;
    dvs::render_float_to_png_bytes(images.data() + k * pixels, width, height, 0, bytes);
    std::ostringstream name;
    name << std::fixed << std::setprecision(3) << "motion_blur_" << win.begin << "_" << win.end << "s.png";
    const auto path = output / name.str();
    if (dvs_io_write_png(path.c_str(), bytes.data(), width, height) != DVS_OK)
      throw std::runtime_error("cannot write " + path.string());
    std::cout << "WINDOW center_s=" << win.center << " begin_s=" << win.begin << " end_s=" << win.end
              << " temporal_samples=" << win.cells << " step_s=" << step << " output=" << path << '\n';
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << "Error: " << e.what() << '\n';
    return 1;
  }
}
