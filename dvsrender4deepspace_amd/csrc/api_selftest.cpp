// Exercises the C++ API (include/dvs_render.hpp over dvs_render_amd.hpp)
// the way the reference's callers use it.
//   dvs_api_selftest --cpu <config.yaml> <out_dir>
//       Config, BVHBuilder, writers, pose conversion, CameraTrajectory, the
//       host physics (advance_pixel, Hapke), EventReadout, the dataset
//       budget (tests/test_dataset_budget.cpp) and frames.h5
//       (tests/test_frames.cpp) (no GPU).
//   dvs_api_selftest --gpu <out_dir>
//       DeviceScene, CamRayCaster and DvsRayCaster on the plane fixture of
//       tests/test_sensor.cu:66-80 (48x36 grid): frame depth, native and
//       adaptive events written to events.bin / events.h5 for comparison;
//       trajectory intervals built on the device driving a DvsRayCaster;
//       EventSensor, OpticalSampler and the diagnostics.
// Prints one line per check and exits non-zero on the first failure.
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "dvs_render.hpp"

namespace {
void expect(bool ok, const char *what) {
  if (!ok) throw std::runtime_error(std::string("check failed: ") + what);
  std::cout << "ok " << what << "\n";
}

dvs::SceneBundle<dvs::float3, dvs::uint3> plane() {
  dvs::SceneBundle<dvs::float3, dvs::uint3> b;
  b.vertices = {{-100, -100, 0}, {100, -100, 0}, {100, 100, 0}, {-100, 100, 0}};
  b.normals.assign(4, dvs::float3{0, 0, 1});
  b.bvh_triangles = {{0, 1, 2}, {0, 2, 3}};
  return b;
}

// tests/test_dataset_budget.cpp: concurrent producers share one allowance,
// sparse files count their logical size, paths stay inside the root, and an
// HDF5 append past the stop is refused before it writes.
void dataset_budget(const std::filesystem::path &root) {
  std::filesystem::create_directories(root);
  dvs::DatasetBudget budget(root, 10000, 9000);
  pid_t children[3];
  for (int i = 0; i < 3; ++i) {
    children[i] = fork();
    if (children[i] < 0) throw std::runtime_error("fork failed");
    if (children[i] == 0) {
      try {
        budget.write_text(root / (std::to_string(i) + ".txt"), std::string(4000, 'x'));
        _exit(0);
      } catch (const dvs::DatasetSizeLimit &) {
        _exit(17);
      } catch (...) {
        _exit(1);
      }
    }
  }
  int accepted = 0, stopped = 0;
  for (pid_t child : children) {
    int status = 0;
    waitpid(child, &status, 0);
    accepted += WIFEXITED(status) && WEXITSTATUS(status) == 0;
    stopped += WIFEXITED(status) && WEXITSTATUS(status) == 17;
  }
  expect(accepted == 2 && stopped == 1, "DatasetBudget: two of three concurrent 4000-byte writers fit 9000");
  expect(budget.logical_bytes() == 8000, "DatasetBudget: aggregate logical bytes");
  bool rejected = false;
  try {
    budget.require_path(root / "../outside.txt");
  } catch (const std::runtime_error &) {
    rejected = true;
  }
  expect(rejected, "DatasetBudget: a path outside the dataset is refused");
  {
    std::ofstream f(root / "sparse.bin");
    f.seekp(4000);
    f.put('x');
  }
  expect(budget.logical_bytes() == 12001, "DatasetBudget: sparse files count their logical size");
  rejected = false;
  try {
    budget.write_text(root / "blocked.txt", "x");
  } catch (const dvs::DatasetSizeLimit &) {
    rejected = true;
  }
  expect(rejected && !std::filesystem::exists(root / "blocked.txt"), "DatasetBudget: nothing is written past the stop");
  std::filesystem::remove_all(root);
  std::filesystem::create_directories(root);
  dvs::DatasetBudget hdf(root, 32000000, 30000000);
  {
    dvs::Hdf5EventStreamWriter writer((root / "events.h5").string(), &hdf);
    writer.append({dvs::Event{.001234, 2, 3, 1, 10}});
    const uint64_t before = hdf.logical_bytes();
    {
      std::ofstream f(root / "filler.bin");
      f.seekp((std::streamoff)(29500000 - before - 1));
      f.put('x');
    }
    const auto size_before = std::filesystem::file_size(root / "events.h5");
    rejected = false;
    try {
      writer.append({dvs::Event{.002, 2, 3, 0, 10}, dvs::Event{.020, 2, 3, 0, 10}, dvs::Event{.030, 2, 3, 0, 10}});
    } catch (const dvs::DatasetSizeLimit &) {
      rejected = true;
    }
    expect(rejected && writer.event_count() == 1 && std::filesystem::file_size(root / "events.h5") == size_before,
           "HDF5 events: an append past the stop is refused unwritten");
    std::filesystem::remove(root / "filler.bin");
    writer.append({dvs::Event{.0021, 2, 3, 0, 10}});
    writer.close(.003);
  }
  int64_t index[8] = {}, count = 0;
  dvs::detail::check_io(dvs_events_h5_ms_to_idx((root / "events.h5").c_str(), index, 8, &count));
  expect(count == 4 && index[0] == 0 && index[1] == 0 && index[2] == 1 && index[3] == 2,
         "HDF5 events: the refused append left the millisecond index intact");
  std::filesystem::remove_all(root);
}

// tests/test_frames.cpp: frames.h5 at a camera's native size, times in
// microseconds, ordered appends, the fixed image encoding, lossless depth.
void frames_file(const std::filesystem::path &root, int width, int height) {
  std::filesystem::create_directories(root);
  dvs::DatasetBudget budget(root);
  const auto path = root / "frames.h5";
  const size_t pixels = (size_t)width * height;
  std::vector<float> light(pixels, 5), depth(pixels, 25);
  depth[0] = -1;
  {
    dvs::Hdf5FrameWriter writer(path, width, height, budget);
    writer.append(0, light, depth);
    writer.append(60, light, depth);
    bool rejected = false;
    try {
      writer.append(59, light, depth);
    } catch (const std::invalid_argument &) {
      rejected = true;
    }
    expect(rejected && writer.count() == 2, "Hdf5FrameWriter refuses an unordered frame");
    writer.close();
  }
  int32_t w = 0, h = 0;
  int64_t n = 0;
  dvs::detail::check_io(dvs_frames_h5_read(path.c_str(), nullptr, nullptr, nullptr, 0, &w, &h, &n));
  std::vector<int64_t> times(2);
  std::vector<uint8_t> images(2 * pixels);
  std::vector<float> read_depth(2 * pixels);
  dvs::detail::check_io(
      dvs_frames_h5_read(path.c_str(), times.data(), images.data(), read_depth.data(), 2, &w, &h, &n));
  expect(n == 2 && w == width && h == height, "frames.h5: native image axes and dimensions");
  expect(times[0] == 0 && times[1] == 60000000, "frames.h5: times in int64 microseconds");
  expect(read_depth[0] == -1 && read_depth[pixels] == -1 && read_depth[1] == 25, "frames.h5: lossless depth");
  std::vector<uint8_t> expected;
  dvs::render_float_to_png_bytes(light.data(), 1, 1, 0, expected);
  expect(images[0] == expected[0] && images[pixels + 1] == expected[0], "frames.h5: the fixed image encoding");
  std::filesystem::remove_all(root);
}

int cpu(const std::string &config_path, const std::string &out) {
  std::string err;
  dvs::Config config(config_path, err);
  expect(err.empty() && config.camera.width == 346 && config.scene.supersample_factor == 16, "Config loads");
  std::string bad;
  dvs::Config missing(out + "/no-such-file.yaml", bad);
  expect(bad.rfind("Cannot open file", 0) == 0, "Config reports a missing file");
  auto b = plane();
  dvs::BVHBuildParams params;
  const int root = dvs::BVHBuilder::build(params, b.bvh_nodes, b.bvh_triangles, b.vertices);
  expect(root == (int)b.bvh_nodes.size() - 1 && b.bvh_nodes[root].count == 2, "BVHBuilder builds one leaf");
  dvs::BVHBuilder::validate(b.bvh_nodes, root);
  bool threw = false;
  try {
    dvs::BVHBuilder::validate(b.bvh_nodes, 7);
  } catch (const std::invalid_argument &) {
    threw = true;
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "BVHBuilder::validate rejects a bad root");
  float q[4];
  const float identity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  dvs::CamRayCaster::quaternion_of(identity, q);
  expect(q[0] == 0 && q[1] == 0 && q[2] == 0 && q[3] == 1, "identity matrix gives the unit quaternion");
  const float half_turn_z[9] = {-1, 0, 0, 0, -1, 0, 0, 0, 1};  // trace -1: largest-diagonal branch
  dvs::CamRayCaster::quaternion_of(half_turn_z, q);
  expect(std::fabs(std::fabs(q[2]) - 1) < 1e-6f && std::fabs(q[3]) < 1e-6f, "half turn about z");
  {
    auto w = dvs::make_event_writer("bin", out);
    w->append({dvs::Event{.001234, 2, 3, 1, 10}, dvs::Event{.0021, 4, 5, 0, 20}});
    w->close(.003);
    expect(w->event_count() == 2, "binary writer counts events");
  }
  {
    auto w = dvs::make_event_writer("hdf5", out);
    w->append({dvs::Event{.001234, 2, 3, 1, 10}, dvs::Event{.0021, 4, 5, 0, 20}});
    bool disorder = false;
    try {
      w->append({dvs::Event{.001, 1, 1, 1, 0}});
    } catch (const std::runtime_error &) {
      disorder = true;
    }
    expect(disorder, "HDF5 writer rejects out-of-order batches");
    w->close(.003);
  }
  threw = false;
  try {
    dvs::make_event_writer("tiff", out);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "make_event_writer rejects unknown formats");
  {
    // tests/test_render.cu:115-130: local mount, world light, seconds
    dvs::FrameData<dvs::Vec3, dvs::Mat3> frame;
    frame.camera_origin = dvs::Vec3(10, 20, 30);
    frame.camera_world(0, 0) = 0, frame.camera_world(1, 0) = 1;  // rotation by pi/2 about z
    frame.camera_world(0, 1) = -1, frame.camera_world(1, 1) = 0;
    frame.light_dir = dvs::Vec3(0, 0, 1);
    frame.light_intensity = 1000;
    frame.timestamp_seconds = .125;
    auto last = frame;
    last.timestamp_seconds = .5;
    dvs::SceneConfig sc;
    sc.fps = 100;
    sc.translation = {1, 0, 0};
    sc.rotation = {3.14159265f / 2, 0, 0};
    dvs::CameraTrajectory trajectory({frame, last}, sc);
    const auto pose = trajectory.evaluate(.125f);
    const float dp = std::fabs(pose.position.x() - 10) + std::fabs(pose.position.y() - 21) +
                     std::fabs(pose.position.z() - 30);
    // Rz(pi/2) Rx(pi/2): columns (0,1,0), (0,0,1), (1,0,0)
    const float expected[9] = {0, 1, 0, 0, 0, 1, 1, 0, 0};
    float dr = 0;
    for (int k = 0; k < 9; ++k) dr += std::fabs(pose.orientation.m[k] - expected[k]);
    expect(dp < 1e-5f && dr < 1e-5f && pose.light_direction.z() == 1 && trajectory.start_time() == .125f &&
               trajectory.end_time() == .5f && trajectory.size() == 2,
           "CameraTrajectory mount, light and seconds");
    threw = false;
    try {
      dvs::CameraTrajectory duplicate({frame, frame}, sc);
    } catch (const std::invalid_argument &e) {
      threw = std::string(e.what()) == "trajectory timestamps must be strictly increasing";
    }
    expect(threw, "CameraTrajectory rejects duplicate timestamps");
  }
  {
    // tests/test_sensor.cu:37-42: the analytic ON crossing of a step from
    // 1000 to 2000 e-/s, charge 10, thresholds 0.2 (TestRng: normal() = 0)
    struct TestRng {
      double normal() { return 0; }
    } rng;
    dvs::PixelParameters p;
    p.threshold_pos = p.threshold_neg = .2;
    p.charge_scale = 10;
    p.refractory_time = 0;
    dvs::PixelState st;
    st.receptor_log = st.reference_log = std::log(1000.0);
    st.threshold_pos = st.threshold_neg = .2;
    dvs::PixelEvent ev[8];
    const int n = dvs::advance_pixel(st, p, 2000, 0, .1, ev, 8, rng);
    const double target = 1000 * std::exp(.2);
    const double expected = -10.0 / 2000 * std::log((1 / target - 1 / 2000.0) / (1 / 1000.0 - 1 / 2000.0));
    expect(n == 3 && ev[0].polarity == 1 && std::fabs(ev[0].time - expected) < 1e-12,
           "advance_pixel: analytic step crossing");
    expect(dvs::advance_pixel(st, p, 2000, .1, .2, ev, 0, rng) >= 0 &&
               std::fabs(dvs::receptor_log_at(std::log(1000.0), 0, 10, 10) - (std::log(1000.0) - std::log1p(1000.0)))
                   < 1e-12,
           "receptor_log_at: zero-current closed form");
    dvs::HapkeConfig h;
    const float ci = std::cos(30 * 3.14159265f / 180);
    const float target_if = 3.14159265f * dvs::hapke_reflectance(ci, 1, ci, .324f, h);
    expect(std::fabs(dvs::invert_hapke_if(30, 0, 30, target_if, h) - .324f) < 2e-6f &&
               std::isnan(dvs::invert_hapke_if(20, 10, 90, .1f, h)),
           "Hapke I/F round trip and invalid geometry");
    const float dir = dvs::hapke_directional_reflectance(dvs::make_float3(0, 0, 1), dvs::make_float3(0, 0, 1),
                                                         dvs::make_float3(0, 0, 1), h.w, h);
    expect(dir > 0 && std::isfinite(dir), "Hapke directional reflectance at opposition");
    // tests/test_sensor.cu:176-177
    dvs::EventReadout bus(100, .001, 2);
    std::vector<dvs::Event> ready;
    bus.submit(std::vector<dvs::Event>(3, dvs::Event{0, 0, 0, 1, 0}), 0, ready);
    bus.drain(ready);
    expect(bus.dropped() == 1 && ready.size() == 2 && std::fabs(ready[0].timestamp - .011) < 1e-12 &&
               bus.peak_queue() == 2,
           "EventReadout drops and delivery time");
  }
  dataset_budget(std::filesystem::path(out) / "budget");
  frames_file(std::filesystem::path(out) / "frames346", config.camera.width, config.camera.height);
  frames_file(std::filesystem::path(out) / "frames1280", 1280, 720);
  return 0;
}

int gpu(const std::string &out) {
  auto bundle = plane();
  dvs::BVHBuildParams params;
  const int root = dvs::BVHBuilder::build(params, bundle.bvh_nodes, bundle.bvh_triangles, bundle.vertices);
  std::string err;
  // The fixture's optics and circuit (48x36 grid of the DAVIS field of view,
  // one sub-ray; thresholds 0.28, 64 events per pixel, no FIFO delay).
  dvs::Config config(std::string(), err);
  const double scale = 48 / 346.0;  // intrinsics scaled in double, rounded once
  config.camera = {48, 36, (float)(226.3802 * scale), (float)(226.1500 * scale), (float)(173.6471 * scale),
                   (float)(133.7327 * scale)};
  config.scene.supersample_factor = 1;
  config.hapke = dvs::HapkeConfig{};
  config.physic.threshold_pos = config.physic.threshold_neg = .28f;
  config.physic.threshold_std = 0;
  config.physic.photon_shot_noise = 0;
  config.physic.max_events_per_step = 64;
  config.physic.readout_bandwidth = 0;
  config.physic.readout_latency = 0;
  config.physic.random_seed = 42;
  auto scene = std::make_shared<dvs::DeviceScene>(bundle, config, root);
  dvs::CamRayCaster cam(scene);
  const float identity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::vector<float> image, depth;
  cam.render_main(dvs::make_float3(0, 0, 10), identity, dvs::make_float3(0, 0, 1), 1000, image, depth);
  const int centre = 18 * 48 + 24;
  expect(std::fabs(depth[centre] - 10) < 0.05f && image[centre] > 0, "nadir frame: depth 10 m at the centre");
  {
    // One exposure of two cells (both the nadir pose, weights 1/4 and 3/4)
    // and one of a single unit cell: the float sums of render_main frames.
    dvs::CamRayCaster::BlurCell cell;
    cell.camera_origin = dvs::make_float3(0, 0, 10);
    std::copy(identity, identity + 9, cell.camera_world);
    cell.light_dir = dvs::make_float3(0, 0, 1);
    cell.light_intensity = 1000;
    std::vector<dvs::CamRayCaster::BlurCell> cells(3, cell);
    cells[0].weight = .25;
    cells[1].weight = .75;
    cells[2].weight = 1;
    std::vector<float> blurred;
    cam.render_blur(cells, {0, 2, 3}, false, blurred);
    bool same = blurred.size() == 2 * image.size();
    for (size_t i = 0; same && i < image.size(); ++i)
      same = blurred[i] == image[i] * .25f + image[i] * .75f && blurred[image.size() + i] == image[i];
    expect(same, "render_blur sums render_main frames");
  }

  dvs::CameraTrajectoryState s;
  s.position.v[2] = 10;
  s.light_intensity = 1000;
  dvs::DvsRayCaster native(scene, config), adaptive(scene, config);
  native.initialize(s, 0.0);
  adaptive.initialize(s, 0.0);
  const dvs::OpticalState pose = dvs::to_optical_state(s);
  dvs::OpticalState start = pose, end = pose;
  start.intensity_rate = 8e6f;
  end.intensity_rate = -8e6f;
  std::vector<dvs::Event> ev_native, ev_adaptive, tail;
  native.advance_many({{pose, pose, 0, .002}}, ev_native);
  expect(ev_native.empty(), "static scene: no startup burst");
  native.advance_many({{start, end, .002, .004}}, ev_native);
  expect(!ev_native.empty(), "Hermite pulse emits events");
  bool ok = adaptive.advance_adaptive({{pose, pose, 0, .002}}, 1e-4, ev_adaptive, true);
  expect(ok && ev_adaptive.empty(), "adaptive static window accepted by the receptor check");
  ok = adaptive.advance_adaptive({{start, end, .002, .004}}, 1e-4, ev_adaptive);
  expect(ok && !ev_adaptive.empty(), "adaptive pulse window emits events");
  const auto stats = adaptive.refinement_statistics();
  expect(stats.traced_pixel_samples >= 5u * 48 * 36 && adaptive.adaptive_cache_capacity() >= 128,
         "refinement statistics");
  native.drain(tail);
  expect(native.dropped_events() == 0, "no FIFO drops");
  bool threw = false;
  try {
    native.advance_many({{pose, pose, .01, .011}}, tail);  // not contiguous with .004
  } catch (const std::invalid_argument &e) {
    threw = std::string(e.what()).find("contiguous") != std::string::npos;
  }
  expect(threw, "non-contiguous intervals throw std::invalid_argument");
  dvs::BinaryEventStreamWriter bin(out + "/native.bin");
  bin.append(ev_native);
  bin.close(.004);
  dvs::Hdf5EventStreamWriter h5(out + "/adaptive.h5");
  h5.append(ev_adaptive);
  h5.close(.004);
  std::cout << "events native " << ev_native.size() << " adaptive " << ev_adaptive.size() << "\n";
  {
    // bin/dvs's loop over a trajectory: device-built intervals at 10 us
    // horizons, advanced in one call, equal to the host evaluation's.
    dvs::FrameData<dvs::Vec3, dvs::Mat3> a, b;
    a.camera_origin = dvs::Vec3(0, 0, 10);
    b.camera_origin = dvs::Vec3(.5f, 0, 10);
    a.light_dir = b.light_dir = dvs::Vec3(0, 0, 1);
    a.light_intensity = b.light_intensity = 1000;
    b.frame_index = 1;
    dvs::SceneConfig sc;
    sc.fps = 100;
    dvs::CameraTrajectory trajectory({a, b}, sc);
    std::vector<double> times;
    for (int k = 0; k <= 200; ++k) times.push_back(std::min<double>(trajectory.end_time(), k * (double)1e-5f));
    const auto ivs = trajectory.intervals(times);
    bool same = ivs.size() == 200;
    for (size_t k = 0; same && k < ivs.size(); k += 37) {
      const auto h = dvs::to_optical_state(trajectory.evaluate((float)times[k]));
      same = std::fabs(h.position.x - ivs[k].start.position.x) < 1e-5f && ivs[k].t0 == times[k] &&
             std::fabs(h.velocity.x - ivs[k].start.velocity.x) < 1e-4f;
    }
    expect(same, "device trajectory intervals equal the host evaluation");
    dvs::DvsRayCaster moving(scene, config);
    moving.initialize(trajectory.evaluate(0), 0.0);
    std::vector<dvs::Event> ev;
    moving.advance_many(ivs, ev);
    moving.drain(ev);
    expect(moving.dropped_events() == 0, "trajectory window advances");
    // two adaptive windows of 100 horizons along the trajectory, the second
    // expected (set up on the device at the end of the first) or not: the
    // same events
    std::vector<std::vector<dvs::Event>> runs;
    for (int form = 0; form < 2; ++form) {
      dvs::DvsRayCaster caster(scene, config);
      caster.initialize(trajectory.evaluate(0), 0.0);
      std::vector<dvs::Event> got;
      const std::vector<double> first(times.begin(), times.begin() + 101), second(times.begin() + 100, times.end());
      if (form == 1) caster.expect_adaptive(trajectory, second, 1e-2);
      bool ok = caster.advance_adaptive(trajectory, first, 1e-2, got);
      ok = ok && caster.advance_adaptive(trajectory, second, 1e-2, got);
      caster.drain(got);
      expect(ok, "adaptive trajectory windows accepted");
      runs.push_back(got);
    }
    bool equal = runs[0].size() == runs[1].size();
    for (size_t k = 0; equal && k < runs[0].size(); ++k)
      equal = runs[0][k].timestamp == runs[1][k].timestamp && runs[0][k].x == runs[1][k].x &&
              runs[0][k].y == runs[1][k].y && runs[0][k].polarity == runs[1][k].polarity;
    expect(equal, "a window set up ahead gives the same events");
  }
  {
    // EventSensor / OpticalSampler on the same scene view, against the
    // casters above: the same frame and the same native events.
    dvs::OpticalSampler sampler(48, 36, 1);
    const std::vector<dvs::OpticalState> poses{pose};
    const dvs::OpticalValue *values = sampler.evaluate(scene->view(), poses);
    std::vector<dvs::OpticalValue> down;
    sampler.download(down);
    expect(down.size() == image.size() && (float)values[centre].radiance == image[centre] &&
               values[centre].depth == depth[centre] && values[centre].samples == 1,
           "OpticalSampler evaluates the frame of CamRayCaster");
    dvs::EventSensor sensor(48, 36, config.physic);
    sensor.initialize(scene->view(), pose, 0.0);
    std::vector<dvs::Event> ev, again;
    sensor.process_many(scene->view(), {{pose, pose, 0, .002}}, ev);
    sensor.process(scene->view(), {start, end, .002, .004}, ev);
    dvs::DvsRayCaster twin(scene, config);
    twin.initialize(s, 0.0);
    twin.advance_many({{pose, pose, 0, .002}}, again);
    twin.advance_many({{start, end, .002, .004}}, again);
    expect(ev.size() == again.size() && !ev.empty() &&
               std::memcmp(ev.data(), again.data(), ev.size() * sizeof(dvs::Event)) == 0,
           "EventSensor gives DvsRayCaster's events");
    dvs::EventSensor adaptive_sensor(48, 36, config.physic);
    adaptive_sensor.enable_numerical_diagnostics();
    adaptive_sensor.initialize(scene->view(), pose, 0.0);
    bool threw_state = false;
    try {
      adaptive_sensor.verify_optics(scene->view(), sampler, {{pose, pose, 0, .002}}, 1e-3);
    } catch (const std::logic_error &) {
      threw_state = true;
    }
    expect(threw_state, "verify_optics before refinement throws std::logic_error");
    const std::vector<dvs::OpticalInterval> window{{pose, pose, 0, .002}};
    expect(adaptive_sensor.process_adaptive(scene->view(), window, sampler, 1e-3, ev) && ev.empty() &&
               sampler.statistics().maximum_knots == 5 && sampler.knot_capacity() >= 128,
           "EventSensor::process_adaptive with an OpticalSampler");
    const auto optics = adaptive_sensor.verify_optics(scene->view(), sampler, window, 1e-3);
    expect(optics.probes == 1024 * 17 && optics.maximum_log_error == 0 && optics.over_tolerance == 0,
           "verify_optics: static knots are exact");
    const auto receptor = adaptive_sensor.verify_receptor(scene->view(), sampler, window);
    // ceil(0.002 s / 1e-6f) cells: the float circuit step is just below 1 us
    expect(receptor.pixels == 1024 && receptor.circuit_cells == 2001 && receptor.maximum_log_voltage_error < 1e-9,
           "verify_receptor after enable_numerical_diagnostics");
    dvs::DvsRayCaster diag(scene, config);
    diag.enable_numerical_diagnostics();
    diag.initialize(s, 0.0);
    expect(diag.advance_adaptive(window, 1e-3, ev) && diag.verify_optics(window, 1e-3).over_tolerance == 0 &&
               diag.verify_receptor(window).pixels == 1024,
           "DvsRayCaster diagnostics");
  }
  return 0;
}
}  // namespace

int main(int argc, char **argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "--cpu" && argc == 4) return cpu(argv[2], argv[3]);
    if (mode == "--gpu" && argc == 3) return gpu(argv[2]);
    std::cerr << "usage: dvs_api_selftest --cpu <config.yaml> <out_dir> | --gpu <out_dir>\n";
    return 2;
  } catch (const std::exception &e) {
    std::cerr << "FAIL: " << e.what() << "\n";
    return 1;
  }
}
