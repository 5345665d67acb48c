// Host I/O of the DVS path (libdvs_io.so): event stream writers and the YAML
// configuration, behind the C ABI of include/dvs_io.h.
//
// Formats follow include/dvs/dvs_writer.hpp (13-byte binary records; HDF5
// x/y/t/p datasets with microsecond times and a millisecond index) and the
// configuration grammar and checks of src/scene_types.cpp.
#include <hdf5.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dvs_io.h"

namespace {

thread_local std::string g_error;

struct IoFailure : std::runtime_error {
  int code;
  IoFailure(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string &msg) { throw IoFailure(code, msg); }

template <class F>
int guarded(F &&f) {
  try {
    f();
    g_error.clear();
    return DVS_OK;
  } catch (const IoFailure &e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_error = e.what();
    return DVS_E_STATE;
  }
}

// ---------------------------------------------------------------------------
// Event stream writers
// ---------------------------------------------------------------------------
class Writer {
 public:
  virtual ~Writer() = default;
  virtual void append(std::vector<dvs_event> events) = 0;
  virtual void close(double end_time_seconds) = 0;
  int64_t count = 0;
};

#pragma pack(push, 1)
struct BinRecord {
  float t;
  int16_t x, y;
  int8_t p;
  float depth;
};
#pragma pack(pop)
static_assert(sizeof(BinRecord) == 13, "binary event record is 13 bytes");
static_assert(sizeof(dvs_event) == 17, "event record is 17 bytes");

class BinWriter final : public Writer {
 public:
  explicit BinWriter(const std::string &path) : path_(path), out_(path, std::ios::binary) {
    if (!out_) fail(DVS_E_STATE, "failed to open output file: " + path);
  }
  void append(std::vector<dvs_event> events) override {
    if (events.empty()) return;
    std::vector<BinRecord> rec(events.size());
    for (size_t i = 0; i < events.size(); ++i) {
      const dvs_event &e = events[i];
      rec[i] = BinRecord{(float)e.timestamp, e.x, e.y, e.polarity, e.depth};
    }
    out_.write(reinterpret_cast<const char *>(rec.data()), (std::streamsize)(rec.size() * sizeof(BinRecord)));
    out_.flush();
    if (!out_) fail(DVS_E_STATE, "failed to write output file: " + path_);
    count += (int64_t)events.size();
  }
  void close(double) override { out_.close(); }

 private:
  std::string path_;
  std::ofstream out_;
};

class H5Writer final : public Writer {
 public:
  explicit H5Writer(const std::string &path) {
    file_ = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file_ < 0) fail(DVS_E_STATE, "failed to create HDF5 file: " + path);
    sets_[0] = column("x", H5T_NATIVE_INT16);
    sets_[1] = column("y", H5T_NATIVE_INT16);
    sets_[2] = column("t", H5T_NATIVE_INT64);
    sets_[3] = column("p", H5T_NATIVE_INT8);
    check_rc(H5Fflush(file_, H5F_SCOPE_GLOBAL), "failed to flush new event file");
  }
  ~H5Writer() override { release(); }

  void append(std::vector<dvs_event> events) override {
    if (closed_) fail(DVS_E_STATE, "cannot append to closed HDF5 event file");
    if (events.empty()) return;
    for (const dvs_event &e : events)
      if (!std::isfinite(e.timestamp) || e.timestamp < 0 ||
          e.timestamp * 1000 > std::numeric_limits<int>::max() - 1)
        fail(DVS_E_STATE, "HDF5 event timestamp is outside millisecond-index bounds");
    std::stable_sort(events.begin(), events.end(),
                     [](const dvs_event &a, const dvs_event &b) { return a.timestamp < b.timestamp; });
    if (events.front().timestamp < last_time_) fail(DVS_E_STATE, "HDF5 events must be globally time ordered");
    const size_t n = events.size();
    std::vector<int16_t> xs(n), ys(n);
    std::vector<int64_t> ts(n);
    std::vector<int8_t> ps(n);
    for (size_t i = 0; i < n; ++i) {
      xs[i] = events[i].x;
      ys[i] = events[i].y;
      ts[i] = std::llround(events[i].timestamp * 1e6);
      ps[i] = events[i].polarity;
    }
    extend(sets_[0], xs.data(), H5T_NATIVE_INT16, n);
    extend(sets_[1], ys.data(), H5T_NATIVE_INT16, n);
    extend(sets_[2], ts.data(), H5T_NATIVE_INT64, n);
    extend(sets_[3], ps.data(), H5T_NATIVE_INT8, n);
    for (size_t i = 0; i < n; ++i) index_until(ts[i] / 1000, count + (int64_t)i);
    last_time_ = events.back().timestamp;
    count += (int64_t)n;
  }

  void close(double end_time) override {
    if (closed_) return;
    if (!std::isfinite(end_time) || end_time * 1000 > std::numeric_limits<int>::max() - 1)
      fail(DVS_E_STATE, "invalid HDF5 observation end time");
    if (end_time >= 0) index_until(std::llround(end_time * 1000), count);
    const hsize_t dims[1] = {(hsize_t)ms_to_idx_.size()};
    const hid_t space = H5Screate_simple(1, dims, nullptr);
    check_id(space, "failed to create HDF5 ms_to_idx dataspace");
    const hid_t set = H5Dcreate2(file_, "ms_to_idx", H5T_NATIVE_INT64, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(space);
    check_id(set, "failed to create HDF5 dataset ms_to_idx");
    if (!ms_to_idx_.empty() &&
        H5Dwrite(set, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, ms_to_idx_.data()) < 0) {
      H5Dclose(set);
      fail(DVS_E_STATE, "failed to write HDF5 dataset ms_to_idx");
    }
    H5Dclose(set);
    check_rc(H5Fflush(file_, H5F_SCOPE_GLOBAL), "failed to flush final event index");
    for (hid_t &s : sets_) {
      if (s >= 0) H5Dclose(s);
      s = -1;
    }
    const herr_t rc = H5Fclose(file_);
    file_ = -1;
    check_rc(rc, "failed to close event file");
    closed_ = true;
  }

 private:
  static void check_rc(herr_t rc, const char *msg) {
    if (rc < 0) fail(DVS_E_STATE, msg);
  }
  static void check_id(hid_t id, const std::string &msg) {
    if (id < 0) fail(DVS_E_STATE, msg);
  }
  hid_t column(const char *name, hid_t type) {
    const hsize_t dims[1] = {0}, max_dims[1] = {H5S_UNLIMITED}, chunk[1] = {kChunk};
    const hid_t space = H5Screate_simple(1, dims, max_dims);
    check_id(space, std::string("failed to create HDF5 dataspace for ") + name);
    const hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
    check_id(plist, std::string("failed to create HDF5 property list for ") + name);
    if (H5Pset_chunk(plist, 1, chunk) < 0) {
      H5Pclose(plist);
      H5Sclose(space);
      fail(DVS_E_STATE, std::string("failed to set HDF5 chunking for ") + name);
    }
    const hid_t set = H5Dcreate2(file_, name, type, space, H5P_DEFAULT, plist, H5P_DEFAULT);
    H5Pclose(plist);
    H5Sclose(space);
    check_id(set, std::string("failed to create HDF5 dataset ") + name);
    return set;
  }
  void extend(hid_t set, const void *data, hid_t type, size_t n) {
    const hsize_t start[1] = {(hsize_t)count}, grown[1] = {(hsize_t)(count + (int64_t)n)}, block[1] = {(hsize_t)n};
    check_rc(H5Dset_extent(set, grown), "failed to extend HDF5 event dataset");
    const hid_t file_space = H5Dget_space(set);
    check_id(file_space, "failed to open HDF5 event dataspace");
    const hid_t mem_space = H5Screate_simple(1, block, nullptr);
    herr_t rc = H5Sselect_hyperslab(file_space, H5S_SELECT_SET, start, nullptr, block, nullptr);
    if (rc >= 0) rc = H5Dwrite(set, type, mem_space, file_space, H5P_DEFAULT, data);
    H5Sclose(mem_space);
    H5Sclose(file_space);
    check_rc(rc, "failed to write HDF5 event dataset");
  }
  // ms_to_idx[m] for every millisecond m up to `ms`: the first event index
  // at or after m ms.
  void index_until(int64_t ms, int64_t event_index) {
    if (ms < 0) return;
    if (ms > (int64_t)std::numeric_limits<int>::max()) fail(DVS_E_STATE, "ms_to_idx would be unreasonably large");
    while ((int64_t)ms_to_idx_.size() <= ms) ms_to_idx_.push_back(event_index);
  }
  void release() {
    for (hid_t &s : sets_) {
      if (s >= 0) H5Dclose(s);
      s = -1;
    }
    if (file_ >= 0) H5Fclose(file_);
    file_ = -1;
  }
  static constexpr hsize_t kChunk = 1 << 16;
  hid_t file_ = -1;
  hid_t sets_[4] = {-1, -1, -1, -1};
  std::vector<int64_t> ms_to_idx_;
  double last_time_ = -std::numeric_limits<double>::infinity();
  bool closed_ = false;
};

// Reads a whole 1-D dataset.
template <class T>
std::vector<T> read_column(hid_t file, const char *name, hid_t type) {
  const hid_t set = H5Dopen2(file, name, H5P_DEFAULT);
  if (set < 0) fail(DVS_E_ARGUMENT, std::string("missing HDF5 dataset ") + name);
  const hid_t space = H5Dget_space(set);
  hsize_t dims[1] = {0};
  H5Sget_simple_extent_dims(space, dims, nullptr);
  std::vector<T> out(dims[0]);
  herr_t rc = 0;
  if (dims[0]) rc = H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data());
  H5Sclose(space);
  H5Dclose(set);
  if (rc < 0) fail(DVS_E_STATE, std::string("failed to read HDF5 dataset ") + name);
  return out;
}

struct H5File {
  hid_t id;
  explicit H5File(const char *path) : id(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) {
    if (id < 0) fail(DVS_E_ARGUMENT, std::string("failed to open HDF5 file: ") + path);
  }
  ~H5File() { H5Fclose(id); }
};

// ---------------------------------------------------------------------------
// Configuration: "section:" lines at column 0 open a mapping; indented
// "key: value" lines fill it; '#' starts a comment outside quotes.
// ---------------------------------------------------------------------------
using Section = std::map<std::string, std::string>;
using Document = std::map<std::string, Section>;

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string strip_comment(const std::string &v) {
  bool single = false, dbl = false;
  for (size_t i = 0; i < v.size(); ++i) {
    const char c = v[i];
    if (single) {
      if (c == '\'') single = false;
    } else if (dbl) {
      if (c == '"') dbl = false;
    } else if (c == '\'') {
      single = true;
    } else if (c == '"') {
      dbl = true;
    } else if (c == '#') {
      return v.substr(0, i);
    }
  }
  return v;
}

Document parse_document(const std::string &path) {
  std::ifstream in(path);
  if (!in.is_open()) fail(DVS_E_ARGUMENT, "Cannot open file: " + path);
  Document doc;
  std::string line, section;
  int number = 0;
  while (std::getline(in, line)) {
    ++number;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    size_t start = 0;
    while (start < line.size() && std::isspace((unsigned char)line[start])) ++start;
    if (start == line.size() || line[start] == '#') continue;
    const size_t colon = line.find(':', start);
    const std::string where = "Line " + std::to_string(number);
    if (colon == std::string::npos) fail(DVS_E_ARGUMENT, where + ": missing colon");
    const std::string key = trim(line.substr(start, colon - start));
    if (key.empty()) fail(DVS_E_ARGUMENT, where + ": empty key");
    const std::string value = trim(strip_comment(line.substr(colon + 1)));
    if (start == 0) {
      if (!value.empty())
        fail(DVS_E_ARGUMENT, where + ": top-level key '" + key + "' has a value, expected a mapping");
      section = key;
      doc[section];
    } else {
      if (section.empty()) fail(DVS_E_ARGUMENT, where + ": key '" + key + "' without a parent section");
      doc[section][key] = value;
    }
  }
  return doc;
}

template <class T>
bool parse_number(const std::string &s, T &out) {
  const char *first = s.data(), *last = s.data() + s.size();
  const auto r = std::from_chars(first, last, out);
  return r.ec == std::errc{} && r.ptr == last;
}

bool parse_triple(const std::string &s, float out[3]) {
  if (s.size() < 2 || s.front() != '[' || s.back() != ']') return false;
  std::stringstream items(s.substr(1, s.size() - 2));
  std::string item;
  int n = 0;
  while (std::getline(items, item, ',')) {
    float v;
    if (!parse_number(trim(item), v)) return false;
    if (n < 3) out[n] = v;
    ++n;
  }
  return n == 3;
}

template <class T>
void required(const Section &sec, const char *section, const char *key, T &out) {
  const auto it = sec.find(key);
  if (it == sec.end()) fail(DVS_E_ARGUMENT, std::string("Missing key '") + key + "' in section '" + section + "'");
  if (!parse_number(it->second, out))
    fail(DVS_E_ARGUMENT, std::string("Invalid value for key '") + key + "' in section '" + section + "': '" +
                             it->second + "'");
}

template <class T>
void optional(const Section &sec, const char *key, T &out) {
  const auto it = sec.find(key);
  if (it != sec.end() && !parse_number(it->second, out))
    fail(DVS_E_ARGUMENT, std::string("invalid physic parameter: ") + key);
}

dvs_physic default_physic() {  // PhysicConfig member defaults (scene_types.hpp:64-79)
  dvs_physic p{};
  p.threshold_pos = .3f;
  p.threshold_neg = .3f;
  p.refractory_time = .0002f;
  p.readout_bandwidth = 5e6f;
  p.readout_latency = 0;
  p.threshold_std = .03f;
  p.threshold_min = .01f;
  p.wavelength_nm = 550;
  p.quantum_efficiency = .3f;
  p.pixel_pitch_um = 15;
  p.fill_factor = .2f;
  p.f_number = 2;
  p.lens_transmission = .8f;
  p.radiance_scale = .001f;
  p.temperature_k = 300;
  p.subthreshold_kappa = .8f;
  p.receptor_capacitance_f = 1e-15f;
  p.reset_capacitance_f = 20e-15f;
  p.dark_current_a = 1e-18f;
  p.reset_leakage_a = 0;
  p.circuit_step = 1e-6f;
  p.photon_shot_noise = 1;
  p.reset_thermal_noise = 0;
  p.random_seed = 114514;
  p.max_events_per_step = 16;
  p.readout_queue_capacity = 1000000;
  return p;
}

void load_config(const std::string &path, dvs_config &c) {
  Document doc = parse_document(path);
  for (const char *s : {"intrinsic", "physic", "scene"})
    if (!doc.count(s)) fail(DVS_E_ARGUMENT, std::string("Missing top-level section '") + s + "'");
  c = dvs_config{};
  const Section &cam = doc["intrinsic"];
  required(cam, "intrinsic", "width", c.camera.width);
  required(cam, "intrinsic", "height", c.camera.height);
  required(cam, "intrinsic", "fx", c.camera.fx);
  required(cam, "intrinsic", "fy", c.camera.fy);
  required(cam, "intrinsic", "cx", c.camera.cx);
  required(cam, "intrinsic", "cy", c.camera.cy);

  dvs_physic &p = c.physic;
  p = default_physic();
  const Section &ph = doc["physic"];
  required(ph, "physic", "threshold_pos", p.threshold_pos);
  required(ph, "physic", "threshold_neg", p.threshold_neg);
  required(ph, "physic", "refractory_time", p.refractory_time);
  required(ph, "physic", "readout_latency", p.readout_latency);
  required(ph, "physic", "readout_bandwidth", p.readout_bandwidth);
  for (const char *key : {"threshold_noise", "dark_current", "hpf_time_constant", "shot_intensity",
                          "photoreceptor_tau", "intensity_reference", "min_bandwidth_fraction", "leak_rate_hz",
                          "shot_noise_rate_hz", "shot_noise_bright_factor"})
    if (ph.count(key))
      fail(DVS_E_ARGUMENT, std::string("removed physic key '") + key +
                               "'; migrate to dimensioned parameters in docs/physical_sensor.md");
  optional(ph, "threshold_std", p.threshold_std);
  optional(ph, "threshold_min", p.threshold_min);
  optional(ph, "wavelength_nm", p.wavelength_nm);
  optional(ph, "quantum_efficiency", p.quantum_efficiency);
  optional(ph, "pixel_pitch_um", p.pixel_pitch_um);
  optional(ph, "fill_factor", p.fill_factor);
  optional(ph, "f_number", p.f_number);
  optional(ph, "lens_transmission", p.lens_transmission);
  optional(ph, "radiance_scale", p.radiance_scale);
  optional(ph, "temperature_k", p.temperature_k);
  optional(ph, "subthreshold_kappa", p.subthreshold_kappa);
  optional(ph, "receptor_capacitance_f", p.receptor_capacitance_f);
  optional(ph, "reset_capacitance_f", p.reset_capacitance_f);
  optional(ph, "dark_current_a", p.dark_current_a);
  optional(ph, "reset_leakage_a", p.reset_leakage_a);
  optional(ph, "circuit_step", p.circuit_step);
  optional(ph, "photon_shot_noise", p.photon_shot_noise);
  optional(ph, "reset_thermal_noise", p.reset_thermal_noise);
  optional(ph, "random_seed", p.random_seed);
  optional(ph, "max_events_per_step", p.max_events_per_step);
  optional(ph, "readout_queue_capacity", p.readout_queue_capacity);
  for (float v : {p.threshold_pos, p.threshold_neg, p.threshold_std, p.threshold_min, p.refractory_time,
                  p.readout_bandwidth, p.readout_latency, p.wavelength_nm, p.quantum_efficiency, p.pixel_pitch_um,
                  p.fill_factor, p.f_number, p.lens_transmission, p.radiance_scale, p.temperature_k,
                  p.subthreshold_kappa, p.receptor_capacitance_f, p.reset_capacitance_f, p.dark_current_a,
                  p.reset_leakage_a, p.circuit_step})
    if (!std::isfinite(v) || v < 0) fail(DVS_E_ARGUMENT, "physic values must be finite and nonnegative");
  if (p.threshold_pos <= 0 || p.threshold_neg <= 0 || p.threshold_min <= 0 || p.wavelength_nm <= 0 ||
      p.pixel_pitch_um <= 0 || p.f_number <= 0 || p.temperature_k <= 0 || p.subthreshold_kappa <= 0 ||
      p.subthreshold_kappa > 1 || p.receptor_capacitance_f <= 0 || p.reset_capacitance_f <= 0 ||
      p.circuit_step <= 0 || p.quantum_efficiency > 1 || p.fill_factor > 1 || p.lens_transmission > 1 ||
      p.photon_shot_noise < 0 || p.photon_shot_noise > 1 || p.reset_thermal_noise < 0 || p.reset_thermal_noise > 1 ||
      p.random_seed < 0 || p.max_events_per_step < 1 || p.max_events_per_step > 64 || p.readout_queue_capacity < 1)
    fail(DVS_E_ARGUMENT, "invalid optical/circuit model parameter range");

  dvs_scene_config &s = c.scene;
  const Section &sc = doc["scene"];
  required(sc, "scene", "frame_count", s.frame_count);
  required(sc, "scene", "fps", s.fps);
  required(sc, "scene", "imu_fps", s.imu_fps);
  required(sc, "scene", "cam_fps", s.cam_fps);
  required(sc, "scene", "time_resolution", s.time_resolution);
  required(sc, "scene", "supersample_factor", s.supersample_factor);
  required(sc, "scene", "blur_resolution", s.blur_resolution);
  for (int k = 0; k < 3; ++k) s.translation[k] = s.rotation[k] = 0;
  for (auto [key, dst] : {std::make_pair("translation", s.translation), std::make_pair("rotation", s.rotation)}) {
    const auto it = sc.find(key);
    if (it != sc.end() && !parse_triple(it->second, dst))
      fail(DVS_E_ARGUMENT, std::string("scene.") + key + " must have three numbers");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(dst[k])) fail(DVS_E_ARGUMENT, "nonfinite extrinsics");
  }
  for (float v : {c.camera.fx, c.camera.fy, c.camera.cx, c.camera.cy, s.fps, s.time_resolution, s.cam_fps, s.imu_fps,
                  s.blur_resolution})
    if (!std::isfinite(v)) fail(DVS_E_ARGUMENT, "nonfinite camera or scene parameter");
  if (c.camera.width <= 0 || c.camera.height <= 0 || c.camera.width > 32767 || c.camera.height > 32767 ||
      c.camera.fx <= 0 || c.camera.fy <= 0 || s.frame_count < 2 || s.fps <= 0 || s.time_resolution <= 0 ||
      s.cam_fps <= 0 || s.imu_fps <= 0 || s.blur_resolution <= 0 || s.supersample_factor < 1)
    fail(DVS_E_ARGUMENT, "invalid scene or camera range");
  const double samples = (double)c.camera.width * c.camera.height * s.supersample_factor * s.supersample_factor;
  if (samples > std::numeric_limits<int>::max()) fail(DVS_E_ARGUMENT, "too many aperture samples");
  c.camera.samples = s.supersample_factor;

  dvs_hapke &h = c.hapke;
  const Section &hk = doc["hapke"];
  required(hk, "hapke", "b", h.b);
  required(hk, "hapke", "c", h.c);
  required(hk, "hapke", "hS", h.hS);
  required(hk, "hapke", "hC", h.hC);
  required(hk, "hapke", "thetap", h.thetap);
  required(hk, "hapke", "w", h.w);
  required(hk, "hapke", "B_S0", h.B_S0);
  required(hk, "hapke", "B_C0", h.B_C0);
  required(hk, "hapke", "phi", h.phi);
  for (float v : {h.b, h.c, h.hS, h.hC, h.thetap, h.w, h.B_S0, h.B_C0, h.phi})
    if (!std::isfinite(v)) fail(DVS_E_ARGUMENT, "nonfinite Hapke parameter");
  constexpr float kHalfPi = 3.14159265358979323846f / 2;
  if (h.b < 0 || h.b >= 1 || h.c < -1 || h.c > 1 || h.hS < 0 || h.hC < 0 || h.thetap < 0 || h.thetap >= kHalfPi ||
      h.w < 0 || h.w > 1 || h.B_S0 < 0 || h.B_C0 < 0 || h.phi < 0 || 1.209f * std::pow(h.phi, 2.0f / 3.0f) >= 1)
    fail(DVS_E_ARGUMENT, "invalid Hapke parameter range");
}

}  // namespace

struct dvs_writer_s {
  std::unique_ptr<Writer> impl;
};

extern "C" {

const char *dvs_io_last_error(void) { return g_error.c_str(); }

int dvs_writer_open(const char *path, int32_t format, dvs_writer *out) {
  return guarded([&] {
    if (!path || !out) fail(DVS_E_ARGUMENT, "invalid event writer arguments");
    *out = nullptr;
    std::unique_ptr<dvs_writer_s> w(new dvs_writer_s());
    if (format == DVS_EVENTS_BIN) w->impl.reset(new BinWriter(path));
    else if (format == DVS_EVENTS_HDF5) w->impl.reset(new H5Writer(path));
    else fail(DVS_E_STATE, "unsupported event format: " + std::to_string(format));
    *out = w.release();
  });
}

int dvs_writer_append(dvs_writer w, const dvs_event *events, int64_t count) {
  return guarded([&] {
    if (!w || count < 0 || (count && !events)) fail(DVS_E_ARGUMENT, "invalid event writer arguments");
    w->impl->append(std::vector<dvs_event>(events, events + count));
  });
}

int dvs_writer_close(dvs_writer w, double end_time_seconds) {
  return guarded([&] {
    if (!w) fail(DVS_E_ARGUMENT, "null event writer");
    w->impl->close(end_time_seconds);
  });
}

int dvs_writer_count(dvs_writer w, int64_t *count) {
  return guarded([&] {
    if (!w || !count) fail(DVS_E_ARGUMENT, "null event writer");
    *count = w->impl->count;
  });
}

int dvs_writer_destroy(dvs_writer w) {
  delete w;
  return DVS_OK;
}

int dvs_events_h5_read(const char *path, int16_t *x, int16_t *y, int64_t *t, int8_t *p, int64_t capacity,
                       int64_t *count) {
  return guarded([&] {
    if (!path || !count) fail(DVS_E_ARGUMENT, "invalid HDF5 read arguments");
    H5File f(path);
    const auto xs = read_column<int16_t>(f.id, "x", H5T_NATIVE_INT16);
    const auto ys = read_column<int16_t>(f.id, "y", H5T_NATIVE_INT16);
    const auto ts = read_column<int64_t>(f.id, "t", H5T_NATIVE_INT64);
    const auto ps = read_column<int8_t>(f.id, "p", H5T_NATIVE_INT8);
    *count = (int64_t)ts.size();
    const size_t n = (size_t)std::min<int64_t>(capacity, (int64_t)ts.size());
    if (x) std::memcpy(x, xs.data(), n * sizeof(int16_t));
    if (y) std::memcpy(y, ys.data(), n * sizeof(int16_t));
    if (t) std::memcpy(t, ts.data(), n * sizeof(int64_t));
    if (p) std::memcpy(p, ps.data(), n * sizeof(int8_t));
  });
}

int dvs_events_h5_ms_to_idx(const char *path, int64_t *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!path || !count) fail(DVS_E_ARGUMENT, "invalid HDF5 read arguments");
    H5File f(path);
    const auto v = read_column<int64_t>(f.id, "ms_to_idx", H5T_NATIVE_INT64);
    *count = (int64_t)v.size();
    if (out) std::memcpy(out, v.data(), (size_t)std::min<int64_t>(capacity, (int64_t)v.size()) * sizeof(int64_t));
  });
}

int dvs_config_load(const char *path, dvs_config *out) {
  return guarded([&] {
    if (!path || !out) fail(DVS_E_ARGUMENT, "invalid configuration arguments");
    load_config(path, *out);
  });
}

}  // extern "C"
