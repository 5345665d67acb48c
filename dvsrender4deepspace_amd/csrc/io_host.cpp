// Host I/O of the DVS path (libdvs_io.so): event stream writers, the YAML
// configuration, the scene bundle and trajectory loaders and the frame image
// files, behind the C ABI of include/dvs_io.h.
//
// Formats follow include/dvs/dvs_writer.hpp (13-byte binary records; HDF5
// x/y/t/p datasets with microsecond times and a millisecond index) and the
// configuration grammar and checks of src/scene_types.cpp.
#include <hdf5.h>
#include <zlib.h>

#include <algorithm>
#include <cctype>
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
#include "raster_host.h"

namespace {

thread_local std::string g_error;

}  // namespace

struct dvs_raster_s {
  dvsio::Raster raster;
  size_t missing = 0;
};

namespace {

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
  // Pushes buffered data to the file, so that its size on disk is current.
  virtual void flush() {}
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
  void flush() override {
    if (file_ >= 0) check_rc(H5Fflush(file_, H5F_SCOPE_GLOBAL), "failed to flush event batch");
  }

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

// Reads a whole dataset (any rank) in row-major order.
template <class T>
std::vector<T> read_column(hid_t file, const char *name, hid_t type) {
  const hid_t set = H5Dopen2(file, name, H5P_DEFAULT);
  if (set < 0) fail(DVS_E_ARGUMENT, std::string("missing HDF5 dataset ") + name);
  const hid_t space = H5Dget_space(set);
  const hssize_t n = H5Sget_simple_extent_npoints(space);
  std::vector<T> out(n > 0 ? (size_t)n : 0);
  herr_t rc = 0;
  if (n > 0) rc = H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data());
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

namespace {
uint8_t encode_pixel(float linear, int x, int y, float dither);  // render_float_to_png_bytes, per pixel
}  // namespace

// frames.h5 (Hdf5FrameWriter, include/dvs/frame_writer.hpp:12-77): t (int64
// microseconds, chunks of 4096), images (uint8, N x H x W) and depth
// (float32, N x H x W), one frame per chunk, byte shuffle and Deflate level
// 1 on the frame datasets when the library has the filter.
class FrameFile {
 public:
  FrameFile(const std::string &path, int width, int height) : width_(width), height_(height) {
    if (width <= 0 || height <= 0) fail(DVS_E_ARGUMENT, "invalid frame dimensions");
    file_ = H5Fcreate(path.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    check(file_, "create frames file");
    times_ = create("t", H5T_STD_I64LE, 1);
    images_ = create("images", H5T_STD_U8LE, 3);
    depths_ = create("depth", H5T_IEEE_F32LE, 3);
    check(H5Fflush(file_, H5F_SCOPE_GLOBAL), "flush frames header");
  }
  ~FrameFile() { release(); }
  void append(double time, const float *radiance, const float *depth) {
    const size_t pixels = (size_t)width_ * height_;
    if (closed_ || !std::isfinite(time) || time <= last_time_ || !radiance || !depth)
      fail(DVS_E_ARGUMENT, "invalid frame timestamp or array size");
    for (size_t i = 0; i < pixels; ++i)
      if (!std::isfinite(radiance[i]) || !std::isfinite(depth[i]) || (depth[i] < 0 && depth[i] != -1))
        fail(DVS_E_ARGUMENT, "invalid radiance or depth value");
    std::vector<uint8_t> image(pixels);
    for (int y = 0; y < height_; ++y)
      for (int x = 0; x < width_; ++x)
        image[(size_t)y * width_ + x] = encode_pixel(radiance[(size_t)y * width_ + x], x, y, 0.0f);
    const int64_t stamp = std::llround(time * 1e6);
    put(times_, H5T_NATIVE_INT64, &stamp, 1);
    put(images_, H5T_NATIVE_UINT8, image.data(), 3);
    put(depths_, H5T_NATIVE_FLOAT, depth, 3);
    check(H5Fflush(file_, H5F_SCOPE_GLOBAL), "flush frame");
    ++count_;
    last_time_ = time;
  }
  void close() {
    if (closed_) return;
    check(H5Fflush(file_, H5F_SCOPE_GLOBAL), "flush final frame");
    release();
    closed_ = true;
  }
  int64_t count() const { return (int64_t)count_; }

 private:
  static void check(int64_t code, const char *message) {
    if (code < 0) fail(DVS_E_STATE, message);
  }
  hid_t create(const char *name, hid_t type, int rank) {
    const hsize_t dims[3] = {0, (hsize_t)height_, (hsize_t)width_};
    const hsize_t maxdims[3] = {H5S_UNLIMITED, (hsize_t)height_, (hsize_t)width_};
    const hsize_t chunks[3] = {rank == 1 ? 4096ULL : 1ULL, (hsize_t)height_, (hsize_t)width_};
    const hid_t space = H5Screate_simple(rank, dims, maxdims), plist = H5Pcreate(H5P_DATASET_CREATE);
    check(space, "create frame dataspace");
    check(plist, "create frame properties");
    check(H5Pset_chunk(plist, rank, chunks), "set frame chunks");
    if (rank == 3 && H5Zfilter_avail(H5Z_FILTER_DEFLATE) > 0) {
      check(H5Pset_shuffle(plist), "set lossless shuffle");
      check(H5Pset_deflate(plist, 1), "set lossless deflate");
    }
    const hid_t set = H5Dcreate2(file_, name, type, space, H5P_DEFAULT, plist, H5P_DEFAULT);
    H5Pclose(plist);
    H5Sclose(space);
    check(set, "create frame dataset");
    return set;
  }
  void put(hid_t set, hid_t type, const void *data, int rank) {
    const hsize_t size[3] = {count_ + 1, (hsize_t)height_, (hsize_t)width_};
    const hsize_t start[3] = {count_, 0, 0}, block[3] = {1, (hsize_t)height_, (hsize_t)width_};
    check(H5Dset_extent(set, size), "extend frame dataset");
    const hid_t space = H5Dget_space(set);
    check(space, "get frame dataspace");
    check(H5Sselect_hyperslab(space, H5S_SELECT_SET, start, nullptr, block, nullptr), "select frame hyperslab");
    const hid_t memory = H5Screate_simple(rank, block, nullptr);
    check(memory, "create frame memory space");
    const herr_t rc = H5Dwrite(set, type, memory, space, H5P_DEFAULT, data);
    H5Sclose(memory);
    H5Sclose(space);
    check(rc, "write frame dataset");
  }
  void release() {
    for (hid_t *p : {&times_, &images_, &depths_})
      if (*p >= 0) {
        H5Dclose(*p);
        *p = -1;
      }
    if (file_ >= 0) {
      H5Fclose(file_);
      file_ = -1;
    }
  }
  int width_, height_;
  hid_t file_ = -1, times_ = -1, images_ = -1, depths_ = -1;
  hsize_t count_ = 0;
  double last_time_ = -std::numeric_limits<double>::infinity();
  bool closed_ = false;
};

struct dvs_writer_s {
  std::unique_ptr<Writer> impl;
};

struct dvs_frames_s {
  std::unique_ptr<FrameFile> impl;
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

int dvs_writer_flush(dvs_writer w) {
  return guarded([&] {
    if (!w) fail(DVS_E_ARGUMENT, "null event writer");
    w->impl->flush();
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

// ---------------------------------------------------------------------------
// Scene bundle and trajectory files (src/scene_data.cpp:19-283), frame
// encode (include/dvs/blenderize.hpp:12-69) and the image files of bin/cam.
// ---------------------------------------------------------------------------
namespace {

// read_binary_f32 / read_binary_u32 (scene_data.cpp:32-90): whole file as
// 4-byte words; the messages name the path.
template <class T>
std::vector<T> read_words(const std::string &path, const char *kind, bool with_data) {
  std::ifstream in(path, std::ios::binary);
  if (!in) fail(DVS_E_ARGUMENT, "failed to open " + path);
  in.seekg(0, std::ios::end);
  const std::streamoff size = in.tellg();
  in.seekg(0, std::ios::beg);
  if (size < 0 || size % (std::streamoff)sizeof(T) != 0)
    fail(DVS_E_ARGUMENT, std::string("invalid ") + kind + " binary size for " + path);
  std::vector<T> words;
  if (!with_data) {
    words.resize((size_t)(size / (std::streamoff)sizeof(T)));  // sizes only; contents unread
    return words;
  }
  words.resize((size_t)(size / (std::streamoff)sizeof(T)));
  in.read(reinterpret_cast<char *>(words.data()), size);
  if (!in) fail(DVS_E_ARGUMENT, "failed to read " + path);
  return words;
}

std::string join(const char *dir, const char *name) {
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name;
}

// Column-major rotation matrix of a quaternion given (w, x, y, z), not
// normalised first (Eigen's toRotationMatrix on the text file's values).
void quaternion_matrix(float w, float x, float y, float z, float m[9]) {
  const float x2 = x + x, y2 = y + y, z2 = z + z;
  const float wx = x2 * w, wy = y2 * w, wz = z2 * w, xx = x2 * x, xy = y2 * x, xz = z2 * x;
  const float yy = y2 * y, yz = z2 * y, zz = z2 * z;
  m[0] = 1 - (yy + zz); m[3] = xy - wz;       m[6] = xz + wy;
  m[1] = xy + wz;       m[4] = 1 - (xx + zz); m[7] = yz - wx;
  m[2] = xz - wy;       m[5] = yz + wx;       m[8] = 1 - (xx + yy);
}

std::string trimmed(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// load_frames_from_text (scene_data.cpp:238-283): one frame per non-blank
// line "frame_idx x y z qx qy qz qw lx ly lz L"; the timestamp stays NaN
// (frame_index / fps).
std::vector<dvs_frame> frames_from_text(const std::string &path) {
  std::ifstream in(path);
  if (!in) {
    std::ifstream probe(path, std::ios::binary);
    fail(DVS_E_ARGUMENT, (probe ? "failed to open " : "file does not exist: ") + path);
  }
  std::vector<dvs_frame> frames;
  std::string line;
  while (std::getline(in, line)) {
    line = trimmed(line);
    if (line.empty()) continue;
    std::istringstream fields(line);
    dvs_frame f{};
    f.timestamp_seconds = std::numeric_limits<double>::quiet_NaN();
    float q[4];  // x, y, z, w
    if (!(fields >> f.frame_index >> f.camera_origin[0] >> f.camera_origin[1] >> f.camera_origin[2] >> q[0] >>
          q[1] >> q[2] >> q[3] >> f.light_dir[0] >> f.light_dir[1] >> f.light_dir[2] >> f.light_intensity))
      fail(DVS_E_ARGUMENT, "failed to parse camera orientation in line: " + line);
    quaternion_matrix(q[3], q[0], q[1], q[2], f.camera_world);
    frames.push_back(f);
  }
  return frames;
}

// load_frames_from_binary (scene_data.cpp:165-198): 24 floats per frame,
// {t, origin[3], camera_world as a row-major 4x4, light_dir[3], intensity}.
std::vector<dvs_frame> frames_from_binary(const char *dir, bool with_data) {
  const auto raw = read_words<float>(join(dir, "frames.bin"), "f32", with_data);
  constexpr size_t kStride = 24;
  if (raw.size() % kStride != 0) fail(DVS_E_ARGUMENT, "frames.bin has invalid stride");
  std::vector<dvs_frame> frames(raw.size() / kStride);
  if (!with_data) return frames;
  for (size_t i = 0; i < frames.size(); ++i) {
    const float *r = raw.data() + kStride * i;
    dvs_frame &f = frames[i];
    f.frame_index = 0;
    f.timestamp_seconds = r[0];
    for (int c = 0; c < 3; ++c) f.camera_origin[c] = r[1 + c];
    for (int row = 0; row < 3; ++row)
      for (int col = 0; col < 3; ++col) f.camera_world[3 * col + row] = r[4 + 4 * row + col];
    for (int c = 0; c < 3; ++c) f.light_dir[c] = r[20 + c];
    f.light_intensity = r[23];
  }
  return frames;
}

// linear_to_srgb, display_exposure, hash dither and to_byte
// (blenderize.hpp:12-69).
uint8_t encode_pixel(float linear, int x, int y, float dither) {
  float v = std::min(std::max(linear / 100, 0.0f), 1.0f);
  v = v < 0.0031308f ? 12.92f * v : 1.055f * std::pow(v, 1.0f / 2.4f) - 0.055f;
  if (dither != 0.0f) {
    const uint32_t ux = (uint32_t)x, uy = (uint32_t)y;
    const uint32_t hx = 1103515245u * ((ux >> 1u) ^ uy), hy = 1103515245u * ((uy >> 1u) ^ ux);
    float h = float(1103515245u * (hx ^ (hy >> 3u))) * (1.0f / float(0xffffffffu));
    h = h * 2.0f - 1.0f;
    v += dither * (std::copysign(1.0f, h) * (1.0f - std::sqrt(1.0f - std::fabs(h))));
  }
  if (!std::isfinite(v) || v <= 0.0f) return 0;
  if (v >= 1.0f) return 255;
  return (uint8_t)(v * 255.0f + 0.5f);
}

void put_be32(std::string &s, uint32_t v) {
  for (int k = 3; k >= 0; --k) s.push_back((char)((v >> (8 * k)) & 0xff));
}
void png_chunk(std::ofstream &out, const char *type, const std::string &data) {
  std::string buf;
  put_be32(buf, (uint32_t)data.size());
  buf.append(type, 4);
  buf += data;
  const uLong crc = crc32(crc32(0L, Z_NULL, 0), reinterpret_cast<const Bytef *>(buf.data() + 4),
                          (uInt)(buf.size() - 4));
  put_be32(buf, (uint32_t)crc);
  out.write(buf.data(), (std::streamsize)buf.size());
}

// 8-bit greyscale PNG (filter 0 on every row, zlib level 6).
void write_png(const std::string &path, const uint8_t *gray, int width, int height) {
  std::string raw;
  raw.reserve((size_t)(width + 1) * height);
  for (int y = 0; y < height; ++y) {
    raw.push_back(0);
    raw.append(reinterpret_cast<const char *>(gray) + (size_t)y * width, (size_t)width);
  }
  uLongf packed_size = compressBound((uLong)raw.size());
  std::string packed(packed_size, '\0');
  if (compress2(reinterpret_cast<Bytef *>(&packed[0]), &packed_size, reinterpret_cast<const Bytef *>(raw.data()),
                (uLong)raw.size(), 6) != Z_OK)
    fail(DVS_E_STATE, "PNG compression failed for " + path);
  packed.resize(packed_size);
  std::ofstream out(path, std::ios::binary);
  if (!out) fail(DVS_E_STATE, "failed to open " + path);
  out.write("\x89PNG\r\n\x1a\n", 8);
  std::string ihdr;
  put_be32(ihdr, (uint32_t)width);
  put_be32(ihdr, (uint32_t)height);
  ihdr += std::string("\x08\x00\x00\x00\x00", 5);  // depth 8, greyscale, deflate, filter 0, no interlace
  png_chunk(out, "IHDR", ihdr);
  png_chunk(out, "IDAT", packed);
  png_chunk(out, "IEND", std::string());
  if (!out) fail(DVS_E_STATE, "failed to write " + path);
}

// Single-strip, uncompressed little-endian TIFF of float32 samples.
void write_tiff_f32(const std::string &path, const float *data, int width, int height) {
  const uint32_t pixels = (uint32_t)width * (uint32_t)height, bytes = pixels * 4;
  struct Entry { uint16_t tag, type; uint32_t count, value; };
  const Entry entries[] = {
      {256, 4, 1, (uint32_t)width},   // ImageWidth
      {257, 4, 1, (uint32_t)height},  // ImageLength
      {258, 3, 1, 32},                // BitsPerSample
      {259, 3, 1, 1},                 // Compression: none
      {262, 3, 1, 1},                 // Photometric: black is zero
      {273, 4, 1, 8},                 // StripOffsets: data right after the header
      {277, 3, 1, 1},                 // SamplesPerPixel
      {278, 4, 1, (uint32_t)height},  // RowsPerStrip
      {279, 4, 1, bytes},             // StripByteCounts
      {284, 3, 1, 1},                 // PlanarConfiguration: chunky
      {339, 3, 1, 3},                 // SampleFormat: IEEE float
  };
  std::ofstream out(path, std::ios::binary);
  if (!out) fail(DVS_E_STATE, "failed to open " + path);
  auto u16 = [&](uint16_t v) { out.write(reinterpret_cast<const char *>(&v), 2); };
  auto u32 = [&](uint32_t v) { out.write(reinterpret_cast<const char *>(&v), 4); };
  out.write("II", 2);
  u16(42);
  u32(8 + bytes);  // first IFD after the pixels
  out.write(reinterpret_cast<const char *>(data), bytes);
  const uint16_t n = (uint16_t)(sizeof(entries) / sizeof(entries[0]));
  u16(n);
  for (const Entry &e : entries) {
    u16(e.tag);
    u16(e.type);
    u32(e.count);
    if (e.type == 3) {  // SHORT values sit left-justified in the 4-byte field
      u16((uint16_t)e.value);
      u16(0);
    } else {
      u32(e.value);
    }
  }
  u32(0);
  if (!out) fail(DVS_E_STATE, "failed to write " + path);
}

}  // namespace

extern "C" {

int dvs_io_mesh_load(const char *dir, float *vertices, float *normals, int64_t vertex_capacity, uint32_t *indices,
                     int64_t triangle_capacity, int64_t *vertex_count, int64_t *triangle_count) {
  return guarded([&] {
    if (!dir || !vertex_count || !triangle_count) fail(DVS_E_ARGUMENT, "invalid mesh arguments");
    const bool with_data = vertices && normals && indices;
    const auto v = read_words<float>(join(dir, "mesh_vertices.bin"), "f32", with_data);
    const auto n = read_words<float>(join(dir, "mesh_normals.bin"), "f32", with_data);
    const auto t = read_words<uint32_t>(join(dir, "mesh_indices.bin"), "u32", with_data);
    if (v.size() % 3 || n.size() % 3 || t.size() % 3) fail(DVS_E_ARGUMENT, "mesh binaries must be Nx3");
    if (v.empty() || t.empty() || v.size() != n.size())
      fail(DVS_E_ARGUMENT, "mesh requires nonempty vertices/triangles and matching normals");
    *vertex_count = (int64_t)(v.size() / 3);
    *triangle_count = (int64_t)(t.size() / 3);
    if (!with_data) return;
    for (uint32_t i : t)
      if (i >= v.size() / 3) fail(DVS_E_ARGUMENT, "mesh index outside vertex array");
    for (float x : v)
      if (!std::isfinite(x)) fail(DVS_E_ARGUMENT, "nonfinite mesh vertex");
    for (float x : n)
      if (!std::isfinite(x)) fail(DVS_E_ARGUMENT, "nonfinite mesh normal");
    if (vertex_capacity < *vertex_count || triangle_capacity < *triangle_count)
      fail(DVS_E_ARGUMENT, "mesh buffers are smaller than the mesh");
    std::memcpy(vertices, v.data(), v.size() * sizeof(float));
    std::memcpy(normals, n.data(), n.size() * sizeof(float));
    std::memcpy(indices, t.data(), t.size() * sizeof(uint32_t));
  });
}

int dvs_io_frames_binary(const char *dir, dvs_frame *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!dir || !count) fail(DVS_E_ARGUMENT, "invalid frame arguments");
    const auto frames = frames_from_binary(dir, out != nullptr);
    *count = (int64_t)frames.size();
    if (out) std::memcpy(out, frames.data(), (size_t)std::min<int64_t>(capacity, *count) * sizeof(dvs_frame));
  });
}

int dvs_io_frames_text(const char *path, dvs_frame *out, int64_t capacity, int64_t *count) {
  return guarded([&] {
    if (!path || !count) fail(DVS_E_ARGUMENT, "invalid frame arguments");
    const auto frames = frames_from_text(path);
    *count = (int64_t)frames.size();
    if (out) std::memcpy(out, frames.data(), (size_t)std::min<int64_t>(capacity, *count) * sizeof(dvs_frame));
  });
}

int dvs_io_frame_bytes(const float *linear, int32_t width, int32_t height, float dither, uint8_t *out) {
  return guarded([&] {
    if (width < 0 || height < 0 || ((int64_t)width * height > 0 && (!linear || !out)))
      fail(DVS_E_ARGUMENT, "invalid frame encode arguments");
    for (int y = 0; y < height; ++y)
      for (int x = 0; x < width; ++x)
        out[(size_t)y * width + x] = encode_pixel(linear[(size_t)y * width + x], x, y, dither);
  });
}

int dvs_frames_open(const char *path, int32_t width, int32_t height, dvs_frames *out) {
  return guarded([&] {
    if (!path || !out) fail(DVS_E_ARGUMENT, "invalid frame writer arguments");
    *out = nullptr;
    auto f = std::make_unique<dvs_frames_s>();
    f->impl = std::make_unique<FrameFile>(path, width, height);
    *out = f.release();
  });
}

int dvs_frames_append(dvs_frames f, double time, const float *radiance, const float *depth) {
  return guarded([&] {
    if (!f) fail(DVS_E_ARGUMENT, "null frame writer");
    f->impl->append(time, radiance, depth);
  });
}

int dvs_frames_close(dvs_frames f) {
  return guarded([&] {
    if (!f) fail(DVS_E_ARGUMENT, "null frame writer");
    f->impl->close();
  });
}

int dvs_frames_count(dvs_frames f, int64_t *count) {
  return guarded([&] {
    if (!f || !count) fail(DVS_E_ARGUMENT, "null frame writer");
    *count = f->impl->count();
  });
}

int dvs_frames_destroy(dvs_frames f) {
  delete f;
  return DVS_OK;
}

int dvs_frames_h5_read(const char *path, int64_t *times, uint8_t *images, float *depth, int64_t capacity,
                       int32_t *width, int32_t *height, int64_t *count) {
  return guarded([&] {
    if (!path || !count || !width || !height) fail(DVS_E_ARGUMENT, "invalid HDF5 read arguments");
    H5File f(path);
    const hid_t set = H5Dopen2(f.id, "images", H5P_DEFAULT);
    if (set < 0) fail(DVS_E_STATE, "failed to open HDF5 dataset images");
    const hid_t space = H5Dget_space(set);
    hsize_t dims[3] = {0, 0, 0};
    const int rank = H5Sget_simple_extent_dims(space, dims, nullptr);
    H5Sclose(space);
    H5Dclose(set);
    if (rank != 3) fail(DVS_E_STATE, "frames images must be N x H x W");
    *count = (int64_t)dims[0];
    *height = (int32_t)dims[1];
    *width = (int32_t)dims[2];
    if (capacity < *count || !times || !images || !depth) return;
    const auto t = read_column<int64_t>(f.id, "t", H5T_NATIVE_INT64);
    const auto im = read_column<uint8_t>(f.id, "images", H5T_NATIVE_UINT8);
    const auto d = read_column<float>(f.id, "depth", H5T_NATIVE_FLOAT);
    std::copy(t.begin(), t.end(), times);
    std::copy(im.begin(), im.end(), images);
    std::copy(d.begin(), d.end(), depth);
  });
}

int dvs_io_write_png(const char *path, const uint8_t *gray, int32_t width, int32_t height) {
  return guarded([&] {
    if (!path || !gray || width <= 0 || height <= 0) fail(DVS_E_ARGUMENT, "invalid PNG arguments");
    write_png(path, gray, width, height);
  });
}

int dvs_io_raster_load(const char *path, dvs_raster *out) {
  return guarded([&] {
    if (!path || !out) fail(DVS_E_ARGUMENT, "invalid raster arguments");
    auto r = std::make_unique<dvs_raster_s>();
    try {
      r->raster = dvsio::read_geotiff(path);
      r->missing = dvsio::finish_imagery(r->raster, path);
    } catch (const std::runtime_error &e) {
      fail(DVS_E_ARGUMENT, e.what());
    }
    *out = r.release();
  });
}

int dvs_io_raster_info(dvs_raster r, dvs_raster_info *out) {
  return guarded([&] {
    if (!r || !out) fail(DVS_E_ARGUMENT, "invalid raster arguments");
    const dvsio::Raster &x = r->raster;
    out->width = x.width;
    out->height = x.height;
    std::copy(x.transform, x.transform + 6, out->transform);
    out->has_nodata = x.has_nodata;
    out->has_scale = x.has_scale;
    out->has_offset = x.has_offset;
    out->nodata = x.nodata;
    out->scale = x.has_scale ? x.scale : 1.0;
    out->offset = x.has_offset ? x.offset : 0.0;
    out->missing = (int64_t)r->missing;
    out->center_raw = x.center_raw;
  });
}

int dvs_io_raster_data(dvs_raster r, float *data, float *x_coords, float *y_coords) {
  return guarded([&] {
    if (!r) fail(DVS_E_ARGUMENT, "invalid raster arguments");
    const dvsio::Raster &x = r->raster;
    if (data) std::copy(x.data.begin(), x.data.end(), data);
    if (x_coords) std::copy(x.x_coords.begin(), x.x_coords.end(), x_coords);
    if (y_coords) std::copy(x.y_coords.begin(), x.y_coords.end(), y_coords);
  });
}

int dvs_io_raster_projection(dvs_raster r, char *out, int64_t capacity, int64_t *length) {
  return guarded([&] {
    if (!r || (capacity > 0 && !out)) fail(DVS_E_ARGUMENT, "invalid raster arguments");
    const std::string &p = r->raster.projection;
    if (length) *length = (int64_t)p.size();
    if (capacity > 0) {
      const size_t n = std::min<size_t>(p.size(), (size_t)capacity - 1);
      std::memcpy(out, p.data(), n);
      out[n] = '\0';
    }
  });
}

int dvs_io_raster_geokey(dvs_raster r, int32_t key, double *value) {
  return guarded([&] {
    if (!r || !value) fail(DVS_E_ARGUMENT, "invalid raster arguments");
    if (!dvsio::geokey_value(r->raster, key, *value)) fail(DVS_E_ARGUMENT, "raster has no GeoKey " + std::to_string(key));
  });
}

int dvs_io_raster_free(dvs_raster r) {
  delete r;
  return DVS_OK;
}

int dvs_io_write_tiff(const char *path, const float *data, int32_t width, int32_t height) {
  return guarded([&] {
    if (!path || !data || width <= 0 || height <= 0) fail(DVS_E_ARGUMENT, "invalid TIFF arguments");
    write_tiff_f32(path, data, width, height);
  });
}

}  // extern "C"
