// Minimal GeoTIFF reader: the part of GDAL that tools::loadImageryData
// (include/tools/loader.hpp:31-100) uses, for single-band elevation and
// albedo rasters.
//
// Reads classic and BigTIFF files in either byte order; strips or tiles;
// no compression, LZW (5) or Deflate (8, 32946); horizontal (2) and
// floating-point (3) predictors; 8/16/32-bit integer and 32/64-bit float
// samples (band 1 of chunky or planar multi-band files).  The geotransform
// comes from ModelTransformationTag, or from ModelTiepointTag with
// ModelPixelScaleTag, shifted by half a pixel for PixelIsPoint rasters as
// GDAL does; NoData from GDAL_NODATA; scale and offset from the band items
// of GDAL_METADATA.  The projection is identified by the GeoKey directory
// and its parameters (GDAL's WKT is not available without GDAL: two rasters
// share a CRS here when their GeoKeys are identical).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dvs_hip.h"
#include "dvs_io.h"
#include "raster_host.h"

namespace dvsio {
namespace {

struct Entry {
  uint16_t type = 0;
  uint64_t count = 0;
  uint64_t value_at = 0;  // file offset of the value bytes
};

class Tiff {
 public:
  explicit Tiff(const std::string &path) : path_(path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to open dataset: " + path);
    bytes_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    if (bytes_.size() < 8) bad("file too short");
    if (bytes_[0] == 'I' && bytes_[1] == 'I') little_ = true;
    else if (bytes_[0] == 'M' && bytes_[1] == 'M') little_ = false;
    else bad("not a TIFF file");
    const uint16_t magic = u16(2);
    uint64_t ifd = 0;
    if (magic == 42) {
      ifd = u32(4);
    } else if (magic == 43) {
      big_ = true;
      if (u16(4) != 8) bad("unsupported BigTIFF offset size");
      ifd = u64(8);
    } else {
      bad("not a TIFF file");
    }
    const uint64_t n = big_ ? u64(ifd) : u16(ifd);
    const uint64_t first = ifd + (big_ ? 8 : 2), width = big_ ? 20 : 12;
    need(first, n * width);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t at = first + i * width;
      Entry e;
      const uint16_t tag = u16(at);
      e.type = u16(at + 2);
      e.count = big_ ? u64(at + 4) : u32(at + 4);
      const uint64_t size = e.count * type_size(e.type);
      const uint64_t inline_room = big_ ? 8 : 4;
      e.value_at = size <= inline_room ? at + (big_ ? 12 : 8) : (big_ ? u64(at + 12) : u32(at + 8));
      need(e.value_at, size);
      tags_[tag] = e;
    }
  }

  bool has(uint16_t tag) const { return tags_.count(tag) != 0; }
  std::vector<double> numbers(uint16_t tag) const {
    const Entry &e = entry(tag);
    std::vector<double> out(e.count);
    const uint64_t w = type_size(e.type);
    for (uint64_t i = 0; i < e.count; ++i) out[i] = value_of(e.type, e.value_at + i * w);
    return out;
  }
  double number(uint16_t tag, double fallback) const {
    if (!has(tag)) return fallback;
    const auto v = numbers(tag);
    return v.empty() ? fallback : v[0];
  }
  std::string ascii(uint16_t tag) const {
    const Entry &e = entry(tag);
    std::string s(reinterpret_cast<const char *>(bytes_.data() + e.value_at), e.count);
    while (!s.empty() && s.back() == '\0') s.pop_back();
    return s;
  }
  const uint8_t *at(uint64_t offset, uint64_t size) const {
    need(offset, size);
    return bytes_.data() + offset;
  }
  bool little() const { return little_; }
  [[noreturn]] void bad(const std::string &what) const {
    throw std::runtime_error("Failed to read raster data from dataset: " + path_ + " (" + what + ")");
  }

 private:
  const Entry &entry(uint16_t tag) const {
    auto it = tags_.find(tag);
    if (it == tags_.end()) bad("missing tag " + std::to_string(tag));
    return it->second;
  }
  static uint64_t type_size(uint16_t type) {
    switch (type) {
      case 1: case 2: case 6: case 7: return 1;
      case 3: case 8: return 2;
      case 4: case 9: case 11: return 4;
      case 5: case 10: case 12: case 16: case 17: case 18: return 8;
      default: return 1;
    }
  }
  void need(uint64_t offset, uint64_t size) const {
    if (offset > bytes_.size() || size > bytes_.size() - offset) bad("truncated file");
  }
  uint64_t raw(uint64_t at, int n) const {
    need(at, (uint64_t)n);
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) {
      const uint64_t b = bytes_[at + (little_ ? i : n - 1 - i)];
      v |= b << (8 * i);
    }
    return v;
  }
  uint16_t u16(uint64_t at) const { return (uint16_t)raw(at, 2); }
  uint32_t u32(uint64_t at) const { return (uint32_t)raw(at, 4); }
  uint64_t u64(uint64_t at) const { return raw(at, 8); }
  double value_of(uint16_t type, uint64_t at) const {
    switch (type) {
      case 1: case 7: return (double)bytes_[at];
      case 6: return (double)(int8_t)bytes_[at];
      case 3: return (double)u16(at);
      case 8: return (double)(int16_t)u16(at);
      case 4: return (double)u32(at);
      case 9: return (double)(int32_t)u32(at);
      case 16: return (double)u64(at);
      case 17: return (double)(int64_t)u64(at);
      case 5: return (double)u32(at) / (double)std::max<uint32_t>(1, u32(at + 4));
      case 10: return (double)(int32_t)u32(at) / (double)std::max<int32_t>(1, (int32_t)u32(at + 4));
      case 11: {
        const uint32_t b = u32(at);
        float f;
        std::memcpy(&f, &b, 4);
        return f;
      }
      case 12: {
        const uint64_t b = u64(at);
        double d;
        std::memcpy(&d, &b, 8);
        return d;
      }
      default: return 0.0;
    }
  }

  std::string path_;
  std::vector<uint8_t> bytes_;
  std::map<uint16_t, Entry> tags_;
  bool little_ = true, big_ = false;
};

// TIFF LZW (MSB-first codes, early change, 256 clear, 257 end).  Each code
// is stored as (prefix code, last byte, length) and written out backwards.
std::vector<uint8_t> lzw_decode(const uint8_t *in, size_t size, size_t expect, const Tiff &t) {
  std::vector<uint8_t> out;
  out.reserve(expect);
  std::vector<int> prefix(4096, -1);
  std::vector<uint8_t> last(4096), first(4096);
  std::vector<int> length(4096, 1);
  for (int i = 0; i < 256; ++i) {
    last[i] = first[i] = (uint8_t)i;
  }
  int next = 258, width = 9, prev = -1;
  size_t bitpos = 0;
  const size_t bits = size * 8;
  auto emit = [&](int code) {
    const size_t at = out.size();
    out.resize(at + length[code]);
    for (int c = code, k = length[code] - 1; k >= 0; c = prefix[c], --k) out[at + k] = last[c];
  };
  while (bitpos + width <= bits) {
    int code = 0;
    for (int i = 0; i < width; ++i) {
      const size_t b = bitpos + i;
      code = (code << 1) | ((in[b >> 3] >> (7 - (b & 7))) & 1);
    }
    bitpos += width;
    if (code == 257) break;
    if (code == 256) {
      next = 258;
      width = 9;
      prev = -1;
      continue;
    }
    if (prev < 0) {
      if (code >= 256) t.bad("corrupt LZW data");
      emit(code);
      prev = code;
      continue;
    }
    if (code > next || next >= 4096) t.bad("corrupt LZW data");
    const uint8_t head = code < next ? first[code] : first[prev];
    prefix[next] = prev;
    last[next] = head;
    first[next] = first[prev];
    length[next] = length[prev] + 1;
    ++next;
    emit(code);
    prev = code;
    if (next + 1 >= 2048) width = 12;  // early change: one code before the table fills
    else if (next + 1 >= 1024) width = 11;
    else if (next + 1 >= 512) width = 10;
  }
  return out;
}

std::vector<uint8_t> inflate_all(const uint8_t *in, size_t size, size_t expect, const Tiff &t) {
  std::vector<uint8_t> out(expect);
  uLongf got = (uLongf)expect;
  const int rc = uncompress(out.data(), &got, in, (uLong)size);
  if (rc != Z_OK && rc != Z_BUF_ERROR) t.bad("corrupt Deflate data");
  out.resize(got);
  return out;
}

struct Layout {
  int width = 0, height = 0, bits = 0, format = 1, samples = 1, planar = 1, compression = 1, predictor = 1;
  bool tiled = false;
  int block_w = 0, block_h = 0;
};

// Sample `i` (band 1) of a decoded block, as float (RasterIO to GDT_Float32).
float sample(const uint8_t *p, const Layout &l, bool little) {
  const int n = l.bits / 8;
  uint64_t v = 0;
  for (int i = 0; i < n; ++i) v |= (uint64_t)p[little ? i : n - 1 - i] << (8 * i);
  if (l.format == 3) {
    if (n == 4) {
      float f;
      const uint32_t b = (uint32_t)v;
      std::memcpy(&f, &b, 4);
      return f;
    }
    double d;
    std::memcpy(&d, &v, 8);
    return (float)d;
  }
  if (l.format == 2) {
    const int64_t s = (int64_t)(v << (64 - l.bits)) >> (64 - l.bits);
    return (float)s;
  }
  return (float)v;
}

// Undoes the predictor on one decoded block of `rows` rows of `cols`
// pixels; integer samples in the file's byte order are summed as values.
void unpredict(std::vector<uint8_t> &block, const Layout &l, int cols, int rows, bool little, const Tiff &t) {
  const int bps = l.bits / 8, spp = l.planar == 1 ? l.samples : 1;
  const size_t row_bytes = (size_t)cols * spp * bps;
  if (block.size() < row_bytes * rows) t.bad("short block");
  if (l.predictor == 2) {
    if (l.format == 3) t.bad("horizontal predictor on float samples");
    for (int r = 0; r < rows; ++r) {
      uint8_t *row = block.data() + r * row_bytes;
      for (int c = spp; c < cols * spp; ++c) {
        uint8_t *cur = row + (size_t)c * bps, *prev = row + (size_t)(c - spp) * bps;
        uint64_t a = 0, b = 0;
        for (int i = 0; i < bps; ++i) {
          a |= (uint64_t)cur[little ? i : bps - 1 - i] << (8 * i);
          b |= (uint64_t)prev[little ? i : bps - 1 - i] << (8 * i);
        }
        const uint64_t s = a + b;
        for (int i = 0; i < bps; ++i) cur[little ? i : bps - 1 - i] = (uint8_t)(s >> (8 * i));
      }
    }
  } else if (l.predictor == 3) {
    if (l.format != 3) t.bad("floating-point predictor on integer samples");
    std::vector<uint8_t> tmp(row_bytes);
    const size_t wc = (size_t)cols * spp;
    for (int r = 0; r < rows; ++r) {
      uint8_t *row = block.data() + r * row_bytes;
      for (size_t i = 1; i < row_bytes; ++i) row[i] = (uint8_t)(row[i] + row[i - 1]);
      std::memcpy(tmp.data(), row, row_bytes);
      // planes hold the most significant byte first; write the file's order
      for (size_t c = 0; c < wc; ++c)
        for (int b = 0; b < bps; ++b) {
          const uint8_t v = tmp[(size_t)b * wc + c];  // byte b from the most significant
          row[c * bps + (little ? bps - 1 - b : b)] = v;
        }
    }
  } else if (l.predictor != 1) {
    t.bad("unsupported predictor " + std::to_string(l.predictor));
  }
}

}  // namespace

Raster read_geotiff(const std::string &path) {
  Tiff t(path);
  Layout l;
  l.width = (int)t.number(256, 0);
  l.height = (int)t.number(257, 0);
  if (l.width <= 0 || l.height <= 0) t.bad("no raster band");
  const auto bits = t.has(258) ? t.numbers(258) : std::vector<double>{1};
  l.bits = (int)bits[0];
  l.format = (int)t.number(339, 1);
  l.samples = (int)t.number(277, 1);
  l.planar = (int)t.number(284, 1);
  l.compression = (int)t.number(259, 1);
  l.predictor = (int)t.number(317, 1);
  if (!(l.bits == 8 || l.bits == 16 || l.bits == 32 || l.bits == 64) || (l.format == 3 && l.bits < 32) ||
      l.format < 1 || l.format > 3)
    t.bad("unsupported sample type");
  if (!(l.compression == 1 || l.compression == 5 || l.compression == 8 || l.compression == 32946))
    t.bad("unsupported compression " + std::to_string(l.compression));
  std::vector<double> offsets, counts;
  if (t.has(322)) {
    l.tiled = true;
    l.block_w = (int)t.number(322, 0);
    l.block_h = (int)t.number(323, 0);
    offsets = t.numbers(324);
    counts = t.numbers(325);
  } else {
    l.block_w = l.width;
    l.block_h = (int)std::min<double>(t.number(278, l.height), l.height);
    offsets = t.numbers(273);
    counts = t.numbers(279);
  }
  if (l.block_w <= 0 || l.block_h <= 0 || offsets.size() != counts.size()) t.bad("bad strip or tile layout");
  const int across = (l.width + l.block_w - 1) / l.block_w, down = (l.height + l.block_h - 1) / l.block_h;
  const size_t per_band = (size_t)across * down;
  if (offsets.size() < per_band) t.bad("missing strips or tiles");
  const int bps = l.bits / 8, spp = l.planar == 1 ? l.samples : 1;
  Raster r;
  r.width = l.width;
  r.height = l.height;
  r.data.assign((size_t)l.width * l.height, 0.0f);
  for (size_t b = 0; b < per_band; ++b) {
    const int bx = (int)(b % across), by = (int)(b / across);
    const int rows = l.tiled ? l.block_h : std::min(l.block_h, l.height - by * l.block_h);
    const size_t expect = (size_t)l.block_w * rows * spp * bps;
    const uint8_t *src = t.at((uint64_t)offsets[b], (uint64_t)counts[b]);
    std::vector<uint8_t> block;
    if (l.compression == 1) block.assign(src, src + std::min<size_t>((size_t)counts[b], expect));
    else if (l.compression == 5) block = lzw_decode(src, (size_t)counts[b], expect, t);
    else block = inflate_all(src, (size_t)counts[b], expect, t);
    if (block.size() < expect) t.bad("short strip or tile");
    unpredict(block, l, l.block_w, rows, t.little(), t);
    for (int y = 0; y < rows; ++y) {
      const int gy = by * l.block_h + y;
      if (gy >= l.height) break;
      for (int x = 0; x < l.block_w; ++x) {
        const int gx = bx * l.block_w + x;
        if (gx >= l.width) break;
        const uint8_t *p = block.data() + ((size_t)y * l.block_w + x) * spp * bps;
        r.data[(size_t)gy * l.width + gx] = sample(p, l, t.little());
      }
    }
  }
  // geotransform (GDAL's GeoTIFF driver)
  if (t.has(34264)) {
    const auto m = t.numbers(34264);
    if (m.size() < 16) t.bad("short ModelTransformationTag");
    const double gt[6] = {m[3], m[0], m[1], m[7], m[4], m[5]};
    std::copy(gt, gt + 6, r.transform);
  } else if (t.has(33922) && t.has(33550)) {
    const auto tie = t.numbers(33922), scale = t.numbers(33550);
    if (tie.size() < 6 || scale.size() < 2) t.bad("short tiepoint or pixel scale");
    r.transform[0] = tie[3] - tie[0] * scale[0];
    r.transform[1] = scale[0];
    r.transform[2] = 0;
    r.transform[3] = tie[4] + tie[1] * scale[1];
    r.transform[4] = 0;
    r.transform[5] = -scale[1];
  } else {
    throw std::runtime_error("Failed to get geotransform from dataset: " + path);
  }
  std::ostringstream proj;
  if (t.has(34735)) {
    const auto keys = t.numbers(34735);
    for (size_t i = 4; i + 3 < keys.size(); i += 4) {
      if ((int)keys[i] == 1025 && (int)keys[i + 1] == 0 && (int)keys[i + 3] == 2) {  // RasterPixelIsPoint
        r.transform[0] -= 0.5 * r.transform[1] + 0.5 * r.transform[2];
        r.transform[3] -= 0.5 * r.transform[4] + 0.5 * r.transform[5];
      }
    }
    r.geokeys = keys;
    proj << "geokeys";
    for (double k : keys) proj << ' ' << (long long)k;
    if (t.has(34736)) {
      r.geodoubles = t.numbers(34736);
      for (double d : r.geodoubles) proj << ' ' << std::hexfloat << d << std::defaultfloat;
    }
    if (t.has(34737)) proj << " '" << t.ascii(34737) << "'";
  }
  r.projection = proj.str();
  if (t.has(42113)) {
    const std::string s = t.ascii(42113);
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end != s.c_str()) {
      r.has_nodata = true;
      r.nodata = v;
    }
  }
  if (t.has(42112)) {  // <Item name="SCALE" sample="0" role="scale">2</Item>
    const std::string xml = t.ascii(42112);
    auto item = [&](const char *role, double &out) {
      const std::string key = std::string("role=\"") + role + "\"";
      size_t at = 0;
      while ((at = xml.find(key, at)) != std::string::npos) {
        const size_t open = xml.rfind("<Item", at);
        const size_t gt = xml.find('>', at);
        if (open == std::string::npos || gt == std::string::npos) break;
        const std::string head = xml.substr(open, gt - open);
        if (head.find("sample=\"0\"") != std::string::npos) {
          out = std::strtod(xml.c_str() + gt + 1, nullptr);
          return true;
        }
        at = gt;
      }
      return false;
    };
    r.has_scale = item("scale", r.scale);
    r.has_offset = item("offset", r.offset);
  }
  return r;
}

bool geokey_value(const Raster &r, int key, double &value) {
  const auto &k = r.geokeys;
  for (size_t i = 4; i + 3 < k.size(); i += 4) {
    if ((int)k[i] != key) continue;
    const int location = (int)k[i + 1], count = (int)k[i + 2], offset = (int)k[i + 3];
    if (location == 0 && count == 1) {
      value = k[i + 3];
      return true;
    }
    if (location == 34736 && offset >= 0 && (size_t)offset < r.geodoubles.size()) {
      value = r.geodoubles[offset];
      return true;
    }
    return false;
  }
  return false;
}

// loadImageryData's post-processing (loader.hpp:68-96): rejects rotated or
// degenerate transforms, maps NoData and non-finite samples to NaN and the
// rest through scale and offset, and builds the pixel-centre coordinates.
size_t finish_imagery(Raster &r, const std::string &path) {
  const double *gt = r.transform;
  if (gt[2] != 0 || gt[4] != 0 || gt[1] == 0 || gt[5] == 0)
    throw std::runtime_error("Rotated/sheared or degenerate rasters require reprojection before planar mesh loading: " +
                             path);
  const float nodata = (float)r.nodata;
  const double scale = r.has_scale ? r.scale : 1.0, offset = r.has_offset ? r.offset : 0.0;
  r.center_raw = r.data[(size_t)r.width * (r.height / 2) + r.width / 2];
  size_t missing = 0;
  for (float &v : r.data) {
    if (!std::isfinite(v) || (r.has_nodata && v == nodata)) {
      v = std::numeric_limits<float>::quiet_NaN();
      ++missing;
    } else {
      v = (float)(v * scale + offset);
    }
  }
  r.x_coords.resize(r.width);
  r.y_coords.resize(r.height);
  for (int j = 0; j < r.width; ++j) r.x_coords[j] = (float)(gt[0] + (j + 0.5) * gt[1] + 0.5 * gt[2]);
  for (int i = 0; i < r.height; ++i) r.y_coords[i] = (float)(gt[3] + 0.5 * gt[4] + (i + 0.5) * gt[5]);
  return missing;
}

}  // namespace dvsio
