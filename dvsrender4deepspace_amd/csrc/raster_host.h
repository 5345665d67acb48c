// Minimal GeoTIFF reader (raster_host.cpp).
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace dvsio {

// Band 1 of a raster as float32 (GDALRasterBand::RasterIO to GDT_Float32),
// with its geotransform, NoData value, scale, offset and projection key.
struct Raster {
  int width = 0, height = 0;
  std::vector<float> data, x_coords, y_coords;
  double transform[6] = {0, 1, 0, 0, 0, 1};
  bool has_nodata = false, has_scale = false, has_offset = false;
  double nodata = 0, scale = 1, offset = 0;
  std::string projection;
  std::vector<double> geokeys, geodoubles;  // GeoKeyDirectoryTag, GeoDoubleParamsTag
  float center_raw = 0;  // the centre pixel's stored sample, before NoData / scale / offset
};

// The value of a GeoKey: its SHORT value or its GeoDoubleParams entry.
bool geokey_value(const Raster &r, int key, double &value);

// Reads the file; throws std::runtime_error with loadImageryData's messages.
Raster read_geotiff(const std::string &path);
// The NoData / scale / offset mapping and pixel-centre coordinates of
// loadImageryData; returns the number of NoData samples.
size_t finish_imagery(Raster &r, const std::string &path);

}  // namespace dvsio
