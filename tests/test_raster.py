"""CPU: the GeoTIFF reader behind tools::loadImageryData
(include/tools/loader.hpp:31-100) without GDAL.

* the known answers of the reference's tests/test_raster.cu:14-20 (a 4x4
  float DEM, geotransform {100, 1, 0, 200, 0, -1}, NoData -9999, scale 2,
  offset 10): NaN at NoData, data[6] == 16, pixel centres 100.5 / 199.5, in
  every layout the reader takes;
* files written by libtiff (through Pillow) -- LZW, Deflate, none, with
  GeoTIFF tags -- decode to the written samples, and libtiff decodes the LZW
  files of the test writer (tests/geotiff.py) to the same samples, so the
  reader's LZW is pinned by an independent codec in both directions;
* PixelIsPoint, model transformation, rejected rotations, a missing
  geotransform, and the projection key.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import geotiff
from dvsrender4deepspace_amd import eventio

LAYOUTS = [dict(), dict(compression="lzw"), dict(compression="deflate"), dict(compression="lzw", predictor=3),
           dict(compression="deflate", predictor=3, endian=">"), dict(tile=(16, 16)),
           dict(bigtiff=True, compression="lzw"), dict(rows_per_strip=1, endian=">"),
           dict(model_transformation=True, compression="deflate")]


def _reference_dem():
    data = np.array([[r + c for c in range(4)] for r in range(4)], np.float32)
    data[0, 0] = -9999
    return data


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda d: ",".join(f"{k}={v}" for k, v in d.items()) or "plain")
def test_reference_known_answers(tmp_path, layout):
    path = tmp_path / "dem.tif"
    geotiff.write(path, _reference_dem(), transform=(100, 1, 0, 200, 0, -1), nodata=-9999, scale=2, offset=10,
                  **layout)
    im = eventio.load_raster(path)
    assert im.width == im.height == 4 and im.missing == 1
    assert np.isnan(im.data[0, 0]) and im.data.ravel()[6] == 16
    assert im.x_coords[0] == 100.5 and im.y_coords[0] == 199.5
    valid = ~np.isnan(im.data)
    want = (np.add.outer(np.arange(4), np.arange(4)) * 2.0 + 10).astype(np.float32)
    assert np.array_equal(im.data[valid], want[valid])
    assert im.transform == (100, 1, 0, 200, 0, -1)


def _pillow_geotiff(path, array, compression, predictor=None):
    from PIL import Image, TiffImagePlugin
    ifd = TiffImagePlugin.ImageFileDirectory_v2()
    for tag, value, kind in ((33550, (2.0, 3.0, 0.0), 12), (33922, (0.0, 0.0, 0.0, 500.0, 900.0, 0.0), 12),
                             (34735, (1, 1, 0, 1, 1024, 0, 1, 1), 3), (42113, "-32768", 2)):
        ifd[tag] = value
        ifd.tagtype[tag] = kind
    if predictor:
        ifd[317] = predictor
        ifd.tagtype[317] = 3
    mode = "F" if array.dtype == np.float32 else None
    Image.fromarray(array, mode=mode).save(path, compression=compression, tiffinfo=ifd)


@pytest.mark.parametrize("compression", ["tiff_lzw", "tiff_adobe_deflate", "raw"])
@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.int16])
def test_libtiff_written_files(tmp_path, compression, dtype):
    rng = np.random.default_rng(7)
    a = (rng.normal(size=(211, 157)) * 900).astype(dtype)
    path = tmp_path / "pil.tif"
    # Pillow's own (uncompressed) writer stores a Predictor tag without predicting
    _pillow_geotiff(path, a, compression, predictor=2 if dtype != np.float32 and compression != "raw" else None)
    im = eventio.load_raster(path)
    assert np.array_equal(im.data, a.astype(np.float32))
    assert im.transform == (500.0, 2.0, 0.0, 900.0, 0.0, -3.0) and im.nodata == -32768
    assert im.x_coords[1] == 503.0 and im.y_coords[1] == np.float32(895.5)


@pytest.mark.parametrize("predictor", [1, 3])
def test_libtiff_decodes_the_test_writers_lzw(tmp_path, predictor):
    from PIL import Image
    a = (np.random.default_rng(8).normal(size=(120, 101)) * 100).astype(np.float32)  # several full LZW tables
    path = tmp_path / "mine.tif"
    geotiff.write(path, a, compression="lzw", predictor=predictor)
    assert np.array_equal(np.array(Image.open(path)), a)
    assert np.array_equal(eventio.load_raster(path).data, a)


@pytest.mark.parametrize("dtype,compression,predictor,tile,endian", [
    combo for combo in itertools.product([np.float32, np.float64, np.int16, np.uint8, np.int32],
                                         ["none", "lzw", "deflate"], [1, 2, 3], [None, (32, 48)], ["<", ">"])
    if not (combo[2] == 3 and combo[0] not in (np.float32, np.float64))
    and not (combo[2] == 2 and combo[0] in (np.float32, np.float64))])
def test_writer_layouts_round_trip(tmp_path, dtype, compression, predictor, tile, endian):
    rng = np.random.default_rng(9)
    a = (rng.normal(size=(41, 37)) * 60).astype(dtype)
    path = tmp_path / "x.tif"
    geotiff.write(path, a, compression=compression, predictor=predictor, tile=tile, endian=endian,
                  rows_per_strip=None if tile else 9)
    assert np.array_equal(eventio.load_raster(path).data, a.astype(np.float32))


def test_pixel_is_point_and_model_transformation(tmp_path):
    a = np.zeros((5, 6), np.float32)
    gt = (1000.0, 2.5, 0.0, 3000.0, 0.0, -2.5)
    for kw in (dict(pixel_is_point=True), dict(model_transformation=True)):
        geotiff.write(tmp_path / "p.tif", a, transform=gt, **kw)
        im = eventio.load_raster(tmp_path / "p.tif")
        assert im.transform == gt
        assert im.x_coords[0] == np.float32(1001.25) and im.y_coords[-1] == np.float32(3000 - 4.5 * 2.5)


def test_rejections(tmp_path):
    from PIL import Image
    a = np.zeros((4, 4), np.float32)
    geotiff.write(tmp_path / "r.tif", a, transform=(0, 1, 0.25, 0, 0, -1), model_transformation=True)
    with pytest.raises(eventio.ConfigError, match="Rotated/sheared or degenerate rasters require reprojection"):
        eventio.load_raster(tmp_path / "r.tif")
    Image.fromarray(a, mode="F").save(tmp_path / "plain.tif")
    with pytest.raises(eventio.ConfigError, match="Failed to get geotransform from dataset"):
        eventio.load_raster(tmp_path / "plain.tif")
    with pytest.raises(eventio.ConfigError, match="Failed to open dataset"):
        eventio.load_raster(tmp_path / "missing.tif")
    (tmp_path / "junk.tif").write_bytes(b"not a tiff at all")
    with pytest.raises(eventio.ConfigError, match="Failed to read raster data"):
        eventio.load_raster(tmp_path / "junk.tif")


def test_projection_key(tmp_path):
    a = np.zeros((3, 3), np.float32)
    geotiff.write(tmp_path / "a.tif", a, geokeys=(1024, 1, 1, 3072, 32, 633))
    geotiff.write(tmp_path / "b.tif", a, geokeys=(1024, 1, 1, 3072, 32, 633), compression="lzw")
    geotiff.write(tmp_path / "c.tif", a, geokeys=(1024, 1, 1, 3072, 32, 634))
    pa, pb, pc = (eventio.load_raster(tmp_path / f"{n}.tif").projection for n in "abc")
    assert pa == pb != pc and pa.startswith("geokeys")
