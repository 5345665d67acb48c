"""GPU: DEM -> mesh on the device (dvs_raster_mesh, restating the planar
path of src/tools/build_from_raster.cu) and the bilinear albedo alignment
(dvs_raster_align, batch_align_kernel) against the oracle: vertices,
triangle indices and albedos bit-identical (both number them in raster
order), normals within 2e-6 (the device's and the host's acosf may differ
in the last bit), aligned albedos bit-identical; the reference's known
answers (tests/test_raster.cu); NoData holes, a decreasing y axis and
projected coordinates of 500 km; and the same mesh on every run.
"""
from __future__ import annotations

import numpy as np
import pytest

from dvsrender4deepspace_amd import capi, terrain

pytestmark = pytest.mark.gpu

NORMAL_TOL = 2e-6


def _assert_same_mesh(got, want):
    v, n, t, a = got
    ov, on, ot, oa = want
    assert v.shape == ov.shape and t.shape == ot.shape
    assert np.array_equal(v, ov) and np.array_equal(t, ot)
    assert np.abs(n - on).max(initial=0) <= NORMAL_TOL
    if oa is not None:
        assert np.array_equal(a, oa, equal_nan=True)


def test_reference_dem(oracle, gpu):
    d = (np.add.outer(np.arange(4), np.arange(4)) * 2.0 + 10).astype(np.float32)
    d[0, 0] = np.nan
    xs = (100.5 + np.arange(4)).astype(np.float32)
    ys = (199.5 - np.arange(4)).astype(np.float32)
    got = gpu.raster_mesh(d, xs, ys)
    assert len(got[0]) == 15 and len(got[2]) == 16
    _assert_same_mesh(got, oracle.raster_mesh(d, xs, ys))
    d[2, 2] = np.nan
    with pytest.raises(capi.DvsError, match="DEM center is NoData"):
        gpu.raster_mesh(d, xs, ys)


def _utm_dem(size=257, seed=11):
    spec = terrain.TerrainSpec(size=size, amplitude=120.0, craters=20, seed=seed, spacing=5.0)
    h = terrain.heightfield(spec).astype(np.float32) + 1500.0
    rng = np.random.default_rng(seed)
    holes = rng.random(h.shape) < 0.01
    holes[size // 2, size // 2] = False
    h[holes] = np.nan
    h[:3, :7] = np.nan  # a NoData corner block
    xs = (512000.0 + 5.0 * (np.arange(size) + 0.5)).astype(np.float32)
    ys = (4100000.0 - 5.0 * (np.arange(size) + 0.5)).astype(np.float32)
    return h, xs, ys


def test_utm_dem_with_holes_and_albedo(oracle, gpu):
    h, xs, ys = _utm_dem()
    # an albedo raster at a coarser pitch covering the DEM, aligned on the device
    sx = (511990.0 + 12.0 * np.arange(110)).astype(np.float32)
    sy = (4100010.0 - 12.0 * np.arange(110)).astype(np.float32)
    src = (0.12 + 0.05 * np.sin(sx[None, :] / 50.0) * np.cos(sy[:, None] / 70.0)).astype(np.float32)
    src[40:44, 50:53] = np.nan
    aligned = gpu.raster_align(xs, ys, src, sx, sy)
    assert np.array_equal(aligned, oracle.raster_align(xs, ys, src, sx, sy), equal_nan=True)
    assert np.isnan(aligned).any() and np.isfinite(aligned).mean() > 0.9
    got = gpu.raster_mesh(h, xs, ys, aligned)
    _assert_same_mesh(got, oracle.raster_mesh(h, xs, ys, aligned))
    again = gpu.raster_mesh(h, xs, ys, aligned)
    for a, b in zip(got, again):
        assert a.tobytes() == b.tobytes()
    print(f"utm DEM: {len(got[0])} vertices, {len(got[2])} triangles", flush=True)


def test_alignment_known_answers(gpu):
    x = np.array([2, 1, 0], np.float32)
    y = np.array([0, 1, 2], np.float32)
    rx = np.array([0, .5, 1.5, 2, -1], np.float32)
    ry = np.array([0, .5, 1.5, 2, 3], np.float32)
    source = (2 * x[None, :] + 3 * y[:, None]).astype(np.float32)
    out = gpu.raster_align(rx, ry, source, x, y)
    inside = np.zeros((5, 5), bool)
    inside[:4, :4] = True
    assert np.isnan(out[~inside]).all()
    assert np.abs(out[inside] - (2 * rx[None, :4] + 3 * ry[:4, None]).ravel()).max() < 1e-6
    source.ravel()[4] = np.nan
    out = gpu.raster_align(rx, ry, source, x, y)
    assert np.isnan(out.ravel()[6]) and out.ravel()[0] == 0


def test_full_size_dem(oracle, gpu):
    """A 2048^2 DEM (8,380,418 triangles without holes), the high-resolution
    request's terrain size."""
    spec = terrain.TerrainSpec(size=2048, amplitude=120.0, craters=100, seed=9)
    h = terrain.heightfield(spec).astype(np.float32)
    axis = (np.arange(2048) + 0.5).astype(np.float32)
    got = gpu.raster_mesh(h, axis, axis[::-1].copy())
    assert len(got[2]) == 2 * 2047 * 2047
    _assert_same_mesh(got, oracle.raster_mesh(h, axis, axis[::-1].copy()))


def test_cam_from_dem_and_albedo_equals_binding(gpu, tmp_path):
    """bin/cam --dem --albedo --traj (the GeoTIFF rasters meshed on the
    device, apps/app_inputs.hpp:10-14) against the same rasters through the
    binding: every frame's PNG identical."""
    import dataclasses
    import subprocess

    from PIL import Image

    import geotiff
    from dvsrender4deepspace_amd import build, eventio, workloads
    from test_gpu_apps import shepperd

    w = workloads.small_variant(workloads.PLAIN_SMALL, size=64, width=32, height=24)
    spec = dataclasses.replace(w.terrain, size=97)
    heights = terrain.heightfield(spec).astype(np.float32)
    heights[5:8, 60:70] = np.nan  # NoData away from the centre
    gt = (-48.0, 1.0, 0.0, 48.0, 0.0, -1.0)  # pixel centres -47.5 .. 47.5, y decreasing
    geotiff.write(tmp_path / "dem.tif", heights, transform=gt, nodata=-32768.0, compression="lzw", predictor=3,
                  geokeys=(1024, 1, 1, 3072, 32, 633))
    alb = (0.1 + 0.2 * np.random.default_rng(4).random((40, 40))).astype(np.float32)
    geotiff.write(tmp_path / "albedo.tif", alb, transform=(-50.0, 2.5, 0.0, 50.0, 0.0, -2.5), compression="deflate",
                  geokeys=(1024, 1, 1, 3072, 32, 633))
    workloads.write_text_trajectory(w, tmp_path / "traj.txt", frames=3)
    cfg_path = tmp_path / "config.yaml"
    workloads.write_config(w, cfg_path, frame_count=3, cam_fps=250.0)
    exe = build.BINDIR / "cam"
    print(f"running bin/cam on {tmp_path / 'dem.tif'}", flush=True)
    out = subprocess.run([str(exe), "--dem", str(tmp_path / "dem.tif"), "--albedo", str(tmp_path / "albedo.tif"),
                          "--traj", str(tmp_path / "traj.txt"), "--config", str(cfg_path), "--output",
                          str(tmp_path / "out")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "DEM dimensions: 97x97" in out.stdout and "Albedo dimensions: 40x40" in out.stdout
    dem, albedo = eventio.load_raster(tmp_path / "dem.tif"), eventio.load_raster(tmp_path / "albedo.tif")
    aligned = gpu.raster_align(dem.x_coords, dem.y_coords, albedo.data, albedo.x_coords, albedo.y_coords)
    v, n, idx, a = gpu.raster_mesh(dem.data, dem.x_coords, dem.y_coords, aligned)
    assert f"Generated {len(v)} vertices, {len(v)} normals, and {len(idx)} triangles." in out.stdout
    nodes, root, leaf_idx = capi.bvh_build(v, idx, 64)
    cfg = eventio.load_config(cfg_path)
    cam = capi.Camera(cfg.camera.width, cfg.camera.height, cfg.camera.fx, cfg.camera.fy, cfg.camera.cx,
                      cfg.camera.cy, cfg.scene.supersample_factor)
    scene = capi.Scene(nodes, root, v, n, leaf_idx, cam, cfg.hapke, albedos=a)
    frames = eventio.load_frames_text(tmp_path / "traj.txt")
    traj = capi.Trajectory(frames, cfg.scene.fps)
    start, end = np.float32(0), np.float32(2 / np.float32(cfg.scene.fps))
    steps = int(np.ceil((float(end) - float(start)) * float(np.float32(cfg.scene.cam_fps))))
    interval = np.float32(1) / np.float32(cfg.scene.cam_fps)
    assert f"Wrote {steps} images" in out.stdout
    lit = 0
    for step in range(steps):
        t = np.float32(start + np.float32(step) * interval)
        s = traj.evaluate([t])[0]
        pose = capi.optical_states(s[None])
        pose["velocity"] = 0
        pose["light_rate"] = 0
        pose["intensity_rate"] = 0
        pose["rotation"][0] = shepperd(s["orientation"])
        img, depth = scene.render(pose)
        with Image.open(tmp_path / "out" / "images" / f"{float(t):f}.png") as im:
            got = np.array(im)
        assert np.array_equal(got, eventio.frame_bytes(img[0]))
        lit += np.count_nonzero(got)
    assert lit > 0
