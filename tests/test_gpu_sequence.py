"""GPU: bin/sequence (apps/main_sequence.cpp) -- a short dataset sequence
rendered from GeoTIFF rasters over the repeated trajectory -- against the
same pipeline driven through the binding: the rasters meshed on the device,
every interval's states from the host trajectory through the back-and-forth
phase map (repeated_trajectory.hpp:15-31), windows of 64 (native) or 200
halved on rejection (optimized), frames at cam_fps and one at the end.
events.h5 and frames.h5 must hold exactly the binding's events and frames;
render_progress.json the counts and the readout's drop figure.
"""
from __future__ import annotations

import json
import math
import os
import subprocess

import numpy as np
import pytest

import geotiff
from dvsrender4deepspace_amd import build, capi, driver, eventio, terrain, workloads
from test_gpu_apps import shepperd

pytestmark = pytest.mark.gpu

SECONDS = 0.03
TURN = 0.5


def _dataset(tmp_path):
    import dataclasses

    w = workloads.small_variant(workloads.PLAIN_SMALL, size=64, width=32, height=24)
    root = tmp_path / "dataset"
    seq = root / "seq0"
    seq.mkdir(parents=True)
    heights = terrain.heightfield(dataclasses.replace(w.terrain, size=97)).astype(np.float32)
    heights[5:8, 60:70] = np.nan
    geotiff.write(root / "dem.tif", heights, transform=(-48.0, 1.0, 0.0, 48.0, 0.0, -1.0), nodata=-32768.0,
                  compression="lzw", predictor=3, geokeys=(1024, 0, 2))  # geographic: axes are lon / lat
    alb = (0.1 + 0.2 * np.random.default_rng(4).random((40, 40))).astype(np.float32)
    geotiff.write(root / "albedo.tif", alb, transform=(-50.0, 2.5, 0.0, 50.0, 0.0, -2.5), compression="deflate",
                  geokeys=(1024, 0, 2))
    workloads.write_text_trajectory(w, seq / "source_trajectory.txt", frames=3)
    workloads.write_config(w, seq / "source_config.yaml", frame_count=3, imu_fps=500.0, cam_fps=250.0)
    return root, seq


def _repeat(t, start, duration):
    """The phase map of repeated_trajectory.hpp:15-31 in double: (float32
    source time, phase rate)."""
    turn = min(TURN, duration / 4)
    r = math.fmod(t, 2 * duration)
    phase = r if r <= duration else 2 * duration - r
    rate = 1.0 if r <= duration else -1.0
    if phase < turn:
        x = phase / turn
        phase = turn * (3 * x * x - 3 * x * x * x + x * x * x * x)
        rate *= 6 * x - 9 * x * x + 4 * x * x * x
    elif duration - phase < turn:
        x = (duration - phase) / turn
        phase = duration - turn * (3 * x * x - 3 * x * x * x + x * x * x * x)
        rate *= 6 * x - 9 * x * x + 4 * x * x * x
    return np.float32(start + phase), rate


def _states(traj, times, start, duration):
    mapped = [_repeat(t, start, duration) for t in times]
    s = traj.evaluate(np.array([m[0] for m in mapped], np.float32))
    f = np.array([np.float32(m[1]) for m in mapped], np.float32)[:, None]
    for k in ("linear_velocity", "angular_velocity", "light_direction_rate"):
        s[k] = s[k] * f
    s["light_intensity_rate"] = (s["light_intensity_rate"].astype(np.float64)
                                 * np.array([m[1] for m in mapped])).astype(np.float32)
    return s


def _binding(root, seq, optimized):
    cfg = eventio.load_config(seq / "source_config.yaml")
    dem, albedo = eventio.load_raster(root / "dem.tif"), eventio.load_raster(root / "albedo.tif")
    aligned = capi.raster_align(dem.x_coords, dem.y_coords, albedo.data, albedo.x_coords, albedo.y_coords)
    v, n, idx, a = capi.raster_mesh(dem.data, dem.x_coords, dem.y_coords, aligned)
    nodes, root_node, leaf_idx = capi.bvh_build(v, idx, 64)
    cam = capi.Camera(cfg.camera.width, cfg.camera.height, cfg.camera.fx, cfg.camera.fy, cfg.camera.cx,
                      cfg.camera.cy, cfg.scene.supersample_factor)
    scene = capi.Scene(nodes, root_node, v, n, leaf_idx, cam, cfg.hapke, albedos=a)
    traj = capi.Trajectory(eventio.load_frames_text(seq / "source_trajectory.txt"), cfg.scene.fps)
    start, end, _ = traj.range()
    duration = float(np.float32(end) - np.float32(start))  # end_time() - start_time() in float
    dt = float(np.float32(cfg.scene.time_resolution))
    times = [0.0]
    step = 0
    while times[-1] < SECONDS:
        step += 1
        times.append(min(SECONDS, step * dt))
    states = capi.optical_states(_states(traj, times, start, duration))
    iv = np.zeros(len(times) - 1, capi.INTERVAL_DTYPE)
    iv["start"], iv["end"] = states[:-1], states[1:]
    iv["t0"], iv["t1"] = times[:-1], times[1:]
    sensor = capi.Sensor(scene, cfg.physic)
    sensor.initialize(states[:1], 0.0)
    batch = driver.batch_size(cfg.physic, cfg.scene.time_resolution, optimized)
    events = driver.drive(sensor, iv, adaptive=optimized, batch=batch)
    events = np.concatenate(events) if events else np.zeros(0, capi.EVENT_DTYPE)
    cam_fps = float(np.float32(cfg.scene.cam_fps))
    frame_times = [k / cam_fps for k in range(int(SECONDS * cam_fps + 1e-9) + 1) if k / cam_fps <= SECONDS]
    if abs(frame_times[-1] - SECONDS) > 1e-9:
        frame_times.append(SECONDS)
    images, depths = [], []
    for t in frame_times:
        s = _states(traj, [t], start, duration)
        pose = capi.optical_states(s)
        pose["velocity"] = pose["light_rate"] = pose["intensity_rate"] = 0
        pose["rotation"][0] = shepperd(s["orientation"][0])
        img, depth = scene.render(pose)
        images.append(eventio.frame_bytes(img[0]))
        depths.append(depth[0])
    return events, sensor, np.array(frame_times), np.array(images), np.array(depths)


@pytest.mark.parametrize("optimized", [False, True], ids=["native", "optimized"])
def test_sequence_equals_binding(gpu, tmp_path, optimized):
    root, seq = _dataset(tmp_path)
    args = [str(build.BINDIR / "sequence")] + (["--optimized"] if optimized else [])
    args += [str(root), str(seq), str(root / "dem.tif"), str(root / "albedo.tif")]
    env = dict(os.environ, DVS_SEQUENCE_SECONDS=str(SECONDS))
    print(f"running {' '.join(args)}", flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "SCENE_GPU_READY" in out.stdout
    events, sensor, frame_times, images, depths = _binding(root, seq, optimized)
    h5 = eventio.read_h5(seq / "events.h5")
    assert len(events) > 0 and len(h5["t"]) == len(events)
    order = np.argsort(events["timestamp"], kind="stable")  # each appended batch is ordered by time
    assert np.array_equal(h5["t"], np.round(events["timestamp"][order] * 1e6).astype(np.int64))
    assert np.array_equal(h5["x"], events["x"][order]) and np.array_equal(h5["y"], events["y"][order])
    assert np.array_equal(h5["p"], events["polarity"][order])
    fr = eventio.read_frames_h5(seq / "frames.h5")
    assert np.array_equal(fr["t"], np.array([round(t * 1e6) for t in frame_times], np.int64))
    assert np.array_equal(fr["images"], images) and np.array_equal(fr["depth"], depths)
    assert np.count_nonzero(images) > 0
    progress = json.loads((seq / "render_progress.json").read_text())
    assert progress["events"] == len(events) and progress["frames"] == len(frame_times)
    assert progress["readout_dropped_events"] == sensor.readout_stats()[0]
    assert progress["numerical_model"] == ("adaptive_optics_native_circuit_cells" if optimized
                                           else "native_rays_and_circuit_cells")
    if optimized:
        model = json.loads((seq / "numerical_model.json").read_text())
        assert model["maximum_window_seconds"] == pytest.approx(200 * float(np.float32(1e-5)))
        assert progress["convergence_checks"] >= 1
    for name in ("poses.csv", "motion.csv", "pose_geographic.csv", "camera_used.yaml"):
        assert (seq / name).exists()
    print(f"sequence ({'optimized' if optimized else 'native'}): {len(events)} events, {len(frame_times)} frames",
          flush=True)


def test_cam_blur_windows_equals_binding(gpu, tmp_path):
    """bin/cam_blur_windows (apps/main_cam_blur_windows.cpp): three 50 ms
    windows of the repeated trajectory, cells of blur_resolution, double
    sums of weight x frame; PNGs equal the binding's render_blur(wide=True)
    of the same cell poses."""
    from PIL import Image

    root, seq = _dataset(tmp_path)
    centers = ("0.03", "0.5", "1.234")
    out_dir = tmp_path / "blur"
    args = [str(build.BINDIR / "cam_blur_windows"), str(seq), str(root / "dem.tif"), str(root / "albedo.tif"),
            str(out_dir), *centers]
    print(f"running {' '.join(args)}", flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("WINDOW center_s=") == 3
    cfg = eventio.load_config(seq / "source_config.yaml")
    dem, albedo = eventio.load_raster(root / "dem.tif"), eventio.load_raster(root / "albedo.tif")
    aligned = capi.raster_align(dem.x_coords, dem.y_coords, albedo.data, albedo.x_coords, albedo.y_coords)
    v, n, idx, a = capi.raster_mesh(dem.data, dem.x_coords, dem.y_coords, aligned)
    nodes, root_node, leaf_idx = capi.bvh_build(v, idx, 64)
    cam = capi.Camera(cfg.camera.width, cfg.camera.height, cfg.camera.fx, cfg.camera.fy, cfg.camera.cx,
                      cfg.camera.cy, cfg.scene.supersample_factor)
    scene = capi.Scene(nodes, root_node, v, n, leaf_idx, cam, cfg.hapke, albedos=a)
    traj = capi.Trajectory(eventio.load_frames_text(seq / "source_trajectory.txt"), cfg.scene.fps)
    start, end, _ = traj.range()
    duration = float(np.float32(end) - np.float32(start))
    step = float(np.float32(cfg.scene.blur_resolution))
    poses, weights, offsets, names = [], [], [0], []
    for c in centers:
        center = float(c)
        begin, finish = center - 0.025, center + 0.025
        for k in range(math.ceil(0.05 / step - 1e-10)):
            left, right = begin + k * step, min(finish, begin + (k + 1) * step)
            if right <= left:
                continue
            s = _states(traj, [(left + right) * 0.5], start, duration)
            pose = capi.optical_states(s)
            pose["velocity"] = pose["light_rate"] = pose["intensity_rate"] = 0
            pose["rotation"][0] = shepperd(s["orientation"][0])
            poses.append(pose[0])
            weights.append((right - left) / 0.05)
        offsets.append(len(poses))
        names.append(f"motion_blur_{begin:.3f}_{finish:.3f}s.png")
    frames = scene.render_blur(np.array(poses), weights, offsets, wide=True)
    for name, frame in zip(names, frames):
        with Image.open(out_dir / name) as im:
            got = np.array(im)
        assert np.array_equal(got, eventio.frame_bytes(frame)), name
        assert np.count_nonzero(got) > 0
