"""GPU parity at the full request sizes: the single DAVIS346 request
(346x260, 4x4 sub-rays, 1024^2 terrain with 2,093,058 triangles, noise on)
and the high-resolution request (1280x720, 4x4 sub-rays, 2048^2 terrain with
8,380,418 triangles, 1 ms horizons, noise on).

Against the oracle:
  * DAVIS346: the request's first adaptive window (200 horizons, 2,000
    circuit cells per pixel, receptor check on) over the full sensor, under
    the adaptive bar of tests/test_gpu_adaptive.py;
  * high resolution: the full 1280x720 frame at two times (per-pixel radiance
    within 1e-4, depth exact, 1 LSB after to_byte), and one native 1 ms
    horizon of a 32x24 sensor window at the frame's corner (the same rays as
    those pixels of the full sensor) with noise on, per-pixel (x, y, p)
    identical and times within 1 us.
On the device alone, at the full 1280x720 size: pixels are independent, so
the corner window's events inside the full sensor's run (noise off, the
same horizon) equal the window sensor's events moved by the window's
offset, and a noise-on run repeats bit for bit.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import scenes
from dvsrender4deepspace_amd import capi, workloads
from test_gpu_adaptive import _assert_adaptive_match, _assert_stats_close
from test_gpu_parity import _assert_events_match, _frame_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def highres(oracle, gpu):
    w = workloads.HIGHRES
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    print(f"highres scene: {len(idx)} triangles", flush=True)
    return w, (v, n, nodes, root, idx)


def test_davis346_first_window_matches_oracle(oracle, gpu):
    """The DAVIS346 request's first bin/dvs window at full size: adaptive,
    200 horizons, receptor check on, shot and reset noise on."""
    w = workloads.DAVIS346
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    g = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    o = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    s = gpu.Sensor(g, w.physic)
    s.initialize(w.initial_state(), w.start_time)
    os_ = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    assert os_.initialize(o, w.initial_state(), w.start_time) == 0
    ivs = w.intervals(0, 200)
    ok, ready = s.process_adaptive(ivs, check_convergence=True)
    code, ook, oraw, oready, ost, oprobe = os_.process_adaptive(o, ivs, check=True, capacity=s.adaptive_capacity())
    assert code == 0 and ok == ook
    print(f"davis346 window 1: accepted {ok}, {len(oraw)} oracle events", flush=True)
    npix = w.camera.width * w.camera.height
    probe = s.last_convergence()
    assert probe["pixels"] == oprobe["pixels"] == 1024
    assert probe["maximum_log_voltage_error"] == pytest.approx(oprobe["maximum_log_voltage_error"], rel=1e-2, abs=1e-9)
    if ok:
        assert len(oraw) > 20
        _assert_stats_close(s.refinement_stats(), ost, npix)
        _assert_adaptive_match(s.raw_events(), oraw, w.camera.width, npix)
        _assert_adaptive_match(ready, oready, w.camera.width, npix)
        d = np.abs(s.states()["receptor_log"] - os_.states()["receptor_log"])
        assert np.median(d) < 1e-6 and np.max(d) < 1e-3


def test_highres_frame_parity(oracle, gpu, highres):
    w, (v, n, nodes, root, idx) = highres
    g = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    o = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    for t in (0.0, 0.5):
        img, dep = _frame_check(oracle, g, o, w.states(np.array([t])))
        assert np.count_nonzero(dep > 0) > 0.9 * dep.size


CORNER = (32, 24)


def _corner_camera(w):
    return dataclasses.replace(w.camera, width=CORNER[0], height=CORNER[1])  # cx, cy unchanged: pixels (0..31, 0..23)


def test_highres_native_horizon_corner_matches_oracle(oracle, gpu, highres):
    """One native 1 ms horizon (1,000 circuit cells, 4x4 sub-rays, noise on,
    thresholds 0.05) of the 32x24 corner of the 1280x720 sensor."""
    w, (v, n, nodes, root, idx) = highres
    cam = _corner_camera(w)
    phys = dataclasses.replace(w.physic, threshold_pos=0.05, threshold_neg=0.05)
    g = gpu.Scene(nodes, root, v, n, idx, cam, w.hapke)
    o = oracle.Scene(nodes, root, v, n, idx, cam, w.hapke)
    s = gpu.Sensor(g, phys)
    s.initialize(w.initial_state(), 0.0)
    os_ = oracle.Sensor(cam.width, cam.height, phys)
    assert os_.initialize(o, w.initial_state(), 0.0) == 0
    ivs = w.intervals(0, 1)
    ready = s.process(ivs)
    code, oraw, oready = os_.process(o, ivs)
    assert code == 0
    print(f"highres corner horizon: {len(oraw)} oracle events", flush=True)
    assert len(oraw) > 100
    _assert_events_match(s.raw_events(), oraw)
    _assert_events_match(ready, oready)


def test_highres_full_sensor_pixel_independence(gpu, highres):
    """The full 1280x720 sensor's native horizon (noise off) restricted to
    the corner equals the corner sensor's; a noise-on full-size horizon
    repeats bit for bit."""
    w, (v, n, nodes, root, idx) = highres
    phys = dataclasses.replace(w.physic, threshold_pos=0.05, threshold_neg=0.05, threshold_std=0.0,
                               photon_shot_noise=0, reset_thermal_noise=0)
    full = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    corner = full.view(_corner_camera(w))
    ivs = w.intervals(0, 1)
    runs = {}
    for name, scene in (("full", full), ("corner", corner)):
        s = gpu.Sensor(scene, phys)
        s.initialize(w.initial_state(), 0.0)
        s.process(ivs)
        runs[name] = s.raw_events()
    f = runs["full"]
    inside = f[(f["x"] < CORNER[0]) & (f["y"] < CORNER[1])]
    print(f"highres full horizon: {len(f)} events, {len(inside)} in the corner", flush=True)
    assert len(f) > 10 * len(inside) > 0
    a = inside[np.lexsort((inside["timestamp"], inside["x"], inside["y"]))]
    b = runs["corner"][np.lexsort((runs["corner"]["timestamp"], runs["corner"]["x"], runs["corner"]["y"]))]
    assert a.tobytes() == b.tobytes()
    noisy = []
    for _ in range(2):
        s = gpu.Sensor(full, w.physic)
        s.initialize(w.initial_state(), 0.0)
        s.process(ivs)
        noisy.append(s.raw_events())
    assert noisy[0].tobytes() == noisy[1].tobytes()
