"""Pins the oracle's adaptive optics (OpticalSampler::refine and
EventSensor::process_adaptive) against the properties the reference's own
sensor test checks (tests/test_sensor.cu:94-158), on the same plane fixture
with a smaller pixel grid (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

import scenes


@pytest.fixture(scope="module")
def fixture(oracle):
    v, n, nodes, root, idx, cam, phys = scenes.sensor_fixture()
    scene = oracle.Scene(nodes, root, v, n, idx, cam, scenes.capi.Hapke())
    return scene, cam, phys


def _pair(oracle, cam, phys, scene, pose):
    a = oracle.Sensor(cam.width, cam.height, phys)
    b = oracle.Sensor(cam.width, cam.height, phys)
    a.initialize(scene, pose, 0.0)
    b.initialize(scene, pose, 0.0)
    return a, b


def _key(ev):
    return np.stack([ev["y"].astype(np.int64), ev["x"].astype(np.int64)], 1)


def _by_pixel(ev, npix, width):
    order = np.lexsort((ev["timestamp"], ev["x"], ev["y"]))
    ev = ev[order]
    pix = ev["y"].astype(np.int64) * width + ev["x"]
    return ev, pix


@pytest.mark.parametrize("shot", [0, 1])
def test_static_optics_reuse_is_exact(oracle, fixture, shot):
    """tests/test_sensor.cu:96-108: a static scene needs the five initial
    probes only and gives the native path's seeded events exactly."""
    scene, cam, phys = fixture
    phys = scenes.dataclasses.replace(phys, photon_shot_noise=shot)
    pose = scenes.nadir_state()
    native, adaptive = _pair(oracle, cam, phys, scene, pose)
    ivs = np.concatenate([scenes.interval(pose, pose, 0, .001), scenes.interval(pose, pose, .001, .002)])
    code, exact, _ = native.process(scene, ivs)
    assert code == 0
    code, ok, cached, _, st, _ = adaptive.process_adaptive(scene, ivs, tolerance=.001)
    assert code == 0 and ok
    assert st["fallback_pixels"] == 0 and st["maximum_knots"] == 5
    assert st["traced_pixel_samples"] == 5 * cam.width * cam.height
    assert len(exact) == len(cached)
    for f in ("timestamp", "x", "y", "polarity"):
        assert np.array_equal(exact[f], cached[f])
    assert np.all(cached["depth"] == -1)
    assert np.array_equal(native.states(), adaptive.states())


def test_interior_pulse_converges(oracle, fixture):
    """tests/test_sensor.cu:110-131: a Hermite light pulse with equal
    endpoints is found; per-pixel counts agree on 99.9% of pixels and matched
    crossings within three circuit cells."""
    scene, cam, phys = fixture
    pose = scenes.nadir_state()
    start, end = pose.copy(), pose.copy()
    start["intensity_rate"], end["intensity_rate"] = 8e6, -8e6
    native, adaptive = _pair(oracle, cam, phys, scene, pose)
    ivs = scenes.interval(start, end, 0, .002)
    _, exact, _ = native.process(scene, ivs)
    code, ok, cached, _, st, _ = adaptive.process_adaptive(scene, ivs, tolerance=1e-4)
    assert code == 0 and ok
    npix = cam.width * cam.height
    assert st["fallback_pixels"] < npix
    assert len(cached) > 0
    a, pa = _by_pixel(exact, npix, cam.width)
    b, pb = _by_pixel(cached, npix, cam.width)
    ca, cb = np.bincount(pa, minlength=npix), np.bincount(pb, minlength=npix)
    assert np.count_nonzero(ca != cb) < npix * .001
    same = np.repeat(ca == cb, ca)
    sa = np.repeat(ca == cb, ca)
    sb = np.repeat(ca == cb, cb)
    assert np.array_equal(a["polarity"][sa], b["polarity"][sb])
    assert np.max(np.abs(a["timestamp"][sa] - b["timestamp"][sb])) < 3 * phys.circuit_step
    assert same.sum() > 0


def test_discontinuity_uses_local_exact_cells(oracle, fixture):
    """tests/test_sensor.cu:132-144: a step in intensity that no knot can
    resolve is traced natively cell by cell without whole-pixel fallback, and
    the events equal the native ones."""
    scene, cam, phys = fixture
    pose = scenes.nadir_state()
    bright = pose.copy()
    bright["intensity"] = 2000
    native, adaptive = _pair(oracle, cam, phys, scene, pose)
    ivs = np.concatenate([scenes.interval(pose, pose, 0, .001), scenes.interval(bright, bright, .001, .002)])
    _, exact, _ = native.process(scene, ivs)
    code, ok, cached, _, st, _ = adaptive.process_adaptive(scene, ivs, tolerance=1e-4)
    assert code == 0 and ok
    assert st["native_cell_queries"] > 0 and st["fallback_pixels"] == 0 and st["exact_segments"] > 0
    npix = cam.width * cam.height
    a, _ = _by_pixel(exact, npix, cam.width)
    b, _ = _by_pixel(cached, npix, cam.width)
    assert len(a) == len(b) > 0
    for f in ("x", "y", "polarity"):
        assert np.array_equal(a[f], b[f])
    assert np.max(np.abs(a["timestamp"] - b["timestamp"])) < 2e-12


def test_convergence_gate_rejects_without_state_change(oracle, fixture):
    """tests/test_sensor.cu:146-159: a flash between the initial probes is
    caught by the receptor check before any pixel or noise state changes; a
    native retry gives the events of a fresh sensor."""
    scene, cam, phys = fixture
    phys = scenes.dataclasses.replace(phys, photon_shot_noise=1)
    pose = scenes.nadir_state()
    flash = pose.copy()
    flash["intensity"] = 20000
    checked, fresh = _pair(oracle, cam, phys, scene, pose)
    ivs = np.concatenate([scenes.interval(pose, pose, 0, .0003), scenes.interval(flash, flash, .0003, .0004),
                          scenes.interval(pose, pose, .0004, .002)])
    code, ok, got, _, _, probe = checked.process_adaptive(scene, ivs, tolerance=.5, check=True)
    assert code == 0 and not ok and len(got) == 0
    thr = min(phys.threshold_pos, phys.threshold_neg)
    assert probe["maximum_log_voltage_error"] > .02 * thr or probe["p99_log_voltage_error"] > .005 * thr
    assert probe["pixels"] == min(1024, cam.width * cam.height)
    _, retried, _ = checked.process(scene, ivs)
    _, exact, _ = fresh.process(scene, ivs)
    assert len(exact) == len(retried) > 0
    assert exact.tobytes() == retried.tobytes()


def test_adaptive_argument_checks(oracle, fixture):
    """process_adaptive rejects a window longer than the event capacity bound
    and a non-positive tolerance (event_sensor.cu:308-312)."""
    scene, cam, phys = fixture
    pose = scenes.nadir_state()
    s, _ = _pair(oracle, cam, phys, scene, pose)
    long = scenes.interval(pose, pose, 0, 63 * phys.refractory_time)
    assert s.process_adaptive(scene, long)[0] == 3
    assert s.process_adaptive(scene, scenes.interval(pose, pose, 0, .001), tolerance=0)[0] == 3
    assert s.process_adaptive(scene, scenes.interval(pose, pose, 0, .001))[0] == 0
