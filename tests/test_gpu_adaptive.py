"""GPU: adaptive optics (wavefront refinement, local exact cells, fallback,
convergence gate) against the oracle's depth-first restatement of
OpticalSampler / EventSensor::process_adaptive, and the native path.

Tolerances.  Against the native path on the device: bit-exact.  Against the
oracle: refinement compares FP64 residuals and voltage errors of FP32
radiance with thresholds, and device and host radiance differ in the last
float bits (the Hapke evaluation; frames agree within 1e-4), so a segment
whose residual sits on the tolerance can resolve differently.  The bar is
the reference's own for adaptive optics (tests/test_sensor.cu:129-130):
per-pixel event sequences differ in at most 0.1% of the pixels (at least
one), matched events have the same polarities and times within three
circuit cells, and at most 0.1% of them differ by more than 1 us;
refinement statistics agree within a few such flips."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import scenes
from dvsrender4deepspace_amd import capi, driver, workloads

pytestmark = pytest.mark.gpu

TIME_TOL = 1e-6


def _per_pixel(events):
    return events[np.lexsort((events["timestamp"], events["x"], events["y"]))]


def _pixel_runs(events, width):
    ev = _per_pixel(events)
    pix = ev["y"].astype(np.int64) * width + ev["x"]
    return ev, pix


def _assert_adaptive_match(dev, ref, width, npix):
    a, pa = _pixel_runs(dev, width)
    b, pb = _pixel_runs(ref, width)
    ca, cb = np.bincount(pa, minlength=npix), np.bincount(pb, minlength=npix)
    same = ca == cb
    sa, sb = np.repeat(same, ca), np.repeat(same, cb)
    assert np.array_equal(a["polarity"][sa], b["polarity"][sb])
    dt = np.abs(a["timestamp"][sa] - b["timestamp"][sb])
    assert np.max(dt, initial=0) < 3e-6
    assert np.count_nonzero(dt >= TIME_TOL) <= max(1, int(len(dt) * 1e-3))
    differing = np.count_nonzero(~same)
    assert differing <= max(1, int(npix * 1e-3)), f"{differing} pixels differ"
    return differing


def _assert_stats_close(st, ost, npix):
    slack = max(4, int(npix * 1e-3) * 4)
    for k in ("fallback_pixels", "maximum_knots", "exact_segments", "native_cell_queries", "traced_pixel_samples"):
        assert abs(int(st[k]) - int(ost[k])) <= slack, (k, st[k], ost[k])


def _plane(oracle, gpu, phys=None):
    v, n, nodes, root, idx, cam, base = scenes.sensor_fixture()
    g = gpu.Scene(nodes, root, v, n, idx, cam, capi.Hapke())
    o = oracle.Scene(nodes, root, v, n, idx, cam, capi.Hapke())
    return g, o, cam, phys or base


@pytest.mark.parametrize("shot", [0, 1])
def test_static_reuse_equals_native(oracle, gpu, shot):
    """tests/test_sensor.cu:96-108 on the device: five probes per pixel and
    the native path's seeded events and states."""
    g, _, cam, phys = _plane(oracle, gpu)
    phys = dataclasses.replace(phys, photon_shot_noise=shot, threshold_pos=.05, threshold_neg=.05)
    pose = scenes.nadir_state()
    native, adaptive = gpu.Sensor(g, phys), gpu.Sensor(g, phys)
    native.initialize(pose, 0.0)
    adaptive.initialize(pose, 0.0)
    ivs = np.concatenate([scenes.interval(pose, pose, 0, .001), scenes.interval(pose, pose, .001, .002)])
    exact = native.process(ivs)
    ok, cached = adaptive.process_adaptive(ivs, tolerance=.001)
    assert ok
    st = adaptive.refinement_stats()
    assert st["fallback_pixels"] == 0 and st["maximum_knots"] == 5
    assert st["traced_pixel_samples"] == 5 * cam.width * cam.height
    assert len(exact) == len(cached)
    for f in ("timestamp", "x", "y", "polarity"):
        assert np.array_equal(exact[f], cached[f])
    assert np.all(cached["depth"] == -1)
    assert native.states().tobytes() == adaptive.states().tobytes()


def _pulse(pose):
    start, end = pose.copy(), pose.copy()
    start["intensity_rate"], end["intensity_rate"] = 8e6, -8e6
    return scenes.interval(start, end, 0, .002)


def _step(pose):
    bright = pose.copy()
    bright["intensity"] = 2000
    return np.concatenate([scenes.interval(pose, pose, 0, .001), scenes.interval(bright, bright, .001, .002)])


@pytest.mark.parametrize("case,tol", [("pulse", 1e-4), ("step", 1e-4), ("pulse", 1e-2)])
def test_plane_cases_match_oracle(oracle, gpu, case, tol):
    g, o, cam, phys = _plane(oracle, gpu)
    pose = scenes.nadir_state()
    ivs = _pulse(pose) if case == "pulse" else _step(pose)
    s = gpu.Sensor(g, phys)
    s.initialize(pose, 0.0)
    os_ = oracle.Sensor(cam.width, cam.height, phys)
    assert os_.initialize(o, pose, 0.0) == 0
    ok, ready = s.process_adaptive(ivs, tolerance=tol)
    code, ook, oraw, oready, ost, _ = os_.process_adaptive(o, ivs, tolerance=tol, capacity=s.adaptive_capacity())
    assert ok and ook and code == 0
    npix = cam.width * cam.height
    st = s.refinement_stats()
    _assert_stats_close(st, ost, npix)
    assert len(oraw) > 0
    _assert_adaptive_match(s.raw_events(), oraw, cam.width, npix)
    _assert_adaptive_match(ready, oready, cam.width, npix)
    if case == "step":
        assert st["native_cell_queries"] > 0


def _terrain(noise: bool):
    base = workloads.small_variant(workloads.DAVIS346_NOISE_OFF, size=192, width=48, height=36, craters=12,
                                   samples=2)
    phys = dataclasses.replace(base.physic, threshold_pos=0.05, threshold_neg=0.05, max_events_per_step=16)
    if noise:
        phys = dataclasses.replace(phys, photon_shot_noise=1, reset_thermal_noise=1, threshold_std=0.03,
                                   random_seed=777)
    return scenes.fast_pass(dataclasses.replace(base, physic=phys), speed=3000.0, yaw_rate_deg=600.0, frames=2,
                            fps=100.0)


@pytest.mark.parametrize("noise", [False, True])
def test_terrain_windows_match_oracle(oracle, gpu, noise):
    """Three consecutive 200-horizon windows over a moving terrain view
    (DAVIS optics at 2x2 sub-rays) and the gate's check of the first."""
    w = _terrain(noise)
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    g = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    o = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    s = gpu.Sensor(g, w.physic)
    s.initialize(w.initial_state(), w.start_time)
    os_ = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    assert os_.initialize(o, w.initial_state(), w.start_time) == 0
    cap = s.adaptive_capacity()
    # The gate rejects the first window (fast motion, thresholds 0.05: its
    # bound is 0.001 in log voltage) on both sides with the same errors; the
    # windows are then integrated unchecked.
    ivs = w.intervals(0, 200)
    ok, _ = s.process_adaptive(ivs, check_convergence=True)
    code, ook, _, _, _, oprobe = os_.process_adaptive(o, ivs, check=True, capacity=cap)
    assert code == 0 and not ok and not ook
    probe = s.last_convergence()
    assert probe["pixels"] == oprobe["pixels"]
    for key in ("maximum_log_voltage_error", "p99_log_voltage_error", "mean_log_voltage_error"):
        # one probe pixel's error: moves with boundary resolutions as above
        assert probe[key] == pytest.approx(oprobe[key], rel=1e-3, abs=1e-12), key
    total = 0
    for k in range(3):
        ivs = w.intervals(200 * k, 200)
        ok, ready = s.process_adaptive(ivs)
        code, ook, oraw, oready, ost, _ = os_.process_adaptive(o, ivs, capacity=cap)
        assert code == 0 and ok and ook
        npix = w.camera.width * w.camera.height
        _assert_stats_close(s.refinement_stats(), ost, npix)
        _assert_adaptive_match(s.raw_events(), oraw, w.camera.width, npix)
        _assert_adaptive_match(ready, oready, w.camera.width, npix)
        total += len(oraw)
    assert total > 0
    # receptor voltages: equal to the float rounding of the radiance except in
    # pixels whose refinement resolved a boundary segment differently, which
    # stay within the refinement's voltage budget (1e-3 x threshold per leaf)
    d = np.abs(s.states()["receptor_log"] - os_.states()["receptor_log"])
    assert np.median(d) < 1e-6 and np.max(d) < 1e-3


def test_gate_rejects_flash_without_state_change(oracle, gpu):
    """tests/test_sensor.cu:146-159 on the device."""
    g, o, cam, phys = _plane(oracle, gpu)
    phys = dataclasses.replace(phys, photon_shot_noise=1)
    pose = scenes.nadir_state()
    flash = pose.copy()
    flash["intensity"] = 20000
    ivs = np.concatenate([scenes.interval(pose, pose, 0, .0003), scenes.interval(flash, flash, .0003, .0004),
                          scenes.interval(pose, pose, .0004, .002)])
    checked, fresh = gpu.Sensor(g, phys), gpu.Sensor(g, phys)
    checked.initialize(pose, 0.0)
    fresh.initialize(pose, 0.0)
    ok, got = checked.process_adaptive(ivs, tolerance=.5, check_convergence=True)
    assert not ok and len(got) == 0
    probe = checked.last_convergence()
    thr = min(phys.threshold_pos, phys.threshold_neg)
    assert probe["maximum_log_voltage_error"] > .02 * thr or probe["p99_log_voltage_error"] > .005 * thr
    retried = checked.process(ivs)
    exact = fresh.process(ivs)
    assert len(exact) == len(retried) > 0
    assert exact.tobytes() == retried.tobytes()
    assert checked.stats()["rejected_windows"] == 1


def test_adaptive_argument_errors(oracle, gpu):
    g, _, cam, phys = _plane(oracle, gpu)
    pose = scenes.nadir_state()
    s = gpu.Sensor(g, phys)
    s.initialize(pose, 0.0)
    with pytest.raises(capi.InvalidArgument, match="event capacity bound"):
        s.process_adaptive(scenes.interval(pose, pose, 0, 63 * phys.refractory_time))
    with pytest.raises(capi.InvalidArgument, match="event capacity bound"):
        s.process_adaptive(scenes.interval(pose, pose, 0, .001), tolerance=0.0)
    with pytest.raises(capi.InvalidArgument, match="contiguous"):
        s.process_adaptive(scenes.interval(pose, pose, .001, .002))
    ok, _ = s.process_adaptive(scenes.interval(pose, pose, 0, .001))
    assert ok


def test_batch_adaptive_policy_equals_driver(oracle, gpu):
    """Segments under the engine's adaptive policy (including a gate
    rejection, halving and a native one-horizon window) give each segment the
    events of a sensor run by the host driver loop."""
    g0, _, cam, phys = _plane(oracle, gpu)
    g = g0.view(dataclasses.replace(cam, samples=2))  # one engine shares one sub-ray grid
    pose = scenes.nadir_state()
    flash = pose.copy()
    flash["intensity"] = 20000
    noisy = dataclasses.replace(phys, photon_shot_noise=1)
    flash_ivs = np.concatenate([scenes.interval(pose, pose, 0, .0003), scenes.interval(flash, flash, .0003, .0004),
                                scenes.interval(pose, pose, .0004, .0008), scenes.interval(pose, pose, .0008, .001)])
    w = _terrain(False)
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    gt = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    terrain_ivs = w.intervals(0, 130)
    specs = [(g, noisy, pose, flash_ivs, 2, .5, 1e-4), (gt, w.physic, w.initial_state(), terrain_ivs, 50, 1e-2, 5e-4)]
    batch = gpu.Batch(horizons_per_step=8)
    ids = []
    for scene, ph, p0, ivs, hz, tol, period in specs:
        sid = batch.add(scene, ph, p0, float(ivs["t0"][0]))
        batch.set_adaptive(sid, tolerance=tol, horizons=hz, check_period=period)
        batch.feed(sid, ivs)
        ids.append(sid)
    got = {sid: [] for sid in ids}
    for _ in range(200):
        pending = batch.step()
        for sid in ids:
            got[sid].append(batch.take(sid))
        if not pending:
            break
    for sid in ids:
        got[sid].append(batch.drain(sid))
    assert batch.stats()["rejected_windows"] >= 1
    total = 0
    for sid, (scene, ph, p0, ivs, hz, tol, period) in zip(ids, specs):
        sensor = gpu.Sensor(scene, ph)
        sensor.initialize(p0, float(ivs["t0"][0]))
        ref = np.concatenate(driver.drive(sensor, ivs, adaptive=True, tolerance=tol, batch=hz, check_period=period))
        mine = np.concatenate(got[sid])
        total += len(ref)
        assert mine.tobytes() == ref.tobytes(), f"segment {sid} differs from the driver loop"
    assert total > 0


def test_readout_fifo_matches_oracle_with_drops(oracle, gpu):
    """The readout FIFO (src/event_readout.cpp:6-34) behind the device's
    comparator: a flash overfills a 40-event queue at 2e5 events/s, later
    windows drain it.  Every call's delivered events, and the dropped count,
    peak and final drain, equal the oracle's FIFO fed the sensor's own
    comparator events of each call (calls both with drops and without)."""
    g, _, cam, phys = _plane(oracle, gpu)
    bw, lat, cap = 2e5, 1e-6, 40
    phys = dataclasses.replace(phys, photon_shot_noise=1, readout_bandwidth=bw, readout_latency=lat,
                               readout_queue_capacity=cap)
    pose = scenes.nadir_state()
    flash = pose.copy()
    flash["intensity"] = 20000
    s = gpu.Sensor(g, phys)
    s.initialize(pose, 0.0)
    # the sensor configuration holds bandwidth and latency as float32
    bus = oracle.Readout(float(np.float32(bw)), float(np.float32(lat)), cap)

    def same(got, want):
        assert len(got) == len(want)
        assert np.array_equal(got["x"], want["x"]) and np.array_equal(got["y"], want["y"])
        assert np.array_equal(got["polarity"], want["polarity"])
        assert np.array_equal(got["timestamp"], want["timestamp"].astype(got["timestamp"].dtype))

    calls = [scenes.interval(pose, flash, 0, .0003), scenes.interval(flash, flash, .0003, .0006),
             scenes.interval(flash, pose, .0006, .0009)] + [scenes.interval(pose, pose, .0009 + .001 * k,
                                                                             .0019 + .001 * k) for k in range(4)]
    submitted = 0
    for ivs in calls:
        got = s.process(ivs)
        raw = s.raw_events()
        ev = np.zeros(len(raw), oracle.EVENT)
        for f in ("timestamp", "x", "y", "polarity", "depth"):
            ev[f] = raw[f]
        submitted += len(ev)
        same(got, bus.submit(ev, float(ivs["t1"][-1])))
    dropped, peak = s.readout_stats()
    assert dropped == bus.dropped > 0 and peak == bus.peak == cap and submitted > cap
    same(s.drain(), bus.drain())
