"""Device-side trajectory intervals (dvs_trajectory_intervals) against the
oracle's CameraTrajectory::evaluate (src/trajectory.cpp:719-787) at the same
float times, as bin/dvs builds its OpticalIntervals
(apps/main_dvs.cpp:103-107).

Tolerance: 1e-5 relative to each field's largest magnitude (float32 curves;
the device's sinf/cosf/atan2f may differ from the host's by a few ulp).
"""
from __future__ import annotations

import numpy as np
import pytest

from dvsrender4deepspace_amd import capi
from test_trajectory import frames, random_frames

pytestmark = pytest.mark.gpu

TOL = 1e-5


def driver_times(start, end, resolution):
    """t_k = min(end, start + k * double(resolution)) up to end
    (apps/main_dvs.cpp:79-108)."""
    res = float(np.float32(resolution))
    k = np.arange(int(np.ceil((end - start) / res)) + 1, dtype=np.float64)
    return np.minimum(end, start + k * res)


def check_against(states, expected):
    for field in states.dtype.names:
        x, y = states[field].astype(np.float64), expected[field].astype(np.float64)
        scale = max(1.0, np.abs(y).max())
        assert np.abs(x - y).max() <= TOL * scale, field


def test_intervals_match_oracle(gpu, oracle):
    args = (100.0, (0.1, -0.2, 0.05), (0.01, 0.02, -0.03))
    o = oracle.Trajectory(random_frames(oracle.FRAME), *args)
    p = gpu.Trajectory(random_frames(gpu.FRAME_DTYPE), *args)
    lo, hi, _ = o.range()
    times = driver_times(float(lo), float(hi), 1e-4)
    times = np.concatenate([[lo - 0.01], times, [hi + 0.02]])  # clamped outside the keys
    print(f"trajectory intervals: {len(times)} times", flush=True)
    iv = p.intervals(times)
    assert len(iv) == len(times) - 1
    assert np.array_equal(iv["t0"], times[:-1]) and np.array_equal(iv["t1"], times[1:])
    expected = gpu.optical_states(o.evaluate(times.astype(np.float32)))
    check_against(iv["start"], expected[:-1])
    check_against(iv["end"], expected[1:])
    # consecutive intervals share the state of their common time
    assert np.array_equal(iv["end"][:-1], iv["start"][1:])
    # the device agrees with the host evaluation of the same library
    check_against(iv["start"], gpu.optical_states(p.evaluate(times[:-1].astype(np.float32))))


def test_static_intervals_are_exact(gpu):
    """tests/test_sensor.cu:25-36 on the device: a static trajectory yields
    identical position, light and intensity at every interval end."""
    f = frames(2, gpu.FRAME_DTYPE, indices=[0, 2], origins=(0, 0, 300), lights=(0.8660254, 0, 0.5))
    p = gpu.Trajectory(f, 100.0)
    times = driver_times(0.0, float(np.float32(0.02)), 2e-5)
    iv = p.intervals(times)
    first = iv["start"][0]
    for field in ("position", "light", "intensity", "rotation"):
        assert np.array_equal(iv["start"][field], np.broadcast_to(first[field], iv["start"][field].shape)), field
        assert np.array_equal(iv["end"][field], np.broadcast_to(first[field], iv["end"][field].shape)), field
    host = gpu.optical_states(p.evaluate([0.0]))[0]
    assert np.array_equal(first["position"], host["position"]) and first["intensity"] == host["intensity"]


def test_many_intervals_and_reuse(gpu, oracle):
    """A 0.2 s window at 1 us horizons (200,001 times) twice on one handle
    (device buffers reused), spot-checked against the oracle."""
    o = oracle.Trajectory(random_frames(oracle.FRAME, n=40, seed=4), 100.0)
    p = gpu.Trajectory(random_frames(gpu.FRAME_DTYPE, n=40, seed=4), 100.0)
    lo, hi, _ = o.range()
    times = driver_times(float(lo), float(lo) + 0.2, 1e-6)
    print(f"trajectory intervals: {len(times)} times", flush=True)
    a = p.intervals(times)
    b = p.intervals(times[:1000])
    assert np.array_equal(a[:999], b)
    pick = np.random.default_rng(1).choice(len(times) - 1, 4000, replace=False)
    check_against(a["start"][pick], gpu.optical_states(o.evaluate(times[pick].astype(np.float32))))


def test_sensor_windows_from_trajectory_times_equal_host_intervals(gpu):
    """dvs_sensor_process_adaptive_times (states evaluated on the device
    inside the window) gives the events of dvs_sensor_process_adaptive fed
    dvs_trajectory_intervals of the same times, byte for byte, over bin/dvs
    windows of the small request's flight with the receptor check on the
    first window; a time grid that does not start where the sensor stands
    is refused."""
    import scenes
    from dvsrender4deepspace_amd import driver, workloads
    w = scenes.fast_pass(workloads.small_variant(workloads.PLAIN_SMALL, size=96, craters=6, samples=2))
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    scene = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    traj = workloads.trajectory(w)
    grids, t, step = [], w.start_time, 0
    for k in range(6):
        times = workloads.horizon_grid(w, t, step, 20)
        grids.append(times)
        t, step = float(times[-1]), step + len(times) - 1
    out = []
    for form in ("intervals", "times"):
        sensor = gpu.Sensor(scene, w.physic)
        sensor.initialize(w.initial_state(), w.start_time)
        got, oks = [], []
        for k, times in enumerate(grids):
            for check in ((True, False) if k == 0 else (False,)):
                if form == "times":
                    ok, ev = sensor.process_adaptive_times(traj, times, driver.DEFAULT_TOLERANCE, check)
                else:
                    ok, ev = sensor.process_adaptive(traj.intervals(times), driver.DEFAULT_TOLERANCE, check)
                oks.append(ok)
                if ok:
                    break
            assert ok
            got.append(ev)
        got.append(sensor.drain())
        out.append((np.concatenate(got), oks))
        if form == "times":
            with pytest.raises(capi.InvalidArgument, match="positive and contiguous"):
                sensor.process_adaptive_times(traj, grids[0], driver.DEFAULT_TOLERANCE, False)
    assert out[0][1] == out[1][1]
    assert len(out[0][0]) > 20
    assert out[0][0].tobytes() == out[1][0].tobytes()


def test_windows_set_up_ahead_equal_plain_windows(gpu):
    """dvs_sensor_expect_adaptive_times: a window whose setup was queued at
    the end of the window before it gives the events of the same window set
    up in its own call, byte for byte, over bin/dvs windows of the small
    request's flight with receptor checks at windows 0 and 5 and a native
    window (which drops a queued setup) at window 4.  An expectation the
    next call does not match (another tolerance) is not used."""
    import scenes
    from dvsrender4deepspace_amd import driver, workloads
    w = scenes.fast_pass(workloads.small_variant(workloads.PLAIN_SMALL, size=96, craters=6, samples=2))
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    scene = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    traj = workloads.trajectory(w)
    grids, t, step = [], w.start_time, 0
    for k in range(10):
        times = workloads.horizon_grid(w, t, step, 20)
        grids.append(times)
        t, step = float(times[-1]), step + len(times) - 1
    tol = driver.DEFAULT_TOLERANCE
    out = {}
    for form in ("plain", "ahead", "other tolerance"):
        print(f"form {form}", flush=True)
        sensor = gpu.Sensor(scene, w.physic)
        sensor.initialize(w.initial_state(), w.start_time)
        got, oks = [], []
        for k, times in enumerate(grids):
            if k == 4:  # a native window between adaptive ones
                got.append(sensor.process(traj.intervals(times)))
                continue
            for check in ((True, False) if k in (0, 5) else (False,)):
                if form != "plain" and k + 1 < len(grids):  # window 4's setup is dropped
                    sensor.expect_adaptive_times(traj, grids[k + 1], tol if form == "ahead" else 2 * tol)
                ok, ev = sensor.process_adaptive_times(traj, times, tol, check)
                oks.append(ok)
                if ok:
                    break
            assert ok
            got.append(ev)
        got.append(sensor.drain())
        out[form] = (np.concatenate(got), oks, sensor.stats()["windows_set_up_ahead"])
        sensor.close()
    print({k: (len(v[0]), v[2]) for k, v in out.items()}, flush=True)
    assert len(out["plain"][0]) > 20
    assert out["plain"][2] == 0 and out["other tolerance"][2] == 0
    assert out["ahead"][2] >= 5
    for form in ("ahead", "other tolerance"):
        assert out[form][1] == out["plain"][1]
        assert out[form][0].tobytes() == out["plain"][0].tobytes()
