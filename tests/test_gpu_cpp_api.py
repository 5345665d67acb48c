"""GPU: the C++ API's DeviceScene, CamRayCaster and DvsRayCaster give the
events of the Python binding on the plane fixture (the same C ABI)."""
from __future__ import annotations

import dataclasses
import subprocess

import numpy as np
import pytest

import scenes
from dvsrender4deepspace_amd import build, capi, eventio

pytestmark = pytest.mark.gpu


def test_cpp_api_matches_python_binding(gpu, tmp_path):
    out = subprocess.run([str(build.API_SELFTEST), "--gpu", str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    v, n, nodes, root, idx, cam, phys = scenes.sensor_fixture()
    g = gpu.Scene(nodes, root, v, n, idx, cam, capi.Hapke())
    pose = scenes.nadir_state()
    start, end = pose.copy(), pose.copy()
    start["intensity_rate"], end["intensity_rate"] = 8e6, -8e6
    native, adaptive = gpu.Sensor(g, phys), gpu.Sensor(g, phys)
    native.initialize(pose, 0.0)
    adaptive.initialize(pose, 0.0)
    native.process(scenes.interval(pose, pose, 0, .002))
    ev = native.process(scenes.interval(start, end, .002, .004))
    rec = eventio.read_bin(tmp_path / "native.bin")
    assert len(rec) == len(ev) > 0
    assert np.array_equal(rec["t"], ev["timestamp"].astype(np.float32))
    assert np.array_equal(rec["x"], ev["x"]) and np.array_equal(rec["y"], ev["y"])
    assert np.array_equal(rec["p"], ev["polarity"]) and np.array_equal(rec["depth"], ev["depth"])
    ok, _ = adaptive.process_adaptive(scenes.interval(pose, pose, 0, .002), 1e-4, True)
    assert ok
    ok, aev = adaptive.process_adaptive(scenes.interval(start, end, .002, .004), 1e-4)
    h5 = eventio.read_h5(tmp_path / "adaptive.h5")
    order = np.argsort(aev["timestamp"], kind="stable")
    assert np.array_equal(h5["t"], np.round(aev["timestamp"][order] * 1e6).astype(np.int64))
    assert np.array_equal(h5["x"], aev["x"][order]) and np.array_equal(h5["p"], aev["polarity"][order])
