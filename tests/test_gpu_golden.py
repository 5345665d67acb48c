"""The device against the committed golden vectors (tests/golden/): the
Hapke table, the circuit table through the device's advance_pixel, the
small plain request's frames, and its noise-off event stream over the first
157 bin/dvs calls (0.1 s).  Tolerances as in tests/test_gpu_parity.py."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import scenes
from dvsrender4deepspace_amd import capi, driver
from test_gpu_parity import _assert_events_match, assert_radiance_close

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as mg  # noqa: E402

pytestmark = pytest.mark.gpu


def test_hapke_table(gpu):
    d = np.load(GOLDEN / "hapke_table.npz")
    g = d["grid"]
    for name, hk in mg.HAPKE_SETS.items():
        dev = gpu.hapke_eval(g[:, 0], g[:, 1], g[:, 2], g[:, 3], hk)
        ref = d[name]
        assert np.max(np.abs(dev - ref) / np.maximum(np.abs(ref), 1e-6)) < 1e-4, name


def test_circuit_table(gpu):
    d = np.load(GOLDEN / "circuit_table.npz")
    cases = d["cases"]
    n = len(cases)
    states = np.zeros(n, capi.PIXEL_STATE_DTYPE)
    states["receptor_log"], states["reference_log"] = cases[:, 0], cases[:, 1]
    states["refractory_until"] = -1
    states["threshold_pos"] = states["threshold_neg"] = cases[:, 4]
    params = np.zeros(n, capi.PIXEL_PARAMS_DTYPE)
    params["threshold_pos"] = params["threshold_neg"] = cases[:, 4]
    params["refractory_time"], params["reference_leak"], params["charge_scale"] = cases[:, 5], cases[:, 6], cases[:, 7]
    counts, events = gpu.advance_pixel_eval(states, params, cases[:, 2], 0.0, cases[:, 3], capacity=256)
    assert np.array_equal(counts, d["counts"])
    times = np.concatenate([events[k, :c]["time"] for k, c in enumerate(counts)])
    pols = np.concatenate([events[k, :c]["polarity"] for k, c in enumerate(counts)])
    assert np.array_equal(pols, d["polarities"])
    assert np.max(np.abs(times - d["times"]) / np.maximum(d["times"], 1e-6)) < 1e-10
    for k, f in enumerate(("receptor_log", "reference_log", "refractory_until")):
        assert np.max(np.abs(states[f] - d["finals"][:, k])) < 1e-9, f


def test_plain_frames(oracle, gpu):
    d = np.load(GOLDEN / "plain_frame.npz")
    w = mg.workloads.PLAIN_SMALL
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    g = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    for k, t in enumerate(d["times"]):
        img, dep, rad = g.render(w.states(np.array([t])), with_radiance=True)
        assert_radiance_close(rad[0], d["radiance"][k])
        assert np.array_equal(dep[0], d["depth"][k])
        b = oracle.render_to_bytes(img[0])
        assert np.max(np.abs(b.astype(int) - d["bytes"][k].astype(int))) <= 1


def test_plain_event_stream(gpu):
    """157 native calls of 64 horizons from the request's start: comparator
    events per pixel identical with times within 1 us, the readout's
    delivered stream likewise, the drained backlog included."""
    d = np.load(GOLDEN / "plain_events.npz")
    w = mg.workloads.PLAIN_SMALL
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    g = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    s = gpu.Sensor(g, w.physic)
    s.initialize(w.initial_state(), w.start_time)
    batch, horizons = int(d["batch"]), int(d["horizons"])
    assert batch == driver.batch_size(w.physic, w.time_resolution, adaptive=False)
    raws, readys = [], []
    for k in range(0, horizons, batch):
        readys.append(s.process(w.intervals(k, min(batch, horizons - k))))
        raws.append(s.raw_events())
    readys.append(s.drain())
    raw, ready = np.concatenate(raws), np.concatenate(readys)
    print(f"plain stream: {len(raw)} comparator events, {len(ready)} delivered", flush=True)
    _assert_events_match(raw, mg.load_events("raw", d))
    _assert_events_match(ready, mg.load_events("ready", d))
