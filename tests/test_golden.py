"""The oracle against the committed golden vectors (tests/golden/, written
by tests/golden/make_golden.py), and the circuit table against the
constant-current closed form of the first crossing."""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as mg  # noqa: E402


def test_hapke_table_reproduces(oracle):
    d = np.load(GOLDEN / "hapke_table.npz")
    assert np.array_equal(d["grid"], mg.hapke_grid())
    for name, hk in mg.HAPKE_SETS.items():
        got = np.array([oracle.hapke_reflectance(*map(float, r[:3]), float(r[3]), hk) for r in d["grid"]], np.float32)
        assert np.array_equal(got, d[name]), name
        assert np.all(np.isfinite(got)) and np.all(got >= 0)


def test_circuit_table_reproduces(oracle):
    d = np.load(GOLDEN / "circuit_table.npz")
    assert np.array_equal(d["cases"], mg.circuit_cases())
    finals, counts, times, pols = mg.run_circuit(d["cases"])
    assert np.array_equal(counts, d["counts"]) and np.array_equal(pols, d["polarities"])
    assert np.array_equal(times, d["times"]) and np.array_equal(finals, d["finals"])
    assert counts.sum() > 100 and (counts == 0).any()


def test_circuit_table_first_crossings_closed_form():
    """Without leakage the receptor is affine in y = exp(-u) at constant rate
    r: y(t) = exp(-r t / C) y0 + (1 - exp(-r t / C)) / r, so the first
    crossing of u = ref + theta (ON) or ref - theta (OFF) from inside the
    band is t = -C / r ln((y* - 1/r) / (y0 - 1/r)), with y* = exp(-(ref +- theta))
    (tests/test_sensor.cu:40-42 for one case)."""
    d = np.load(GOLDEN / "circuit_table.npz")
    offsets = np.concatenate([[0], np.cumsum(np.maximum(d["counts"], 0))])
    checked = 0
    for k, (u0, ref, rate, dt, thr, refractory, leak, charge) in enumerate(d["cases"]):
        if leak != 0 or rate <= 0 or abs(u0 - ref) >= thr or d["counts"][k] <= 0:
            continue
        pol = d["polarities"][offsets[k]]
        ystar = math.exp(-(ref + thr)) if pol == 1 else math.exp(-(ref - thr))
        t = -charge / rate * math.log((ystar - 1 / rate) / (math.exp(-u0) - 1 / rate))
        assert pol == (1 if math.log(rate) > u0 else 0)
        assert abs(d["times"][offsets[k]] - t) <= 1e-12 + 1e-9 * t, k
        checked += 1
    assert checked > 30


def test_plain_frame_reproduces(oracle):
    d = np.load(GOLDEN / "plain_frame.npz")
    w, o = mg.plain_scene()
    for k, t in enumerate(d["times"]):
        code, img, dep, rad = o.render(w.states(np.array([t])))
        assert code == 0
        assert np.array_equal(rad, d["radiance"][k]) and np.array_equal(img, d["image"][k])
        assert np.array_equal(dep, d["depth"][k]) and np.array_equal(oracle.render_to_bytes(img), d["bytes"][k])
    for dep in d["depth"]:
        assert np.count_nonzero(dep > 0) > 0.3 * dep.size  # the 127 m terrain fills part of the view


def test_plain_events_prefix_reproduces(oracle):
    """The oracle's first eight 64-horizon calls give the golden stream's
    events up to their end time (the native solver is batch invariant)."""
    d = np.load(GOLDEN / "plain_events.npz")
    w, o = mg.plain_scene()
    batch = int(d["batch"])
    s = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    assert s.initialize(o, w.initial_state(), w.start_time) == 0
    raws = []
    for k in range(0, 8 * batch, batch):
        code, raw, _ = s.process(o, w.intervals(k, batch))
        assert code == 0
        raws.append(raw)
    raw = np.concatenate(raws)
    end = w.intervals(7 * batch, batch)["t1"][-1]
    gold = mg.load_events("raw", d)
    prefix = gold[gold["timestamp"] <= end]
    assert len(raw) == len(prefix) > 40
    assert raw.tobytes() == prefix.tobytes()
