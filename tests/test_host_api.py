"""The reference's host-callable pieces of the drop-in API on the CPU:
hapke_reflectance / hapke_directional_reflectance / invert_hapke_if
(include/dvs/hapke.hpp:131-172), receptor_log_at and advance_pixel
(include/dvs/pixel_model.hpp:18-102) compiled from the library's own
__host__ __device__ source, and the standalone EventReadout
(src/event_readout.cpp:6-34), each against the oracle."""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

from dvsrender4deepspace_amd import capi, workloads

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as mg  # noqa: E402

HAPKES = [capi.Hapke(), workloads.WORKLOAD_HAPKE, capi.Hapke(phi=0, thetap=0, hC=0, B_C0=0),
          capi.Hapke(b=0.24, c=0.305, hS=0.08, hC=1.0, thetap=0.453786, w=0.324, B_S0=2.42, B_C0=0.5, phi=0.3)]


@pytest.mark.parametrize("k", range(len(HAPKES)))
def test_host_hapke_equals_oracle(oracle, k):
    hp = HAPKES[k]
    rng = np.random.default_rng(k)
    for _ in range(3000):
        ci, ce, cg, w = (float(np.float32(x)) for x in (*rng.uniform(-0.1, 1, 3), rng.uniform(0, 1)))
        a, b = capi.hapke_reflectance(ci, ce, cg, w, hp), oracle.hapke_reflectance(ci, ce, cg, w, hp)
        assert (math.isnan(a) and math.isnan(b)) or a == b
    for _ in range(200):
        n, l, v = (rng.normal(size=3).astype(np.float32) for _ in range(3))
        a, b = capi.hapke_directional(n, l, v, 0.3, hp), oracle.hapke_directional(n, l, v, 0.3, hp)
        assert (math.isnan(a) and math.isnan(b)) or a == b


def test_host_invert_hapke_equals_oracle(oracle):
    hp = HAPKES[3]
    for inc in (0.0, 15.0, 45.0, 80.0):
        for emi in (0.0, 30.0, 60.0):
            for phase in (abs(inc - emi), 0.5 * (abs(inc - emi) + inc + emi), inc + emi, inc + emi + 1):
                for target in (0.0, 0.02, 0.08, 0.3, 5.0):
                    a = capi.invert_hapke_if(inc, emi, phase, target, hp)
                    b = oracle.invert_hapke_if(inc, emi, phase, target, hp)
                    assert (math.isnan(a) and math.isnan(b)) or a == b
    # round trip (tests/test_render.cu:107-112)
    ci = float(np.float32(math.cos(30 * np.float32(math.pi) / 180)))
    target = float(np.float32(math.pi)) * capi.hapke_reflectance(ci, 1, ci, 0.324, hp)
    assert abs(capi.invert_hapke_if(30, 0, 30, target, hp) - 0.324) < 2e-6


def test_host_receptor_equals_oracle(oracle):
    rng = np.random.default_rng(9)
    for _ in range(2000):
        u, rate, dt = rng.uniform(0, 15), 10 ** rng.uniform(0, 7) * (rng.uniform() > .1), 10 ** rng.uniform(-8, -1)
        a, b = capi.receptor_log_at(u, rate, dt, 202.0), oracle.receptor_log_at(u, rate, dt, 202.0)
        assert abs(a - b) <= 1e-15 * abs(b)


def _host_advance(state, params, rate, t0, t1, capacity):
    s = np.zeros(1, capi.PIXEL_STATE_DTYPE)
    for f in ("receptor_log", "reference_log", "refractory_until", "threshold_pos", "threshold_neg"):
        s[f] = state[f]
    s["reset_pending"] = state["reset_pending"]
    p = np.zeros(1, capi.PIXEL_PARAMS_DTYPE)
    for f in capi.PIXEL_PARAMS_DTYPE.names:
        p[f] = getattr(params, f)
    n, ev = capi.advance_pixel(s, p, rate, t0, t1, capacity)
    for f in ("receptor_log", "reference_log", "refractory_until"):
        state[f] = s[f]
    state["reset_pending"] = s["reset_pending"]
    out = np.zeros(max(n, 0), [("time", "<f8"), ("polarity", "<i4"), ("pad", "<i4")])
    if n > 0:
        out["time"], out["polarity"] = ev["time"], ev["polarity"]
    return n, out


def test_host_advance_pixel_reproduces_circuit_table():
    d = np.load(Path(__file__).resolve().parent / "golden" / "circuit_table.npz")
    finals, counts, times, pols = mg.run_circuit(d["cases"], advance=_host_advance)
    assert np.array_equal(counts, d["counts"]) and np.array_equal(pols, d["polarities"])
    assert np.array_equal(times, d["times"]) and np.array_equal(finals, d["finals"])


def _random_events(rng, n, t0, t1, width=32, height=24):
    ev = np.zeros(n, capi.EVENT_DTYPE)
    ev["timestamp"] = np.round(rng.uniform(t0, t1, n), 7)  # rounded: ties in t order by (y, x)
    ev["x"], ev["y"] = rng.integers(0, width, n), rng.integers(0, height, n)
    ev["polarity"] = rng.integers(0, 2, n)
    ev["depth"] = rng.uniform(1, 100, n).astype(np.float32)
    return ev


@pytest.mark.parametrize("bandwidth,capacity", [(2e5, 40), (1e6, 100000), (0.0, 5)])
def test_readout_equals_oracle(oracle, bandwidth, capacity):
    rng = np.random.default_rng(int(bandwidth) + capacity)
    mine, ref = capi.Readout(bandwidth, 1e-6, capacity), oracle.Readout(bandwidth, 1e-6, capacity)
    for k in range(8):
        ev = _random_events(rng, int(rng.integers(0, 400)), k * 1e-3, (k + 1) * 1e-3)
        oev = np.zeros(len(ev), oracle.EVENT)
        for f in ("timestamp", "x", "y", "polarity", "depth"):
            oev[f] = ev[f]
        a, b = mine.submit(ev, (k + 1) * 1e-3), ref.submit(oev, (k + 1) * 1e-3)
        assert len(a) == len(b) and a.tobytes() == b.astype(capi.EVENT_DTYPE).tobytes()
    assert mine.stats() == (ref.dropped, ref.peak)
    assert mine.drain().tobytes() == ref.drain().astype(capi.EVENT_DTYPE).tobytes()
    if bandwidth and capacity < 100:
        assert ref.dropped > 0


def test_readout_known_answer():
    """tests/test_sensor.cu:176-177: EventReadout(100, 0.001, 2) drops one of
    three simultaneous events and delivers the first at 0.011 s."""
    r = capi.Readout(100, 0.001, 2)
    ev = np.zeros(3, capi.EVENT_DTYPE)
    out = r.submit(ev, 0.0)
    assert len(out) == 0 and r.stats()[0] == 1
    got = r.drain()
    assert len(got) == 2 and abs(got[0]["timestamp"] - 0.011) < 1e-12
