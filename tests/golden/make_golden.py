"""Writes the golden vectors under tests/golden/ with the CPU oracle.

    python tests/golden/make_golden.py [hapke] [circuit] [frame] [events]

The oracle is itself pinned (tests/test_oracle*.py) by the reference's known
answers, closed forms, brute force and the published Hapke form; these files
freeze its outputs so that the device (tests/test_gpu_golden.py) and later
oracle changes (tests/test_golden.py) are checked against fixed data:

* hapke_table.npz: hapke_reflectance (include/dvs/hapke.hpp:131-151) over a
  grid of incidence, emission, phase and albedo, for the workloads' Hapke
  parameters (theta-bar 20 deg) and the reference's defaults;
* circuit_table.npz: advance_pixel (include/dvs/pixel_model.hpp:35-102) on
  seeded constant-current pieces, with and without refractory hold and
  reference leakage: crossing times, polarities and the final state;
* plain_frame.npz: the small plain request's 64x48 frame (128^2 terrain,
  one sub-ray) at t = 0 and t = 0.25 s: radiance, image, depth, 8-bit;
* plain_events.npz: the small plain request's noise-off event stream over
  its first 157 bin/dvs calls (10,048 horizons of 10 us, 0.1 s, 1 us
  circuit cells) through the native solver: comparator events and the
  readout's delivered events (the backlog drained at the end).
"""
from __future__ import annotations

import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_ctypes as oracle  # noqa: E402
import scenes  # noqa: E402
from dvsrender4deepspace_amd import capi, driver, workloads  # noqa: E402

HAPKE_SETS = {"workload": workloads.WORKLOAD_HAPKE, "default": capi.Hapke()}


def hapke_grid():
    """(cos i, cos e, cos g, w) float32 over i, e in 0..85 deg, g at six
    points of [|i - e|, i + e] and four albedos."""
    rows = []
    for i in range(0, 90, 5):
        for e in range(0, 90, 5):
            lo, hi = abs(i - e), min(i + e, 180)
            for g in np.linspace(lo, hi, 6):
                for w in (0.05, 0.3, 0.7, 1.0):
                    rows.append((math.cos(math.radians(i)), math.cos(math.radians(e)), math.cos(math.radians(g)), w))
    return np.array(rows, np.float32)


def write_hapke():
    grid = hapke_grid()
    out = {"grid": grid}
    for name, hk in HAPKE_SETS.items():
        out[name] = np.array([oracle.hapke_reflectance(*map(float, r[:3]), float(r[3]), hk) for r in grid], np.float32)
    np.savez_compressed(HERE / "hapke_table.npz", **out)


CIRCUIT_CASES = 400


def circuit_cases():
    """Seeded pieces: initial receptor/reference log currents, a constant
    rate, a duration, thresholds, refractory hold and reference leakage."""
    rng = np.random.default_rng(20260104)
    cases = []
    for k in range(CIRCUIT_CASES):
        u0 = math.log(10 ** rng.uniform(2, 6))
        ref = u0 + rng.uniform(-0.15, 0.15)
        rate = 0.0 if k % 25 == 0 else math.exp(u0) * math.exp(rng.uniform(-1.5, 1.5))
        dt = 10 ** rng.uniform(-5, -1)
        thr = rng.uniform(0.05, 0.4)
        refractory = 0.0 if k % 3 else 10 ** rng.uniform(-6, -3)
        leak = 0.0 if k % 4 else rng.uniform(0.01, 2.0)
        charge = rng.uniform(50, 400)
        cases.append((u0, ref, rate, dt, thr, refractory, leak, charge))
    return np.array(cases)


def run_circuit(cases, advance=None):
    """Runs each case through `advance(state, params, rate, t0, t1, capacity)`
    (the oracle's by default); returns the final states and events."""
    advance = advance or oracle.advance_pixel
    finals, counts, times, pols = [], [], [], []
    for u0, ref, rate, dt, thr, refractory, leak, charge in cases:
        p = oracle.PixelParams()
        p.threshold_pos = p.threshold_neg = thr
        p.charge_scale, p.refractory_time, p.reference_leak, p.reset_noise_std = charge, refractory, leak, 0.0
        s = np.zeros(1, oracle.PIXEL_STATE)
        s["receptor_log"], s["reference_log"] = u0, ref
        s["refractory_until"] = -1
        s["threshold_pos"] = s["threshold_neg"] = thr
        n, ev = advance(s, p, rate, 0.0, dt, 256)
        counts.append(n)
        if n > 0:
            times.extend(ev["time"][:n])
            pols.extend(ev["polarity"][:n])
        finals.append([s["receptor_log"][0], s["reference_log"][0], s["refractory_until"][0]])
    return np.array(finals), np.array(counts, np.int32), np.array(times), np.array(pols, np.int32)


def write_circuit():
    cases = circuit_cases()
    finals, counts, times, pols = run_circuit(cases)
    np.savez_compressed(HERE / "circuit_table.npz", cases=cases, finals=finals, counts=counts, times=times,
                        polarities=pols)


FRAME_TIMES = (0.0, 0.25)


def plain_scene():
    w = workloads.PLAIN_SMALL
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    return w, oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)


def write_frame():
    w, o = plain_scene()
    rads, imgs, deps, bytes_ = [], [], [], []
    for t in FRAME_TIMES:
        code, img, dep, rad = o.render(w.states(np.array([t])))
        assert code == 0
        rads.append(rad)
        imgs.append(img)
        deps.append(dep)
        bytes_.append(oracle.render_to_bytes(img))
    np.savez_compressed(HERE / "plain_frame.npz", times=np.array(FRAME_TIMES), radiance=np.array(rads),
                        image=np.array(imgs), depth=np.array(deps), bytes=np.array(bytes_))


EVENT_FIELDS = ("timestamp", "x", "y", "polarity", "depth")
EVENT_CALLS = 157


def write_events():
    w, o = plain_scene()
    batch = driver.batch_size(w.physic, w.time_resolution, adaptive=False)
    s = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    assert s.initialize(o, w.initial_state(), w.start_time) == 0
    raws, readys = [], []
    t0 = time.time()
    horizons = min(w.total_steps, EVENT_CALLS * batch)
    for k in range(0, horizons, batch):
        code, raw, ready = s.process(o, w.intervals(k, min(batch, horizons - k)))
        assert code == 0
        raws.append(raw)
        readys.append(ready)
        if k % (64 * 100) == 0:
            print(f"horizon {k}/{horizons}: {sum(map(len, raws))} events, {time.time() - t0:.0f} s", flush=True)
    readys.append(s.drain())
    raw, ready = np.concatenate(raws), np.concatenate(readys)
    np.savez_compressed(HERE / "plain_events.npz", batch=batch, horizons=horizons,
                        **{f"raw_{f}": raw[f] for f in EVENT_FIELDS}, **{f"ready_{f}": ready[f] for f in EVENT_FIELDS})


def load_events(prefix: str, data) -> np.ndarray:
    out = np.zeros(len(data[f"{prefix}_timestamp"]), oracle.EVENT)
    for f in EVENT_FIELDS:
        out[f] = data[f"{prefix}_{f}"]
    return out


if __name__ == "__main__":
    which = sys.argv[1:] or ["hapke", "circuit", "frame", "events"]
    for name in which:
        t = time.time()
        {"hapke": write_hapke, "circuit": write_circuit, "frame": write_frame, "events": write_events}[name]()
        print(f"{name}: {time.time() - t:.1f} s", flush=True)
