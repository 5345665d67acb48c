#!/usr/bin/env python3
"""DVS hot-path benchmark on one MI355X.

Workload: the single DAVIS346 request.
- 346x260 sensor, fx=fy=300, 4x4 sub-rays per pixel.
- 1024^2 heightfield (2,093,058 triangles, 200 craters).
- 100 fps trajectory (250 frames, 2.49 s), 10 us host horizons of 1 us
  circuit cells, sun 30 deg above the horizon at intensity 1000.
- Shot and reset noise on (Philox seed).
- Adaptive optics under bin/dvs's solver policy (apps/main_dvs.cpp:96-116):
  windows of 200 horizons, the receptor check at the first window and every
  0.25 s, a rejected window halved.

One step is one twentieth of the flight: the driver windows (bin/dvs calls)
that start in that twentieth, each with its trajectory states evaluated on
the device (CameraTrajectory over the flight's frames) and
EventSensor::process_adaptive: refinement probes and bisection, local exact
cells, the photocurrent with shot noise, the FP64 circuit of every pixel,
device compaction/sort and the readout FIFO.  `--steps 20` times the whole
flight; more steps fly it again on a re-initialised sensor.  Warm-up steps
run on a separate sensor.  Inputs (BVH, mesh, pixel state) are resident in
HBM before timing starts.

Prints one JSON line: comparator events/s over the timed steps (value);
Mrays/s traced and native-equivalent; trajectory seconds per wall second;
per-window p50/p99 latency; the roofline of the dominant kernel (the
refinement's probe trace, VALU instruction issue) and of
the other kernels of a window; the native solver on the same request; the
CPU oracle and the GPU timed on the same first windows; the served form
against the plain per-request path on a stated request mix; and the
batched form (64 uneven segments in one engine) against the same segments
run one by one.  `--mode served|batched` runs the full served and batched
forms alone.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s + events/s at DAVIS346; p50/p99 latency, served vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Wave64 VALU instructions per second at the published FP32 vector peak
# (157.3 TFLOP/s / (64 lanes x 2 flops per FMA)).  32-bit integer VALU
# instructions reach 1,150 G/s on the box (a microbenchmark of independent
# chains); FP32 and FP64 FMAs without packed math issue at half that rate.
VALU_ISSUE_PEAK_G = 157.3e12 / 128 / 1e9
PAIR_BYTES, LEAF_BYTES, TRIANGLE_BYTES = 64, 8, 48
# FP32 operations of the traversal: a node-pair slab test (12 subtractions,
# 12 multiplications), a Moller-Trumbore test (46 adds/multiplies/one divide
# on its full path), and the inverse direction of a ray (3 divides)
PAIR_FLOPS, TRIANGLE_FLOPS, RAY_FLOPS = 24, 46, 3
SLICES = 20

# Per-unit counter figures of this build's kernels on this workload, from
# rocprofv3 --pmc runs of `python3 bench.py --steps 1 --warmup 0
# --no-cpu-baseline --no-served --no-batched` (the flight's first twentieth; the
# summary with its commands is profiles/adaptive_window_davis346_counters.md):
# VALU instructions (SQ_INSTS_VALU), VALUBusy, lane utilisation
# (VALUUtilization), L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)) and HBM
# bytes (FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0: a 1 GiB device copy reads
# FETCH_SIZE 0.5 GiB and WRITE_SIZE 1 GiB), per unit of each kernel's work.
PROFILE = {
    "probe": {"unit": "traced pixel sample", "valu_insts": 1698.36, "valu_busy": 1.260, "lane_util": 0.864,
              "l2_hit": 0.871, "fetch": 293.39, "write": 23.70},
    "query": {"unit": "traced pixel sample", "valu_insts": 2019.82, "valu_busy": 0.568, "lane_util": 0.778,
              "l2_hit": 0.830, "fetch": 659.44, "write": 12.00},
    "rates": {"unit": "pixel cell", "valu_insts": 2.2363, "valu_busy": 0.817, "lane_util": 0.903,
              "l2_hit": 0.683, "fetch": 0.1520, "write": 1.0010},
    "integrate": {"unit": "pixel cell", "valu_insts": 0.3773, "valu_busy": 0.249, "lane_util": 0.508,
                  "l2_hit": 0.787, "fetch": 1.1097, "write": 0.0366},
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def bytes_per_ray(stats: dict) -> tuple[float, float]:
    """Algorithmic bytes per primary and per shadow ray of the SoA layout."""
    def per(prefix):
        rays = max(stats[f"{prefix}_rays"], 1)
        total = (stats[f"{prefix}_pairs"] * PAIR_BYTES + stats[f"{prefix}_leaves"] * LEAF_BYTES +
                 stats[f"{prefix}_triangles"] * TRIANGLE_BYTES)
        return total / rays
    return per("primary"), per("shadow")


def flops_per_ray(stats: dict) -> tuple[float, float]:
    def per(prefix):
        rays = max(stats[f"{prefix}_rays"], 1)
        return (stats[f"{prefix}_pairs"] * PAIR_FLOPS + stats[f"{prefix}_triangles"] * TRIANGLE_FLOPS) / rays + \
            RAY_FLOPS
    return per("primary"), per("shadow")


def valu_roofline(kernel: str, key: str, launch_ms: float, units_per_launch: float, extra: dict) -> dict:
    """VALU-issue-bound kernel: wave64 VALU instructions per launch (the
    profile's per-unit count x the launch's units, counted live) over the
    launch's live HIP-event duration, against VALU_ISSUE_PEAK_G."""
    prof = PROFILE[key]
    achieved = prof["valu_insts"] * units_per_launch / max(launch_ms * 1e-3, 1e-30) / 1e9
    out = {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 2), "peak": round(VALU_ISSUE_PEAK_G, 1),
           "unit": "G VALU instr/s", "frac": round(achieved / VALU_ISSUE_PEAK_G, 4),
           "traffic": int((prof["fetch"] + prof["write"]) * units_per_launch),
           "avg_launch_ms": round(launch_ms, 4), "units_per_launch": round(units_per_launch, 1),
           "unit_of_work": prof["unit"], "valu_instr_per_unit": prof["valu_insts"],
           "valu_busy": prof["valu_busy"], "lane_utilisation": prof["lane_util"], "l2_hit_rate": prof["l2_hit"]}
    out.update(extra)
    return out


def hbm_roofline(kernel: str, key: str, launch_ms: float, bytes_per_launch: float, units_per_launch: float,
                 extra: dict) -> dict:
    prof = PROFILE[key]
    achieved = bytes_per_launch / max(launch_ms * 1e-3, 1e-30) / 1e9
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int((prof["fetch"] + prof["write"]) * units_per_launch),
           "avg_launch_ms": round(launch_ms, 4), "bytes_per_launch": int(bytes_per_launch),
           "valu_busy": prof["valu_busy"], "lane_utilisation": prof["lane_util"], "l2_hit_rate": prof["l2_hit"]}
    out.update(extra)
    return out


def cpu_baseline(w, mesh, windows: list, gpu_windows: dict) -> dict:
    """The CPU oracle (restatement of the reference path) on the host cores
    over the request's first driver windows `windows` (intervals, and
    whether the receptor check ran), from the same initial state as the
    GPU; `gpu_windows` is the GPU's time and events on the same windows."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ctypes as oracle

    v, n, nodes, root, idx = mesh
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scene = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    sensor = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    sensor.initialize(scene, windows[0][0]["start"][:1], w.start_time, threads)
    events, wall = 0, 0.0
    notes = []
    for ivs, check in windows:
        t0 = time.perf_counter()
        code, ok, raw, _, st, _ = sensor.process_adaptive(scene, ivs, tolerance=1e-2, check=check, capacity=512,
                                                          threads=threads)
        if code == 0 and not ok:
            code, raw, _ = sensor.process(scene, ivs, threads)
        wall += time.perf_counter() - t0
        if code:
            raise RuntimeError(f"oracle failed with status {code}")
        events += len(raw)
        notes.append(f"{len(ivs)} horizons{' checked' if check else ''}{'' if ok else ' (rejected, native)'}: "
                     f"{len(raw)} events")
    sim = float(windows[-1][0]["t1"][-1] - windows[0][0]["t0"][0])
    return {"value": round(events / wall, 1), "unit": "events/s", "cores": threads, "kind": "port",
            "wall_s": round(wall, 3), "events": events, "trajectory_s_per_wall_s": round(sim / wall, 6),
            "sample": f"the request's first {len(windows)} driver windows from t=0 on the full "
                      f"{w.camera.width}x{w.camera.height} sensor ({'; '.join(notes)}), initialisation excluded",
            "gpu_same_sample": {"value": round(gpu_windows["events"] / gpu_windows["wall_s"], 1),
                                "unit": "events/s", "events": gpu_windows["events"],
                                "wall_s": round(gpu_windows["wall_s"], 4),
                                "trajectory_s_per_wall_s": round(sim / gpu_windows["wall_s"], 4)}}


def _latency(reqs) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    return {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2)}


def _same_streams(a, b) -> bool:
    return all(x.event_array().tobytes() == y.event_array().tobytes() for x, y in zip(a, b))


def _tune(serve, geometries, reqs, adaptive: bool) -> dict:
    """The autotuned launch configuration on the mix's first requests: native
    steps over (horizons per step, graph replay); adaptive windows (their
    length set by the driver policy) over the photocurrent run length and
    the trace block size."""
    probe = [dataclasses.replace(r, horizons=min(r.horizons, 800)) for r in reqs[: min(6, len(reqs))]]
    if adaptive:
        cands = tuple({"rate_cells": rc, "trace_threads": tt} for rc in (16, 32, 64, 128) for tt in (128, 256))
    else:
        cands = ((16, False), (32, False), (32, True), (64, False))
    tune = serve.autotune(geometries, probe, candidates=cands)
    log("autotune: " + ", ".join(f"{ {k: v for k, v in c.items() if k != 'probe_s'} }" for c in tune["candidates"]))
    return tune["best"]


def served_compare(args, requests: int, horizons: tuple) -> dict:
    """Served form against the plain per-request path on the same requests:
    Poisson arrivals at args.rate over two terrains (512^2 and 1024^2),
    resolutions 64x48 / 173x130 / 346x260, per-request seed, trajectory and
    noise flag, continuous batching under the HBM admission budget.  The
    plain path runs each request alone as bin/dvs does (BVH built and
    uploaded per request), and again with the scenes kept resident."""
    from dvsrender4deepspace_amd import serve, terrain, workloads

    specs = {"terrain512": terrain.TerrainSpec(size=512, amplitude=20.0, craters=100, seed=512),
             "terrain1024": workloads.DAVIS346.terrain}
    base = workloads.DAVIS346
    t = time.perf_counter()
    meshes = serve.terrain_meshes(specs)
    log(f"served mix: meshes in {time.perf_counter() - t:.1f} s")
    adaptive = args.solver == "adaptive"
    reqs = serve.request_mix(base, requests, args.rate, horizons, specs, seed=args.seed, speed=tuple(args.speed),
                             adaptive=adaptive, trajectories=True)
    geometries = {k: serve.build_scene(meshes[k], base, base.camera) for k in specs}
    best = _tune(serve, geometries, reqs, adaptive)
    served = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, budget_bytes=args.budget_gb * 1e9,
                          horizons_per_step=best["horizons_per_step"], use_graph=best["use_graph"],
                          tuning={"rate_cells": best["rate_cells"], "trace_threads": best["trace_threads"]})
    wall_served = server.run(served, realtime=True)
    log(f"served: {wall_served:.2f} s, {server.steps} steps, peak {server.peak_active} active")
    resident = [dataclasses.replace(r, events=[]) for r in reqs]
    wall_resident = serve.run_plain(resident, meshes, realtime=True, scenes_resident=True, geometries=geometries)
    log(f"plain, scenes resident: {wall_resident:.2f} s")
    plain = [dataclasses.replace(r, events=[]) for r in reqs]
    wall_plain = serve.run_plain(plain, meshes, realtime=True)
    log(f"plain per request: {wall_plain:.2f} s")
    events = sum(len(r.event_array()) for r in served)
    noise_off = [i for i, r in enumerate(reqs) if not r.workload.physic.photon_shot_noise]
    return {
        "config": {"workload": "served mix", "requests": len(reqs), "arrival_rate_hz": args.rate,
                   "horizons": list(horizons), "resolutions": serve.RESOLUTIONS,
                   "terrains": {k: v.triangles for k, v in specs.items()}, "admission_budget_gb": args.budget_gb,
                   "supersample": base.camera.samples, "solver": args.solver, "autotuned": best,
                   "feed": "horizon times of each request's CameraTrajectory; the engine evaluates the states "
                           "on the device in each step"},
        "events": events,
        "served": {"wall_s": round(wall_served, 3), "events_per_s": round(events / wall_served, 1),
                   "latency_ms": _latency(served), "steps": server.steps, "peak_active": server.peak_active},
        "plain_per_request": {"wall_s": round(wall_plain, 3), "events_per_s": round(events / wall_plain, 1),
                              "latency_ms": _latency(plain)},
        "plain_scenes_resident": {"wall_s": round(wall_resident, 3), "events_per_s": round(events / wall_resident, 1),
                                  "latency_ms": _latency(resident)},
        "speedup_vs_plain": round(wall_plain / wall_served, 2),
        "speedup_vs_resident_plain": round(wall_resident / wall_served, 2),
        "identical_event_streams": {"all": _same_streams(served, plain) and _same_streams(served, resident),
                                    "noise_off_requests": len(noise_off)},
    }


def served_mode(args) -> dict:
    out = served_compare(args, args.requests, (args.min_horizons, args.max_horizons))
    return {"metric": METRIC, "value": out["served"]["events_per_s"], "unit": "events/s", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic", "mode": "served", **out}


def batched_compare(args, segments: int, horizons: tuple) -> dict:
    """Batched form: `segments` segments of uneven size (resolutions 64x48 /
    173x130 / 346x260, per-segment seed, trajectory and noise flag) over a
    shared 1024^2 terrain advanced together in one engine (autotuned step
    size and launch shapes, trajectories evaluated on the device in each
    step), against the same segments run one by one on the resident scene."""
    from dvsrender4deepspace_amd import serve, workloads

    base = workloads.DAVIS346
    specs = {"terrain1024": base.terrain}
    meshes = serve.terrain_meshes(specs)
    adaptive = args.solver == "adaptive"
    reqs = serve.request_mix(base, segments, None, horizons, specs,
                             seed=args.seed, speed=tuple(args.speed), adaptive=adaptive, trajectories=True)
    geometries = {"terrain1024": serve.build_scene(meshes["terrain1024"], base, base.camera)}
    best = _tune(serve, geometries, reqs, adaptive)
    batched = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, horizons_per_step=best["horizons_per_step"], use_graph=best["use_graph"],
                          tuning={"rate_cells": best["rate_cells"], "trace_threads": best["trace_threads"]})
    wall_batched = server.run(batched, realtime=False)
    log(f"batched: {wall_batched:.2f} s, {server.steps} steps")
    st = server.batch.stats()
    alone = [dataclasses.replace(r, events=[]) for r in reqs]
    wall_alone = serve.run_plain(alone, meshes, realtime=False, scenes_resident=True, geometries=geometries)
    log(f"one by one, scene resident: {wall_alone:.2f} s")
    events = sum(len(r.event_array()) for r in batched)
    rays = st["primary_rays"] + st["shadow_rays"] + st["event_rays"]
    return {
        "config": {"workload": f"{len(reqs)} uneven segments", "segments": len(reqs),
                   "resolutions": serve.RESOLUTIONS, "horizons": list(horizons),
                   "terrain_triangles": base.terrain.triangles, "solver": args.solver, "autotuned": best,
                   "feed": "horizon times of each segment's CameraTrajectory, evaluated on the device"},
        "events": events, "batched_wall_s": round(wall_batched, 3), "one_by_one_wall_s": round(wall_alone, 3),
        "batched_events_per_s": round(events / wall_batched, 1),
        "one_by_one_events_per_s": round(events / wall_alone, 1),
        "batched_mrays_per_s": round(rays / wall_batched / 1e6, 1),
        "speedup_vs_one_by_one": round(wall_alone / wall_batched, 2),
        "identical_event_streams": _same_streams(batched, alone), "steps": server.steps,
    }


def batched_mode(args) -> dict:
    out = batched_compare(args, args.segments, (args.min_horizons, args.max_horizons))
    return {"metric": METRIC, "value": out["batched_events_per_s"], "unit": "events/s", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic", "mode": "batched", **out}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--mode", choices=["request", "served", "batched", "blur"], default="request",
                    help="request: the single DAVIS346 request with the served comparison (default bench line); "
                         "served / batched: the full served and batched forms alone; blur: the request's "
                         "motion-blur exposures alone")
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--segments", type=int, default=64)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--min-horizons", type=int, default=2000)
    ap.add_argument("--max-horizons", type=int, default=20000)
    ap.add_argument("--solver", choices=["adaptive", "native"], default="adaptive",
                    help="event solver of the served/batched requests")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--speed", type=float, nargs=2, default=(100.0, 400.0),
                    help="ground speed range (m/s) of the served/batched mix")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=SLICES, help="timed twentieths of the flight (request mode)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="davis346-terrain1024-ss4")
    ap.add_argument("--horizons", type=int, default=64, help="host horizons per native-solver step")
    ap.add_argument("--cpu-windows", type=int, default=3, help="driver windows of the CPU baseline sample")
    ap.add_argument("--rate-cells", type=int, default=None,
                    help="photocurrent cells per thread run of the adaptive solver (library default 128)")
    ap.add_argument("--trace-threads", type=int, default=None,
                    help="lanes per refinement trace block: 64, 128 or 256 (library default 128)")
    ap.add_argument("--traversal", type=int, choices=[0, 1], default=None,
                    help="0: per-lane traversal stacks, 1: wave packets (the library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the served-vs-plain comparison")
    ap.add_argument("--served-requests", type=int, default=30)
    ap.add_argument("--served-horizons", type=int, nargs=2, default=(2000, 6000))
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-one-by-one comparison")
    ap.add_argument("--batched-segments", type=int, default=64)
    ap.add_argument("--batched-horizons", type=int, nargs=2, default=(2000, 6000))
    ap.add_argument("--blur-exposures", type=int, default=16, help="1/30 s exposures of the motion-blur leg")
    ap.add_argument("--blur-cpu-cells", type=int, default=4, help="cells the CPU oracle renders for its sample")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout "
              "(pixel-row shards, event merge on rank 0) is not launched in this build", file=sys.stderr)
        return 2

    from dvsrender4deepspace_amd import capi

    capi.load_library()
    if capi.device_count() < 1:
        raise RuntimeError("no HIP device visible")
    if args.mode == "blur":
        print(json.dumps(blur_mode(args)), flush=True)
        return 0
    if args.mode != "request":
        print(json.dumps(served_mode(args) if args.mode == "served" else batched_mode(args)), flush=True)
        return 0
    print(json.dumps(request_mode(args)), flush=True)
    return 0


def copy_probe() -> int:
    """A 1 GiB device-to-device copy through the library's HIP runtime."""
    from dvsrender4deepspace_amd import capi

    n = 1 << 30
    a, b = capi.DeviceBuffer(n), capi.DeviceBuffer(n)
    hip = capi.hip_runtime()
    hip.hipMemcpy(b.ptr, a.ptr, n, 3)  # device to device
    hip.hipDeviceSynchronize()
    print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n}))
    return 0


class Flight:
    """bin/dvs's window loop over the flight on one sensor (apps/main_dvs.cpp:
    96-116): windows of `batch` horizons along the flight's CameraTrajectory,
    the receptor check at the start and every 0.25 s, a rejected window
    halved (the smaller window is kept), a rejected one-horizon window
    integrated natively."""

    def __init__(self, w, traj, sensor, batch, ahead=True):
        from dvsrender4deepspace_amd import driver
        self.w, self.traj, self.sensor, self.batch, self.driver = w, traj, sensor, batch, driver
        self.ahead = ahead  # queue each expected next window's setup (dvs_sensor_expect_adaptive_times)
        self.step, self.t_prev, self.next_check = 0, w.start_time, w.start_time
        self.windows = []  # (seconds, comparator events, intervals, checked)

    def window(self):
        w, d = self.w, self.driver
        from dvsrender4deepspace_amd import workloads
        t0 = time.perf_counter()
        while True:
            # the window's horizon times; the sensor evaluates their states
            # on the device inside the window (dvs_sensor_process_adaptive_times)
            times = workloads.horizon_grid(w, self.t_prev, self.step, self.batch)
            check = float(times[0]) >= self.next_check
            # the next window if this one is accepted: its setup is queued
            # on the device before this call returns
            step_next = self.step + len(times) - 1
            if self.ahead and step_next < w.total_steps:
                nxt = workloads.horizon_grid(w, float(times[-1]), step_next, self.batch)
                if len(nxt) > 1:
                    self.sensor.expect_adaptive_times(self.traj, nxt, d.DEFAULT_TOLERANCE)
            ok, ready = self.sensor.process_adaptive_times(self.traj, times, d.DEFAULT_TOLERANCE, check)
            if ok:
                break
            if self.batch > 1:
                self.batch = max(1, self.batch // 2)
                continue
            ready = self.sensor.process(self.traj.intervals(times))
            break
        if check:
            self.next_check = float(times[-1]) + d.CHECK_PERIOD
        dt = time.perf_counter() - t0
        self.step += len(times) - 1
        self.t_prev = float(times[-1])
        self.windows.append((dt, self.sensor.stats()["comparator_events"], times, check))
        return dt

    def run_until(self, horizon: int):
        while self.step < min(horizon, self.w.total_steps):
            self.window()


def blur_leg(w, scene, traj, mesh, exposures: int, cpu_cells: int) -> dict:
    """bin/cam_blur on the request (apps/main_cam_blur.cpp:64-83): 1/30 s
    exposures of 1 ms cells along the flight, every cell of `exposures`
    exposures rendered by one multi-pose launch and summed on the device,
    against the reference's form (one render_main call per cell, summed on
    the host) on the same GPU and, for `cpu_cells` cells, the CPU oracle."""
    from dvsrender4deepspace_amd import capi, driver

    sched = driver.blur_schedule(w.start_time, w.end_time, 30.0, 1e-3)[:exposures]
    times = np.concatenate([x[2] for x in sched])
    weights = np.concatenate([x[3] for x in sched])
    offsets = np.concatenate([[0], np.cumsum([len(x[2]) for x in sched])]).astype(np.int32)
    poses = capi.optical_states(traj.evaluate(times))
    for f in ("velocity", "light_rate", "intensity_rate"):
        poses[f] = 0
    scene.render_blur(poses[:offsets[1]], weights[:offsets[1]], offsets[:2])  # warm-up
    before = scene.stats()
    t0 = time.perf_counter()
    batched = scene.render_blur(poses, weights, offsets)
    wall = time.perf_counter() - t0
    after = scene.stats()
    launches = after["trace_launches"] - before["trace_launches"]
    trace_ms = after["trace_ms"] - before["trace_ms"]
    rays = (after["primary_rays"] - before["primary_rays"]) + (after["shadow_rays"] - before["shadow_rays"])
    t0 = time.perf_counter()
    for e in range(len(sched)):
        acc = np.zeros((w.camera.height, w.camera.width), np.float32)
        for c in range(offsets[e], offsets[e + 1]):
            acc = acc + scene.render(poses[c:c + 1])[0][0] * np.float32(weights[c])
        if e == 0 and not np.array_equal(acc, batched[0]):
            raise RuntimeError("batched motion blur differs from the per-cell sum")
    wall_cells = time.perf_counter() - t0
    cells = int(offsets[-1])
    out = {"exposures": len(sched), "cells": cells, "cells_per_exposure": int(offsets[1]),
           "exposures_per_s": round(len(sched) / wall, 2),
           "mrays_per_s": round(rays / wall / 1e6, 1),
           "trace_launches": int(launches), "trace_ms_per_launch": round(trace_ms / max(launches, 1), 3),
           "trace_mrays_per_s": round(rays / max(trace_ms * 1e-3, 1e-30) / 1e6, 1),
           "per_cell_form": {"exposures_per_s": round(len(sched) / wall_cells, 2),
                             "note": "one render_main call per cell and a host float sum, as the reference"}}
    if cpu_cells > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_ctypes as oracle

        v, n, nodes, root, idx = mesh
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
        t0 = time.perf_counter()
        for c in range(cpu_cells):
            code = o.render(poses[c:c + 1], threads)[0]
            if code:
                raise RuntimeError(f"oracle failed with status {code}")
        cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cpu_cells / cpu / offsets[1], 4), "unit": "exposures/s",
                               "cores": threads, "kind": "port", "wall_s": round(cpu, 3),
                               "sample": f"the oracle's frames of the first {cpu_cells} cells of the first "
                                         f"exposure ({int(offsets[1])} cells per exposure), the sum excluded"}
    return out


def terrain_leg(w, heights, cpu: bool) -> dict:
    """Terrain ingest (tools::load_from_raster, src/tools/build_from_raster.cu):
    the request's DEM meshed on the device, host buffers in and out, against
    the CPU oracle's restatement (one core)."""
    from dvsrender4deepspace_amd import capi

    n = heights.shape[0]
    axis = (np.arange(n) * w.terrain.spacing + 0.5).astype(np.float32)
    ys = axis[::-1].copy()
    capi.raster_mesh(heights[:64, :64], axis[:64], ys[:64])  # warm-up
    t0 = time.perf_counter()
    v, _, tri, _ = capi.raster_mesh(heights, axis, ys)
    wall = time.perf_counter() - t0
    out = {"dem": f"{n}x{n}", "vertices": int(len(v)), "triangles": int(len(tri)), "ms": round(wall * 1e3, 2),
           "mtriangles_per_s": round(len(tri) / wall / 1e6, 2),
           "note": "host DEM in, host mesh out (uploads, flags, two prefix sums, vertices, normals, triangles, "
                   "downloads)"}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_ctypes as oracle

        t0 = time.perf_counter()
        oracle.raster_mesh(heights, axis, ys)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(len(tri) / cpu_s / 1e6, 3), "unit": "Mtriangles/s", "cores": 1,
                               "kind": "port", "ms": round(cpu_s * 1e3, 1), "sample": f"the whole {n}x{n} DEM"}
    return out


def blur_mode(args) -> dict:
    from dvsrender4deepspace_amd import capi, terrain, workloads

    w = workloads.WORKLOADS[args.workload]
    v, n, idx = terrain.mesh(w.terrain)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, w.leaf_size)
    scene = capi.Scene(nodes, root, v, n, leaf_idx, w.camera, w.hapke)
    out = blur_leg(w, scene, workloads.trajectory(w), (v, n, nodes, root, leaf_idx), args.blur_exposures,
                   0 if args.no_cpu_baseline else args.blur_cpu_cells)
    return {"metric": "motion-blur exposures/s at DAVIS346, 1 GPU", "value": out["exposures_per_s"],
            "unit": "exposures/s", "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "mode": "blur",
            "config": {"workload": w.name, "exposure_s": 1 / 30, "cell_s": 1e-3}, **out}


def request_mode(args) -> dict:
    """The single DAVIS346 request over its whole flight (the headline line),
    the native solver on the same request, the CPU oracle on the first
    windows, and the served comparison."""
    from dvsrender4deepspace_amd import capi, driver, terrain, workloads

    w = workloads.WORKLOADS[args.workload]
    t = time.perf_counter()
    v, n, idx = terrain.mesh(w.terrain)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, w.leaf_size)
    log(f"scene: {len(idx)} triangles, {len(nodes)} BVH nodes, built in {time.perf_counter() - t:.1f} s")
    scene = capi.Scene(nodes, root, v, n, leaf_idx, w.camera, w.hapke)
    info = scene.info()
    traj = workloads.trajectory(w)
    first = workloads.horizon_intervals(w, traj, w.start_time, 0, 1)
    initial = first["start"][:1]
    trav = scene.traversal_stats(initial)
    b_primary, b_shadow = bytes_per_ray(trav)
    f_primary, f_shadow = flops_per_ray(trav)
    log(f"scene on device: {info}; per ray: primary {b_primary:.0f} B {f_primary:.0f} flops, "
        f"shadow {b_shadow:.0f} B {f_shadow:.0f} flops")
    npix = w.camera.width * w.camera.height
    R = w.camera.samples ** 2
    batch = driver.batch_size(w.physic, w.time_resolution, adaptive=True)
    H = w.total_steps
    bounds = [k * H // SLICES for k in range(SLICES + 1)]

    # warm-up slices on a separate sensor
    warm = capi.Sensor(scene, w.physic)
    warm.initialize(initial, w.start_time)
    wf = Flight(w, traj, warm, batch)
    for k in range(args.warmup):
        wf.run_until(bounds[(k % SLICES) + 1])
    warm.close()
    log(f"warmup done ({args.warmup} twentieths of the flight on a separate sensor)")

    sensor = capi.Sensor(scene, w.physic)
    tuning = None
    if args.rate_cells or args.trace_threads or args.traversal is not None:
        tuning = {"rate_cells": args.rate_cells or 128, "trace_threads": args.trace_threads or 128,
                  "traversal": 1 if args.traversal is None else args.traversal}
        sensor.set_tuning(**tuning)
        if args.traversal is not None:
            scene.set_traversal(args.traversal)
    sensor.initialize(initial, w.start_time)
    sensor.reset_stats()
    flights = []
    flight = Flight(w, traj, sensor, batch)
    t_start = time.perf_counter()
    for k in range(args.steps):
        if k and k % SLICES == 0:  # the next flight, from a re-initialised sensor
            flights.append(flight)
            sensor.initialize(initial, w.start_time)
            flight = Flight(w, traj, sensor, batch)
        flight.run_until(bounds[(k % SLICES) + 1])
        if (k + 1) % 5 == 0:
            log(f"step {k + 1}: t = {flight.t_prev:.3f} s, {len(flight.windows)} windows, "
                f"{sensor.stats()['comparator_events']} events")
    wall = time.perf_counter() - t_start
    flights.append(flight)
    st = sensor.stats()
    refine = sensor.refinement_stats()
    windows = [x for f in flights for x in f.windows]
    traced = st["primary_rays"] + st["shadow_rays"]
    horizons = sum(f.step for f in flights)
    sim_s = horizons * float(np.float32(w.time_resolution))
    lat = np.array([x[0] for x in windows]) * 1e3
    cells_per_window = int(round(batch * float(np.float32(w.time_resolution)) / float(np.float32(w.physic.circuit_step))))

    # roofline of the dominant kernel (the probe trace), then the others
    probe_launches = max(st["probe_launches"], 1)
    probe_ms = st["probe_ms"] / probe_launches
    probe_units = st["probe_samples"] / probe_launches
    shadow_frac = st["shadow_rays"] / max(st["primary_rays"], 1)
    traversal_flops = probe_units * R * (f_primary + shadow_frac * f_shadow)
    roofline = valu_roofline("adapt_trace (five initial probes per pixel)", "probe", probe_ms, probe_units, {
        "fp32_flops": {"algorithmic_per_launch": int(traversal_flops),
                       "achieved_tflops": round(traversal_flops / (probe_ms * 1e-3) / 1e12, 3),
                       "peak_tflops": FP32_PEAK_TFLOPS,
                       "frac": round(traversal_flops / (probe_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
                       "per_ray": {"primary": round(f_primary, 1), "shadow": round(f_shadow, 1)},
                       "note": "slab tests and triangle tests only: compares, address arithmetic, stack "
                               "and branch instructions, which set the issue time, are not FLOPs"},
        "algorithmic_bytes_per_ray": {"primary": round(b_primary, 1), "shadow": round(b_shadow, 1),
                                      "note": "the 186 MB scene is served from L2 and MALL, so the algorithmic "
                                              "bytes are not the HBM traffic (traffic: counted HBM bytes)"}})
    q_launches = max(st["query_launches"], 1)
    secondary = [
        valu_roofline("adapt_trace (bisection midpoints)", "query", st["query_ms"] / q_launches,
                      st["query_samples"] / q_launches, {"launches_per_window": round(st["query_launches"] /
                                                                                        max(len(windows), 1), 2)}),
    ]
    circuit_launches = max(st["circuit_launches"], 1)
    cells = st["pixel_cells"] / circuit_launches
    state_bytes = capi.PIXEL_STATE_DTYPE.itemsize
    integrate_bytes = cells * 1 + npix * (2 * state_bytes + 2 * 4) + st["comparator_events"] / circuit_launches * 16
    secondary.append(hbm_roofline("adapt_integrate (FP64 circuit)", "integrate", st["circuit_ms"] / circuit_launches,
                                  integrate_bytes, cells,
                                  {"note": "one lane per pixel, cells in series: latency bound; in: one "
                                           "photocurrent code byte per pixel cell, pixel state; out: state, "
                                           "crossings (16 B)"}))
    secondary.append(valu_roofline("adapt_rates (photocurrent, shot noise)", "rates",
                                   st["rates_ms"] / circuit_launches, cells, {}))

    result = {
        "metric": METRIC,
        "value": round(st["comparator_events"] / wall, 1),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 rays, fp64 circuit/timestamps",
        "data": "synthetic",
        "config": {"workload": w.name, "sensor": f"{w.camera.width}x{w.camera.height}",
                   "supersample": w.camera.samples, "terrain_triangles": int(len(idx)),
                   "flight_s": round(w.end_time - w.start_time, 4), "frames": w.frames, "fps": w.fps,
                   "step": f"one twentieth of the flight ({H // SLICES} horizons)",
                   "solver": "adaptive-optics", "window_horizons": batch,
                   "tolerance": driver.DEFAULT_TOLERANCE, "check_period_s": driver.CHECK_PERIOD,
                   "time_resolution_s": w.time_resolution, "circuit_step_s": w.physic.circuit_step,
                   "sun_elevation_deg": w.sun_elevation_deg, "intensity": w.intensity,
                   "noise": "shot+reset on", "bvh_leaf_size": w.leaf_size, "device_leaf_size": 4,
                   "launch_tuning": tuning or "library default"},
        "events": st["comparator_events"],
        "delivered_events_per_s": round(st["delivered_events"] / wall, 1),
        "mrays_per_s": round(traced / wall / 1e6, 3),
        "native_equivalent_mrays_per_s": round(st["native_equivalent_rays"] / wall / 1e6, 3),
        "trajectory_s_per_wall_s": round(sim_s / wall, 5),
        "windows": len(windows),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "breakdown_ms_per_window": {k: round(v / max(len(windows), 1), 4) for k, v in (
            ("probe_trace", st["probe_ms"]), ("query_trace", st["query_ms"]),
            ("refinement_other", st["trace_ms"] - st["probe_ms"] - st["query_ms"]),
            ("photocurrent", st["rates_ms"]), ("circuit", st["circuit_ms"]), ("events", st["events_ms"]),
            ("wall", wall * 1e3))},
        "refinement_last_window": refine,
        "rejected_windows": int(st["rejected_windows"]),
        "roofline": roofline,
        "roofline_other_kernels": secondary,
        "totals": {k: st[k] for k in ("probe_samples", "query_samples", "pixel_cells", "probe_launches",
                                      "query_launches", "circuit_launches", "primary_rays", "shadow_rays")},
        "readout": {"form": "device FIFO: min-plus / max-plus scans, serial recurrence where they do not apply",
                    "windows": int(st["readout_tasks"]), "serial_windows": int(st["readout_serial_tasks"]),
                    "dropped": int(sensor.readout_stats()[0]), "peak_queue": int(sensor.readout_stats()[1]),
                    "bandwidth_events_per_s": w.physic.readout_bandwidth,
                    "queue_capacity": w.physic.readout_queue_capacity},
    }
    sensor.close()

    # --- native solver on the same request (Mrays/s of the cell trace)
    nat = capi.Sensor(scene, w.physic)
    nat.initialize(initial, w.start_time)
    ivs = workloads.horizon_intervals(w, traj, w.start_time, 0, args.horizons)
    nat.process(ivs)
    nat.reset_stats()
    t0 = time.perf_counter()
    nsteps = 2
    for k in range(nsteps):
        ivs = workloads.horizon_intervals(w, traj, float(ivs["t1"][-1]), args.horizons * (k + 1), args.horizons)
        nat.process(ivs)
    nwall = time.perf_counter() - t0
    ns = nat.stats()
    result["native_solver"] = {
        "mrays_per_s": round((ns["primary_rays"] + ns["shadow_rays"] + ns["event_rays"]) / nwall / 1e6, 3),
        "ms_per_step": round(nwall / nsteps * 1e3, 3), "horizons_per_step": args.horizons,
        "trajectory_s_per_wall_s": round(nsteps * args.horizons * w.time_resolution / nwall, 6),
        "trace_cells_ms_per_launch": round(ns["trace_ms"] / max(ns["trace_launches"], 1), 3)}
    nat.close()
    log("terrain ingest ...")
    result["terrain_ingest"] = terrain_leg(w, terrain.heightfield(w.terrain).astype(np.float32),
                                           not args.no_cpu_baseline)
    log("motion blur ...")
    result["motion_blur"] = blur_leg(w, scene, traj, (v, n, nodes, root, leaf_idx), args.blur_exposures,
                                     0 if args.no_cpu_baseline else args.blur_cpu_cells)
    if not args.no_cpu_baseline:
        k = min(args.cpu_windows, len(flights[0].windows))
        first_windows = flights[0].windows[:k]
        events_before = 0
        gpu = {"wall_s": sum(x[0] for x in first_windows), "events": first_windows[-1][1] - events_before}
        log(f"cpu baseline (oracle) on the first {k} windows ...")
        result["cpu_baseline"] = cpu_baseline(w, (v, n, nodes, root, leaf_idx),
                                              [(traj.intervals(x[2]), x[3]) for x in first_windows], gpu)
    if not args.no_served:
        log("served vs plain ...")
        result["served_vs_plain"] = served_compare(args, args.served_requests, tuple(args.served_horizons))
    if not args.no_batched:
        log("batched vs one by one ...")
        result["batched_vs_one_by_one"] = batched_compare(args, args.batched_segments,
                                                          tuple(args.batched_horizons))
    return result


if __name__ == "__main__":
    sys.exit(main())
