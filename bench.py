#!/usr/bin/env python3
"""DVS hot-path benchmark on one MI355X.

Workload: the single DAVIS346 request.
- 346x260 sensor, fx=fy=300, 4x4 sub-rays per pixel.
- 1024^2 heightfield (2,093,058 triangles, 200 craters).
- 100 fps trajectory, 10 us host horizons of 1 us circuit cells.
- Shot and reset noise on (Philox seed).
- Adaptive optics under bin/dvs's solver policy (apps/main_dvs.cpp:96-116):
  windows of 200 horizons, the receptor check at the first window and every
  0.25 s.

One step is one driver call as bin/dvs makes it: the window's trajectory
states (CameraTrajectory over the flight's 100 fps frames, evaluated on the
device) and EventSensor::process_adaptive over the window.  It covers the
refinement probes and bisection, local exact cells, the photocurrent with
shot noise, the FP64 circuit of 2,000 cells x 89,960 pixels, device
compaction/sort and the readout FIFO.  Warm-up windows run on
a separate sensor; the timed sensor starts at t=0.  Inputs (BVH, mesh, pixel
state) are resident in HBM before timing starts.

Prints one JSON line: comparator events/s (value); Mrays/s traced and
native-equivalent; trajectory seconds per wall second; per-window p50/p99
latency; the roofline of the dominant kernel (the serial circuit).  Next to
them: the native solver on the same request (Mrays/s and the cell-trace
roofline, with algorithmic bytes per ray from the product's traversal
counters) and the CPU oracle timed on the host on the request's first
window.  `--mode served|batched` runs the served and batched forms against
the plain per-request path.
"""
from __future__ import annotations

import argparse
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
PAIR_BYTES, LEAF_BYTES, TRIANGLE_BYTES = 64, 8, 48


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


def cpu_baseline(w, mesh, horizons: int, ivs) -> dict:
    """The CPU oracle (restatement of the reference path) on the host cores:
    the request's first adaptive window (intervals `ivs`), checked by the
    receptor gate as the driver checks it."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ctypes as oracle

    v, n, nodes, root, idx = mesh
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scene = oracle.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    sensor = oracle.Sensor(w.camera.width, w.camera.height, w.physic)
    sensor.initialize(scene, ivs["start"][:1], w.start_time, threads)
    t0 = time.perf_counter()
    code, ok, raw, _, st, _ = sensor.process_adaptive(scene, ivs, tolerance=1e-2, check=True, capacity=512,
                                                      threads=threads)
    if code == 0 and not ok:
        code, raw, _ = sensor.process(scene, ivs, threads)
    wall = time.perf_counter() - t0
    if code:
        raise RuntimeError(f"oracle failed with status {code}")
    sim = float(ivs["t1"][-1] - ivs["t0"][0])
    return {"value": round(len(raw) / wall, 1), "unit": "events/s", "cores": threads, "kind": "port",
            "trajectory_s_per_wall_s": round(sim / wall, 6),
            "sample": f"the request's first {horizons}-horizon window from t=0 on the full "
                      f"{w.camera.width}x{w.camera.height} sensor with the receptor check, initialisation "
                      f"excluded ({'accepted' if ok else 'rejected, then native'}): {len(raw)} events in "
                      f"{wall:.2f} s; refinement {st}"}


def _latency(reqs) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    return {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2)}


def _same_streams(a, b) -> bool:
    return all(x.event_array().tobytes() == y.event_array().tobytes() for x, y in zip(a, b))


def _tune(serve, geometries, reqs, adaptive: bool) -> dict:
    """Native steps: the autotuned (horizons per step, graph replay).  Adaptive
    segments step by the driver's windows, so there is nothing to tune."""
    if adaptive:
        return {"horizons_per_step": 64, "use_graph": False, "note": "adaptive windows of the driver policy"}
    tune = serve.autotune(geometries, reqs[: min(6, len(reqs))],
                          candidates=((16, False), (32, False), (32, True), (64, False)))
    log(f"autotune: {tune}")
    return tune["best"]


def served_mode(args) -> dict:
    """Served form: Poisson request arrivals over two resident terrains,
    continuous batching, HBM-footprint admission; against the plain
    per-request path (scene built and uploaded per request) and the plain path
    with the scenes kept resident.  Same requests, same seeds."""
    import dataclasses

    from dvsrender4deepspace_amd import capi, serve, terrain, workloads

    specs = {"terrain512": terrain.TerrainSpec(size=512, amplitude=20.0, craters=100, seed=512),
             "terrain1024": workloads.DAVIS346.terrain}
    base = workloads.DAVIS346
    t = time.perf_counter()
    meshes = serve.terrain_meshes(specs)
    log(f"meshes in {time.perf_counter() - t:.1f} s")
    adaptive = args.solver == "adaptive"
    reqs = serve.request_mix(base, args.requests, args.rate, (args.min_horizons, args.max_horizons), specs,
                             seed=args.seed, speed=tuple(args.speed), adaptive=adaptive)
    geometries = {k: serve.build_scene(meshes[k], base, base.camera) for k in specs}
    best = _tune(serve, geometries, reqs, adaptive)
    served = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, budget_bytes=args.budget_gb * 1e9,
                          horizons_per_step=best["horizons_per_step"], use_graph=best["use_graph"])
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
        "metric": METRIC, "value": round(events / wall_served, 1), "unit": "events/s", "n_gpus": 1,
        "higher_is_better": True, "data": "synthetic", "mode": "served",
        "config": {"workload": "served mix", "requests": len(reqs), "arrival_rate_hz": args.rate,
                   "horizons": [args.min_horizons, args.max_horizons], "resolutions": serve.RESOLUTIONS,
                   "terrains": {k: v.triangles for k, v in specs.items()}, "admission_budget_gb": args.budget_gb,
                   "supersample": base.camera.samples, "solver": args.solver, "autotuned": best},
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


def batched_mode(args) -> dict:
    """Batched form: 64 segments of uneven size over a shared 1024^2 terrain
    advanced together (graph-replayed steps, autotuned step size), checked
    against the same segments run one by one."""
    import dataclasses

    from dvsrender4deepspace_amd import serve, terrain, workloads

    base = workloads.DAVIS346
    specs = {"terrain1024": base.terrain}
    meshes = serve.terrain_meshes(specs)
    adaptive = args.solver == "adaptive"
    reqs = serve.request_mix(base, args.segments, None, (args.min_horizons, args.max_horizons), specs,
                             seed=args.seed, speed=tuple(args.speed), adaptive=adaptive)
    geometries = {"terrain1024": serve.build_scene(meshes["terrain1024"], base, base.camera)}
    best = _tune(serve, geometries, reqs, adaptive)
    batched = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, horizons_per_step=best["horizons_per_step"], use_graph=best["use_graph"])
    wall_batched = server.run(batched, realtime=False)
    st = server.batch.stats()
    alone = [dataclasses.replace(r, events=[]) for r in reqs]
    wall_alone = serve.run_plain(alone, meshes, realtime=False, scenes_resident=True, geometries=geometries)
    events = sum(len(r.event_array()) for r in batched)
    rays = st["primary_rays"] + st["shadow_rays"] + st["event_rays"]
    return {
        "metric": METRIC, "value": round(events / wall_batched, 1), "unit": "events/s", "n_gpus": 1,
        "higher_is_better": True, "data": "synthetic", "mode": "batched",
        "config": {"workload": "64 uneven segments", "segments": len(reqs), "resolutions": serve.RESOLUTIONS,
                   "horizons": [args.min_horizons, args.max_horizons], "terrain_triangles": base.terrain.triangles,
                   "solver": args.solver, "autotuned": best},
        "events": events, "batched_wall_s": round(wall_batched, 3), "one_by_one_wall_s": round(wall_alone, 3),
        "batched_mrays_per_s": round(rays / wall_batched / 1e6, 1),
        "speedup_vs_one_by_one": round(wall_alone / wall_batched, 2),
        "identical_event_streams": _same_streams(batched, alone), "steps": server.steps,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--mode", choices=["request", "served", "batched"], default="request",
                    help="request: the single DAVIS346 request (default bench line); served / batched: the "
                         "served and batched forms against the plain per-request path")
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
    ap.add_argument("--steps", type=int, default=40, help="timed driver windows (request mode)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="davis346-terrain1024-ss4")
    ap.add_argument("--horizons", type=int, default=64, help="host horizons per native-solver step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout "
              "(pixel-row shards, event merge on rank 0) is not launched in this build", file=sys.stderr)
        return 2

    from dvsrender4deepspace_amd import capi, terrain, workloads

    capi.load_library()
    if capi.device_count() < 1:
        raise RuntimeError("no HIP device visible")
    if args.mode != "request":
        print(json.dumps(served_mode(args) if args.mode == "served" else batched_mode(args)), flush=True)
        return 0
    print(json.dumps(request_mode(args)), flush=True)
    return 0


def request_mode(args) -> dict:
    """The single DAVIS346 request: adaptive optics under the driver's window
    policy (the headline line), then a short native-solver run of the same
    request for the cell-trace roofline, then the CPU oracle on a sample."""
    from dvsrender4deepspace_amd import capi, driver, terrain, workloads

    w = workloads.WORKLOADS[args.workload]
    t = time.perf_counter()
    v, n, idx = terrain.mesh(w.terrain)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, w.leaf_size)
    log(f"scene: {len(idx)} triangles, {len(nodes)} BVH nodes, built in {time.perf_counter() - t:.1f} s")
    scene = capi.Scene(nodes, root, v, n, leaf_idx, w.camera, w.hapke)
    info = scene.info()
    trav = scene.traversal_stats(workloads.horizon_intervals(w, workloads.trajectory(w), w.start_time, 0, 1)
                                 ["start"][:1])
    b_primary, b_shadow = bytes_per_ray(trav)
    log(f"scene on device: {info}; bytes/ray primary {b_primary:.0f}, shadow {b_shadow:.0f}")
    npix = w.camera.width * w.camera.height
    R = w.camera.samples ** 2

    # --- adaptive optics: one step = one driver call (window of `batch`
    # horizons) flown along the flight's CameraTrajectory (frames at the
    # workload's fps), the window's states evaluated on the device
    batch = driver.batch_size(w.physic, w.time_resolution, adaptive=True)
    traj = workloads.trajectory(w)
    first = workloads.horizon_intervals(w, traj, w.start_time, 0, 1)
    initial = first["start"][:1]

    def windows(sensor, count, timed):
        step, next_check, hz, lat, ready_n, rejected = 0, w.start_time, batch, [], 0, 0
        done, t_prev = 0, w.start_time
        while done < count:
            t0 = time.perf_counter()
            ivs = workloads.horizon_intervals(w, traj, t_prev, step, hz)
            check = float(ivs["t0"][0]) >= next_check
            ok, ready = sensor.process_adaptive(ivs, driver.DEFAULT_TOLERANCE, check)
            if not ok:
                rejected += 1
                if hz > 1:
                    hz = max(1, hz // 2)
                    continue
                ready = sensor.process(ivs)
            if check:
                next_check = float(ivs["t1"][-1]) + driver.CHECK_PERIOD
            lat.append(time.perf_counter() - t0)
            ready_n += len(ready)
            step += len(ivs)
            t_prev = float(ivs["t1"][-1])
            done += 1
            if timed and done % 10 == 0:
                log(f"window {done}: {lat[-1] * 1e3:.1f} ms, {len(ready)} events delivered")
        return step, lat, ready_n, rejected

    warm = capi.Sensor(scene, w.physic)
    warm.initialize(initial, w.start_time)
    windows(warm, max(args.warmup, 1), False)
    warm.close()
    log(f"warmup done ({args.warmup} windows on a separate sensor)")
    sensor = capi.Sensor(scene, w.physic)
    sensor.initialize(initial, w.start_time)
    sensor.reset_stats()
    t_start = time.perf_counter()
    horizons, latencies, delivered, rejected = windows(sensor, args.steps, True)
    wall = time.perf_counter() - t_start
    st = sensor.stats()
    refine = sensor.refinement_stats()
    traced = st["primary_rays"] + st["shadow_rays"]
    sim_s = horizons * float(np.float32(w.time_resolution))
    cells = horizons * int(round(float(np.float32(w.time_resolution)) / float(np.float32(w.physic.circuit_step))))
    lat = np.array(latencies) * 1e3
    # dominant kernel: the serial circuit (adapt_integrate), HIP-event timed
    launches = max(st["circuit_launches"], 1)
    circuit_s = st["circuit_ms"] / 1e3
    state_bytes = capi.PIXEL_STATE_DTYPE.itemsize
    bytes_integrate = (npix * cells * 8            # photocurrent in, one double per pixel cell
                       + st["circuit_launches"] * npix * (2 * state_bytes + 2 * 4)  # pixel state and counts in/out
                       + st["comparator_events"] * 16)                               # crossings out
    achieved = bytes_integrate / max(circuit_s, 1e-30) / 1e9
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
                   "solver": "adaptive-optics", "window_horizons": batch,
                   "tolerance": driver.DEFAULT_TOLERANCE, "check_period_s": driver.CHECK_PERIOD,
                   "time_resolution_s": w.time_resolution, "circuit_step_s": w.physic.circuit_step,
                   "noise": "shot+reset on", "bvh_leaf_size": w.leaf_size, "device_leaf_size": 4},
        "events": st["comparator_events"],
        "delivered_events_per_s": round(delivered / wall, 1),
        "mrays_per_s": round(traced / wall / 1e6, 3),
        "native_equivalent_mrays_per_s": round(st["native_equivalent_rays"] / wall / 1e6, 3),
        "trajectory_s_per_wall_s": round(sim_s / wall, 5),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "breakdown_ms": {"refine_and_probes": round(st["trace_ms"], 3), "photocurrent": round(st["rates_ms"], 3),
                         "circuit": round(st["circuit_ms"], 3), "events": round(st["events_ms"], 3),
                         "wall": round(wall * 1e3, 3)},
        "refinement_last_window": refine,
        "rejected_windows": rejected,
        "roofline": {"bound": "hbm", "kernel": "adapt_integrate", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "avg_launch_ms": round(st["circuit_ms"] / launches, 3),
                     "bytes_per_launch": int(bytes_integrate / launches),
                     "note": "one lane per pixel, cells in series (FP64 inverse receptor): latency bound at "
                             f"{npix / 64:.0f} waves, one launch per window after the window's photocurrent "
                             "kernel; photocurrent in is 8 B per pixel cell"},
    }
    sensor.close()

    # --- native solver on the same request: the cell-trace roofline
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
    trace_s = ns["trace_ms"] / 1e3
    algorithmic = ns["primary_rays"] * b_primary + ns["shadow_rays"] * b_shadow
    nachieved = algorithmic / max(trace_s, 1e-30) / 1e9
    result["native_solver"] = {
        "mrays_per_s": round((ns["primary_rays"] + ns["shadow_rays"] + ns["event_rays"]) / nwall / 1e6, 3),
        "ms_per_step": round(nwall / nsteps * 1e3, 3), "horizons_per_step": args.horizons,
        "trajectory_s_per_wall_s": round(nsteps * args.horizons * w.time_resolution / nwall, 6),
        "roofline": {"bound": "hbm", "kernel": "trace_cells_multi", "achieved": round(nachieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(nachieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_ms": round(ns["trace_ms"] / max(ns["trace_launches"], 1), 3),
                     "bytes_per_ray": {"primary": round(b_primary, 1), "shadow": round(b_shadow, 1)},
                     "note": "the 186 MB scene is served from L2/MALL: algorithmic bytes exceed HBM peak; "
                             "the kernel is VALU bound"},
    }
    nat.close()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        result["cpu_baseline"] = cpu_baseline(w, (v, n, nodes, root, leaf_idx), batch,
                                              workloads.horizon_intervals(w, traj, w.start_time, 0, batch))
    return result


if __name__ == "__main__":
    sys.exit(main())
