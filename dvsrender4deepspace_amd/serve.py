"""Batched and served forms of the DVS path on one GPU.

``Server`` is a continuous-batching request loop over the multi-segment
engine (``capi.Batch``): requests wait until their HBM footprint fits the
admission budget, join the running batch at the next engine step, share every
launch with the requests already running, keep their circuit state resident
between steps and leave when their last horizon has been read out.  Terrain
BVHs stay resident across requests; each request renders through a camera
view of its terrain.  Render (frame) jobs run through the same loop.

``run_plain`` is the plain per-request path of the reference's driver
(apps/main_dvs.cpp:60-129): per request the scene is built (SAH BVH) and
uploaded, a sensor is created and initialised, the horizons go through the
driver loop (``driver.drive``: native calls of 64 horizons, or adaptive
windows with the receptor check) and the readout is drained.
``scenes_resident=True`` keeps the scenes uploaded between requests
(isolating the launch-sharing gain from the setup gain).  Served requests
of the adaptive solver run under the same window policy inside the engine,
so both paths give every request the same events.

``autotune`` times short probes of the actual segment mix per candidate
launch configuration: horizons per step and graph replay for native steps;
for adaptive windows (whose length the driver policy fixes) the photocurrent
cells per thread run and the refinement trace's lanes per block (pixel
samples per block at 4x4 sub-rays: 4, 8 or 16).
"""
from __future__ import annotations

import dataclasses
import heapq
import time
from collections import deque

import numpy as np

from . import capi, driver, terrain
from .workloads import Workload


@dataclasses.dataclass
class Request:
    """One event job: a trajectory segment of `horizons` host horizons over a
    terrain, seen by `camera`, with its own physic (seed, noise flags).  With
    `frames` > 0 it is a render job of that many frames instead."""
    id: int
    arrival: float
    terrain: str
    camera: capi.Camera
    workload: Workload
    horizons: int = 0
    frames: int = 0
    adaptive: bool = False
    # CameraTrajectory of the flight: event jobs with one are fed as horizon
    # times whose states the engine evaluates on the device
    trajectory: object = None
    # filled by the runners
    events: list = dataclasses.field(default_factory=list)
    images: object = None
    admitted: float = float("nan")
    finished: float = float("nan")

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    def event_array(self) -> np.ndarray:
        return np.concatenate(self.events) if self.events else np.zeros(0, capi.EVENT_DTYPE)


def horizon_grid(w: Workload, horizons: int) -> np.ndarray:
    """The request's horizon times from its start (workloads.horizon_grid)."""
    from .workloads import horizon_grid as grid
    return grid(w, w.start_time, 0, horizons)


def footprint_bytes(req: Request, horizons_per_step: int) -> int:
    """Device bytes a request holds while admitted: pixel states, crossing
    slots, counts/offsets, its share of the step's cell radiance scratch and
    event buffers (upper bound: every slot filled)."""
    npix = req.camera.width * req.camera.height
    if req.frames:
        return npix * req.frames * (4 + 4 + 8) + (1 << 20)
    w = req.workload
    step = driver.batch_size(w.physic, w.time_resolution, True) if req.adaptive else horizons_per_step
    cells = int(np.ceil(step * w.time_resolution / w.physic.circuit_step))
    slots = npix * w.physic.max_events_per_step
    knots = 64 * 16 if req.adaptive else 0  # typical refinement knots per pixel
    return npix * (56 + 8 + cells * 8 + knots) + slots * (16 + 64) + (1 << 20)


class Server:
    """Continuous-batching request loop with HBM-footprint admission."""

    def __init__(self, geometries: dict, budget_bytes: float = 200e9, horizons_per_step: int = 16,
                 use_graph: bool = False, tuning: dict | None = None):
        self.geometries = geometries          # terrain key -> capi.Scene (resident BVH + mesh)
        self.budget = budget_bytes
        self.horizons_per_step = horizons_per_step
        self.batch = capi.Batch(horizons_per_step=horizons_per_step, use_graph=use_graph)
        if tuning:
            self.batch.set_tuning(**tuning)
        self.views = {}
        self.waiting = deque()
        self.active = {}                      # segment id -> (request, bytes)
        self.used = 0
        self.steps = 0
        self.peak_active = 0
        self.peak_used = 0

    def view(self, req: Request) -> capi.Scene:
        key = (req.terrain, req.camera)
        if key not in self.views:
            self.views[key] = self.geometries[req.terrain].view(req.camera, req.workload.hapke)
        return self.views[key]

    def submit(self, req: Request) -> None:
        self.waiting.append(req)

    def _admit(self, clock) -> list:
        done = []
        while self.waiting:
            req = self.waiting[0]
            need = footprint_bytes(req, self.horizons_per_step)
            if self.used + need > self.budget and (self.active or self.used):
                break
            self.waiting.popleft()
            req.admitted = clock()
            view = self.view(req)
            if req.frames:
                times = np.linspace(req.workload.start_time, req.workload.end_time, req.frames)
                req.images = view.render(req.workload.states(times))
                req.finished = clock()
                done.append(req)
                continue
            w = req.workload
            if req.trajectory is not None:
                times = horizon_grid(w, req.horizons)
                first = req.trajectory.intervals(times[:2])["start"][:1]
                seg = self.batch.add(view, w.physic, first, w.start_time)
            else:
                seg = self.batch.add(view, w.physic, w.initial_state(), w.start_time)
            if req.adaptive:
                self.batch.set_adaptive(seg, driver.DEFAULT_TOLERANCE,
                                        driver.batch_size(w.physic, w.time_resolution, True), driver.CHECK_PERIOD)
            if req.trajectory is not None:
                self.batch.feed_trajectory(seg, req.trajectory, times)
            else:
                self.batch.feed(seg, w.intervals(0, req.horizons))
            self.active[seg] = (req, need)
            self.used += need
        self.peak_active = max(self.peak_active, len(self.active))
        self.peak_used = max(self.peak_used, self.used)
        return done

    def step(self, clock) -> list:
        """Admits what fits, runs one engine step, returns finished requests."""
        done = self._admit(clock)
        if not self.active:
            return done
        self.batch.step()
        self.steps += 1
        for seg in list(self.active):
            req, need = self.active[seg]
            req.events.append(self.batch.take(seg))
            if self.batch.info(seg)["queued"] == 0:
                req.events.append(self.batch.drain(seg))
                req.finished = clock()
                self.batch.remove(seg)
                del self.active[seg]
                self.used -= need
                done.append(req)
        return done

    @property
    def idle(self) -> bool:
        return not self.waiting and not self.active

    def run(self, requests: list, realtime: bool = True) -> float:
        """Serves requests arriving at their `arrival` times (seconds after the
        start; all at once when realtime is False).  Returns the wall time."""
        pending = [(r.arrival, r.id, r) for r in requests]
        heapq.heapify(pending)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        finished = 0
        while finished < len(requests):
            now = clock()
            while pending and (not realtime or pending[0][0] <= now):
                self.submit(heapq.heappop(pending)[2])
            if self.idle and pending:
                time.sleep(max(0.0, pending[0][0] - clock()))
                continue
            finished += len(self.step(clock))
        return clock()


def build_scene(mesh, w: Workload, camera: capi.Camera, device_leaf_size=None):
    """Reference flow per request: SAH BVH build (leaf 64) and upload."""
    v, n, idx = mesh
    nodes, root, leaf_idx = capi.bvh_build(v, idx, w.leaf_size)
    return capi.Scene(nodes, root, v, n, leaf_idx, camera, w.hapke, device_leaf_size=device_leaf_size)


def run_plain(requests: list, meshes: dict, realtime: bool = True, scenes_resident: bool = False,
              geometries: dict | None = None) -> float:
    """Processes requests one at a time in arrival order, as separate runs of
    the reference driver.  Returns the wall time."""
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    for req in sorted(requests, key=lambda r: (r.arrival, r.id)):
        if realtime and req.arrival > clock():
            time.sleep(req.arrival - clock())
        req.admitted = clock()
        w = req.workload
        if scenes_resident:
            scene = geometries[req.terrain].view(req.camera, w.hapke)
        else:
            scene = build_scene(meshes[req.terrain], w, req.camera)
        if req.frames:
            times = np.linspace(w.start_time, w.end_time, req.frames)
            req.images = scene.render(w.states(times))
        else:
            if req.trajectory is not None:
                iv = req.trajectory.intervals(horizon_grid(w, req.horizons))
            else:
                iv = w.intervals(0, req.horizons)
            sensor = capi.Sensor(scene, w.physic)
            sensor.initialize(iv["start"][:1], w.start_time)
            batch = driver.batch_size(w.physic, w.time_resolution, req.adaptive)
            req.events.extend(driver.drive(sensor, iv, adaptive=req.adaptive, batch=batch))
            sensor.close()
        scene.close()
        req.finished = clock()
    return clock()


NATIVE_CANDIDATES = ((8, False), (16, False), (32, False), (16, True), (32, True), (64, False))
ADAPTIVE_CANDIDATES = tuple({"rate_cells": rc, "trace_threads": tt} for rc in (16, 32, 64, 128) for tt in (64, 128, 256))


def autotune(geometries: dict, requests: list, candidates=None) -> dict:
    """Runs the probe requests through a Server twice per candidate, the
    first pass as warm-up (allocations, graph capture), and returns the
    candidate whose second pass traced the most rays per second.  A
    candidate is (horizons per step, graph replay) or a dict of Server
    settings and launch tuning (horizons_per_step, use_graph, rate_cells,
    trace_threads); by default NATIVE_CANDIDATES, or ADAPTIVE_CANDIDATES when
    the probe requests are adaptive."""
    if candidates is None:
        candidates = ADAPTIVE_CANDIDATES if any(r.adaptive for r in requests) else NATIVE_CANDIDATES
    results = []
    for cand in candidates:
        cfg = {"horizons_per_step": 16, "use_graph": False, "rate_cells": 128, "trace_threads": 128}
        if isinstance(cand, dict):
            cfg.update(cand)
        else:
            cfg["horizons_per_step"], cfg["use_graph"] = cand
        hps, graph = cfg["horizons_per_step"], cfg["use_graph"]
        tuning = {"rate_cells": cfg["rate_cells"], "trace_threads": cfg["trace_threads"]}
        server = Server(geometries, horizons_per_step=hps, use_graph=graph, tuning=tuning)
        for rep in range(2):
            probe = [dataclasses.replace(r, events=[], arrival=0.0) for r in requests]
            server.batch.reset_stats()
            t = time.perf_counter()
            server.run(probe, realtime=False)
            dt = time.perf_counter() - t
        st = server.batch.stats()
        rays = st["primary_rays"] + st["shadow_rays"] + st["event_rays"]
        results.append({**cfg, "mrays_per_s": rays / dt / 1e6, "probe_s": dt})
        server.batch.close()
    best = max(results, key=lambda r: r["mrays_per_s"])
    return {"best": best, "candidates": results}


RESOLUTIONS = ((64, 48), (173, 130), (346, 260))


def request_mix(base: Workload, count: int, rate_hz: float | None, horizons: tuple, terrains: dict, seed: int,
                noise_fraction: float = 0.5, render_fraction: float = 0.0, speed: tuple = (20.0, 80.0),
                adaptive: bool = False, trajectories: bool = False) -> list:
    """Seeded request mix: resolution from RESOLUTIONS (fx scaled with width),
    length U(horizons), terrain uniform over `terrains`, per-request seed and
    noise flag, ground speed U(speed) and yaw rate U(-5, 5) deg/s seeded;
    Poisson arrivals at rate_hz (all at t=0 when None); event jobs use the
    adaptive solver when `adaptive`, and with `trajectories` fly a
    CameraTrajectory through the flight's two end frames, evaluated on the
    device (served: inside the engine's steps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    keys = sorted(terrains)
    for i in range(count):
        if rate_hz:
            t += rng.exponential(1.0 / rate_hz)
        wpx, hpx = RESOLUTIONS[rng.integers(len(RESOLUTIONS))]
        f = base.camera.fx * wpx / base.camera.width
        cam = capi.Camera(wpx, hpx, f, f, wpx / 2.0, hpx / 2.0, base.camera.samples)
        key = keys[rng.integers(len(keys))]
        spec = terrains[key]
        noisy = rng.uniform() < noise_fraction
        phys = dataclasses.replace(base.physic, random_seed=int(rng.integers(1, 2 ** 31 - 1)),
                                   photon_shot_noise=int(noisy), reset_thermal_noise=int(noisy),
                                   threshold_std=0.03 if noisy else 0.0)
        # seeded start pose and velocity; altitude below the terrain's half size
        half = (spec.size - 1) * spec.spacing / 2
        render = rng.uniform() < render_fraction
        n_h = 0 if render else int(rng.integers(horizons[0], horizons[1] + 1))
        n_f = int(rng.integers(2, 6)) if render else 0
        # the pass is centred on the terrain and spans this request only:
        # render jobs n_f frames at the base frame rate, event jobs n_h horizons
        fps = base.fps if render else 1.0 / (n_h * base.time_resolution)
        w = dataclasses.replace(base, name=f"{base.name}-req{i}", terrain=spec, physic=phys,
                                frames=n_f if render else 2, fps=fps,
                                speed=float(rng.uniform(*speed)), yaw_rate_deg=float(rng.uniform(-5, 5)),
                                altitude=float(min(base.altitude, 0.8 * half * f / wpx * 2)))
        traj = None
        if trajectories and not render:
            from .workloads import trajectory
            traj = trajectory(w)
        out.append(Request(id=i, arrival=t, terrain=key, camera=cam, workload=w, horizons=n_h, frames=n_f,
                           adaptive=adaptive and not render, trajectory=traj))
    return out


def terrain_meshes(specs: dict) -> dict:
    return {k: terrain.mesh(s) for k, s in specs.items()}
