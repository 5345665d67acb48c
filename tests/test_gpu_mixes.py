"""GPU: the batched and served request mixes at full resolution on the
DAVIS346 request's terrain.

* Batched: 64 uneven segments admitted together, resolutions 64x48 /
  173x130 / 346x260 at 4x4 sub-rays on the 1024^2 terrain (2,093,058
  triangles), 600-2,400 horizons of 10 us each under the adaptive policy,
  shot and reset noise on in half of them, every segment fed as horizon
  times whose states the engine evaluates on the device.
* Served: 40 requests arriving as a Poisson process at 20 requests/s over the
  512^2 and 1024^2 terrains, admitted under an HBM budget that holds about
  a dozen at once, joining and leaving the running batch.

Each request's event stream must equal, byte for byte, the plain driver
loop's for that request alone (scenes resident).  The mixes' horizon counts
are smaller than the 0.5-5 s segments of the full mixes so that the test
runs in seconds; resolutions, terrains, sub-ray grid, solver and noise are
the full mixes'.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from dvsrender4deepspace_amd import serve, terrain, workloads

pytestmark = pytest.mark.gpu


def _plain_equal(reqs, meshes, geometries, served):
    plain = [dataclasses.replace(r, events=[]) for r in reqs]
    serve.run_plain(plain, meshes, realtime=False, scenes_resident=True, geometries=geometries)
    total = 0
    for a, b in zip(served, plain):
        ea, eb = a.event_array(), b.event_array()
        total += len(ea)
        assert ea.tobytes() == eb.tobytes(), f"request {a.id} ({a.camera.width}x{a.camera.height}) differs"
    return total


def test_batched_64_segments_full_resolution(gpu):
    base = workloads.DAVIS346
    specs = {"terrain1024": base.terrain}
    reqs = serve.request_mix(base, 64, None, (600, 2400), specs, seed=31, adaptive=True, trajectories=True)
    assert {r.camera.width for r in reqs} == {64, 173, 346} and all(r.camera.samples == 4 for r in reqs)
    meshes = serve.terrain_meshes(specs)
    geometries = {k: serve.build_scene(meshes[k], base, base.camera) for k in specs}
    print(f"batched mix: {len(reqs)} segments, {sum(r.horizons for r in reqs)} horizons", flush=True)
    served = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, budget_bytes=200e9)
    server.run(served, realtime=False)
    assert server.peak_active == 64
    total = _plain_equal(reqs, meshes, geometries, served)
    print(f"batched mix: {total} events, {server.steps} steps", flush=True)
    assert total > 1000


def test_served_poisson_mix_full_resolution(gpu):
    base = workloads.DAVIS346
    specs = {"terrain512": dataclasses.replace(base.terrain, size=512, craters=50, seed=77),
             "terrain1024": base.terrain}
    reqs = serve.request_mix(base, 40, 20.0, (600, 2400), specs, seed=32, adaptive=True, trajectories=True)
    meshes = serve.terrain_meshes(specs)
    geometries = {k: serve.build_scene(meshes[k], base, base.camera) for k in specs}
    budget = 12 * np.median([serve.footprint_bytes(r, 16) for r in reqs])
    served = [dataclasses.replace(r, events=[]) for r in reqs]
    server = serve.Server(geometries, budget_bytes=budget)
    wall = server.run(served, realtime=True)
    assert server.peak_used <= budget and 1 < server.peak_active < len(reqs)
    lat = np.array([r.latency for r in served])
    assert np.all(lat > 0)
    total = _plain_equal(reqs, meshes, geometries, served)
    print(f"served mix: {total} events in {wall:.2f} s, peak {server.peak_active} active, "
          f"p50 latency {np.percentile(lat, 50) * 1e3:.0f} ms", flush=True)
    assert total > 1000
