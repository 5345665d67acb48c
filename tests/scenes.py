"""Scene fixtures shared by the tests: workload meshes, BVHs and plane fixtures."""
from __future__ import annotations

import dataclasses
import functools

import numpy as np

from dvsrender4deepspace_amd import capi, terrain, workloads


@functools.lru_cache(maxsize=8)
def workload_mesh(w: workloads.Workload):
    """(vertices, normals, nodes, root, indices in leaf order) for a workload."""
    v, n, idx = terrain.mesh(w.terrain)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, w.leaf_size)
    return v, n, nodes, root, leaf_idx


def fast_pass(w: workloads.Workload, **kw) -> workloads.Workload:
    """Faster flight over the same scene, so short windows emit many events."""
    base = dict(speed=4000.0, yaw_rate_deg=300.0, frames=2, fps=100.0)
    base.update(kw)
    return dataclasses.replace(w, name=w.name + "-fast", **base)


def plane_fixture():
    """Two-triangle 200x100 m half-plane on +Y at z=0 (tests/test_render.cu:37-41)."""
    v = np.array([[-100, 0, 0], [100, 0, 0], [100, 100, 0], [-100, 100, 0]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, 8)
    return v, n, nodes, root, leaf_idx


def nadir_state(z=10.0, intensity=1000.0, light=(0.0, 0.0, 1.0)):
    s = np.zeros(1, capi.STATE_DTYPE)
    s["position"] = (0.0, 0.0, z)
    s["rotation"] = (0.0, 0.0, 0.0, 1.0)
    s["light"] = light
    s["intensity"] = intensity
    return s


# The sensor fixture of tests/test_sensor.cu:66-73: a 200 x 200 m plane at
# z=0 as two triangles, nadir pose at 10 m with the light overhead at
# intensity 1000, the DAVIS346 profile's circuit (config/davis346.yaml) with
# mismatch, shot noise and FIFO delay off, 64 events per pixel, seed 42.  The
# camera keeps the profile's field of view on a 48 x 36 grid (one sub-ray).
DAVIS_PROFILE = capi.Physic(threshold_pos=0.28, threshold_neg=0.28, threshold_std=0.03, threshold_min=0.01,
                            refractory_time=200e-6, photon_shot_noise=1, reset_thermal_noise=0,
                            readout_bandwidth=5e6, readout_latency=18e-9, max_events_per_step=16,
                            random_seed=114514)


def sensor_fixture(width=48, height=36):
    v = np.array([[-100, -100, 0], [100, -100, 0], [100, 100, 0], [-100, 100, 0]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    nodes, root, leaf_idx = capi.bvh_build(v, idx, 8)
    scale = width / 346.0
    cam = capi.Camera(width, height, 226.3802 * scale, 226.1500 * scale, 173.6471 * scale, 133.7327 * scale, 1)
    phys = dataclasses.replace(DAVIS_PROFILE, threshold_std=0.0, photon_shot_noise=0, max_events_per_step=64,
                               readout_bandwidth=0.0, readout_latency=0.0, random_seed=42)
    return v, n, nodes, root, leaf_idx, cam, phys


def interval(start, end, t0, t1):
    iv = np.zeros(1, capi.INTERVAL_DTYPE)
    iv["start"], iv["end"], iv["t0"], iv["t1"] = start, end, t0, t1
    return iv
