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
