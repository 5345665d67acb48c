"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared entry point, data layouts agree with the oracle, the host BVH
build equals the oracle's restatement of the reference build, argument
validation, the terrain generator and workload horizons."""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import pytest

from dvsrender4deepspace_amd import capi, terrain, workloads


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    names = capi.exported_symbols()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/dvs_hip.h but not exported: {missing}"


def test_io_library_exports_every_declared_symbol():
    from dvsrender4deepspace_amd import eventio
    lib = eventio.load_library()
    names = capi.exported_symbols("dvs_io.h")
    assert len(names) >= 9
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/dvs_io.h but not exported: {missing}"


def test_layouts_match_oracle(oracle):
    assert capi.NODE_DTYPE == oracle.NODE
    assert capi.STATE_DTYPE == oracle.STATE
    assert capi.INTERVAL_DTYPE == oracle.INTERVAL
    assert capi.EVENT_DTYPE == oracle.EVENT
    assert ctypes.sizeof(capi._PhysicC) == ctypes.sizeof(oracle.Physic)
    assert [f[0] for f in capi._PhysicC._fields_] == [f[0] for f in oracle.Physic._fields_]
    assert [f[0] for f in capi._HapkeC._fields_] == [f[0] for f in oracle.Hapke._fields_]


@pytest.mark.parametrize("leaf", [1, 4, 8, 64])
@pytest.mark.parametrize("seed,size", [(101, 128), (7, 96)])
def test_host_bvh_equals_oracle_build(oracle, leaf, seed, size):
    spec = terrain.TerrainSpec(size=size, craters=6, seed=seed)
    v, n, idx = terrain.mesh(spec)
    on, oroot, oidx = oracle.bvh_build(v, idx, leaf)
    pn, proot, pidx = capi.bvh_build(v, idx, leaf)
    assert oroot == proot == len(pn) - 1
    assert on.tobytes() == pn.tobytes(), "node arrays differ"
    assert np.array_equal(oidx, pidx), "leaf-ordered triangles differ"
    capi.bvh_validate(pn, proot)


def test_host_bvh_sah_costs(oracle):
    """Non-default SAH costs change the tree identically in both builds."""
    v, n, idx = terrain.mesh(terrain.TerrainSpec(size=64, craters=2, seed=3))
    on, oroot, _ = oracle.bvh_build(v, idx, 4, 2.5, 0.75)
    pn, proot, _ = capi.bvh_build(v, idx, 4, 2.5, 0.75)
    assert on.tobytes() == pn.tobytes() and oroot == proot


def test_bvh_degenerate_inputs(oracle):
    # all triangles share one centroid: binning cannot split; median fallback or leaf
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.tile(np.array([[0, 1, 2]], np.uint32), (40, 1))
    on, oroot, oidx = oracle.bvh_build(v, idx, 2)
    pn, proot, pidx = capi.bvh_build(v, idx, 2)
    assert on.tobytes() == pn.tobytes() and oroot == proot
    # a single triangle is a root leaf
    nodes, root, _ = capi.bvh_build(v, idx[:1], 8)
    assert len(nodes) == 1 and root == 0 and nodes[0]["count"] == 1


def test_bvh_validation_errors():
    v, n, idx = terrain.mesh(terrain.TerrainSpec(size=16, craters=0))
    nodes, root, _ = capi.bvh_build(v, idx, 4)
    with pytest.raises(capi.InvalidArgument, match="root index out of range"):
        capi.bvh_validate(nodes, len(nodes))
    bad = nodes.copy()
    inner = np.flatnonzero(bad["count"] == 0)[0]
    bad[inner]["left"] = inner
    with pytest.raises(capi.InvalidArgument, match="cycle"):
        capi.bvh_validate(bad, root)
    bad = nodes.copy()
    bad[inner]["right"] = len(nodes) + 5
    with pytest.raises(capi.InvalidArgument, match="out of range"):
        capi.bvh_validate(bad, root)
    with pytest.raises(capi.InvalidArgument, match="outside vertex array"):
        capi.bvh_build(v[:10], idx, 4)


def test_terrain_generator_properties():
    spec = terrain.TerrainSpec(size=128, amplitude=20.0, craters=10, seed=101)
    v, n, idx = terrain.mesh(spec)
    assert len(idx) == spec.triangles == 32258
    assert len(v) == 128 * 128
    assert np.allclose(v[:, 0].min(), -63.5) and np.allclose(v[:, 1].max(), 63.5)
    assert np.all(n[:, 2] > 0) and np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    assert np.abs(v[:, 2]).max() < 2 * spec.amplitude
    # deterministic per seed, different across seeds
    v2, _, _ = terrain.mesh(spec)
    v3, _, _ = terrain.mesh(dataclasses.replace(spec, seed=102))
    assert np.array_equal(v, v2) and not np.array_equal(v, v3)
    # winding of the raster convention: T1 = (v1, v3, v2), T2 = (v2, v3, v4)
    assert list(idx[0]) == [0, 128, 1] and list(idx[1]) == [1, 128, 129]


def test_craters_are_parabolic_bowls():
    flat = terrain.TerrainSpec(size=64, amplitude=0.0, craters=1, seed=5, crater_radius=(10.0, 10.0))
    h = terrain.heightfield(flat)
    assert h.min() == pytest.approx(-0.2 * 10.0, abs=0.05)
    assert np.count_nonzero(h < 0) == pytest.approx(math.pi * 100, rel=0.1)


def test_workload_horizons_follow_driver_loop():
    w = workloads.PLAIN_SMALL
    res = float(np.float32(w.time_resolution))
    assert w.total_steps == math.ceil(float(np.float32(0.49)) / res)
    iv = w.intervals(0, 64)
    assert iv["t0"][0] == 0.0 and np.all(iv["t0"][1:] == iv["t1"][:-1])
    assert iv["t1"][5] == 6 * res
    nxt = w.intervals(64, 3)
    assert nxt["t0"][0] == iv["t1"][-1]
    last = w.intervals(w.total_steps - 1, 5)
    assert len(last) == 1 and last["t1"][-1] == float(np.float32(w.end_time))
    s = w.states(np.array([0.0, 0.1]))
    assert np.allclose(np.linalg.norm(s["rotation"], axis=1), 1, atol=1e-6)
    assert s["position"][1][0] - s["position"][0][0] == pytest.approx(w.speed * 0.1, rel=1e-5)


def test_error_codes_without_gpu_paths():
    """Argument errors are reported before any device work."""
    with pytest.raises(capi.InvalidArgument):
        capi.bvh_build(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))


def test_expected_window_argument_errors_without_gpu_paths():
    """dvs_sensor_expect_adaptive_times refuses a missing sensor and a
    negative length before any device work (error 3, invalid argument)."""
    lib = capi.load_library()
    times = np.array([0.0, 1e-5, 2e-5])
    ptr = times.ctypes.data_as(ctypes.c_void_p)
    assert lib.dvs_sensor_expect_adaptive_times(None, None, ptr, 3, 1e-2) == 3
    assert "null sensor" in lib.dvs_last_error().decode()
    assert lib.dvs_sensor_expect_adaptive_times(None, None, ptr, -1, 1e-2) == 3
