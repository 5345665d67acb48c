"""CPU: the oracle's DEM -> mesh and raster alignment (oracle/dvs_oracle.cpp,
orc_raster_mesh / orc_raster_align, restating src/tools/build_from_raster.cu)
pinned by

* the reference's own known answers (tests/test_raster.cu): the 4x4 DEM
  with a NoData corner gives 15 vertices and 16 triangles, finite vertices
  and normals, compact indices; a DEM with a NoData centre is rejected; the
  bilinear alignment of 2x + 3y over decreasing x / increasing y axes, NaN
  outside and next to NoData, 0 kept at a valid zero;
* closed forms: a planar DEM's normals are its unit plane normal, its
  vertices are the recentred grid;
* an independent float64 formulation: the angle-weighted normals of a full
  grid computed cell by cell with numpy (terrain.grid_normals, faces turned
  to +z), within float32 rounding;
* a planted mistake: a swapped triangle order fails the normals check.
"""
from __future__ import annotations

import numpy as np
import pytest

from dvsrender4deepspace_amd import terrain


def _reference_dem():
    d = (np.add.outer(np.arange(4), np.arange(4)) * 2.0 + 10).astype(np.float32)
    d[0, 0] = np.nan
    xs = (100 + np.arange(4) + 0.5).astype(np.float32)
    ys = (200 - np.arange(4) - 0.5).astype(np.float32)
    return d, xs, ys


def test_reference_mesh_counts(oracle):
    d, xs, ys = _reference_dem()
    v, n, t, _ = oracle.raster_mesh(d, xs, ys)
    assert len(v) == 15 and len(t) == 16
    assert np.isfinite(v).all() and np.isfinite(n).all() and t.max() < len(v)
    # recentred on the centre pixel (row 2, col 2): x 102.5, y 197.5, z 2*4+10
    assert np.array_equal(v[0], [101.5 - 102.5, 199.5 - 197.5, 12 - 18])


def test_nodata_centre_rejected(oracle):
    d, xs, ys = _reference_dem()
    d[2, 2] = np.nan
    assert oracle.raster_mesh(d, xs, ys) is None


def test_plane_normals_closed_form(oracle):
    h, w = 9, 11
    xs = (5000 + 2.0 * np.arange(w)).astype(np.float32)
    ys = (9000 - 3.0 * np.arange(h)).astype(np.float32)
    a, b = 0.3, -0.2
    d = (a * (xs[None, :] - 5000) + b * (ys[:, None] - 9000) + 100).astype(np.float32)
    v, n, t, _ = oracle.raster_mesh(d, xs, ys)
    want = np.array([-a, -b, 1.0]) / np.linalg.norm([-a, -b, 1.0])
    assert np.abs(n - want).max() < 2e-5
    gx, gy = np.meshgrid(xs - xs[w // 2], ys - ys[h // 2])
    assert np.array_equal(v[:, 0], gx.ravel().astype(np.float32)) and np.array_equal(v[:, 1], gy.ravel())
    assert len(t) == 2 * (w - 1) * (h - 1)


def _grid(spec):
    heights = terrain.heightfield(spec).astype(np.float32)
    n = spec.size
    axis = (np.arange(n) * spec.spacing).astype(np.float32)
    return heights, axis, axis


def test_full_grid_normals_match_float64_cells(oracle):
    """terrain.grid_normals takes triangles (v1, v3, v2), (v2, v3, v4) with
    v3 one row further in the grid's y: with y_coords increasing, the same
    cells and corner angles as the raster path."""
    spec = terrain.TerrainSpec(size=48, amplitude=15.0, craters=4, seed=3)
    heights, xs, ys = _grid(spec)
    v, n, t, _ = oracle.raster_mesh(heights, xs, ys)
    grid = np.empty((spec.size, spec.size, 3), np.float32)
    grid[..., 0] = xs[None, :]
    grid[..., 1] = ys[:, None]
    grid[..., 2] = heights
    want = terrain.grid_normals(grid)
    assert np.abs(n - want).max() < 3e-6


def test_planted_triangle_order_fails(oracle):
    """The cell split the other way ((v1, v3, v4), (v1, v4, v2)) changes the
    normals beyond the check above."""
    spec = terrain.TerrainSpec(size=32, amplitude=15.0, craters=3, seed=5)
    heights, xs, ys = _grid(spec)
    _, n, _, _ = oracle.raster_mesh(heights, xs, ys)
    h = heights.astype(np.float64)
    p = np.stack(np.broadcast_arrays(xs[None, :].astype(np.float64), ys[:, None].astype(np.float64), h), -1)
    acc = np.zeros_like(p)
    v1, v2, v3, v4 = p[:-1, :-1], p[:-1, 1:], p[1:, :-1], p[1:, 1:]
    face, (a1, a3, a4) = terrain._corner_weights(v1, v3, v4)
    acc[:-1, :-1] += face * a1
    acc[1:, :-1] += face * a3
    acc[1:, 1:] += face * a4
    face, (b1, b4, b2) = terrain._corner_weights(v1, v4, v2)
    acc[:-1, :-1] += face * b1
    acc[1:, 1:] += face * b4
    acc[:-1, 1:] += face * b2
    wrong = (acc / np.linalg.norm(acc, axis=-1, keepdims=True)).reshape(-1, 3)
    assert np.abs(n - wrong).max() > 1e-3


def test_alignment_known_answers(oracle):
    """tests/test_raster.cu:29-42."""
    x = np.array([2, 1, 0], np.float32)
    y = np.array([0, 1, 2], np.float32)
    rx = np.array([0, .5, 1.5, 2, -1], np.float32)
    ry = np.array([0, .5, 1.5, 2, 3], np.float32)
    source = (2 * x[None, :] + 3 * y[:, None]).astype(np.float32)
    out = oracle.raster_align(rx, ry, source, x, y)
    for row in range(5):
        for col in range(5):
            if row == 4 or col == 4:
                assert np.isnan(out[row, col])
            else:
                assert abs(out[row, col] - 2 * rx[col] - 3 * ry[row]) < 1e-6
    source.ravel()[4] = np.nan
    out = oracle.raster_align(rx, ry, source, x, y)
    assert np.isnan(out.ravel()[6]) and out.ravel()[0] == 0
