"""Seeded synthetic planetary terrain: fBm heightfield plus parabolic craters.

The reference ingests terrain from Blender bundles or GDAL rasters
(src/scene_data.cpp, src/tools/build_from_raster.cu); neither is available
here, so workloads are generated.  The mesh follows the raster-to-mesh
conventions of src/tools/build_from_raster.cu:333-470: vertex (col, row) at
the grid point, two triangles per cell (v1, v3, v2) and (v2, v3, v4) with
v1=(c,r), v2=(c+1,r), v3=(c,r+1), v4=(c+1,r+1), angle-weighted vertex normals
oriented to +Z, coordinates recentred on the grid centre.
"""
from __future__ import annotations

import dataclasses

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK
    return x ^ (x >> np.uint64(31))


def _lattice(seed: int, octave: int, ix: np.ndarray, iy: np.ndarray) -> np.ndarray:
    """Uniform [-1, 1) value at integer lattice points."""
    with np.errstate(over="ignore"):
        h = _mix(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(octave))
        h = _mix(h ^ (ix.astype(np.int64).astype(np.uint64) * np.uint64(0xD6E8FEB86659FD93)))
        h = _mix(h ^ (iy.astype(np.int64).astype(np.uint64) * np.uint64(0xA0761D6478BD642F)))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def _value_noise(seed: int, octave: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Value noise on the grid x (columns) by y (rows), both 1-D: the lattice
    is hashed once per lattice point and the quintic fades once per column
    and per row (the grid coordinates are separable)."""
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = x - ix, y - iy
    sx = (fx * fx * fx * (fx * (fx * 6 - 15) + 10))[None, :]  # quintic fade
    sy = (fy * fy * fy * (fy * (fy * 6 - 15) + 10))[:, None]
    x0, y0 = int(ix.min()), int(iy.min())
    lx = np.arange(x0, int(ix.max()) + 2, dtype=np.float64)
    ly = np.arange(y0, int(iy.max()) + 2, dtype=np.float64)
    table = _lattice(seed, octave, lx[None, :], ly[:, None])  # [row, col] over the lattice
    cx = (ix - x0).astype(np.int64)
    cy = (iy - y0).astype(np.int64)
    top, bottom = table[cy], table[cy + 1]
    a, b = top[:, cx], top[:, cx + 1]
    c, d = bottom[:, cx], bottom[:, cx + 1]
    return (a + (b - a) * sx) * (1 - sy) + (c + (d - c) * sx) * sy


@dataclasses.dataclass(frozen=True)
class TerrainSpec:
    size: int = 128                 # grid points per side
    spacing: float = 1.0            # metres between grid points
    amplitude: float = 20.0         # fBm amplitude, metres
    octaves: int = 6
    hurst: float = 0.8              # per-octave gain 2^-H, lacunarity 2
    base_wavelength: float = 64.0   # metres, first octave
    craters: int = 10
    crater_radius: tuple = (3.0, 15.0)
    depth_ratio: float = 0.2        # crater depth / radius
    seed: int = 1

    @property
    def triangles(self) -> int:
        return 2 * (self.size - 1) ** 2


def heightfield(spec: TerrainSpec) -> np.ndarray:
    """(size, size) heights [row, col] in metres, float64."""
    n = spec.size
    coords = np.arange(n, dtype=np.float64) * spec.spacing
    h = np.zeros((n, n))
    gain, norm = 1.0, 0.0
    freq = 1.0 / spec.base_wavelength
    for o in range(spec.octaves):
        h += gain * _value_noise(spec.seed, o, coords * freq, coords * freq)
        norm += gain
        gain *= 2.0 ** -spec.hurst
        freq *= 2.0
    h *= spec.amplitude / norm
    rng = np.random.default_rng(np.random.SeedSequence([spec.seed, 0xC4A7E5]))
    extent = (n - 1) * spec.spacing
    for _ in range(spec.craters):
        r = rng.uniform(*spec.crater_radius)
        cx, cy = rng.uniform(r, extent - r, size=2) if extent > 2 * r else (extent / 2, extent / 2)
        # parabolic bowl of depth depth_ratio * r, applied on its bounding window
        c0, c1 = max(int(np.floor((cx - r) / spec.spacing)), 0), min(int(np.ceil((cx + r) / spec.spacing)) + 1, n)
        r0, r1 = max(int(np.floor((cy - r) / spec.spacing)), 0), min(int(np.ceil((cy + r) / spec.spacing)) + 1, n)
        wx, wy = coords[None, c0:c1], coords[r0:r1, None]
        d2 = ((wx - cx) ** 2 + (wy - cy) ** 2) / (r * r)
        h[r0:r1, c0:c1] -= np.where(d2 < 1.0, spec.depth_ratio * r * (1.0 - d2), 0.0)
    return h


def mesh(spec: TerrainSpec, heights: np.ndarray | None = None):
    """Indexed mesh of the heightfield: vertices (N,3) f32, normals (N,3) f32, indices (T,3) u32."""
    h = heightfield(spec) if heights is None else heights
    n = spec.size
    half = (n - 1) * spec.spacing / 2.0
    axis = np.arange(n) * spec.spacing - half
    grid = np.empty((n, n, 3), np.float32)
    grid[..., 0] = axis[None, :]
    grid[..., 1] = axis[:, None]
    grid[..., 2] = h
    c, r = np.meshgrid(np.arange(n - 1, dtype=np.uint32), np.arange(n - 1, dtype=np.uint32))
    v1 = (r * np.uint32(n) + c).ravel()
    v2 = v1 + 1
    v3 = v1 + n
    v4 = v3 + 1
    tris = np.empty((2 * len(v1), 3), np.uint32)
    tris[0::2] = np.stack([v1, v3, v2], axis=1)
    tris[1::2] = np.stack([v2, v3, v4], axis=1)
    return grid.reshape(-1, 3), grid_normals(grid), tris


def _corner_weights(pa, pb, pc):
    """Unit face normal (oriented to +Z) of triangle (pa, pb, pc), float64
    arrays (..., 3), and its angle at each corner."""
    face = np.cross(pb - pa, pc - pa)
    face *= np.where(face[..., 2:3] < 0, -1.0, 1.0)
    length = np.linalg.norm(face, axis=-1, keepdims=True)
    face = np.divide(face, length, out=np.zeros_like(face), where=length > 0)
    corners = (pa, pb, pc)
    angles = []
    for k in range(3):
        e1 = corners[(k + 1) % 3] - corners[k]
        e2 = corners[(k + 2) % 3] - corners[k]
        cosang = np.einsum("...i,...i->...", e1, e2) / np.maximum(
            np.linalg.norm(e1, axis=-1) * np.linalg.norm(e2, axis=-1), 1e-30)
        angles.append(np.arccos(np.clip(cosang, -1.0, 1.0))[..., None])
    return face, angles


def grid_normals(grid: np.ndarray) -> np.ndarray:
    """Angle-weighted vertex normals of the grid mesh (faces oriented to +Z):
    each vertex sums face normal x corner angle over the (up to six)
    triangles around it, cell by cell over the whole grid at once."""
    p = grid.astype(np.float64)
    v1, v2, v3, v4 = p[:-1, :-1], p[:-1, 1:], p[1:, :-1], p[1:, 1:]
    acc = np.zeros_like(p)
    face, (a1, a3, a2) = _corner_weights(v1, v3, v2)  # T1 = (v1, v3, v2)
    acc[:-1, :-1] += face * a1
    acc[1:, :-1] += face * a3
    acc[:-1, 1:] += face * a2
    face, (b2, b3, b4) = _corner_weights(v2, v3, v4)  # T2 = (v2, v3, v4)
    acc[:-1, 1:] += face * b2
    acc[1:, :-1] += face * b3
    acc[1:, 1:] += face * b4
    acc = acc.reshape(-1, 3)
    length = np.linalg.norm(acc, axis=1, keepdims=True)
    normals = np.divide(acc, length, out=np.zeros_like(acc), where=length > 1e-12)
    return normals.astype(np.float32)


def vertex_normals(verts: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """Angle-weighted vertex normals of an indexed mesh, faces oriented to +Z."""
    p = verts.astype(np.float64)[tris]  # (T, 3 corners, 3)
    face, angles = _corner_weights(p[:, 0], p[:, 1], p[:, 2])
    acc = np.zeros((len(verts), 3))
    for k in range(3):
        w = face * angles[k]
        for a in range(3):
            acc[:, a] += np.bincount(tris[:, k], weights=w[:, a], minlength=len(verts))
    length = np.linalg.norm(acc, axis=1, keepdims=True)
    normals = np.divide(acc, length, out=np.zeros_like(acc), where=length > 1e-12)
    return normals.astype(np.float32)
