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
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = x - ix, y - iy
    sx = fx * fx * fx * (fx * (fx * 6 - 15) + 10)  # quintic fade
    sy = fy * fy * fy * (fy * (fy * 6 - 15) + 10)
    a = _lattice(seed, octave, ix, iy)
    b = _lattice(seed, octave, ix + 1, iy)
    c = _lattice(seed, octave, ix, iy + 1)
    d = _lattice(seed, octave, ix + 1, iy + 1)
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
    x, y = np.meshgrid(coords, coords)  # x = col, y = row
    h = np.zeros((n, n))
    gain, norm = 1.0, 0.0
    freq = 1.0 / spec.base_wavelength
    for o in range(spec.octaves):
        h += gain * _value_noise(spec.seed, o, x * freq, y * freq)
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
        wx, wy = x[r0:r1, c0:c1], y[r0:r1, c0:c1]
        d2 = ((wx - cx) ** 2 + (wy - cy) ** 2) / (r * r)
        h[r0:r1, c0:c1] -= np.where(d2 < 1.0, spec.depth_ratio * r * (1.0 - d2), 0.0)
    return h


def mesh(spec: TerrainSpec, heights: np.ndarray | None = None):
    """Indexed mesh of the heightfield: vertices (N,3) f32, normals (N,3) f32, indices (T,3) u32."""
    h = heightfield(spec) if heights is None else heights
    n = spec.size
    half = (n - 1) * spec.spacing / 2.0
    cols, rows = np.meshgrid(np.arange(n), np.arange(n))
    verts = np.stack([cols * spec.spacing - half, rows * spec.spacing - half, h], axis=-1).reshape(-1, 3)
    verts = verts.astype(np.float32)
    c, r = np.meshgrid(np.arange(n - 1), np.arange(n - 1))
    v1 = (r * n + c).ravel()
    v2 = v1 + 1
    v3 = v1 + n
    v4 = v3 + 1
    tris = np.empty((2 * len(v1), 3), np.uint32)
    tris[0::2] = np.stack([v1, v3, v2], axis=1)
    tris[1::2] = np.stack([v2, v3, v4], axis=1)
    return verts, vertex_normals(verts, tris), tris


def vertex_normals(verts: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """Angle-weighted vertex normals, faces oriented to +Z."""
    p = verts.astype(np.float64)[tris]  # (T, 3 corners, 3)
    face = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    face *= np.where(face[:, 2:3] < 0, -1.0, 1.0)
    length = np.linalg.norm(face, axis=1, keepdims=True)
    face = np.divide(face, length, out=np.zeros_like(face), where=length > 0)
    acc = np.zeros((len(verts), 3))
    for k in range(3):
        e1 = p[:, (k + 1) % 3] - p[:, k]
        e2 = p[:, (k + 2) % 3] - p[:, k]
        cosang = np.einsum("ij,ij->i", e1, e2) / np.maximum(
            np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1), 1e-30)
        ang = np.arccos(np.clip(cosang, -1.0, 1.0))
        w = face * ang[:, None]
        for a in range(3):
            acc[:, a] += np.bincount(tris[:, k], weights=w[:, a], minlength=len(verts))
    length = np.linalg.norm(acc, axis=1, keepdims=True)
    normals = np.divide(acc, length, out=np.zeros_like(acc), where=length > 1e-12)
    return normals.astype(np.float32)
