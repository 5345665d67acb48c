"""Seeded synthetic workloads of the DVS path: scene, camera, flight, sensor.

Each workload fixes a terrain (terrain.TerrainSpec), a pinhole camera, a
straight-pass flight and the physic/Hapke parameters.  Trajectory states are
evaluated in float32 like CameraTrajectory::evaluate(float)
(include/dvs/trajectory.hpp:100) and host horizons follow the driver loop of
apps/main_dvs.cpp:79-108: t_k = min(end, start + k * double(time_resolution)).

Light intensity is 1000 (the reference's own sensor fixtures use 1000,
tests/test_sensor.cu:28,72) so that photocurrents are ~1e4-1e5 e-/s and the
receptor time constant (charge scale / current) is milliseconds: with
intensity 1.0 the 0.5 s passes would emit almost no events.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .capi import FRAME_DTYPE, INTERVAL_DTYPE, STATE_DTYPE, Camera, Hapke, Physic
from .terrain import TerrainSpec

# Hapke parameters of the workloads: w=0.3, B0=1.0, h=0.06, theta-bar=20 deg,
# b=0.25, c=0.3; hC, B_C0 and phi keep the reference defaults.
WORKLOAD_HAPKE = Hapke(b=0.25, c=0.3, hS=0.06, hC=1.0, thetap=0.349066, w=0.3, B_S0=1.0, B_C0=0.0, phi=0.3)


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    camera: Camera
    terrain: TerrainSpec
    frames: int
    fps: float
    altitude: float = 200.0
    altitude_end: float | None = None  # linear descent to this altitude (None: level)
    speed: float = 50.0
    yaw_rate_deg: float = 1.0
    roll_rate_deg: float = 0.0
    sun_elevation_deg: float = 30.0
    intensity: float = 1000.0
    time_resolution: float = 1e-5
    physic: Physic = dataclasses.field(default_factory=Physic)
    hapke: Hapke = WORKLOAD_HAPKE
    leaf_size: int = 64  # BVHBuildParams max_leaf_size used by bin/dvs and bin/cam

    @property
    def start_time(self) -> float:
        return 0.0

    @property
    def end_time(self) -> float:
        # frame_time of the last frame, as float (apps/main_dvs.cpp:79)
        return float(np.float32((self.frames - 1) / self.fps))

    @property
    def total_steps(self) -> int:
        res = float(np.float32(self.time_resolution))
        return int(math.ceil((self.end_time - self.start_time) / res))

    def states(self, times) -> np.ndarray:
        """Trajectory states at float32 times: straight pass along +x centred on
        the terrain at constant speed, level or descending linearly, with the
        camera (nadir-looking at t=0) yawing and rolling at constant rates."""
        t = np.asarray(times, dtype=np.float32).astype(np.float64)
        duration = max(self.end_time - self.start_time, 1e-30)
        s = t - self.start_time
        out = np.zeros(t.shape, STATE_DTYPE)
        out["position"][..., 0] = -0.5 * self.speed * duration + self.speed * s
        climb = 0.0 if self.altitude_end is None else (self.altitude_end - self.altitude) / duration
        out["position"][..., 2] = self.altitude + climb * s
        out["velocity"][..., 0] = self.speed
        out["velocity"][..., 2] = climb
        half_yaw = 0.5 * np.deg2rad(self.yaw_rate_deg) * s
        half_roll = 0.5 * np.deg2rad(self.roll_rate_deg) * s
        # q = q_yaw(z) * q_roll(x), quaternion (x, y, z, w)
        cy, sy, cr, sr = np.cos(half_yaw), np.sin(half_yaw), np.cos(half_roll), np.sin(half_roll)
        out["rotation"][..., 0] = cy * sr
        out["rotation"][..., 1] = sy * sr
        out["rotation"][..., 2] = sy * cr
        out["rotation"][..., 3] = cy * cr
        el = np.deg2rad(self.sun_elevation_deg)
        out["light"][..., 0] = np.cos(el)
        out["light"][..., 2] = np.sin(el)
        out["intensity"] = self.intensity
        return out

    def horizon_times(self, first_step: int, count: int) -> np.ndarray:
        """Host horizon ends t_k for k = first_step+1 .. first_step+count."""
        res = float(np.float32(self.time_resolution))
        k = np.arange(first_step + 1, first_step + count + 1, dtype=np.float64)
        return np.minimum(float(np.float32(self.end_time)), self.start_time + k * res)

    def intervals(self, first_step: int, count: int) -> np.ndarray:
        """OpticalIntervals for horizons [first_step, first_step + count)."""
        res = float(np.float32(self.time_resolution))
        t0 = self.start_time + first_step * res if first_step else self.start_time
        ends = self.horizon_times(first_step, count)
        starts = np.concatenate([[t0], ends[:-1]])
        keep = ends > starts
        ends, starts = ends[keep], starts[keep]
        iv = np.zeros(len(ends), INTERVAL_DTYPE)
        iv["start"] = self.states(starts)
        iv["end"] = self.states(ends)
        iv["t0"] = starts
        iv["t1"] = ends
        return iv

    def initial_state(self) -> np.ndarray:
        return self.states(np.array([self.start_time]))


def _noise_off(**kw) -> Physic:
    base = dict(threshold_pos=0.2, threshold_neg=0.2, threshold_std=0.0, photon_shot_noise=0,
                reset_thermal_noise=0, refractory_time=200e-6, readout_bandwidth=5e6, readout_latency=18e-9,
                circuit_step=1e-6, max_events_per_step=16, random_seed=114514)
    base.update(kw)
    return Physic(**base)


def _noise_on(**kw) -> Physic:
    base = dict(threshold_pos=0.2, threshold_neg=0.2, threshold_std=0.03, photon_shot_noise=1,
                reset_thermal_noise=1, refractory_time=200e-6, readout_bandwidth=5e6, readout_latency=18e-9,
                circuit_step=1e-6, max_events_per_step=16, random_seed=20260101)
    base.update(kw)
    return Physic(**base)


# Plain method, small: 64x48, fx=fy=60, 128^2 terrain (32,258 triangles),
# 10 craters, 50 frames at 100 fps, noise off, supersample 1.
PLAIN_SMALL = Workload(
    name="plain-64x48-terrain128",
    camera=Camera(64, 48, 60.0, 60.0, 32.0, 24.0, 1),
    terrain=TerrainSpec(size=128, amplitude=20.0, craters=10, seed=101),
    frames=50, fps=100.0, physic=_noise_off())

# Single DAVIS346 request: 346x260, fx=fy=300, 1024^2 terrain (2,093,058
# triangles), 200 craters, 250 frames at 100 fps, supersample 4, noise on.
DAVIS346 = Workload(
    name="davis346-terrain1024-ss4",
    camera=Camera(346, 260, 300.0, 300.0, 173.0, 130.0, 4),
    terrain=TerrainSpec(size=1024, amplitude=20.0, craters=200, seed=202),
    frames=250, fps=100.0, physic=_noise_on())

# Same request with noise off (exact-parity form).
DAVIS346_NOISE_OFF = dataclasses.replace(DAVIS346, name="davis346-terrain1024-ss4-noiseoff", physic=_noise_off())

# Single high-resolution request: 1280x720, fx=fy=1000, 2048^2 terrain
# (8,380,418 triangles), 800 craters, amplitude 40 m, 1 kHz horizons.
HIGHRES = Workload(
    name="hd1280x720-terrain2048-ss4",
    camera=Camera(1280, 720, 1000.0, 1000.0, 640.0, 360.0, 4),
    terrain=TerrainSpec(size=2048, amplitude=40.0, craters=800, seed=303),
    frames=1001, fps=1000.0, altitude=500.0, altitude_end=100.0, speed=20.0, yaw_rate_deg=0.0,
    roll_rate_deg=5.0, physic=_noise_on(), time_resolution=1e-3)

WORKLOADS = {w.name: w for w in (PLAIN_SMALL, DAVIS346, DAVIS346_NOISE_OFF, HIGHRES)}


def small_variant(w: Workload, size: int = 64, width: int = 32, height: int = 24, samples: int | None = None,
                  craters: int = 4) -> Workload:
    """Shrunken copy of a workload for fast parity tests (same generator)."""
    cam = w.camera
    scale = width / cam.width
    return dataclasses.replace(
        w, name=f"{w.name}-small{size}",
        camera=Camera(width, height, cam.fx * scale, cam.fy * scale, width / 2.0, height / 2.0,
                      cam.samples if samples is None else samples),
        terrain=dataclasses.replace(w.terrain, size=size, craters=craters),
        altitude=min(w.altitude, 0.9 * (size - 1) * w.terrain.spacing * cam.fx / max(cam.width, 1)))


def quaternion_matrices(q: np.ndarray) -> np.ndarray:
    """Row-major rotation matrices (..., 3, 3) of unit quaternions (x, y, z, w)."""
    q = np.asarray(q, np.float64)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def flight_frames(w: Workload, count: int | None = None) -> np.ndarray:
    """Trajectory frames of the workload's flight at its fps (FRAME_DTYPE):
    frame k at t = k / fps with the flight's position, orientation (the
    rotation matrix of its quaternion), light and intensity."""
    n = w.frames if count is None else count
    t = (np.arange(n) / w.fps).astype(np.float32)
    s = w.states(t)
    f = np.zeros(n, FRAME_DTYPE)
    f["frame_index"] = np.arange(n)
    f["timestamp_seconds"] = t.astype(np.float64)
    f["camera_origin"] = s["position"]
    f["camera_world"] = np.transpose(quaternion_matrices(s["rotation"]), (0, 2, 1)).reshape(n, 9)  # column major
    f["light_dir"] = s["light"]
    f["light_intensity"] = s["intensity"]
    return f


def trajectory(w: Workload) -> "Trajectory":
    """CameraTrajectory over the workload's frames (src/trajectory.cpp)."""
    from .capi import Trajectory
    return Trajectory(flight_frames(w), w.fps)


def horizon_grid(w: Workload, start_time: float, first_step: int, count: int) -> np.ndarray:
    """`start_time` (the end of the previous horizon) and the ends of host
    horizons first_step+1 .. first_step+count, a horizon that does not
    advance dropped (apps/main_dvs.cpp:100-107)."""
    ends = w.horizon_times(first_step, count)
    times = np.concatenate([[start_time], ends])
    keep = np.concatenate([[True], np.diff(times) > 0])
    return times[keep]


def horizon_intervals(w: Workload, traj, start_time: float, first_step: int, count: int) -> np.ndarray:
    """The OpticalIntervals of host horizons first_step+1 .. first_step+count
    flown along `traj`, from `start_time`: bin/dvs's window, each state
    evaluated once on the device."""
    return traj.intervals(horizon_grid(w, start_time, first_step, count))


def write_bundle(w: Workload, directory, frames: int | None = None) -> None:
    """A scene bundle in the reference's format (src/scene_data.cpp:93-236):
    mesh_vertices.bin / mesh_normals.bin (float32 N x 3), mesh_indices.bin
    (uint32 M x 3) of the workload's terrain, and frames.bin with 24 floats
    per frame {t, origin[3], camera_world as a row-major 4x4, light_dir[3],
    intensity} sampled from the workload's flight at its fps."""
    from pathlib import Path

    from . import terrain as terrain_mod

    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    v, n, idx = terrain_mod.mesh(w.terrain)
    np.ascontiguousarray(v, np.float32).tofile(d / "mesh_vertices.bin")
    np.ascontiguousarray(n, np.float32).tofile(d / "mesh_normals.bin")
    np.ascontiguousarray(idx, np.uint32).tofile(d / "mesh_indices.bin")
    f = flight_frames(w, frames)
    count = len(f)
    rec = np.zeros((count, 24), np.float32)
    rec[:, 0] = f["timestamp_seconds"]
    rec[:, 1:4] = f["camera_origin"]
    m = np.zeros((count, 4, 4))
    m[:, :3, :3] = np.transpose(f["camera_world"].reshape(count, 3, 3), (0, 2, 1))
    m[:, :3, 3] = f["camera_origin"]
    m[:, 3, 3] = 1
    rec[:, 4:20] = m.reshape(count, 16)
    rec[:, 20:23] = f["light_dir"]
    rec[:, 23] = f["light_intensity"]
    rec.tofile(d / "frames.bin")


def write_text_trajectory(w: Workload, path, frames: int | None = None) -> None:
    """The workload's flight as a --traj text file (load_frames_from_text,
    src/scene_data.cpp:238-283): "frame_idx x y z qx qy qz qw lx ly lz L"
    per frame at t = frame_idx / fps."""
    from pathlib import Path

    n = w.frames if frames is None else frames
    s = w.states((np.arange(n) / w.fps).astype(np.float32))
    lines = []
    for k in range(n):
        vals = [*s["position"][k], *s["rotation"][k], *s["light"][k], s["intensity"][k]]
        lines.append(f"{k} " + " ".join(repr(float(v)) for v in vals))
    Path(path).write_text("\n".join(lines) + "\n")


def write_config(w: Workload, path, frame_count: int | None = None, imu_fps: float = 200.0, cam_fps: float = 30.0,
                 physic: Physic | None = None, blur_resolution: float = 1e-3) -> None:
    """The workload as a configuration file (Config's grammar,
    src/scene_types.cpp): intrinsics, every physic field, scene timing at
    the workload's fps and time resolution, and its Hapke parameters."""
    from pathlib import Path

    p = physic or w.physic
    cam, h = w.camera, w.hapke
    lines = ["intrinsic:"] + [f"  {k}: {getattr(cam, k)!r}" for k in ("width", "height", "fx", "fy", "cx", "cy")]
    lines += ["", "physic:"] + [f"  {f.name}: {getattr(p, f.name)!r}" for f in dataclasses.fields(p)]
    lines += ["", "scene:", f"  frame_count: {w.frames if frame_count is None else frame_count}",
              f"  fps: {w.fps!r}", f"  imu_fps: {imu_fps!r}", f"  cam_fps: {cam_fps!r}",
              f"  time_resolution: {w.time_resolution!r}", f"  supersample_factor: {cam.samples}",
              f"  blur_resolution: {blur_resolution!r}"]
    lines += ["", "hapke:"] + [f"  {f.name}: {getattr(h, f.name)!r}" for f in dataclasses.fields(h)]
    Path(path).write_text("\n".join(lines) + "\n")
