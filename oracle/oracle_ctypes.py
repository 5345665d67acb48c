"""ctypes binding of the CPU oracle (oracle/dvs_oracle.h) — TEST INFRASTRUCTURE.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  Structured dtypes are declared here independently of
the product package; tests assert the layouts agree.
"""
from __future__ import annotations

import ctypes
import dataclasses
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "_build" / "liboracle_dvs.so"

NODE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,)), ("left", "<i4"), ("right", "<i4"),
                 ("start", "<i4"), ("count", "<i4")])
STATE = np.dtype([("position", "<f4", (3,)), ("velocity", "<f4", (3,)), ("light", "<f4", (3,)),
                  ("light_rate", "<f4", (3,)), ("rotation", "<f4", (4,)), ("intensity", "<f4"),
                  ("intensity_rate", "<f4")])
INTERVAL = np.dtype([("start", STATE), ("end", STATE), ("t0", "<f8"), ("t1", "<f8")])
EVENT = np.dtype([("timestamp", "<f8"), ("x", "<i2"), ("y", "<i2"), ("polarity", "i1"), ("depth", "<f4")])
PIXEL_STATE = np.dtype([("receptor_log", "<f8"), ("reference_log", "<f8"), ("refractory_until", "<f8"),
                        ("threshold_pos", "<f8"), ("threshold_neg", "<f8"), ("reset_pending", "<i4"),
                        ("pad", "<u4")])
PIXEL_EVENT = np.dtype([("time", "<f8"), ("polarity", "<i4"), ("pad", "<i4")])
FRAME = np.dtype([("frame_index", "<i4"), ("timestamp_seconds", "<f8"), ("camera_origin", "<f4", (3,)),
                  ("camera_world", "<f4", (9,)), ("light_dir", "<f4", (3,)), ("light_intensity", "<f4")], align=True)
TRAJECTORY_STATE = np.dtype([("timestamp", "<f4"), ("position", "<f4", (3,)), ("orientation", "<f4", (9,)),
                             ("quaternion", "<f4", (4,)), ("linear_velocity", "<f4", (3,)),
                             ("angular_velocity", "<f4", (3,)), ("light_direction", "<f4", (3,)),
                             ("light_direction_rate", "<f4", (3,)), ("light_intensity", "<f4"),
                             ("light_intensity_rate", "<f4")])


class RefinementStats(ctypes.Structure):
    _fields_ = [("traced_pixel_samples", ctypes.c_uint64), ("fallback_pixels", ctypes.c_int32),
                ("maximum_knots", ctypes.c_int32), ("exact_segments", ctypes.c_int32),
                ("native_cell_queries", ctypes.c_int32)]


class ReceptorProbe(ctypes.Structure):
    _fields_ = [("maximum_log_voltage_error", ctypes.c_double), ("p99_log_voltage_error", ctypes.c_double),
                ("mean_log_voltage_error", ctypes.c_double), ("pixels", ctypes.c_int32),
                ("circuit_cells", ctypes.c_int32)]


class Hapke(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("b", "c", "hS", "hC", "thetap", "w", "B_S0", "B_C0", "phi")]


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("samples", ctypes.c_int32)]


_PHYSIC_FIELDS = ["threshold_pos", "threshold_neg", "refractory_time", "readout_bandwidth", "readout_latency",
                  "threshold_std", "threshold_min", "wavelength_nm", "quantum_efficiency", "pixel_pitch_um",
                  "fill_factor", "f_number", "lens_transmission", "radiance_scale", "temperature_k",
                  "subthreshold_kappa", "receptor_capacitance_f", "reset_capacitance_f", "dark_current_a",
                  "reset_leakage_a", "circuit_step"]
_PHYSIC_INTS = ["photon_shot_noise", "reset_thermal_noise", "random_seed", "max_events_per_step",
                "readout_queue_capacity"]


class Physic(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in _PHYSIC_FIELDS] + [(n, ctypes.c_int32) for n in _PHYSIC_INTS]


class PixelParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("threshold_pos", "threshold_neg", "charge_scale", "refractory_time",
                                                "reference_leak", "reset_noise_std")]


class PhotoParams(ctypes.Structure):
    _fields_ = [("electrons_per_radiance", ctypes.c_double), ("dark_electrons", ctypes.c_double),
                ("integration_step", ctypes.c_double), ("shot_noise", ctypes.c_int32), ("pad", ctypes.c_int32)]


_P = ctypes.c_void_p
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIBRARY_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIBRARY_PATH))
        I, I64, D, F = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_float
        sig = {
            "orc_bvh_build": (I, [I, F, F, _P, I, _P, I, _P, ctypes.POINTER(I)]),
            "orc_bvh_validate": (I, [_P, I, I]),
            "orc_scene_create": (_P, [_P, I, I, _P, _P, I, _P, I, _P, Camera, Hapke]),
            "orc_scene_destroy": (None, [_P]),
            "orc_trace_nearest": (I, [_P, _P, _P, _P, _P, _P, _P]),
            "orc_trace_nearest_counted": (I, [_P, _P, _P, _P]),
            "orc_trace_shadow": (I, [_P, _P, _P, I, F, F]),
            "orc_shadow_origin": (None, [_P, I, F, F, _P, _P]),
            "orc_sample_pixel": (I, [_P, I, _P, _P, _P]),
            "orc_optical_state_at": (None, [_P, D, _P]),
            "orc_render": (I, [_P, _P, _P, _P, _P, I]),
            "orc_hapke_reflectance": (F, [F, F, F, F, Hapke]),
            "orc_hapke_directional": (F, [_P, _P, _P, F, Hapke]),
            "orc_invert_hapke_if": (F, [F, F, F, F, Hapke]),
            "orc_receptor_log_at": (D, [D, D, D, D]),
            "orc_advance_pixel": (I, [_P, ctypes.POINTER(PixelParams), D, D, D, _P, I, D]),
            "orc_parameters": (None, [ctypes.POINTER(Physic), I, ctypes.POINTER(PixelParams),
                                      ctypes.POINTER(PhotoParams)]),
            "orc_philox4x64": (None, [_P, _P, _P]),
            "orc_poisson_series": (None, [ctypes.c_uint64, ctypes.c_uint64, D, I, _P]),
            "orc_normal_series": (None, [ctypes.c_uint64, ctypes.c_uint64, I, _P]),
            "orc_shot_counts": (None, [ctypes.c_uint64, ctypes.c_uint64, D, ctypes.c_uint64, I, _P]),
            "orc_poisson_inversion": (D, [D, D]),
            "orc_sensor_create": (_P, [_P, I, ctypes.POINTER(Physic), I, I]),
            "orc_sensor_destroy": (None, [_P]),
            "orc_sensor_initialize": (I, [_P, _P, _P, D, I]),
            "orc_sensor_process": (I, [_P, _P, _P, I, _P, I64, ctypes.POINTER(I64), _P, I64, ctypes.POINTER(I64), I]),
            "orc_sensor_process_adaptive": (I, [_P, _P, _P, I, ctypes.c_double, I, I, ctypes.POINTER(I), _P, _P,
                                                _P, I64, ctypes.POINTER(I64), _P, I64, ctypes.POINTER(I64), I]),
            "orc_sensor_drain": (I, [_P, _P, I64, ctypes.POINTER(I64)]),
            "orc_sensor_verify_optics": (I, [_P, _P, D, _P, _P, I]),
            "orc_sensor_dropped": (I64, [_P]),
            "orc_sensor_peak_queue": (I64, [_P]),
            "orc_sensor_states": (I, [_P, _P]),
            "orc_sensor_shadow_rays": (I64, [_P]),
            "orc_readout_create": (_P, [D, D, I64]),
            "orc_readout_destroy": (None, [_P]),
            "orc_readout_submit": (I64, [_P, _P, I64, D, _P, I64]),
            "orc_readout_drain": (I64, [_P, _P, I64]),
            "orc_readout_dropped": (I64, [_P]),
            "orc_readout_peak": (I64, [_P]),
            "orc_render_to_bytes": (None, [_P, I, I, F, _P]),
            "orc_raster_mesh": (I, [_P, _P, _P, _P, I, I, _P, _P, _P, ctypes.POINTER(I64), _P, ctypes.POINTER(I64)]),
            "orc_raster_align": (None, [_P, _P, I, I, _P, _P, _P, I, I, _P]),
            "orc_trajectory_create": (_P, [_P, I, F, _P, _P, ctypes.c_char_p, I]),
            "orc_trajectory_destroy": (None, [_P]),
            "orc_trajectory_range": (None, [_P, ctypes.POINTER(F), ctypes.POINTER(F), ctypes.POINTER(I)]),
            "orc_trajectory_evaluate": (None, [_P, _P, I64, _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


PLANTS = range(1, 9)


def planted_hapke(n: int):
    """hapke_reflectance of the oracle built with planted mistake ORC_PLANT=n
    (a wrong term of the rough-surface branch), for the tests that show
    their checks reject it."""
    path = HERE / "_build" / f"liboracle_plant{n}.so"
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-j8", "planted"], check=True)
    L = ctypes.CDLL(str(path))
    fn = L.orc_hapke_reflectance
    fn.restype = ctypes.c_float
    fn.argtypes = [ctypes.c_float] * 4 + [Hapke]
    return lambda ci, ce, cg, w, h: fn(ci, ce, cg, w, hapke_struct(h))


def ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _fields(obj):
    return dataclasses.asdict(obj) if dataclasses.is_dataclass(obj) else obj


def hapke_struct(h) -> Hapke:
    return Hapke(**_fields(h))


def camera_struct(c) -> Camera:
    return Camera(**_fields(c))


def physic_struct(p) -> Physic:
    return Physic(**_fields(p))


def bvh_build(vertices, indices, max_leaf=8, node_cost=1.0, prim_cost=1.0):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    idx = np.array(indices, np.uint32, copy=True).reshape(-1, 3)
    nodes = np.zeros(max(1, 2 * len(idx)), NODE)
    count = ctypes.c_int(0)
    root = lib().orc_bvh_build(max_leaf, node_cost, prim_cost, ptr(v), len(v), ptr(idx), len(idx), ptr(nodes),
                               ctypes.byref(count))
    return nodes[: count.value].copy(), root, idx


class Scene:
    def __init__(self, nodes, root, vertices, normals, indices, camera, hapke, albedos=None):
        self._keep = [np.ascontiguousarray(nodes, NODE), np.ascontiguousarray(vertices, np.float32).reshape(-1, 3),
                      np.ascontiguousarray(normals, np.float32).reshape(-1, 3),
                      np.ascontiguousarray(indices, np.uint32).reshape(-1, 3),
                      None if albedos is None else np.ascontiguousarray(albedos, np.float32)]
        n, v, nr, idx, alb = self._keep
        self.camera = camera
        self.handle = lib().orc_scene_create(ptr(n), len(n), root, ptr(v), ptr(nr), len(v), ptr(idx), len(idx),
                                             ptr(alb), camera_struct(camera), hapke_struct(hapke))

    @property
    def pixels(self):
        return self.camera.width * self.camera.height

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_scene_destroy(self.handle)
            self.handle = None

    def render(self, pose, threads=0):
        p = np.ascontiguousarray(np.atleast_1d(pose)[:1], STATE)
        img = np.empty(self.pixels, np.float32)
        dep = np.empty(self.pixels, np.float32)
        rad = np.empty(self.pixels, np.float64)
        code = lib().orc_render(self.handle, ptr(p), ptr(img), ptr(dep), ptr(rad), threads)
        shape = (self.camera.height, self.camera.width)
        return code, img.reshape(shape), dep.reshape(shape), rad.reshape(shape)

    def trace(self, origin, direction):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        tri, t, u, v = ctypes.c_int(), ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        hit = lib().orc_trace_nearest(self.handle, ptr(o), ptr(d), ctypes.byref(tri), ctypes.byref(t),
                                      ctypes.byref(u), ctypes.byref(v))
        return hit, tri.value, t.value, u.value, v.value

    def trace_counted(self, origin, direction):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        counts = np.zeros(3, np.int64)
        hit = lib().orc_trace_nearest_counted(self.handle, ptr(o), ptr(d), ptr(counts))
        return hit, counts

    def shadow(self, origin, direction, self_tri, t_min=1e-6, t_max=1e30):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        return lib().orc_trace_shadow(self.handle, ptr(o), ptr(d), int(self_tri), t_min, t_max)

    def shadow_origin(self, tri, u, v, light):
        out = np.empty(3, np.float32)
        lib().orc_shadow_origin(self.handle, int(tri), float(u), float(v), ptr(np.ascontiguousarray(light, np.float32)),
                                ptr(out))
        return out


class Sensor:
    def __init__(self, width, height, physic, effects=True, aer=True):
        self.width, self.height = width, height
        self.handle = lib().orc_sensor_create(width, height, ctypes.byref(physic_struct(physic)), int(effects),
                                              int(aer))
        if not self.handle:
            raise ValueError("invalid sensor capacity")

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_sensor_destroy(self.handle)
            self.handle = None

    def initialize(self, scene: Scene, pose, t0, threads=0):
        p = np.ascontiguousarray(np.atleast_1d(pose)[:1], STATE)
        return lib().orc_sensor_initialize(self.handle, scene.handle, ptr(p), float(t0), threads)

    def process(self, scene: Scene, intervals, threads=0, capacity=None):
        iv = np.ascontiguousarray(np.atleast_1d(intervals), INTERVAL)
        cap = capacity or self.width * self.height * 64
        raw = np.empty(cap, EVENT)
        ready = np.empty(cap, EVENT)
        nr, nd = ctypes.c_int64(), ctypes.c_int64()
        code = lib().orc_sensor_process(self.handle, scene.handle, ptr(iv), len(iv), ptr(raw), cap, ctypes.byref(nr),
                                        ptr(ready), cap, ctypes.byref(nd), threads)
        return code, raw[: nr.value].copy(), ready[: nd.value].copy()

    def process_adaptive(self, scene: Scene, intervals, tolerance=1e-2, check=False, capacity=512, threads=0,
                         events_capacity=None):
        """process_adaptive: returns (code, accepted, raw, ready, refinement stats dict, probe dict)."""
        iv = np.ascontiguousarray(np.atleast_1d(intervals), INTERVAL)
        cap = events_capacity or self.width * self.height * 64
        raw = np.empty(cap, EVENT)
        ready = np.empty(cap, EVENT)
        nr, nd, acc = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        st, pr = RefinementStats(), ReceptorProbe()
        code = lib().orc_sensor_process_adaptive(self.handle, scene.handle, ptr(iv), len(iv), float(tolerance),
                                                 int(check), int(capacity), ctypes.byref(acc), ctypes.byref(st),
                                                 ctypes.byref(pr), ptr(raw), cap, ctypes.byref(nr), ptr(ready), cap,
                                                 ctypes.byref(nd), threads)
        as_dict = lambda c: {name: getattr(c, name) for name, _ in c._fields_}
        return code, bool(acc.value), raw[: nr.value].copy(), ready[: nd.value].copy(), as_dict(st), as_dict(pr)

    def verify_optics(self, scene: Scene, tolerance=1e-2, threads=0):
        """OpticalSampler::verify of the last adaptive window: (code, dict)."""
        out = np.zeros(3)
        iout = np.zeros(3, np.int32)
        code = lib().orc_sensor_verify_optics(self.handle, scene.handle, float(tolerance), ptr(out), ptr(iout), threads)
        keys = ("maximum_log_error", "p99_log_error", "mean_log_error")
        return code, {**dict(zip(keys, map(float, out))),
                      **dict(zip(("probes", "over_tolerance", "exact_fallback_probes"), map(int, iout)))}

    def drain(self):
        cap = self.width * self.height * 64 + 1024
        out = np.empty(cap, EVENT)
        n = ctypes.c_int64()
        lib().orc_sensor_drain(self.handle, ptr(out), cap, ctypes.byref(n))
        return out[: n.value].copy()

    def states(self):
        out = np.empty(self.width * self.height, PIXEL_STATE)
        lib().orc_sensor_states(self.handle, ptr(out))
        return out

    @property
    def dropped(self):
        return lib().orc_sensor_dropped(self.handle)

    @property
    def shadow_rays(self):
        return lib().orc_sensor_shadow_rays(self.handle)


def parameters(physic, effects=True):
    p, ph = PixelParams(), PhotoParams()
    lib().orc_parameters(ctypes.byref(physic_struct(physic)), int(effects), ctypes.byref(p), ctypes.byref(ph))
    return p, ph


def advance_pixel(state: np.ndarray, params: PixelParams, rate, t0, t1, capacity=64, fixed_normal=0.0):
    """Runs advance_pixel on a one-element PIXEL_STATE array (modified in place)."""
    out = np.zeros(max(capacity, 1), PIXEL_EVENT)
    n = lib().orc_advance_pixel(ptr(state), ctypes.byref(params), float(rate), float(t0), float(t1), ptr(out),
                                int(capacity), float(fixed_normal))
    return n, out[: max(n, 0)].copy()


def receptor_log_at(u, rate, dt, charge):
    return lib().orc_receptor_log_at(u, rate, dt, charge)


def philox4x64(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    out = np.empty(4, np.uint64)
    lib().orc_philox4x64(ptr(c), ptr(k), ptr(out))
    return out


def poisson_series(seed, pixel, lam, n):
    out = np.empty(n, np.float64)
    lib().orc_poisson_series(seed, pixel, lam, n, ptr(out))
    return out


def shot_counts(seed, pixel, lam, serial0, n):
    out = np.empty(n, np.float64)
    lib().orc_shot_counts(seed, pixel, lam, serial0, n, ptr(out))
    return out


def poisson_inversion(lam, u):
    f = lib().orc_poisson_inversion
    return np.array([f(float(a), float(b)) for a, b in zip(np.atleast_1d(lam), np.atleast_1d(u))])


def normal_series(seed, pixel, n):
    out = np.empty(n, np.float64)
    lib().orc_normal_series(seed, pixel, n, ptr(out))
    return out


def hapke_reflectance(cosi, cose, cosg, albedo, hapke):
    return lib().orc_hapke_reflectance(cosi, cose, cosg, albedo, hapke_struct(hapke))


def hapke_directional(n, l, v, albedo, hapke):
    f = lambda x: np.ascontiguousarray(x, np.float32)
    n, l, v = f(n), f(l), f(v)
    return lib().orc_hapke_directional(ptr(n), ptr(l), ptr(v), albedo, hapke_struct(hapke))


def invert_hapke_if(inc, emi, phase, target, hapke):
    return lib().orc_invert_hapke_if(inc, emi, phase, target, hapke_struct(hapke))


class Readout:
    def __init__(self, bandwidth, latency, capacity):
        self.handle = lib().orc_readout_create(bandwidth, latency, capacity)

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_readout_destroy(self.handle)
            self.handle = None

    def submit(self, events, now):
        ev = np.ascontiguousarray(events, EVENT)
        cap = len(ev) + (1 << 16)
        out = np.empty(cap, EVENT)
        n = lib().orc_readout_submit(self.handle, ptr(ev), len(ev), now, ptr(out), cap)
        return out[:n].copy()

    def drain(self):
        cap = 1 << 20
        out = np.empty(cap, EVENT)
        n = lib().orc_readout_drain(self.handle, ptr(out), cap)
        return out[:n].copy()

    @property
    def dropped(self):
        return lib().orc_readout_dropped(self.handle)

    @property
    def peak(self):
        return lib().orc_readout_peak(self.handle)


def blur_exposures(frames, weights, cell_offsets, wide=False):
    """The motion-blur sum over sub-exposure frames (cells, H, W) float32:
    exposure e adds cells [offsets[e], offsets[e+1]) in cell order.
    wide=False: apps/main_cam_blur.cpp:68,76 -- float accumulator starting at
    0, acc = float(acc + float(frame * float(weight))) (the cv::Mat
    `accumulated += image * float(w)`: the scaled image is rounded to float,
    then the sum); wide=True: apps/main_cam_blur_windows.cpp:69,77,81 --
    double accumulator, acc += weight * double(frame), cast to float."""
    frames = np.asarray(frames, np.float32)
    offsets = np.asarray(cell_offsets, np.int64)
    out = np.empty((len(offsets) - 1,) + frames.shape[1:], np.float32)
    for e in range(len(offsets) - 1):
        if wide:
            acc = np.zeros(frames.shape[1:], np.float64)
            for c in range(offsets[e], offsets[e + 1]):
                acc = acc + np.float64(weights[c]) * frames[c].astype(np.float64)
            out[e] = acc.astype(np.float32)
        else:
            acc = np.zeros(frames.shape[1:], np.float32)
            for c in range(offsets[e], offsets[e + 1]):
                acc = acc + frames[c] * np.float32(weights[c])
            out[e] = acc
    return out


def raster_mesh(dem, x_coords, y_coords, albedo=None):
    """load_from_raster's mesh of a DEM (src/tools/build_from_raster.cu,
    planar path): (vertices (N,3), normals (N,3), triangles (M,3) uint32,
    albedos (N,) or None), or None for a NoData centre pixel."""
    d = np.ascontiguousarray(dem, np.float32)
    h, w = d.shape
    xs = np.ascontiguousarray(x_coords, np.float32)
    ys = np.ascontiguousarray(y_coords, np.float32)
    a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    v = np.empty((w * h, 3), np.float32)
    n = np.empty((w * h, 3), np.float32)
    alb = np.empty(w * h, np.float32) if a is not None else None
    t = np.empty((max(2 * (w - 1) * (h - 1), 1), 3), np.uint32)
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    if lib().orc_raster_mesh(ptr(d), ptr(a), ptr(xs), ptr(ys), w, h, ptr(v), ptr(n), ptr(alb), ctypes.byref(nv),
                             ptr(t), ctypes.byref(nt)):
        return None
    return v[:nv.value], n[:nv.value], t[:nt.value], None if alb is None else alb[:nv.value]


def raster_align(ref_x, ref_y, source, source_x, source_y):
    """batch_align_kernel (src/tools/build_from_raster.cu:653-682)."""
    rx, ry = np.ascontiguousarray(ref_x, np.float32), np.ascontiguousarray(ref_y, np.float32)
    src = np.ascontiguousarray(source, np.float32)
    sx, sy = np.ascontiguousarray(source_x, np.float32), np.ascontiguousarray(source_y, np.float32)
    out = np.empty((len(ry), len(rx)), np.float32)
    lib().orc_raster_align(ptr(rx), ptr(ry), len(rx), len(ry), ptr(src), ptr(sx), ptr(sy), src.shape[1], src.shape[0],
                           ptr(out))
    return out


def render_to_bytes(linear, dither=0.0):
    a = np.ascontiguousarray(linear, np.float32)
    h, w = a.shape
    out = np.empty((h, w), np.uint8)
    lib().orc_render_to_bytes(ptr(a), w, h, dither, ptr(out))
    return out


class Trajectory:
    """CameraTrajectory (src/trajectory.cpp) over FRAME records; raises
    ValueError with the reference's message where it throws
    std::invalid_argument."""

    def __init__(self, frames, fps, translation=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0)):
        frames = np.ascontiguousarray(frames, FRAME)
        tr = np.ascontiguousarray(translation, np.float32)
        ro = np.ascontiguousarray(rotation, np.float32)
        err = ctypes.create_string_buffer(256)
        self._h = lib().orc_trajectory_create(ptr(frames), len(frames), fps, ptr(tr), ptr(ro), err, 256)
        if not self._h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_trajectory_destroy(self._h)
            self._h = None

    def range(self):
        a, b, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        lib().orc_trajectory_range(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n))
        return a.value, b.value, n.value

    def evaluate(self, times):
        t = np.ascontiguousarray(np.atleast_1d(times), np.float32)
        out = np.zeros(len(t), TRAJECTORY_STATE)
        lib().orc_trajectory_evaluate(self._h, ptr(t), len(t), ptr(out))
        return out
