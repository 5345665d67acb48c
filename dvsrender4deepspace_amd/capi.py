"""ctypes binding of the gfx950 library's C ABI (include/dvs_hip.h).

Mirrors the reference's C++ objects used on the hot path:
``Scene`` ~ DeviceScene + CamRayCaster (include/dvs/device_scene.hpp,
include/dvs/raycast.hpp:64-86), ``Sensor`` ~ EventSensor
(include/dvs/event_sensor.hpp:40-79), ``bvh_build`` ~ BVHBuilder::build
(include/bvh/build.h:19-24).  Errors are raised as the reference's exception
classes would be: ``InvalidArgument`` (std::invalid_argument) and
``DvsError`` (std::runtime_error), with the same messages.

There is no CPU fallback: if the HIP library is missing, importing
``load_library`` raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "lib" / "libdvs_hip.so"

DVS_OK, DVS_E_CAPACITY, DVS_E_OPTICAL, DVS_E_ARGUMENT, DVS_E_HIP, DVS_E_STATE = range(6)

NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,)), ("left", "<i4"), ("right", "<i4"),
                       ("start", "<i4"), ("count", "<i4")])
STATE_DTYPE = np.dtype([("position", "<f4", (3,)), ("velocity", "<f4", (3,)), ("light", "<f4", (3,)),
                        ("light_rate", "<f4", (3,)), ("rotation", "<f4", (4,)), ("intensity", "<f4"),
                        ("intensity_rate", "<f4")])
INTERVAL_DTYPE = np.dtype([("start", STATE_DTYPE), ("end", STATE_DTYPE), ("t0", "<f8"), ("t1", "<f8")])
EVENT_DTYPE = np.dtype([("timestamp", "<f8"), ("x", "<i2"), ("y", "<i2"), ("polarity", "i1"),
                        ("depth", "<f4")])
PIXEL_STATE_DTYPE = np.dtype([("receptor_log", "<f8"), ("reference_log", "<f8"), ("refractory_until", "<f8"),
                              ("threshold_pos", "<f8"), ("threshold_neg", "<f8"), ("reset_pending", "<i4"),
                              ("reset_block", "<u4"), ("shot_block", "<u8"), ("receptor_inverse", "<f8")])
FRAME_DTYPE = np.dtype([("frame_index", "<i4"), ("timestamp_seconds", "<f8"), ("camera_origin", "<f4", (3,)),
                        ("camera_world", "<f4", (9,)), ("light_dir", "<f4", (3,)), ("light_intensity", "<f4")],
                       align=True)
TRAJECTORY_STATE_DTYPE = np.dtype([("timestamp", "<f4"), ("position", "<f4", (3,)), ("orientation", "<f4", (9,)),
                                   ("quaternion", "<f4", (4,)), ("linear_velocity", "<f4", (3,)),
                                   ("angular_velocity", "<f4", (3,)), ("light_direction", "<f4", (3,)),
                                   ("light_direction_rate", "<f4", (3,)), ("light_intensity", "<f4"),
                                   ("light_intensity_rate", "<f4")])
PIXEL_PARAMS_DTYPE = np.dtype([(n, "<f8") for n in ("threshold_pos", "threshold_neg", "charge_scale",
                                                     "refractory_time", "reference_leak", "reset_noise_std")])
CROSSING_DTYPE = np.dtype([("time", "<f8"), ("polarity", "<i4"), ("pad", "<i4")])
assert NODE_DTYPE.itemsize == 40 and STATE_DTYPE.itemsize == 72 and INTERVAL_DTYPE.itemsize == 160
assert EVENT_DTYPE.itemsize == 17 and PIXEL_STATE_DTYPE.itemsize == 64
assert FRAME_DTYPE.itemsize == 80 and TRAJECTORY_STATE_DTYPE.itemsize == 124


class DvsError(RuntimeError):
    """Failure reported by the library (std::runtime_error in the reference)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class InvalidArgument(DvsError, ValueError):
    """Rejected input (std::invalid_argument in the reference)."""


@dataclasses.dataclass(frozen=True)
class Hapke:
    """HapkeConfig (include/dvs/scene_types.hpp:89-99) with its defaults."""
    b: float = 0.24
    c: float = 0.305
    hS: float = 0.08
    hC: float = 1.0
    thetap: float = 26.0 * np.float32(np.pi) / 180.0
    w: float = 0.324
    B_S0: float = 2.42
    B_C0: float = 0.0
    phi: float = 0.3


@dataclasses.dataclass(frozen=True)
class Camera:
    """CameraConfig (scene_types.hpp:59-62) plus scene.supersample_factor."""
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    samples: int = 1


@dataclasses.dataclass(frozen=True)
class Physic:
    """PhysicConfig (scene_types.hpp:64-79), same fields and defaults."""
    threshold_pos: float = 0.3
    threshold_neg: float = 0.3
    refractory_time: float = 0.0002
    readout_bandwidth: float = 5e6
    readout_latency: float = 0.0
    threshold_std: float = 0.03
    threshold_min: float = 0.01
    wavelength_nm: float = 550
    quantum_efficiency: float = 0.3
    pixel_pitch_um: float = 15
    fill_factor: float = 0.2
    f_number: float = 2
    lens_transmission: float = 0.8
    radiance_scale: float = 0.001
    temperature_k: float = 300
    subthreshold_kappa: float = 0.8
    receptor_capacitance_f: float = 1e-15
    reset_capacitance_f: float = 20e-15
    dark_current_a: float = 1e-18
    reset_leakage_a: float = 0.0
    circuit_step: float = 1e-6
    photon_shot_noise: int = 1
    reset_thermal_noise: int = 0
    random_seed: int = 114514
    max_events_per_step: int = 16
    readout_queue_capacity: int = 1000000


class _HapkeC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("b", "c", "hS", "hC", "thetap", "w", "B_S0", "B_C0", "phi")]


class _CameraC(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("samples", ctypes.c_int32)]


_PHYSIC_INT = {"photon_shot_noise", "reset_thermal_noise", "random_seed", "max_events_per_step",
               "readout_queue_capacity"}


class _PhysicC(ctypes.Structure):
    _fields_ = [(f.name, ctypes.c_int32 if f.name in _PHYSIC_INT else ctypes.c_float)
                for f in dataclasses.fields(Physic)]


class _TuningC(ctypes.Structure):
    _fields_ = [("rate_cells", ctypes.c_int32), ("trace_threads", ctypes.c_int32), ("traversal", ctypes.c_int32)]


class _StatsC(ctypes.Structure):
    _fields_ = [("trace_ms", ctypes.c_double), ("circuit_ms", ctypes.c_double), ("events_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("trace_launches", ctypes.c_double),
                ("primary_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("event_rays", ctypes.c_uint64), ("pixel_cells", ctypes.c_uint64),
                ("comparator_events", ctypes.c_uint64), ("delivered_events", ctypes.c_uint64),
                ("native_equivalent_rays", ctypes.c_uint64), ("adaptive_windows", ctypes.c_uint64),
                ("rejected_windows", ctypes.c_uint64), ("rates_ms", ctypes.c_double),
                ("circuit_launches", ctypes.c_double), ("probe_ms", ctypes.c_double),
                ("query_ms", ctypes.c_double), ("probe_launches", ctypes.c_double),
                ("query_launches", ctypes.c_double), ("probe_samples", ctypes.c_uint64),
                ("query_samples", ctypes.c_uint64), ("readout_tasks", ctypes.c_uint64),
                ("readout_serial_tasks", ctypes.c_uint64), ("windows_set_up_ahead", ctypes.c_uint64)]


class _RefineC(ctypes.Structure):
    _fields_ = [("traced_pixel_samples", ctypes.c_uint64), ("fallback_pixels", ctypes.c_int32),
                ("maximum_knots", ctypes.c_int32), ("exact_segments", ctypes.c_int32),
                ("native_cell_queries", ctypes.c_int32)]


class _ProbeC(ctypes.Structure):
    _fields_ = [("maximum_log_voltage_error", ctypes.c_double), ("p99_log_voltage_error", ctypes.c_double),
                ("mean_log_voltage_error", ctypes.c_double), ("pixels", ctypes.c_int32),
                ("circuit_cells", ctypes.c_int32)]


class _OpticalProbeC(ctypes.Structure):
    _fields_ = [("maximum_log_error", ctypes.c_double), ("p99_log_error", ctypes.c_double),
                ("mean_log_error", ctypes.c_double), ("probes", ctypes.c_int32), ("over_tolerance", ctypes.c_int32),
                ("exact_fallback_probes", ctypes.c_int32)]


def _as_dict(c) -> dict:
    return {name: getattr(c, name) for name, _ in c._fields_}


def _struct(cls, obj):
    return cls(**dataclasses.asdict(obj))


_lib = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32


def load_library() -> ctypes.CDLL:
    """Loads libdvs_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(f"gfx950 library not built: {LIBRARY_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIBRARY_PATH))
    sig = {
        "dvs_last_error": (ctypes.c_char_p, []),
        "dvs_device_count": (ctypes.c_int, []),
        "dvs_bvh_build": (ctypes.c_int, [ctypes.c_int, ctypes.c_float, ctypes.c_float, _P, _I64, _P, _I64, _P,
                                         ctypes.POINTER(_I64), ctypes.POINTER(_I32)]),
        "dvs_bvh_validate": (ctypes.c_int, [_P, _I64, _I32]),
        "dvs_scene_create": (ctypes.c_int, [_P, _I64, _I32, _P, _P, _I64, _P, _I64, _P, ctypes.POINTER(_CameraC),
                                            ctypes.POINTER(_HapkeC), ctypes.POINTER(_P)]),
        "dvs_scene_create_ex": (ctypes.c_int, [_P, _I64, _I32, _P, _P, _I64, _P, _I64, _P,
                                               ctypes.POINTER(_CameraC), ctypes.POINTER(_HapkeC), _I32,
                                               ctypes.POINTER(_P)]),
        "dvs_scene_destroy": (ctypes.c_int, [_P]),
        "dvs_scene_camera": (ctypes.c_int, [_P, ctypes.POINTER(_CameraC)]),
        "dvs_scene_info": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                          ctypes.POINTER(_I64)]),
        "dvs_frame_render": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P]),
        "dvs_frame_render_device": (ctypes.c_int, [_P, _P, _I32, _P, _P]),
        "dvs_frame_render_device_noalloc": (ctypes.c_int, [_P, _P, _I32, _P, _P]),
        "dvs_frame_blur": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P]),
        "dvs_scan_eval": (ctypes.c_int, [_P, _I64, _P]),
        "dvs_raster_mesh": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _P, ctypes.POINTER(_I64), _P,
                                           ctypes.POINTER(_I64)]),
        "dvs_raster_align": (ctypes.c_int, [_P, _P, _I32, _I32, _P, _P, _P, _I32, _I32, _P]),
        "dvs_sort_eval": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P]),
        "dvs_sort_bucketed_eval": (ctypes.c_int, [_P, _P, _I64, ctypes.c_uint64, _I32, _I32, _P, _P, _P]),
        "dvs_scene_stats": (ctypes.c_int, [_P, ctypes.POINTER(_StatsC)]),
        "dvs_frame_traversal_stats": (ctypes.c_int, [_P, _P, _P]),
        "dvs_sensor_create": (ctypes.c_int, [_P, ctypes.POINTER(_PhysicC), _I32, _I32, ctypes.POINTER(_P)]),
        "dvs_sensor_destroy": (ctypes.c_int, [_P]),
        "dvs_sensor_initialize": (ctypes.c_int, [_P, _P, ctypes.c_double]),
        "dvs_sensor_process": (ctypes.c_int, [_P, _P, _I32, ctypes.POINTER(_I64)]),
        "dvs_sensor_process_adaptive": (ctypes.c_int, [_P, _P, _I32, ctypes.c_double, _I32, ctypes.POINTER(_I32),
                                                       ctypes.POINTER(_I64)]),
        "dvs_sensor_adaptive_capacity": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
        "dvs_sensor_refinement_stats": (ctypes.c_int, [_P, ctypes.POINTER(_RefineC)]),
        "dvs_sensor_last_convergence": (ctypes.c_int, [_P, ctypes.POINTER(_ProbeC)]),
        "dvs_sensor_take": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_sensor_raw_events": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_sensor_drain": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
        "dvs_sensor_readout_stats": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
        "dvs_sensor_states": (ctypes.c_int, [_P, _P]),
        "dvs_sensor_stats": (ctypes.c_int, [_P, ctypes.POINTER(_StatsC)]),
        "dvs_sensor_reset_stats": (ctypes.c_int, [_P]),
        "dvs_scene_view": (ctypes.c_int, [_P, ctypes.POINTER(_CameraC), _P, ctypes.POINTER(_P)]),
        "dvs_scene_set_traversal": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I32)]),
        "dvs_batch_create": (ctypes.c_int, [_I32, _I32, ctypes.POINTER(_P)]),
        "dvs_batch_destroy": (ctypes.c_int, [_P]),
        "dvs_batch_add": (ctypes.c_int, [_P, _P, ctypes.POINTER(_PhysicC), _I32, _I32, _P, ctypes.c_double,
                                         ctypes.POINTER(_I32)]),
        "dvs_batch_feed": (ctypes.c_int, [_P, _I32, _P, _I64]),
        "dvs_batch_feed_trajectory": (ctypes.c_int, [_P, _I32, _P, _P, _I64]),
        "dvs_sensor_process_adaptive_times": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.c_double, _I32,
                                                             ctypes.POINTER(_I32), ctypes.POINTER(_I64)]),
        "dvs_sensor_expect_adaptive_times": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.c_double]),
        "dvs_batch_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(_TuningC)]),
        "dvs_batch_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(_TuningC)]),
        "dvs_sensor_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(_TuningC)]),
        "dvs_batch_step": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
        "dvs_batch_take": (ctypes.c_int, [_P, _I32, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_batch_drain": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I64)]),
        "dvs_batch_segment_info": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                                  ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
        "dvs_batch_remove": (ctypes.c_int, [_P, _I32]),
        "dvs_batch_set_adaptive": (ctypes.c_int, [_P, _I32, ctypes.c_double, _I32, ctypes.c_double]),
        "dvs_batch_refinement_stats": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_RefineC)]),
        "dvs_batch_stats": (ctypes.c_int, [_P, ctypes.POINTER(_StatsC)]),
        "dvs_batch_reset_stats": (ctypes.c_int, [_P]),
        "dvs_hapke_eval": (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.POINTER(_HapkeC), _P]),
        "dvs_receptor_eval": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _I64, _P]),
        "dvs_noise_eval": (ctypes.c_int, [ctypes.c_uint64, _I64, ctypes.c_double, _I32, _P, _P]),
        "dvs_poisson_eval": (ctypes.c_int, [_P, _P, _I64, _P]),
        "dvs_advance_pixel_eval": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I32, ctypes.c_double, _P, _P]),
        "dvs_invert_hapke_eval": (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.POINTER(_HapkeC), _P]),
        "dvs_hapke_reflectance": (ctypes.c_float, [ctypes.c_float] * 4 + [ctypes.POINTER(_HapkeC)]),
        "dvs_hapke_directional": (ctypes.c_float, [_P, _P, _P, ctypes.c_float, ctypes.POINTER(_HapkeC)]),
        "dvs_invert_hapke_if": (ctypes.c_float, [ctypes.c_float] * 4 + [ctypes.POINTER(_HapkeC)]),
        "dvs_receptor_log_at": (ctypes.c_double, [ctypes.c_double] * 4),
        "dvs_advance_pixel": (_I32, [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P, _I32, _P, _P]),
        "dvs_sensor_enable_diagnostics": (ctypes.c_int, [_P]),
        "dvs_sensor_verify_receptor": (ctypes.c_int, [_P, _P, _I32, ctypes.POINTER(_ProbeC)]),
        "dvs_sensor_verify_optics": (ctypes.c_int, [_P, _P, _I32, ctypes.c_double, ctypes.POINTER(_OpticalProbeC)]),
        "dvs_readout_create": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, _I64, ctypes.POINTER(_P)]),
        "dvs_readout_destroy": (ctypes.c_int, [_P]),
        "dvs_readout_submit": (ctypes.c_int, [_P, _P, _I64, ctypes.c_double, ctypes.POINTER(_I64)]),
        "dvs_readout_drain": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
        "dvs_readout_take": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_readout_stats": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
        "dvs_trace_rays": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, _P, _P]),
        "dvs_trace_shadow_rays": (ctypes.c_int, [_P, _P, _P, _P, _I64, _P]),
        "dvs_trajectory_create": (ctypes.c_int, [_P, _I32, ctypes.c_float, _P, _P, ctypes.POINTER(_P)]),
        "dvs_trajectory_destroy": (ctypes.c_int, [_P]),
        "dvs_trajectory_range": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                                ctypes.POINTER(_I32)]),
        "dvs_trajectory_evaluate": (ctypes.c_int, [_P, _P, _I64, _P]),
        "dvs_trajectory_intervals": (ctypes.c_int, [_P, _P, _I64, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols(header: str = "dvs_hip.h") -> list[str]:
    """Names of every C entry point declared in include/<header>."""
    text = (Path(__file__).resolve().parent.parent / "include" / header).read_text()
    import re
    return sorted(set(re.findall(r"\b(dvs_[a-z0-9_]+)\s*\(", text)))


def _check(code: int) -> None:
    if code == DVS_OK:
        return
    message = load_library().dvs_last_error().decode()
    if code == DVS_E_ARGUMENT:
        raise InvalidArgument(code, message)
    raise DvsError(code, message)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    return int(load_library().dvs_device_count())


def bvh_build(vertices: np.ndarray, indices: np.ndarray, max_leaf_size: int = 8, node_cost: float = 1.0,
              primitive_cost: float = 1.0):
    """BVHBuilder::build: returns (nodes, root, indices in leaf order)."""
    lib = load_library()
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    idx = np.array(indices, dtype=np.uint32, copy=True).reshape(-1, 3)
    nodes = np.zeros(max(1, 2 * len(idx)), dtype=NODE_DTYPE)
    count, root = _I64(0), _I32(0)
    _check(lib.dvs_bvh_build(int(max_leaf_size), float(node_cost), float(primitive_cost), _ptr(v), len(v), _ptr(idx),
                             len(idx), _ptr(nodes), ctypes.byref(count), ctypes.byref(root)))
    return nodes[: count.value].copy(), int(root.value), idx


def bvh_validate(nodes: np.ndarray, root: int) -> None:
    _check(load_library().dvs_bvh_validate(_ptr(np.ascontiguousarray(nodes)), len(nodes), int(root)))


class Stats(dict):
    pass


def _stats(c: _StatsC) -> Stats:
    return Stats({name: getattr(c, name) for name, _ in c._fields_})


class Scene:
    """DeviceScene + CamRayCaster over one uploaded BVH and mesh."""

    def __init__(self, nodes, root, vertices, normals, indices, camera: Camera, hapke: Hapke | None = None,
                 albedos=None, device_leaf_size: int | None = None):
        lib = load_library()
        self._keep = [np.ascontiguousarray(nodes, dtype=NODE_DTYPE),
                      np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3),
                      np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3),
                      np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3),
                      None if albedos is None else np.ascontiguousarray(albedos, dtype=np.float32)]
        n, v, nr, idx, alb = self._keep
        self.camera = camera
        self.hapke = hapke or Hapke()
        handle = _P()
        cam = _struct(_CameraC, camera)
        hk = _struct(_HapkeC, self.hapke)
        if device_leaf_size is None:
            _check(lib.dvs_scene_create(_ptr(n), len(n), int(root), _ptr(v), _ptr(nr), len(v), _ptr(idx), len(idx),
                                        _ptr(alb), ctypes.byref(cam), ctypes.byref(hk), ctypes.byref(handle)))
        else:
            _check(lib.dvs_scene_create_ex(_ptr(n), len(n), int(root), _ptr(v), _ptr(nr), len(v), _ptr(idx),
                                           len(idx), _ptr(alb), ctypes.byref(cam), ctypes.byref(hk),
                                           int(device_leaf_size), ctypes.byref(handle)))
        self.handle = handle
        self._keep = None

    @property
    def pixels(self) -> int:
        return self.camera.width * self.camera.height

    def close(self):
        if getattr(self, "handle", None):
            load_library().dvs_scene_destroy(self.handle)
            self.handle = None

    __del__ = close

    def info(self) -> dict:
        d, i, l, b = _I32(), _I64(), _I64(), _I64()
        _check(load_library().dvs_scene_info(self.handle, ctypes.byref(d), ctypes.byref(i), ctypes.byref(l),
                                             ctypes.byref(b)))
        return {"max_depth": d.value, "internal_nodes": i.value, "leaves": l.value, "device_bytes": b.value}

    def render(self, poses: np.ndarray, with_radiance: bool = False):
        """render_main over poses: (image, depth[, radiance]), each (n, H, W)."""
        p = np.ascontiguousarray(np.atleast_1d(poses), dtype=STATE_DTYPE)
        shape = (len(p), self.camera.height, self.camera.width)
        img = np.empty(shape, np.float32)
        dep = np.empty(shape, np.float32)
        rad = np.empty(shape, np.float64) if with_radiance else None
        _check(load_library().dvs_frame_render(self.handle, _ptr(p), len(p), _ptr(img), _ptr(dep), _ptr(rad)))
        return (img, dep, rad) if with_radiance else (img, dep)

    def render_blur(self, poses: np.ndarray, weights, cell_offsets, wide: bool = False) -> np.ndarray:
        """Motion-blur exposures (apps/main_cam_blur.cpp:64-83): exposure e is
        the weighted sum, in cell order, of the frames of poses[offsets[e] :
        offsets[e+1]]; float sums of frame x float(weight), or with `wide`
        double sums cast at the end (apps/main_cam_blur_windows.cpp:77,81).
        Returns (exposures, H, W) float32."""
        p = np.ascontiguousarray(np.atleast_1d(poses), dtype=STATE_DTYPE)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        o = np.ascontiguousarray(cell_offsets, dtype=np.int32)
        if o.ndim != 1 or len(o) < 2 or o[-1] != len(p) or len(w) != len(p):
            raise ValueError("cell_offsets must end at the number of poses, one weight per pose")
        out = np.empty((len(o) - 1, self.camera.height, self.camera.width), np.float32)
        _check(load_library().dvs_frame_blur(self.handle, _ptr(p), _ptr(w), _ptr(o), len(o) - 1, int(wide),
                                             _ptr(out)))
        return out

    def render_device(self, device_poses_ptr: int, count: int, device_image_ptr: int, device_depth_ptr: int):
        _check(load_library().dvs_frame_render_device(self.handle, ctypes.c_void_p(device_poses_ptr), count,
                                                      ctypes.c_void_p(device_image_ptr),
                                                      ctypes.c_void_p(device_depth_ptr)))

    def render_device_noalloc(self, device_poses_ptr: int, count: int, device_image_ptr: int,
                              device_depth_ptr: int):
        """render_device without any device allocation (at most 4096 poses)."""
        _check(load_library().dvs_frame_render_device_noalloc(self.handle, ctypes.c_void_p(device_poses_ptr), count,
                                                              ctypes.c_void_p(device_image_ptr),
                                                              ctypes.c_void_p(device_depth_ptr)))

    def stats(self) -> Stats:
        s = _StatsC()
        _check(load_library().dvs_scene_stats(self.handle, ctypes.byref(s)))
        return _stats(s)

    def set_traversal(self, form: int) -> int:
        """Traversal form of this scene's frame, depth and ray launches
        (dvs_scene_set_traversal): 1 wave packets, 0 per-lane stacks; returns
        the form in use."""
        eff = _I32()
        _check(load_library().dvs_scene_set_traversal(self.handle, int(form), ctypes.byref(eff)))
        return eff.value

    def view(self, camera: Camera, hapke: Hapke | None = None) -> "Scene":
        """Another camera on this scene's resident geometry (shared BVH and mesh)."""
        out = Scene.__new__(Scene)
        out.camera = camera
        out.hapke = hapke or self.hapke
        out._keep = None
        out._base = self
        handle = _P()
        cam = _struct(_CameraC, camera)
        hk = _struct(_HapkeC, out.hapke)
        _check(load_library().dvs_scene_view(self.handle, ctypes.byref(cam), ctypes.byref(hk), ctypes.byref(handle)))
        out.handle = handle
        return out

    def traversal_stats(self, pose: np.ndarray) -> dict:
        """Per-frame traversal work (node-pair visits, leaf visits, triangle tests)."""
        p = np.ascontiguousarray(np.atleast_1d(pose)[:1], dtype=STATE_DTYPE)
        t = np.zeros(8, np.uint64)
        _check(load_library().dvs_frame_traversal_stats(self.handle, _ptr(p), _ptr(t)))
        keys = ("primary_rays", "primary_pairs", "primary_leaves", "primary_triangles",
                "shadow_rays", "shadow_pairs", "shadow_leaves", "shadow_triangles")
        return {k: int(v) for k, v in zip(keys, t)}

    def trace(self, origins, directions):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        n = len(o)
        hit, tri = np.empty(n, np.int32), np.empty(n, np.int32)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        _check(load_library().dvs_trace_rays(self.handle, _ptr(o), _ptr(d), n, _ptr(hit), _ptr(tri), _ptr(t),
                                             _ptr(u), _ptr(v)))
        return hit, tri, t, u, v

    def trace_shadow(self, origins, directions, self_triangles):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        s = np.ascontiguousarray(self_triangles, np.int32)
        occ = np.empty(len(o), np.int32)
        _check(load_library().dvs_trace_shadow_rays(self.handle, _ptr(o), _ptr(d), _ptr(s), len(o), _ptr(occ)))
        return occ


class Sensor:
    """EventSensor on a Scene (native solver)."""

    def __init__(self, scene: Scene, physic: Physic, physic_effects: bool = True, aer_delay: bool = True):
        lib = load_library()
        self.scene = scene
        self.physic = physic
        handle = _P()
        ph = _struct(_PhysicC, physic)
        _check(lib.dvs_sensor_create(scene.handle, ctypes.byref(ph), int(physic_effects), int(aer_delay),
                                     ctypes.byref(handle)))
        self.handle = handle

    def close(self):
        if getattr(self, "handle", None):
            load_library().dvs_sensor_destroy(self.handle)
            self.handle = None

    __del__ = close

    def initialize(self, pose: np.ndarray, timestamp: float) -> None:
        p = np.ascontiguousarray(np.atleast_1d(pose)[:1], dtype=STATE_DTYPE)
        _check(load_library().dvs_sensor_initialize(self.handle, _ptr(p), float(timestamp)))

    def _take(self, count: int) -> np.ndarray:
        out = np.empty(count, EVENT_DTYPE)
        got = _I64()
        _check(load_library().dvs_sensor_take(self.handle, _ptr(out), count, ctypes.byref(got)))
        return out[: got.value]

    def process(self, intervals: np.ndarray) -> np.ndarray:
        """process_many: returns the readout-delivered events of this call."""
        iv = np.ascontiguousarray(np.atleast_1d(intervals), dtype=INTERVAL_DTYPE)
        ready = _I64()
        _check(load_library().dvs_sensor_process(self.handle, _ptr(iv), len(iv), ctypes.byref(ready)))
        return self._take(ready.value)

    def process_adaptive(self, intervals: np.ndarray, tolerance: float = 1e-2,
                         check_convergence: bool = False) -> tuple[bool, np.ndarray]:
        """process_adaptive over the intervals as one window: (accepted,
        readout-delivered events).  Not accepted: the receptor check rejected
        the window and nothing changed."""
        iv = np.ascontiguousarray(np.atleast_1d(intervals), dtype=INTERVAL_DTYPE)
        ready, ok = _I64(), _I32()
        _check(load_library().dvs_sensor_process_adaptive(self.handle, _ptr(iv), len(iv), float(tolerance),
                                                          int(check_convergence), ctypes.byref(ok),
                                                          ctypes.byref(ready)))
        return bool(ok.value), self._take(ready.value)

    def process_adaptive_times(self, trajectory: "Trajectory", times, tolerance: float = 1e-2,
                               check_convergence: bool = False) -> tuple[bool, np.ndarray]:
        """process_adaptive over the intervals between consecutive `times`,
        their states evaluated from `trajectory` on the device inside the
        window (dvs_sensor_process_adaptive_times): the events
        process_adaptive(trajectory.intervals(times)) gives."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        ready, ok = _I64(), _I32()
        _check(load_library().dvs_sensor_process_adaptive_times(self.handle, trajectory.handle, _ptr(t), len(t),
                                                                float(tolerance), int(check_convergence),
                                                                ctypes.byref(ok), ctypes.byref(ready)))
        return bool(ok.value), self._take(ready.value)

    def expect_adaptive_times(self, trajectory: "Trajectory", times, tolerance: float = 1e-2) -> None:
        """The window the next process_adaptive_times call will most likely
        run (dvs_sensor_expect_adaptive_times).  When the current call's window
        is accepted, this window's setup (its horizon states, cell tables and
        initial probes) is queued on the device before the call returns; a
        next call with other times, trajectory or tolerance runs its own.
        Events are the same either way.  `times` of length 0 forgets it."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        _check(load_library().dvs_sensor_expect_adaptive_times(self.handle, trajectory.handle if len(t) else None,
                                                                 _ptr(t), len(t), float(tolerance)))

    def enable_diagnostics(self) -> None:
        """EventSensor::enable_numerical_diagnostics: every later adaptive
        window also measures its receptor check, for verify_receptor."""
        _check(load_library().dvs_sensor_enable_diagnostics(self.handle))

    def verify_receptor(self, intervals: np.ndarray) -> dict:
        """EventSensor::verify_receptor of the last adaptive window."""
        ivs = np.ascontiguousarray(np.atleast_1d(intervals), dtype=INTERVAL_DTYPE)
        r = _ProbeC()
        _check(load_library().dvs_sensor_verify_receptor(self.handle, _ptr(ivs), len(ivs), ctypes.byref(r)))
        return _as_dict(r)

    def verify_optics(self, intervals: np.ndarray, tolerance: float = 1e-2) -> dict:
        """EventSensor::verify_optics: the last adaptive window's knots against
        native traces of 1024 hashed pixels at 17 times."""
        ivs = np.ascontiguousarray(np.atleast_1d(intervals), dtype=INTERVAL_DTYPE)
        r = _OpticalProbeC()
        _check(load_library().dvs_sensor_verify_optics(self.handle, _ptr(ivs), len(ivs), float(tolerance),
                                                       ctypes.byref(r)))
        return _as_dict(r)

    def adaptive_capacity(self) -> int:
        c = _I32()
        _check(load_library().dvs_sensor_adaptive_capacity(self.handle, ctypes.byref(c)))
        return c.value

    def refinement_stats(self) -> dict:
        r = _RefineC()
        _check(load_library().dvs_sensor_refinement_stats(self.handle, ctypes.byref(r)))
        return _as_dict(r)

    def last_convergence(self) -> dict:
        r = _ProbeC()
        _check(load_library().dvs_sensor_last_convergence(self.handle, ctypes.byref(r)))
        return _as_dict(r)

    def raw_events(self) -> np.ndarray:
        """Comparator events of the last process call, (t, y, x) order, before readout."""
        lib = load_library()
        n = _I64()
        # Size query: capacity 0 reports the count (and an argument error if any exist).
        code = lib.dvs_sensor_raw_events(self.handle, None, 0, ctypes.byref(n))
        if code not in (DVS_OK, DVS_E_ARGUMENT):
            _check(code)
        out = np.empty(n.value, EVENT_DTYPE)
        _check(lib.dvs_sensor_raw_events(self.handle, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def drain(self) -> np.ndarray:
        ready = _I64()
        _check(load_library().dvs_sensor_drain(self.handle, ctypes.byref(ready)))
        return self._take(ready.value)

    def set_tuning(self, rate_cells: int = 128, trace_threads: int = 128, traversal: int = 1) -> None:
        """Launch shapes and traversal form of the sensor's kernels
        (dvs_sensor_set_tuning); results do not depend on them."""
        _check(load_library().dvs_sensor_set_tuning(
            self.handle, ctypes.byref(_TuningC(rate_cells, trace_threads, traversal))))

    def readout_stats(self) -> tuple[int, int]:
        d, p = _I64(), _I64()
        _check(load_library().dvs_sensor_readout_stats(self.handle, ctypes.byref(d), ctypes.byref(p)))
        return d.value, p.value

    def states(self) -> np.ndarray:
        out = np.empty(self.scene.pixels, PIXEL_STATE_DTYPE)
        _check(load_library().dvs_sensor_states(self.handle, _ptr(out)))
        return out

    def stats(self) -> Stats:
        s = _StatsC()
        _check(load_library().dvs_sensor_stats(self.handle, ctypes.byref(s)))
        return _stats(s)

    def reset_stats(self) -> None:
        _check(load_library().dvs_sensor_reset_stats(self.handle))


class Trajectory:
    """CameraTrajectory (src/trajectory.cpp:604-790) over FRAME_DTYPE records:
    mount translation and ZYX Euler rotation in the camera frame, as
    SceneConfig's translation/rotation.  Invalid frames raise InvalidArgument
    with the reference's std::invalid_argument messages."""

    def __init__(self, frames: np.ndarray, fps: float, translation=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0)):
        lib = load_library()
        f = np.ascontiguousarray(np.atleast_1d(frames), dtype=FRAME_DTYPE)
        tr = np.ascontiguousarray(translation, dtype=np.float32)
        ro = np.ascontiguousarray(rotation, dtype=np.float32)
        if tr.shape != (3,) or ro.shape != (3,):
            raise InvalidArgument(DVS_E_ARGUMENT,
                                  "trajectory requires positive fps and three-component mount transforms")
        handle = _P()
        _check(lib.dvs_trajectory_create(_ptr(f), len(f), float(fps), _ptr(tr), _ptr(ro), ctypes.byref(handle)))
        self.handle = handle

    def close(self):
        if getattr(self, "handle", None):
            load_library().dvs_trajectory_destroy(self.handle)
            self.handle = None

    __del__ = close

    def range(self) -> tuple[float, float, int]:
        """(start_time(), end_time(), size())."""
        a, b, n = ctypes.c_float(), ctypes.c_float(), _I32()
        _check(load_library().dvs_trajectory_range(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def evaluate(self, times) -> np.ndarray:
        """evaluate(float t) at each time, on the host."""
        t = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float32)
        out = np.empty(len(t), TRAJECTORY_STATE_DTYPE)
        _check(load_library().dvs_trajectory_evaluate(self.handle, _ptr(t), len(t), _ptr(out)))
        return out

    def intervals(self, times) -> np.ndarray:
        """The len(times) - 1 optical intervals between consecutive FP64 times,
        each endpoint state evaluated once at float(t), on the device."""
        t = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float64)
        out = np.empty(max(len(t) - 1, 0), INTERVAL_DTYPE)
        _check(load_library().dvs_trajectory_intervals(self.handle, _ptr(t), len(t), _ptr(out)))
        return out


def optical_states(states: np.ndarray) -> np.ndarray:
    """to_optical_state (include/dvs/raycast.hpp:17-22) of trajectory states."""
    s = np.atleast_1d(states)
    out = np.zeros(len(s), STATE_DTYPE)
    out["position"] = s["position"]
    out["velocity"] = s["linear_velocity"]
    out["light"] = s["light_direction"]
    out["light_rate"] = s["light_direction_rate"]
    out["rotation"] = s["quaternion"]
    out["intensity"] = s["light_intensity"]
    out["intensity_rate"] = s["light_intensity_rate"]
    return out


class Batch:
    """Multi-segment engine: many sensors advancing together in shared launches."""

    def __init__(self, horizons_per_step: int = 64, use_graph: bool = False):
        handle = _P()
        _check(load_library().dvs_batch_create(int(horizons_per_step), int(use_graph), ctypes.byref(handle)))
        self.handle = handle
        self.horizons_per_step = horizons_per_step
        self._scenes = {}
        self._trajectories = {}

    def close(self):
        if getattr(self, "handle", None):
            load_library().dvs_batch_destroy(self.handle)
            self.handle = None

    __del__ = close

    def add(self, scene: Scene, physic: Physic, initial_pose: np.ndarray, t0: float, physic_effects: bool = True,
            aer_delay: bool = True) -> int:
        seg = _I32()
        ph = _struct(_PhysicC, physic)
        p = np.ascontiguousarray(np.atleast_1d(initial_pose)[:1], dtype=STATE_DTYPE)
        _check(load_library().dvs_batch_add(self.handle, scene.handle, ctypes.byref(ph), int(physic_effects),
                                            int(aer_delay), _ptr(p), float(t0), ctypes.byref(seg)))
        self._scenes[seg.value] = scene
        return seg.value

    def feed(self, segment: int, intervals: np.ndarray) -> None:
        iv = np.ascontiguousarray(np.atleast_1d(intervals), dtype=INTERVAL_DTYPE)
        _check(load_library().dvs_batch_feed(self.handle, int(segment), _ptr(iv), len(iv)))

    def set_tuning(self, rate_cells: int = 128, trace_threads: int = 128, traversal: int = 1) -> None:
        """Launch shapes of the adaptive kernels and the traversal form
        (dvs_batch_set_tuning): photocurrent cells per thread run, trace lanes
        per block, wave packets (1) or per-lane stacks (0)."""
        _check(load_library().dvs_batch_set_tuning(
            self.handle, ctypes.byref(_TuningC(rate_cells, trace_threads, traversal))))

    def tuning(self) -> dict:
        t = _TuningC()
        _check(load_library().dvs_batch_get_tuning(self.handle, ctypes.byref(t)))
        return {"rate_cells": t.rate_cells, "trace_threads": t.trace_threads, "traversal": t.traversal}

    def feed_trajectory(self, segment: int, trajectory: "Trajectory", times) -> None:
        """Queues the intervals between consecutive `times`, their states
        evaluated from `trajectory` on the device when their step runs (the
        states Trajectory.intervals gives).  The batch keeps the trajectory
        alive while the segment exists."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        _check(load_library().dvs_batch_feed_trajectory(self.handle, int(segment), trajectory.handle, _ptr(t),
                                                        len(t)))
        self._trajectories.setdefault(int(segment), []).append(trajectory)

    def step(self) -> int:
        pending = _I32()
        _check(load_library().dvs_batch_step(self.handle, ctypes.byref(pending)))
        return pending.value

    def info(self, segment: int) -> dict:
        q, d, e, b = _I64(), _I64(), _I64(), _I64()
        _check(load_library().dvs_batch_segment_info(self.handle, int(segment), ctypes.byref(q), ctypes.byref(d),
                                                     ctypes.byref(e), ctypes.byref(b)))
        return {"queued": q.value, "done": d.value, "events": e.value, "device_bytes": b.value}

    def take(self, segment: int) -> np.ndarray:
        lib = load_library()
        n = _I64()
        code = lib.dvs_batch_take(self.handle, int(segment), None, 0, ctypes.byref(n))
        if code not in (DVS_OK, DVS_E_ARGUMENT):
            _check(code)
        out = np.empty(n.value, EVENT_DTYPE)
        _check(lib.dvs_batch_take(self.handle, int(segment), _ptr(out), n.value, ctypes.byref(n)))
        return out[: n.value]

    def drain(self, segment: int) -> np.ndarray:
        _check(load_library().dvs_batch_drain(self.handle, int(segment), ctypes.byref(_I64())))
        return self.take(segment)

    def remove(self, segment: int) -> None:
        _check(load_library().dvs_batch_remove(self.handle, int(segment)))
        self._scenes.pop(segment, None)
        self._trajectories.pop(segment, None)

    def set_adaptive(self, segment: int, tolerance: float = 1e-2, horizons: int = 200,
                     check_period: float = 0.25) -> None:
        """Adaptive windows for a segment under the driver policy of
        apps/main_dvs.cpp:96-116 (see dvs_batch_set_adaptive)."""
        _check(load_library().dvs_batch_set_adaptive(self.handle, int(segment), float(tolerance), int(horizons),
                                                     float(check_period)))

    def refinement_stats(self, segment: int) -> dict:
        r = _RefineC()
        _check(load_library().dvs_batch_refinement_stats(self.handle, int(segment), ctypes.byref(r)))
        return _as_dict(r)

    def stats(self) -> Stats:
        s = _StatsC()
        _check(load_library().dvs_batch_stats(self.handle, ctypes.byref(s)))
        return _stats(s)

    def reset_stats(self) -> None:
        _check(load_library().dvs_batch_reset_stats(self.handle))


_hip = None


def hip_runtime() -> ctypes.CDLL:
    """The HIP runtime libdvs_hip.so is linked against (libamdhip64.so.7 of
    /opt/rocm), for device buffers handed to the *_device entry points.
    (A second HIP runtime in the process, such as the one bundled with
    torch, cannot share the device with it.)"""
    global _hip
    if _hip is None:
        load_library()
        hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipDeviceSynchronize.argtypes = []
        _hip = hip
    return _hip


class DeviceBuffer:
    """Device memory of `nbytes` from hipMalloc, freed on close()."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        if hip_runtime().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise DvsError(DVS_E_HIP, f"hipMalloc of {self.nbytes} bytes failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if hip_runtime().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, self.H2D) != 0:
            raise DvsError(DVS_E_HIP, "hipMemcpy to device failed")
        return self

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        if hip_runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, self.D2H) != 0:
            raise DvsError(DVS_E_HIP, "hipMemcpy to host failed")
        return out

    def close(self) -> None:
        if self.ptr.value:
            hip_runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hapke_eval(cosi, cose, cosg, albedo, hapke: Hapke) -> np.ndarray:
    a, b, c, d = (np.ascontiguousarray(x, np.float32) for x in np.broadcast_arrays(cosi, cose, cosg, albedo))
    a, b, c, d = (np.ascontiguousarray(x) for x in (a, b, c, d))
    out = np.empty(a.shape, np.float32)
    hk = _struct(_HapkeC, hapke)
    _check(load_library().dvs_hapke_eval(_ptr(a), _ptr(b), _ptr(c), _ptr(d), a.size, ctypes.byref(hk), _ptr(out)))
    return out


def receptor_eval(u, rate, dt, charge: float) -> np.ndarray:
    a, b, c = (np.ascontiguousarray(x, np.float64) for x in np.broadcast_arrays(u, rate, dt))
    a, b, c = (np.ascontiguousarray(x) for x in (a, b, c))
    out = np.empty(a.shape, np.float64)
    _check(load_library().dvs_receptor_eval(_ptr(a), _ptr(b), _ptr(c), float(charge), a.size, _ptr(out)))
    return out


class Readout:
    """EventReadout (include/dvs/event_sensor.hpp:24-37): the bounded FIFO
    readout on its own (host side)."""

    def __init__(self, bandwidth: float, latency: float, capacity: int):
        self.handle = _P()
        _check(load_library().dvs_readout_create(float(bandwidth), float(latency), int(capacity),
                                                 ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            load_library().dvs_readout_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _take(self, count: int) -> np.ndarray:
        out = np.empty(count, EVENT_DTYPE)
        n = _I64()
        _check(load_library().dvs_readout_take(self.handle, _ptr(out), count, ctypes.byref(n)))
        return out[: n.value]

    def submit(self, events: np.ndarray, now: float) -> np.ndarray:
        ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        n = _I64()
        _check(load_library().dvs_readout_submit(self.handle, _ptr(ev), len(ev), float(now), ctypes.byref(n)))
        return self._take(n.value)

    def drain(self) -> np.ndarray:
        n = _I64()
        _check(load_library().dvs_readout_drain(self.handle, ctypes.byref(n)))
        return self._take(n.value)

    def stats(self) -> tuple[int, int]:
        d, p = _I64(), _I64()
        _check(load_library().dvs_readout_stats(self.handle, ctypes.byref(d), ctypes.byref(p)))
        return d.value, p.value


def hapke_reflectance(cosi: float, cose: float, cosg: float, albedo: float, hapke: Hapke) -> float:
    """hapke_reflectance (include/dvs/hapke.hpp:131-151) on the host."""
    hk = _struct(_HapkeC, hapke)
    return load_library().dvs_hapke_reflectance(cosi, cose, cosg, albedo, ctypes.byref(hk))


def hapke_directional(normal, light, view, albedo: float, hapke: Hapke) -> float:
    n, l, v = (np.ascontiguousarray(x, np.float32) for x in (normal, light, view))
    hk = _struct(_HapkeC, hapke)
    return load_library().dvs_hapke_directional(_ptr(n), _ptr(l), _ptr(v), albedo, ctypes.byref(hk))


def invert_hapke_if(incidence: float, emission: float, phase: float, target: float, hapke: Hapke) -> float:
    """invert_hapke_if (include/dvs/hapke.hpp:160-172) on the host."""
    hk = _struct(_HapkeC, hapke)
    return load_library().dvs_invert_hapke_if(incidence, emission, phase, target, ctypes.byref(hk))


def invert_hapke_eval(incidence, emission, phase, target, hapke: Hapke) -> np.ndarray:
    """invert_hapke_if over arrays, on the device."""
    a, b, c, d = (np.ascontiguousarray(x, np.float32) for x in np.broadcast_arrays(incidence, emission, phase, target))
    a, b, c, d = (np.ascontiguousarray(x) for x in (a, b, c, d))
    out = np.empty(a.shape, np.float32)
    hk = _struct(_HapkeC, hapke)
    _check(load_library().dvs_invert_hapke_eval(_ptr(a), _ptr(b), _ptr(c), _ptr(d), a.size, ctypes.byref(hk),
                                                _ptr(out)))
    return out


def receptor_log_at(u: float, rate: float, dt: float, charge: float) -> float:
    return load_library().dvs_receptor_log_at(u, rate, dt, charge)


def advance_pixel(state: np.ndarray, params: np.ndarray, rate: float, t0: float, t1: float, capacity: int = 64):
    """advance_pixel (include/dvs/pixel_model.hpp:35-102) on the host for one
    pixel (state: one PIXEL_STATE_DTYPE entry, updated in place); reset-noise
    draws are 0.  Returns (count, crossings)."""
    assert state.dtype == PIXEL_STATE_DTYPE and state.size == 1 and state.flags.c_contiguous
    p = np.ascontiguousarray(params, PIXEL_PARAMS_DTYPE).reshape(1)
    out = np.zeros(max(capacity, 1), CROSSING_DTYPE)
    n = load_library().dvs_advance_pixel(_ptr(state), _ptr(p), float(rate), float(t0), float(t1), _ptr(out),
                                         int(capacity), None, None)
    return n, out[: max(n, 0)]


def advance_pixel_eval(states: np.ndarray, params: np.ndarray, rates, t0, t1, capacity: int = 64,
                       normal: float = 0.0):
    """advance_pixel (include/dvs/pixel_model.hpp:35-102) of n independent
    pixels on the device.  `states` (PIXEL_STATE_DTYPE) is updated in place;
    returns (counts, crossings[n, capacity]) with count -1 on overflow."""
    n = len(states)
    assert states.dtype == PIXEL_STATE_DTYPE and states.flags.c_contiguous
    params = np.ascontiguousarray(params, PIXEL_PARAMS_DTYPE)
    r, a, b = (np.ascontiguousarray(np.broadcast_to(x, (n,)), np.float64) for x in (rates, t0, t1))
    counts = np.empty(n, np.int32)
    events = np.zeros((n, max(capacity, 1)), CROSSING_DTYPE)
    _check(load_library().dvs_advance_pixel_eval(_ptr(states), _ptr(params), _ptr(r), _ptr(a), _ptr(b), n,
                                                 int(capacity), float(normal), _ptr(counts), _ptr(events)))
    return counts, events[:, :capacity]


def poisson_eval(lam, u) -> np.ndarray:
    """The device's Poisson inversion count of each uniform u at lambda."""
    lam = np.ascontiguousarray(lam, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty(len(lam), np.float64)
    _check(load_library().dvs_poisson_eval(_ptr(lam), _ptr(u), len(lam), _ptr(out)))
    return out


def noise_eval(seed: int, pixels: int, lam: float, draws: int):
    p = np.empty((pixels, draws), np.float64)
    q = np.empty((pixels, draws), np.float64)
    _check(load_library().dvs_noise_eval(int(seed), int(pixels), float(lam), int(draws), _ptr(p), _ptr(q)))
    return p, q


def scan_eval(values) -> np.ndarray:
    """The event compaction's device inclusive prefix sum of int32 values."""
    a = np.ascontiguousarray(values, dtype=np.int32)
    out = np.empty_like(a)
    _check(load_library().dvs_scan_eval(_ptr(a), len(a), _ptr(out)))
    return out


def sort_eval(keys, values, begin_bit: int = 0, end_bit: int = 64):
    """The device's stable radix sort of (uint64 key, uint32 value) pairs by
    key bits [begin_bit, end_bit)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(values, dtype=np.uint32)
    if len(k) != len(v):
        raise ValueError("one value per key")
    ko, vo = np.empty_like(k), np.empty_like(v)
    _check(load_library().dvs_sort_eval(_ptr(k), _ptr(v), len(k), begin_bit, end_bit, _ptr(ko), _ptr(vo)))
    return ko, vo


def sort_bucketed_eval(keys, values, key_lo: int, key_bits: int, bucket_bits: int):
    """The event sort's bucketed form: (keys, values, overflow) for keys in
    [key_lo, key_lo + 2**key_bits), sorted stably by the top bucket_bits of
    key - key_lo, then by the whole key inside each bucket; overflow is True
    when a bucket held more than 64 pairs."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(values, dtype=np.uint32)
    if len(k) != len(v):
        raise ValueError("one value per key")
    ko, vo = np.empty_like(k), np.empty_like(v)
    flag = np.zeros(1, dtype=np.int32)
    _check(load_library().dvs_sort_bucketed_eval(_ptr(k), _ptr(v), len(k), int(key_lo), key_bits, bucket_bits,
                                                 _ptr(ko), _ptr(vo), _ptr(flag)))
    return ko, vo, bool(flag[0])


def raster_mesh(dem, x_coords, y_coords, albedo=None):
    """The mesh of a DEM on the device (tools::build_from_raster, planar
    path, and load_from_raster's recentring): (vertices (N,3), normals (N,3),
    triangles (M,3) uint32, albedos (N,) or None), in raster order."""
    d = np.ascontiguousarray(dem, dtype=np.float32)
    h, w = d.shape
    xs = np.ascontiguousarray(x_coords, dtype=np.float32)
    ys = np.ascontiguousarray(y_coords, dtype=np.float32)
    a = None if albedo is None else np.ascontiguousarray(albedo, dtype=np.float32)
    v = np.empty((w * h, 3), np.float32)
    n = np.empty((w * h, 3), np.float32)
    alb = np.empty(w * h, np.float32) if a is not None else None
    tri = np.empty((max(2 * (w - 1) * (h - 1), 1), 3), np.uint32)
    nv, nt = _I64(), _I64()
    _check(load_library().dvs_raster_mesh(_ptr(d), _ptr(a), _ptr(xs), _ptr(ys), w, h, _ptr(v), _ptr(n), _ptr(alb),
                                          ctypes.byref(nv), _ptr(tri), ctypes.byref(nt)))
    return v[:nv.value], n[:nv.value], tri[:nt.value], None if alb is None else alb[:nv.value]


def raster_align(ref_x, ref_y, source, source_x, source_y) -> np.ndarray:
    """batch_align_kernel (src/tools/build_from_raster.cu:653-682) on the device."""
    rx, ry = np.ascontiguousarray(ref_x, dtype=np.float32), np.ascontiguousarray(ref_y, dtype=np.float32)
    src = np.ascontiguousarray(source, dtype=np.float32)
    sx, sy = np.ascontiguousarray(source_x, dtype=np.float32), np.ascontiguousarray(source_y, dtype=np.float32)
    out = np.empty((len(ry), len(rx)), np.float32)
    _check(load_library().dvs_raster_align(_ptr(rx), _ptr(ry), len(rx), len(ry), _ptr(src), _ptr(sx), _ptr(sy),
                                           src.shape[1], src.shape[0], _ptr(out)))
    return out
