"""Event stream files, the YAML configuration, the scene bundle and
trajectory loaders and the frame images of bin/dvs and bin/cam, over
libdvs_io.so (include/dvs_io.h).

EventWriter writes events.bin (13-byte records) or events.h5 (x, y, t in
microseconds, p and ms_to_idx), as BinaryEventStreamWriter and
Hdf5EventStreamWriter do (include/dvs/dvs_writer.hpp).  load_config reads a
configuration file with Config's grammar and checks (src/scene_types.cpp).
"""
from __future__ import annotations

import ctypes
import dataclasses
from pathlib import Path

import numpy as np

from . import capi

LIBRARY_PATH = Path(__file__).resolve().parent / "lib" / "libdvs_io.so"
BIN_RECORD_DTYPE = np.dtype([("t", "<f4"), ("x", "<i2"), ("y", "<i2"), ("p", "i1"), ("depth", "<f4")])
assert BIN_RECORD_DTYPE.itemsize == 13
FORMATS = {"bin": 0, "hdf5": 1, "h5": 1, "binary": 0}

_lib = None
_P, _I32, _I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


class _SceneC(ctypes.Structure):
    _fields_ = [("frame_count", ctypes.c_int32), ("fps", ctypes.c_float), ("imu_fps", ctypes.c_float),
                ("cam_fps", ctypes.c_float), ("time_resolution", ctypes.c_float),
                ("blur_resolution", ctypes.c_float), ("supersample_factor", ctypes.c_int32),
                ("translation", ctypes.c_float * 3), ("rotation", ctypes.c_float * 3)]


class _ConfigC(ctypes.Structure):
    _fields_ = [("camera", capi._CameraC), ("physic", capi._PhysicC), ("scene", _SceneC), ("hapke", capi._HapkeC)]


@dataclasses.dataclass(frozen=True)
class SceneConfig:
    """SceneConfig (include/dvs/scene_types.hpp:80-86)."""
    frame_count: int
    fps: float
    imu_fps: float
    cam_fps: float
    time_resolution: float
    blur_resolution: float
    supersample_factor: int
    translation: tuple
    rotation: tuple


@dataclasses.dataclass(frozen=True)
class Config:
    camera: capi.Camera
    physic: capi.Physic
    scene: SceneConfig
    hapke: capi.Hapke


class IoError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class ConfigError(ValueError):
    pass


class _RasterInfoC(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("transform", ctypes.c_double * 6),
                ("has_nodata", ctypes.c_int32), ("has_scale", ctypes.c_int32), ("has_offset", ctypes.c_int32),
                ("nodata", ctypes.c_double), ("scale", ctypes.c_double), ("offset", ctypes.c_double),
                ("missing", ctypes.c_int64), ("center_raw", ctypes.c_float)]


@dataclasses.dataclass
class Imagery:
    """tools::Imagery (include/tools/loader.hpp:17-22): band 1 as float32
    (NoData as NaN, scale and offset applied), pixel-centre coordinates, the
    geotransform and the projection key."""
    data: np.ndarray
    x_coords: np.ndarray
    y_coords: np.ndarray
    transform: tuple
    projection: str
    nodata: float | None
    scale: float
    offset: float
    missing: int

    @property
    def width(self) -> int:
        return self.data.shape[1]

    @property
    def height(self) -> int:
        return self.data.shape[0]


def load_raster(path) -> Imagery:
    """tools::loadImageryData (include/tools/loader.hpp:31-100) without GDAL."""
    lib = load_library()
    handle = _P()
    _check(lib.dvs_io_raster_load(str(path).encode(), ctypes.byref(handle)))
    try:
        info = _RasterInfoC()
        _check(lib.dvs_io_raster_info(handle, ctypes.byref(info)))
        data = np.empty((info.height, info.width), np.float32)
        xs = np.empty(info.width, np.float32)
        ys = np.empty(info.height, np.float32)
        _check(lib.dvs_io_raster_data(handle, data.ctypes.data, xs.ctypes.data, ys.ctypes.data))
        length = _I64()
        _check(lib.dvs_io_raster_projection(handle, None, 0, ctypes.byref(length)))
        buf = ctypes.create_string_buffer(length.value + 1)
        _check(lib.dvs_io_raster_projection(handle, buf, length.value + 1, ctypes.byref(length)))
        return Imagery(data, xs, ys, tuple(info.transform), buf.value.decode(),
                       info.nodata if info.has_nodata else None, info.scale, info.offset, int(info.missing))
    finally:
        lib.dvs_io_raster_free(handle)


def load_library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(f"I/O library not built: {LIBRARY_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIBRARY_PATH))
    sig = {
        "dvs_io_last_error": (ctypes.c_char_p, []),
        "dvs_writer_open": (ctypes.c_int, [ctypes.c_char_p, _I32, ctypes.POINTER(_P)]),
        "dvs_writer_append": (ctypes.c_int, [_P, _P, _I64]),
        "dvs_writer_close": (ctypes.c_int, [_P, ctypes.c_double]),
        "dvs_writer_count": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
        "dvs_writer_destroy": (ctypes.c_int, [_P]),
        "dvs_events_h5_read": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_events_h5_ms_to_idx": (ctypes.c_int, [ctypes.c_char_p, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_config_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_ConfigC)]),
        "dvs_io_mesh_load": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _I64, _P, _I64, ctypes.POINTER(_I64),
                                            ctypes.POINTER(_I64)]),
        "dvs_io_frames_binary": (ctypes.c_int, [ctypes.c_char_p, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_io_frames_text": (ctypes.c_int, [ctypes.c_char_p, _P, _I64, ctypes.POINTER(_I64)]),
        "dvs_io_frame_bytes": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_float, _P]),
        "dvs_io_write_png": (ctypes.c_int, [ctypes.c_char_p, _P, _I32, _I32]),
        "dvs_io_write_tiff": (ctypes.c_int, [ctypes.c_char_p, _P, _I32, _I32]),
        "dvs_io_raster_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
        "dvs_io_raster_info": (ctypes.c_int, [_P, ctypes.POINTER(_RasterInfoC)]),
        "dvs_io_raster_data": (ctypes.c_int, [_P, _P, _P, _P]),
        "dvs_io_raster_projection": (ctypes.c_int, [_P, ctypes.c_char_p, _I64, ctypes.POINTER(_I64)]),
        "dvs_io_raster_free": (ctypes.c_int, [_P]),
        "dvs_io_raster_geokey": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double)]),
        "dvs_frames_open": (ctypes.c_int, [ctypes.c_char_p, _I32, _I32, ctypes.POINTER(_P)]),
        "dvs_frames_append": (ctypes.c_int, [_P, ctypes.c_double, _P, _P]),
        "dvs_frames_close": (ctypes.c_int, [_P]),
        "dvs_frames_count": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
        "dvs_frames_destroy": (ctypes.c_int, [_P]),
        "dvs_frames_h5_read": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _I64, ctypes.POINTER(_I32),
                                              ctypes.POINTER(_I32), ctypes.POINTER(_I64)]),
        "dvs_writer_flush": (ctypes.c_int, [_P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code: int) -> None:
    if code == capi.DVS_OK:
        return
    message = load_library().dvs_io_last_error().decode()
    if code == capi.DVS_E_ARGUMENT:
        raise ConfigError(message)
    raise IoError(code, message)


class EventWriter:
    """EventStreamWriter: append(events) in time order, close(end_time)."""

    def __init__(self, path, fmt: str = "bin"):
        if fmt not in FORMATS:
            raise IoError(capi.DVS_E_STATE, f"unsupported event format: {fmt}")
        handle = _P()
        _check(load_library().dvs_writer_open(str(path).encode(), FORMATS[fmt], ctypes.byref(handle)))
        self.handle = handle
        self.path = Path(path)

    def append(self, events: np.ndarray) -> None:
        ev = np.ascontiguousarray(np.atleast_1d(events), dtype=capi.EVENT_DTYPE)
        _check(load_library().dvs_writer_append(self.handle, ev.ctypes.data, len(ev)))

    def close(self, end_time_seconds: float) -> None:
        _check(load_library().dvs_writer_close(self.handle, float(end_time_seconds)))

    @property
    def count(self) -> int:
        n = _I64()
        _check(load_library().dvs_writer_count(self.handle, ctypes.byref(n)))
        return n.value

    def release(self):
        if getattr(self, "handle", None):
            load_library().dvs_writer_destroy(self.handle)
            self.handle = None

    __del__ = release


def read_bin(path) -> np.ndarray:
    return np.fromfile(path, dtype=BIN_RECORD_DTYPE)


def read_h5(path) -> dict:
    lib = load_library()
    n = _I64()
    _check(lib.dvs_events_h5_read(str(path).encode(), None, None, None, None, 0, ctypes.byref(n)))
    out = {"x": np.empty(n.value, np.int16), "y": np.empty(n.value, np.int16), "t": np.empty(n.value, np.int64),
           "p": np.empty(n.value, np.int8)}
    _check(lib.dvs_events_h5_read(str(path).encode(), out["x"].ctypes.data, out["y"].ctypes.data,
                                  out["t"].ctypes.data, out["p"].ctypes.data, n.value, ctypes.byref(n)))
    m = _I64()
    _check(lib.dvs_events_h5_ms_to_idx(str(path).encode(), None, 0, ctypes.byref(m)))
    idx = np.empty(m.value, np.int64)
    _check(lib.dvs_events_h5_ms_to_idx(str(path).encode(), idx.ctypes.data, m.value, ctypes.byref(m)))
    out["ms_to_idx"] = idx
    return out


def load_config(path) -> Config:
    """Config(path): raises ConfigError with the reference's message."""
    c = _ConfigC()
    _check(load_library().dvs_config_load(str(path).encode(), ctypes.byref(c)))
    fields = lambda s: {name: getattr(s, name) for name, _ in s._fields_}
    cam = fields(c.camera)
    phys = fields(c.physic)
    scene = fields(c.scene)
    scene["translation"] = tuple(c.scene.translation)
    scene["rotation"] = tuple(c.scene.rotation)
    return Config(camera=capi.Camera(**cam), physic=capi.Physic(**phys), scene=SceneConfig(**scene),
                  hapke=capi.Hapke(**fields(c.hapke)))


def load_mesh(directory) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """load_mesh_data (src/scene_data.cpp:93-163): (vertices (N,3) f32,
    normals (N,3) f32, indices (M,3) u32); raises ConfigError with the
    reference's message."""
    lib = load_library()
    d = str(directory).encode()
    nv, nt = _I64(), _I64()
    _check(lib.dvs_io_mesh_load(d, None, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(nt)))
    v = np.empty((nv.value, 3), np.float32)
    n = np.empty((nv.value, 3), np.float32)
    t = np.empty((nt.value, 3), np.uint32)
    _check(lib.dvs_io_mesh_load(d, v.ctypes.data, n.ctypes.data, nv.value, t.ctypes.data, nt.value,
                                ctypes.byref(nv), ctypes.byref(nt)))
    return v, n, t


def _frames(fn, path) -> np.ndarray:
    p = str(path).encode()
    count = _I64()
    _check(fn(p, None, 0, ctypes.byref(count)))
    out = np.empty(count.value, capi.FRAME_DTYPE)
    _check(fn(p, out.ctypes.data, count.value, ctypes.byref(count)))
    return out


def load_frames_binary(directory) -> np.ndarray:
    """load_frames_from_binary (src/scene_data.cpp:165-198): FRAME_DTYPE records."""
    return _frames(load_library().dvs_io_frames_binary, directory)


def load_frames_text(path) -> np.ndarray:
    """load_frames_from_text (src/scene_data.cpp:238-283): FRAME_DTYPE records."""
    return _frames(load_library().dvs_io_frames_text, path)


def frame_bytes(linear: np.ndarray, dither: float = 0.0) -> np.ndarray:
    """render_float_to_png_bytes (include/dvs/blenderize.hpp:50-69) of an
    (H, W) radiance frame."""
    img = np.ascontiguousarray(linear, dtype=np.float32)
    out = np.empty(img.shape, np.uint8)
    _check(load_library().dvs_io_frame_bytes(img.ctypes.data, img.shape[1], img.shape[0], float(dither),
                                             out.ctypes.data))
    return out


def write_png(path, gray: np.ndarray) -> None:
    g = np.ascontiguousarray(gray, dtype=np.uint8)
    _check(load_library().dvs_io_write_png(str(path).encode(), g.ctypes.data, g.shape[1], g.shape[0]))


def write_tiff(path, data: np.ndarray) -> None:
    d = np.ascontiguousarray(data, dtype=np.float32)
    _check(load_library().dvs_io_write_tiff(str(path).encode(), d.ctypes.data, d.shape[1], d.shape[0]))


class FrameWriter:
    """Hdf5FrameWriter (include/dvs/frame_writer.hpp:12-77) without the
    dataset budget: frames.h5 with t (int64 us), images (uint8 sRGB preview)
    and depth (float32)."""

    def __init__(self, path, width: int, height: int):
        self.width, self.height = width, height
        self._h = _P()
        _check(load_library().dvs_frames_open(str(path).encode(), width, height, ctypes.byref(self._h)))

    def append(self, time: float, radiance, depth) -> None:
        r = np.ascontiguousarray(radiance, dtype=np.float32).reshape(-1)
        d = np.ascontiguousarray(depth, dtype=np.float32).reshape(-1)
        if r.size != self.width * self.height or d.size != r.size:
            raise ConfigError("invalid frame timestamp or array size")
        _check(load_library().dvs_frames_append(self._h, float(time), r.ctypes.data, d.ctypes.data))

    def count(self) -> int:
        n = _I64()
        _check(load_library().dvs_frames_count(self._h, ctypes.byref(n)))
        return n.value

    def close(self) -> None:
        _check(load_library().dvs_frames_close(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            load_library().dvs_frames_destroy(self._h)
            self._h = None


def read_frames_h5(path) -> dict:
    """frames.h5 back: t (N,) int64, images (N, H, W) uint8, depth (N, H, W) float32."""
    lib = load_library()
    w, h, n = _I32(), _I32(), _I64()
    _check(lib.dvs_frames_h5_read(str(path).encode(), None, None, None, 0, ctypes.byref(w), ctypes.byref(h),
                                  ctypes.byref(n)))
    t = np.empty(n.value, np.int64)
    im = np.empty((n.value, h.value, w.value), np.uint8)
    d = np.empty((n.value, h.value, w.value), np.float32)
    _check(lib.dvs_frames_h5_read(str(path).encode(), t.ctypes.data, im.ctypes.data, d.ctypes.data, n.value,
                                  ctypes.byref(w), ctypes.byref(h), ctypes.byref(n)))
    return {"t": t, "images": im, "depth": d}
