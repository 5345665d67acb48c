"""Host driver of the event path: the horizon loop of bin/dvs
(apps/main_dvs.cpp:92-127) on one sensor.

Horizons are grouped into calls of `batch` intervals: 64 for the native
solver, 200 for adaptive optics, capped by the event capacity bound
(max_events_per_step - 2) * refractory_time / time_resolution, and 1 when the
refractory time is zero (:96-98).  An adaptive window is checked by the
receptor convergence gate at the first window and then every 0.25 s of
trajectory time (:99,110,115); a rejected window is retried at half the
batch, and a rejected one-horizon window is integrated natively (:111-114).
"""
from __future__ import annotations

import numpy as np

from . import capi

NATIVE_BATCH = 64
ADAPTIVE_BATCH = 200
DEFAULT_TOLERANCE = 1e-2   # kDefaultOpticalLogTolerance (include/dvs/optical_sampler.hpp:6)
CHECK_PERIOD = 0.25


def batch_size(physic: capi.Physic, time_resolution: float, adaptive: bool) -> int:
    """Intervals per sensor call (apps/main_dvs.cpp:96-98), with the
    config's float fields widened to double as the driver does."""
    batch = ADAPTIVE_BATCH if adaptive else NATIVE_BATCH
    refractory = float(np.float32(physic.refractory_time))
    if refractory <= 0:
        return 1
    res = float(np.float32(time_resolution))
    cap = int(max(1.0, (physic.max_events_per_step - 2) * refractory / res))
    return max(1, min(batch, cap))


def drive(sensor: capi.Sensor, intervals: np.ndarray, adaptive: bool = False,
          tolerance: float = DEFAULT_TOLERANCE, batch: int | None = None,
          check_period: float = CHECK_PERIOD, drain: bool = True) -> list:
    """Runs every interval through `sensor` in driver-sized calls; returns the
    delivered event arrays in order (the drained backlog last).  A negative
    check_period never checks."""
    ivs = np.ascontiguousarray(np.atleast_1d(intervals), dtype=capi.INTERVAL_DTYPE)
    if batch is None:
        raise ValueError("batch is required (see batch_size)")
    out = []
    step, total = 0, len(ivs)
    next_check = float(ivs["t0"][0]) if total else 0.0
    while step < total:
        window = ivs[step: step + batch]
        if adaptive:
            check = check_period >= 0 and float(window["t0"][0]) >= next_check
            ok, events = sensor.process_adaptive(window, tolerance, check)
            if not ok:
                if batch > 1:
                    batch = max(1, batch // 2)
                    continue
                events = sensor.process(window)
            if check:
                next_check = float(window["t1"][-1]) + check_period
        else:
            events = sensor.process(window)
        out.append(events)
        step += len(window)
    if drain:
        out.append(sensor.drain())
    return out


def blur_schedule(start_time: float, end_time: float, cam_fps: float, blur_resolution: float):
    """bin/cam_blur's exposures (apps/main_cam_blur.cpp:56-72): one per
    1/cam_fps from the first to the last frame time (float), each cut into
    cells of blur_resolution seconds.  Returns a list of (begin, end, times,
    weights): the cells' midpoint times as float32 (the trajectory is
    evaluated at float times) and their weights float((right - left) /
    (end - begin))."""
    start, stop = float(np.float32(start_time)), float(np.float32(end_time))
    interval = 1.0 / float(np.float32(cam_fps))
    res = float(np.float32(blur_resolution))
    out = []
    for exposure in range(int(np.ceil((stop - start) / interval))):
        begin = start + exposure * interval
        end = min(stop, start + (exposure + 1) * interval)
        if end <= begin:
            break
        cells = int(np.ceil((end - begin) / res))
        left = begin + np.arange(cells) * res
        right = np.minimum(end, begin + (np.arange(cells) + 1) * res)
        times = ((left + right) * 0.5).astype(np.float32)
        weights = ((right - left) / (end - begin)).astype(np.float32).astype(np.float64)
        out.append((begin, end, times, weights))
    return out


def blur_window(center: float, blur_resolution: float, exposure: float = 0.050):
    """bin/cam_blur_windows' exposure around `center` seconds
    (apps/main_cam_blur_windows.cpp:55-78): [center - exposure/2, center +
    exposure/2] in ceil(exposure / step - 1e-10) cells, empty cells dropped,
    weights (right - left) / exposure in double.  Returns (begin, end,
    midpoint times (float64), weights)."""
    res = float(np.float32(blur_resolution))
    begin, end = center - exposure * .5, center + exposure * .5
    cells = int(np.ceil(exposure / res - 1e-10))
    left = begin + np.arange(cells) * res
    right = np.minimum(end, begin + (np.arange(cells) + 1) * res)
    keep = right > left
    left, right = left[keep], right[keep]
    return begin, end, (left + right) * 0.5, (right - left) / exposure
