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
