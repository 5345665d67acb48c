// Trajectory intervals written straight into device memory (trajectory.hip),
// for the engine's step tables.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "dvs_hip.h"

namespace dvsamd {
// The n - 1 intervals between consecutive times (n >= 2), as
// dvs_trajectory_intervals computes them, written to device memory `out` by
// launches on `stream`.  The trajectory's keys go to the device on first use.
// Throws on invalid arguments or a HIP error.
void trajectory_intervals_on(dvs_trajectory trajectory, const double *times, int64_t n, dvs_optical_interval *out,
                             hipStream_t stream);
}  // namespace dvsamd
