// The reference's include/dvs/raycast.hpp (DvsRayCaster, CamRayCaster,
// to_optical_state; raycast.hpp:17-86) under its own path, for callers that
// include it directly: the same classes of the drop-in API.
#pragma once
#include "dvs_render_amd.hpp"
