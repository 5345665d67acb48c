// Umbrella header of the drop-in API, named as the reference's
// include/dvs_render.hpp (BugBubbles/DvsRender4DeepSpace, include/dvs_render.hpp:
// 1-18) so that a caller switches by pointing its include path here.  It
// brings the classes and functions of the DVS path: BVHBuilder, the scene and
// configuration types, Config, DeviceScene, DvsRayCaster, CamRayCaster,
// EventSensor, OpticalSampler, EventReadout, CameraTrajectory,
// advance_pixel / receptor_log_at / contrast_at, hapke_reflectance /
// hapke_directional_reflectance / invert_hapke_if, the scene bundle and
// trajectory loaders, render_float_to_png_bytes and the event stream
// writers, all in namespace dvs over the C ABIs dvs_hip.h (libdvs_hip.so)
// and dvs_io.h (libdvs_io.so).
#pragma once
#include "dvs_render_amd.hpp"
