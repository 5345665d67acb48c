"""MI355X (gfx950) DVS rendering path: BVH, ray cast + Hapke shading,
photocurrent, log-receptor circuit, comparator crossings, Philox noise and
event compaction/readout, behind the C ABI declared in include/dvs_hip.h."""
from .capi import (EVENT_DTYPE, INTERVAL_DTYPE, NODE_DTYPE, PIXEL_STATE_DTYPE, STATE_DTYPE, Camera, DvsError,
                   Hapke, InvalidArgument, Physic, Scene, Sensor, bvh_build, bvh_validate, device_count,
                   load_library)

__all__ = ["EVENT_DTYPE", "INTERVAL_DTYPE", "NODE_DTYPE", "PIXEL_STATE_DTYPE", "STATE_DTYPE", "Camera", "DvsError",
           "Hapke", "InvalidArgument", "Physic", "Scene", "Sensor", "bvh_build", "bvh_validate", "device_count",
           "load_library"]
