"""dvs_frame_render_device_noalloc (no reference counterpart: the device
frame render for callers that cannot allocate) against
dvs_frame_render_device through the C ABI: bit-identical image and depth on
two calls in a row, the second with fewer poses, and the pose limit."""
from __future__ import annotations

import numpy as np
import pytest

import scenes
from dvsrender4deepspace_amd import capi, workloads

pytestmark = pytest.mark.gpu


def test_noalloc_render_equals_render_device(gpu):
    w = workloads.small_variant(workloads.DAVIS346, size=256, width=96, height=72, craters=20)
    v, n, nodes, root, idx = scenes.workload_mesh(w)
    scene = gpu.Scene(nodes, root, v, n, idx, w.camera, w.hapke)
    npix = w.camera.width * w.camera.height
    fill = np.full(6 * npix, -7.0, np.float32)
    for count in (6, 2):
        poses = w.states(np.linspace(0.0, 0.5 * w.end_time, count))
        print(f"noalloc render: {count} poses of {w.camera.width}x{w.camera.height}", flush=True)
        dposes = gpu.DeviceBuffer(poses.nbytes).upload(poses)
        out = {}
        for name in ("render_device", "render_device_noalloc"):
            img = gpu.DeviceBuffer(fill.nbytes).upload(fill)
            dep = gpu.DeviceBuffer(fill.nbytes).upload(fill)
            getattr(scene, name)(dposes.address, count, img.address, dep.address)
            out[name] = (img.download(np.float32, count * npix), dep.download(np.float32, count * npix))
        (ia, da), (ib, db) = out["render_device"], out["render_device_noalloc"]
        assert ia.tobytes() == ib.tobytes() and da.tobytes() == db.tobytes()
        assert np.count_nonzero(db > 0) > npix // 2 and np.all(ib >= 0)
        host_img, host_dep = scene.render(poses)
        assert host_img.ravel().tobytes() == ib.tobytes() and host_dep.ravel().tobytes() == db.tobytes()
    with pytest.raises(capi.InvalidArgument, match="at most 4096"):
        scene.render_device_noalloc(dposes.address, 4097, img.address, dep.address)
