"""Event stream files and the configuration loader (libdvs_io.so, no GPU).

Known answers from the reference's writer test (tests/test_writer.cpp:11-19):
two binary records take 26 bytes; HDF5 times are integer microseconds and
ms_to_idx has the stated boundaries.  h5dump (the HDF5 distribution's own
reader) checks the file independently of this library's reader."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dvsrender4deepspace_amd import capi, eventio

GOLDEN = Path(__file__).resolve().parent / "golden"


def _events(rows):
    ev = np.zeros(len(rows), capi.EVENT_DTYPE)
    for k, (t, x, y, p, d) in enumerate(rows):
        ev[k] = (t, x, y, p, d)
    return ev


def test_binary_writer_record_layout(tmp_path):
    path = tmp_path / "events.bin"
    w = eventio.EventWriter(path, "bin")
    w.append(_events([(.001234, 2, 3, 1, 10), (.0021, 4, 5, 0, 20)]))
    w.close(.003)
    assert path.stat().st_size == 26
    rec = eventio.read_bin(path)
    assert rec["t"].tolist() == [np.float32(.001234), np.float32(.0021)]
    assert rec["x"].tolist() == [2, 4] and rec["y"].tolist() == [3, 5] and rec["p"].tolist() == [1, 0]
    assert rec["depth"].tolist() == [10.0, 20.0]
    assert w.count == 2


def test_hdf5_writer_microseconds_and_ms_index(tmp_path):
    path = tmp_path / "events.h5"
    w = eventio.EventWriter(path, "hdf5")
    w.append(_events([(.001234, 2, 3, 1, 10), (.0021, 4, 5, 0, 20)]))
    w.append(_events([(50.000001, 2, 3, 1, 10), (50.000002, 2, 3, 1, 10)]))
    w.close(50.001)
    d = eventio.read_h5(path)
    assert d["t"].tolist() == [1234, 2100, 50000001, 50000002]
    idx = d["ms_to_idx"]
    assert idx[0] == 0 and idx[1] == 0 and idx[2] == 1 and idx[3] == 2 and idx[50000] == 2 and idx[50001] == 4
    assert len(idx) == 50002
    assert d["x"].tolist() == [2, 4, 2, 2] and d["p"].tolist() == [1, 0, 1, 1]


@pytest.mark.skipif(not Path("/opt/conda/bin/h5dump").exists(), reason="h5dump not available")
def test_hdf5_file_read_by_h5dump(tmp_path):
    path = tmp_path / "events.h5"
    w = eventio.EventWriter(path, "hdf5")
    # a batch out of order is sorted (stably) by time before writing
    w.append(_events([(.0021, 4, 5, 0, 20), (.001234, 2, 3, 1, 10)]))
    w.close(.003)
    out = subprocess.run(["/opt/conda/bin/h5dump", "-d", "/t", "-d", "/x", "-d", "/ms_to_idx", str(path)],
                         capture_output=True, text=True, check=True).stdout
    assert "H5T_STD_I64LE" in out and "1234, 2100" in out
    assert "(0): 2, 4" in out
    assert "(0): 0, 0, 1, 2" in out


def test_hdf5_writer_rejects_disorder_and_late_appends(tmp_path):
    w = eventio.EventWriter(tmp_path / "a.h5", "hdf5")
    w.append(_events([(.002, 1, 1, 1, 0)]))
    with pytest.raises(eventio.IoError, match="globally time ordered"):
        w.append(_events([(.001, 1, 1, 1, 0)]))
    with pytest.raises(eventio.IoError, match="millisecond-index bounds"):
        w.append(_events([(-1.0, 1, 1, 1, 0)]))
    w.close(.01)
    w.close(.01)  # closing twice is a no-op
    with pytest.raises(eventio.IoError, match="closed HDF5"):
        w.append(_events([(.02, 1, 1, 1, 0)]))
    with pytest.raises(eventio.IoError, match="unsupported event format"):
        eventio.EventWriter(tmp_path / "b.x", "tiff")


def test_config_profile_values():
    c = eventio.load_config(GOLDEN / "davis346_profile.yaml")
    assert c.camera.width == 346 and c.camera.height == 260 and c.camera.samples == 16
    assert c.camera.fx == np.float32(226.3802) and c.camera.cy == np.float32(133.7327)
    p = c.physic
    assert p.threshold_pos == np.float32(0.28) and p.refractory_time == np.float32(200e-6)
    assert p.readout_latency == np.float32(18e-9) and p.max_events_per_step == 16 and p.random_seed == 114514
    assert p.photon_shot_noise == 1 and p.reset_thermal_noise == 0
    s = c.scene
    assert s.frame_count == 201 and s.fps == 24.0 and s.time_resolution == np.float32(1e-5)
    assert s.translation == (0.0, 0.0, 0.0) and s.rotation == (0.0, 0.0, 0.0)
    assert c.hapke.B_S0 == np.float32(2.42) and c.hapke.thetap == np.float32(0.453786)


def _variant(tmp_path, edit):
    text = (GOLDEN / "davis346_profile.yaml").read_text()
    path = tmp_path / "c.yaml"
    path.write_text(edit(text))
    return path


@pytest.mark.parametrize("edit,message", [
    (lambda t: t.replace("scene:\n", "scenery:\n"), "Missing top-level section 'scene'"),
    (lambda t: t.replace("  fx: 226.3802\n", ""), "Missing key 'fx' in section 'intrinsic'"),
    (lambda t: t.replace("width: 346", "width: 34x6"), "Invalid value for key 'width' in section 'intrinsic': '34x6'"),
    (lambda t: t.replace("  threshold_min: 0.01\n", "  threshold_min: 0.01\n  leak_rate_hz: 1\n"),
     "removed physic key 'leak_rate_hz'"),
    (lambda t: t.replace("max_events_per_step: 16", "max_events_per_step: 65"),
     "invalid optical/circuit model parameter range"),
    (lambda t: t.replace("dark_current_a: 1e-18", "dark_current_a: -1"), "physic values must be finite and nonnegative"),
    (lambda t: t.replace("frame_count: 201", "frame_count: 1"), "invalid scene or camera range"),
    (lambda t: t.replace("  blur_resolution: 1e-3\n", "  blur_resolution: 1e-3\n  rotation: [1, 2]\n"),
     "scene.rotation must have three numbers"),
    (lambda t: t.replace("phi: 0.3", "phi: 0.9"), "invalid Hapke parameter range"),
    (lambda t: t.replace("intrinsic:\n", "intrinsic: 3\n"), "top-level key 'intrinsic' has a value"),
    (lambda t: t.replace("  width: 346\n", "  width 346\n"), "missing colon"),
    (lambda t: t.replace("  circuit_step: 1e-6\n", "  circuit_step: fast\n"), "invalid physic parameter: circuit_step"),
])
def test_config_errors_match_reference_messages(tmp_path, edit, message):
    with pytest.raises(eventio.ConfigError, match=message):
        eventio.load_config(_variant(tmp_path, edit))


def test_config_extrinsics_and_comments(tmp_path):
    path = _variant(tmp_path, lambda t: t.replace(
        "  blur_resolution: 1e-3\n", "  blur_resolution: 1e-3   # '#' starts a comment\n"
        "  translation: [1.5, -2, 3e2]\n  rotation: [0, 0, 90]\n"))
    c = eventio.load_config(path)
    assert c.scene.translation == (1.5, -2.0, 300.0) and c.scene.rotation == (0.0, 0.0, 90.0)
    assert c.scene.blur_resolution == np.float32(1e-3)
