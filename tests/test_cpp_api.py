"""The C++ API (include/dvs_render_amd.hpp) through its self-test program:
Config, BVHBuilder, the writers and pose conversion, without a GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np

from dvsrender4deepspace_amd import build, eventio

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_cpp_api_host_parts(tmp_path):
    exe = build.API_SELFTEST
    assert exe.exists(), "run __graft_entry__.build() first"
    out = subprocess.run([str(exe), "--cpu", str(GOLDEN / "davis346_profile.yaml"), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("ok ") == 33
    assert (tmp_path / "events.bin").stat().st_size == 26
    assert eventio.read_h5(tmp_path / "events.h5")["t"].tolist() == [1234, 2100]
    rec = eventio.read_bin(tmp_path / "events.bin")
    assert np.array_equal(rec["x"], [2, 4]) and np.array_equal(rec["p"], [1, 0])
