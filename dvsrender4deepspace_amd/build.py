"""Builds the gfx950 HIP library (and, for tests, the CPU oracle) in-tree.

The library is compiled with hipcc for gfx950 only, with FP contraction off
(ray geometry keeps the IEEE operation sequence of its source expressions)
and packed-FP32 code generation disabled.
"""
from __future__ import annotations

import concurrent.futures as futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBDIR = PACKAGE / "lib"
OBJDIR = PACKAGE / "_build"
LIBRARY = LIBDIR / "libdvs_hip.so"

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_SOURCES = ["dvs_kernels.hip", "adaptive.hip", "dvs_capi.hip", "engine.hip", "scan_sort.hip", "readout.hip",
               "trajectory.hip", "raster.hip"]
HOST_SOURCES = ["bvh_host.cpp"]
# Host I/O library (event writers, configuration): HDF5 1.10 C library of the
# image's conda prefix, searched after the system headers.
IO_LIBRARY = LIBDIR / "libdvs_io.so"
HDF5_PREFIX = Path("/opt/conda")


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def build_library(jobs: int = 4) -> Path:
    """Compiles every HIP/C++ source of the library and links libdvs_hip.so."""
    hipcc = _hipcc()
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    steps = []
    objects = []
    for src in HIP_SOURCES:
        obj = OBJDIR / (src + ".o")
        objects.append(obj)
        steps.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJDIR / (src + ".o")
        objects.append(obj)
        steps.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        for f in [pool.submit(_run, s) for s in steps]:
            f.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)])
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_io() -> Path:
    """Compiles libdvs_io.so (event writers, configuration, scene and
    trajectory loaders, GeoTIFF rasters, frame images; host C++)."""
    LIBDIR.mkdir(exist_ok=True)
    if not (HDF5_PREFIX / "include" / "hdf5.h").exists():
        raise RuntimeError(f"HDF5 headers not found under {HDF5_PREFIX}; events.h5 cannot be built")
    tmp = IO_LIBRARY.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{ROOT / 'include'}",
          f"-idirafter{HDF5_PREFIX / 'include'}", str(CSRC / "io_host.cpp"), str(CSRC / "raster_host.cpp"), "-o",
          str(tmp),
          f"-L{HDF5_PREFIX / 'lib'}", "-lhdf5", "-lz", f"-Wl,-rpath,{HDF5_PREFIX / 'lib'}"])
    os.replace(tmp, IO_LIBRARY)
    return IO_LIBRARY


BINDIR = PACKAGE / "bin"
API_SELFTEST = BINDIR / "dvs_api_selftest"


# Programs on the C++ API: the self-test and the bin/dvs and bin/cam front
# ends (apps/main_dvs.cpp, apps/main_cam.cpp).
PROGRAMS = {"dvs_api_selftest": "api_selftest.cpp", "dvs": "main_dvs.cpp", "cam": "main_cam.cpp",
            "cam_blur": "main_cam_blur.cpp", "cam_blur_windows": "main_cam_blur_windows.cpp",
            "sequence": "main_sequence.cpp"}


def _link_program(name: str) -> Path:
    BINDIR.mkdir(exist_ok=True)
    rocm_lib = Path("/opt/rocm/lib")
    exe = BINDIR / name
    tmp = exe.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{CSRC}", str(CSRC / PROGRAMS[name]), "-o",
          str(tmp), f"-L{LIBDIR}", "-ldvs_hip", "-ldvs_io", f"-Wl,-rpath,$ORIGIN/../lib",
          f"-Wl,-rpath-link,{rocm_lib}", f"-Wl,-rpath-link,{HDF5_PREFIX / 'lib'}", f"-Wl,-rpath,{rocm_lib}"])
    os.replace(tmp, exe)
    return exe


def build_api_selftest() -> Path:
    """Links the C++ API self-test (include/dvs_render_amd.hpp) against both
    libraries; run by the tests."""
    return _link_program("dvs_api_selftest")


def build_apps() -> list[Path]:
    """Links bin/dvs, bin/cam, bin/cam_blur, bin/cam_blur_windows and bin/sequence."""
    return [_link_program(name) for name in ("dvs", "cam", "cam_blur", "cam_blur_windows", "sequence")]


def build_oracle() -> Path:
    """Builds the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "liboracle_dvs.so"


if __name__ == "__main__":
    print(build_library(), file=sys.stderr)
    print(build_io(), file=sys.stderr)
    print(build_api_selftest(), file=sys.stderr)
    print(build_apps(), file=sys.stderr)
    if "--with-oracle" in sys.argv:
        print(build_oracle(), file=sys.stderr)
