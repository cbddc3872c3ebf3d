"""Build step of libct_hip.so: hipcc for gfx950, in-tree.

Run as `python -m qtctreconstruction_amd.build` or through
__graft_entry__.build(). Objects go to qtctreconstruction_amd/_build/, the
shared library to qtctreconstruction_amd/libct_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libct_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: index arithmetic must round exactly like the reference's
# unfused float expressions; FMAs are written explicitly where allowed.
# -packed-fp32-ops off: packed FP32 VALU forms are avoided on this target.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-command-line-argument", "-Wno-unknown-pragmas",
          "-I", os.path.join(os.path.dirname(PKG), "include")]

SOURCES = ["phantom.hip", "radon.hip", "filter.hip", "filter_wave.hip", "backproject.hip", "pointcloud.hip",
           "filter_gemm.hip", "filter_design.cpp", "metrics.hip", "fileio.cpp", "capi.cpp",
           "batch.cpp", "serve.cpp", "hostpath.cpp", "noalloc.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(warn, file=sys.stderr)
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    return LIB


HOST = os.path.join(PKG, "host")
DRIVER = os.path.join(PKG, "hip_backend_driver")


def build_driver() -> str:
    """The C++ IReconstructionBackend adapter + its driver (host code only, g++)."""
    srcs = [os.path.join(HOST, f) for f in ("HipBackend.cpp", "HipVolumeGenerator.cpp",
                                             "BackendFactory.cpp", "FileIO.cpp",
                                             "hip_backend_driver.cpp")]
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", DRIVER + ".tmp", *srcs,
           "-L", PKG, "-lct_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build())
