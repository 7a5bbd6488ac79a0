"""Builds the native libraries of the package in-tree.

  lib/libprunejuice_hip.so   HIP kernels + device C ABI (include/prunejuice.h), hipcc for gfx950
  lib/libpj_host.so          host C++ (formats, generators, constraints, step machine; include/prunejuice_host.h)
  lib/patternmatching_hip    command-line driver (patternmatching_binary equivalent)

Run as ``python -m pattenmatching_sharedmemory_amd.build`` or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_ARCH = os.environ.get("PJ_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={HIP_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-result",
    # Packed FP32 instructions stay off in every HIP build of this project.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-result"]

HOST_SOURCES = ["pj_graph.cc", "pj_constraints.cc", "pj_effectiveness.cc", "pj_driver.cc", "pj_host_capi.cc"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {res.returncode})")


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libprunejuice_hip.so"
    tmp = LIB / "libprunejuice_hip.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", str(CSRC / "hip" / "pj_device.hip"),
          str(CSRC / "hip" / "pj_compact.hip"), str(CSRC / "hip" / "pj_stat.hip"), str(CSRC / "hip" / "pj_batch.hip"),
          "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libpj_host.so"
    tmp = LIB / "libpj_host.so.tmp"
    srcs = [str(CSRC / "host" / s) for s in HOST_SOURCES]
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{INCLUDE}", *srcs, f"-L{LIB}", "-lprunejuice_hip",
          "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    cli = LIB / "patternmatching_hip"
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", str(CSRC / "host" / "patternmatching_binary.cc"), f"-L{LIB}",
          "-lpj_host", "-lprunejuice_hip", "-Wl,-rpath,$ORIGIN", "-o", str(cli)])
    return out


def build_all() -> None:
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
