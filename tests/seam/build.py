"""Builds tests/seam/libpj_seam.so (test infrastructure): the step classes of
pj_driver.h driven through stand-ins of the reference's seam types, linked
against the host and device libraries."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIB = ROOT / "pattenmatching_sharedmemory_amd" / "lib"


def build() -> Path:
    out = HERE / "libpj_seam.so"
    tmp = HERE / "libpj_seam.so.tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{ROOT / 'include'}",
           str(HERE / "seam_driver.cc"), f"-L{LIB}", "-lpj_host", "-lprunejuice_hip", f"-Wl,-rpath,{LIB}",
           "-Wl,-rpath,$ORIGIN/../../pattenmatching_sharedmemory_amd/lib", "-o", str(tmp)]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("seam check build failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
