"""Builds the oracle (test infrastructure): oracle/libpj_oracle.so from
oracle/pj_oracle.cc with g++ -O3 -fopenmp.

The reference itself is not built here: its pipeline sources include CUDA
and TBB headers (e.g. src/patternmatching/multiple_label/
multiple_label_lcc0_cpu.tpp:4 includes <cuda.h>, src/totem/totem_comdef.h
includes TBB) that this image does not have, so a reference build would need
stand-in headers and is treated as unbuildable.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


def build() -> Path:
    out = HERE / "libpj_oracle.so"
    tmp = HERE / "libpj_oracle.so.tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", f"-I{ROOT / 'include'}",
           str(HERE / "pj_oracle.cc"), "-o", str(tmp)]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
