"""Build of the C++ HMC host program (su2_amd/host/su2_hmc_gpu) linked against
libsu2mi355x.so through the adapter header include/su2_mi355x/hmc_gpu.hh."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
EXE = os.path.join(HERE, "su2_hmc_gpu")


def build(verbose: bool = False) -> str:
    from su2_amd import build as b
    if not os.path.exists(b.LIB):
        b.build(verbose)
    cmd = [b.hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "hmc_main.cc"), "-L", PKG, "-lsu2mi355x",
           "-Wl,-rpath,$ORIGIN/..", "-o", EXE + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("host program build failed")
    os.replace(EXE + ".tmp", EXE)
    return EXE
