"""In-tree build of libsu2mi355x.so for gfx950 (python -m su2_amd.build).

Every source file is compiled to its own object in parallel (one hipcc
process each), then the objects are linked into the shared library."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsu2mi355x.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["su2_kernels.hip", "su2_flow.hip", "su2_batch.hip", "su2_sweep.hip", "su2_capi.cpp",
           "su2_batch.cpp", "su2_noalloc.hip"]
HEADERS = ["su2_device.hpp", "su2_launch.hpp"]

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in any device code
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}): {cmd[-1]}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile the HIP kernels and the C ABI into su2_amd/libsu2mi355x.so."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJDIR, s + ".o") for s in srcs]
    inc = ["-I", os.path.join(ROOT, "include")]
    cmds = [[hipcc()] + HIPCC_FLAGS + inc + ["-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(srcs, objs)]
    n = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 16)))
    with cf.ThreadPoolExecutor(max_workers=n) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
