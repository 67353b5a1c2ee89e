"""Known-answer test of the Philox4x32-10 block function (su2_launch.hpp) that
draws the device momenta and the Metropolis random numbers: the three
published test vectors of Salmon et al., SC11 (Random123's kat_vectors for
philox4x32_10).  Compiles tests/native/philox_kat.cpp (host code only runs)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

KAT = ["6627e8d5 e169c58d bc57ac4c 9b00dbd8",   # counter 0, key 0
       "408f276d 41c83b0e a20bc7c6 6d5451fd",   # all ones
       "d16cfe09 94fdcceb 5001e420 24126ea1"]   # digits of pi


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_philox4x32_10_known_answers(tmp_path):
    exe = tmp_path / "philox_kat"
    src = os.path.join(HERE, "native", "philox_kat.cpp")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-o", str(exe), src], check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split("\n")[:3] == KAT
