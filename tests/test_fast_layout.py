"""CPU check of the launch layout of the specialised SU(2) d=4 MD step: the
(block, thread) -> link map is a bijection onto the 4 V links with periodic
lexicographic neighbours, and the multiply-high division is exact over its
stated range.  Compiles tests/native/fast_layout_check.cpp (host code only runs)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_fast_step_layout(tmp_path):
    exe = tmp_path / "fast_layout_check"
    src = os.path.join(HERE, "native", "fast_layout_check.cpp")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-o", str(exe), src], check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ALL OK")
    assert "32x32x32x32 tblock=1 xwalk=0 wmu=0: ok" in out.stdout
    assert "32x32x32x32 tblock=1 xwalk=1 wmu=0: ok" in out.stdout
    assert "32x32x32x32 tblock=1 xwalk=0 wmu=1: ok" in out.stdout
    assert "32x32x32x32 tblock=0 xwalk=0 wmu=1: ok" in out.stdout
    assert "48x48x48x48 tblock=1 xwalk=1 wmu=0: ok" in out.stdout
    assert "48x48x48x48 tblock=1 xwalk=0 wmu=1: ok" in out.stdout
    assert "8x8x8x8 tblock=1 xwalk=0 wmu=0: ok" in out.stdout
    assert "8x8x8x8 tblock=1 xwalk=0 wmu=1: ok" in out.stdout
    assert "24x24x24x24 tblock=1 xwalk=0 wmu=0: ok" in out.stdout
    # the t-walk shape at the benchmark sizes, at Lt = 8 (one slice per XCD)
    # and on the 32 x 8^3 lattice of its GPU parity test
    for ext in ("32x32x32x32", "48x48x48x48", "24x24x24x24", "8x8x8x8", "32x8x8x8"):
        assert f"{ext} t-walk: ok" in out.stdout
    # the general-exchange t-walk (one direction per wave), Lz = 48, 24 and 8
    for ext in ("48x48x48x48", "24x24x24x24", "32x8x8x8"):
        assert f"{ext} t-walk wmu: ok" in out.stdout
    # both walks with C t-chunks of Lt / C slices and 8 / C 3-volume parts
    for ext in ("32x32x32x32", "48x48x48x48", "64x16x16x16", "40x16x8x8", "32x8x16x16"):
        for wmu in ("", " wmu"):
            for c in (2, 4):
                assert f"{ext} t-walk{wmu} {c} chunks: ok" in out.stdout
    for ext in ("32x32x32x32", "48x48x48x48", "64x16x16x16"):
        for wmu in ("", " wmu"):
            assert f"{ext} t-walk{wmu} 1 chunks: ok" in out.stdout
    # the plaquette t-walk, with the +z lane exchange where 64 % Lz == 0
    for ext in ("32x32x32x32", "8x8x8x8", "32x8x8x8", "16x16x16x16"):
        assert f"{ext} plaquette t-walk z-exchange: ok" in out.stdout
    for ext in ("48x48x48x48", "24x24x24x24"):
        assert f"{ext} plaquette t-walk: ok" in out.stdout
    # the Lt = 8 layout (Tc = 1) with the truncated divide-by-1 constant goes
    # out of range; the library codes that divisor separately
    assert "8x8x8x8 truncated d=1 multiplier" in out.stdout
