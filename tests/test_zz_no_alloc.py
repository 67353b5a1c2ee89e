"""GPU test of su2gpu_retr_sum_plaquettes_noalloc: the plaquette sum of a handle
written to device memory with the scratch of a static device array, against
su2gpu_retr_sum_plaquettes (retr_sum_Wplaquettes, flat-gauge_energy.hpp:31-94)
on the same links.  Two calls in a row (the second on a smaller lattice, so
the scratch holds the first call's larger partial array), both forms,
isotropic / anisotropic / spatial-only, and the size limit."""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 13526463


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime (test-side
    device memory for the output)."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.lib = C.CDLL(name)
                break
            except OSError:
                continue
        else:
            pytest.skip("HIP runtime library not found")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        return p.value

    def read_double(self, p):
        out = C.c_double(math.nan)
        assert self.lib.hipMemcpy(C.byref(out), C.c_void_p(p), 8, 2) == 0  # device to host
        return out.value

    def write_double(self, p, v):
        src = C.c_double(v)
        assert self.lib.hipMemcpy(C.c_void_p(p), C.byref(src), 8, 1) == 0  # host to device

    def free(self, p):
        self.lib.hipFree(C.c_void_p(p))


@pytest.fixture(scope="module")
def S():
    import su2_amd
    su2_amd.load_library()
    return su2_amd


def _one(S, hip, out, grp, dims, prec, xi, aniso, spatial):
    with S.Lattice(grp, dims, 2.4, xi=xi, precision=prec) as lat:
        lat.hotstart(SEED, 0.7)
        want = lat.retr_sum_plaquettes(xi, aniso, spatial)
        hip.write_double(out, math.nan)
        lat.retr_sum_plaquettes_noalloc(out, xi, aniso, spatial)
        lat.synchronize()
        got = hip.read_double(out)
    return got, want


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_noalloc_plaquette_sum_two_calls(S, prec):
    hip = _Hip()
    out = hip.malloc(8)
    try:
        cases = [
            (1, (16, 16, 16, 16, 4), 1.0, False, False),  # larger first
            (1, (8, 8, 8, 8, 4), 1.0, False, False),      # then smaller
            (1, (8, 8, 8, 8, 4), 1.3, True, False),
            (1, (6, 4, 5, 3, 4), 1.0, False, True),       # generic kernel, spatial only
            (0, (12, 1, 1, 10, 2), 1.0, False, False),    # U(1) d = 2
        ]
        for grp, dims, xi, aniso, spatial in cases:
            got, want = _one(S, hip, out, grp, dims, prec, xi, aniso, spatial)
            # both sums are the same compensated reduction of the same partials;
            # the result differs at most in the rounding of hi + lo
            assert math.isfinite(got), (dims, got)
            assert got == pytest.approx(want, rel=3e-16, abs=1e-300), (dims, got, want)
    finally:
        hip.free(out)


def test_noalloc_rejects_beyond_scratch(S):
    hip = _Hip()
    out = hip.malloc(8)
    try:
        # 8192 x 8192 sites in d = 2: 262144 blocks of partials, beyond the
        # 262142 the scratch array holds
        with S.Lattice(S.U1, (8192, 1, 1, 8192, 2), 1.0, precision=S.FP32) as lat:
            with pytest.raises(S.SU2GPUError, match="too large"):
                lat.retr_sum_plaquettes_noalloc(out)
    finally:
        hip.free(out)
