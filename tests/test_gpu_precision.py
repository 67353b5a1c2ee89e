"""GPU tests of the FP32 storage emulation of the FP64 form
(su2gpu_set_fp32_storage_emulation), the tool of the FP32 precision-floor
study: the emulated fields are exactly FP32 values after a trajectory, the
emulation changes dH by no more than FP32 storage can, and switching it off
restores the plain FP64 form bit for bit."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 4713


@pytest.fixture(scope="module")
def S():
    import su2_amd
    su2_amd.load_library()
    assert su2_amd.device_count() >= 1
    return su2_amd


def _bound(dims):
    # 20 sqrt(N) 2^-24 over the N = 7 per link stored FP32 degrees of freedom
    V = dims[0] * dims[1] * dims[2] * dims[3]
    return 20.0 * math.sqrt(V * dims[4] * 7) * 2.0 ** -24


def _is_f32(a):
    return bool(np.all(a == a.astype(np.float32).astype(np.float64)))


_START = {}


def _start(S, dims):
    """Links after a Haar start and three FP64 OMF2(8) trajectories."""
    if dims not in _START:
        with S.Lattice(S.SU2, dims, 2.4, precision=S.FP64) as lat:
            lat.haarstart(SEED)
            for i in range(3):
                lat.md_update(SEED + 10 + i, S.OMF2, 8, momenta=S.MOMENTA_DEVICE)
            _START[dims] = lat.get_links()
    return _START[dims]


def _traj(S, dims, prec=None, emu=(False, False), revtest=False):
    """One OMF2(10) trajectory from the common start, Philox momenta."""
    prec = S.FP64 if prec is None else prec
    with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
        lat.set_links(_start(S, dims))
        if emu != (False, False):
            lat.set_fp32_storage_emulation(*emu)
        r = lat.md_update(SEED + 99, S.OMF2, 10, momenta=S.MOMENTA_DEVICE, revtest=revtest)
        return r, lat.get_links(), lat.get_momenta()


@pytest.mark.parametrize("dims", [(8, 8, 8, 8, 4), (4, 8, 6, 12, 4)])
def test_emulated_fields_are_fp32(S, dims):
    r64, U64, P64 = _traj(S, dims)
    rb, Ub, Pb = _traj(S, dims, emu=(True, True))
    rl, Ul, Pl = _traj(S, dims, emu=(True, False))
    rp, Up, Pp = _traj(S, dims, emu=(False, True))
    # the start links are rounded in place, so the links are FP32 values
    # whether the trajectory was accepted or not
    assert _is_f32(Pb) and _is_f32(Pp) and _is_f32(Ub) and _is_f32(Ul)
    assert not _is_f32(P64) and not _is_f32(Pl) and not _is_f32(U64) and not _is_f32(Up)
    b = _bound(dims)
    for r in (rb, rl, rp):
        d = abs(r["deltaH"] - r64["deltaH"])
        print(dims, "emulated dH", r["deltaH"], "FP64", r64["deltaH"], "diff", d, "bound", b)
        assert d < b
    # rounding happened: the emulated trajectories are not the FP64 one
    assert rb["deltaH"] != r64["deltaH"]


def test_emulation_against_fp32_form(S):
    """Links and momenta both rounded: dH within the FP32 bound of the FP32 form
    (whose staple products are FP32) and of the FP64 form; the reversibility
    violation ddH of the emulated form is of FP32 size, not FP64's."""
    dims = (8, 8, 8, 16, 4)
    r64, _, _ = _traj(S, dims, revtest=True)
    rb, _, _ = _traj(S, dims, emu=(True, True), revtest=True)
    r32, _, _ = _traj(S, dims, prec=S.FP32, revtest=True)
    b = _bound(dims)
    print("dH64", r64["deltaH"], "dH emulated", rb["deltaH"], "dH32", r32["deltaH"])
    print("ddH64", r64["deltadeltaH"], "ddH emulated", rb["deltadeltaH"], "ddH32", r32["deltadeltaH"])
    assert abs(rb["deltaH"] - r64["deltaH"]) < b
    assert abs(rb["deltaH"] - r32["deltaH"]) < 2 * b
    assert abs(r64["deltadeltaH"]) < 1e-9
    assert 1e-9 < abs(rb["deltadeltaH"]) < b


def test_emulation_off_is_plain_fp64(S):
    dims = (8, 8, 8, 8, 4)
    r64, U64, P64 = _traj(S, dims)
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP64) as lat:
        lat.set_links(_start(S, dims))
        lat.set_fp32_storage_emulation(True, True)
        lat.set_fp32_storage_emulation(False, False)
        r = lat.md_update(SEED + 99, S.OMF2, 10, momenta=S.MOMENTA_DEVICE)
        assert r["deltaH"] == r64["deltaH"]
        assert np.array_equal(lat.get_links(), U64)


def test_emulation_needs_fp64(S):
    with S.Lattice(S.SU2, (8, 8, 8, 8, 4), 2.4, precision=S.FP32) as lat:
        with pytest.raises(Exception):
            lat.set_fp32_storage_emulation(True, False)
