"""Equilibrium physics of the GPU HMC chain against the reference's published
plaquette values (run on an MI355X: pytest -m gpu).

doc/reference-values.txt:6-9 quotes, for SU(2) in 3+1 d on a 10^4 lattice,
<P> = 0.24320(5) at beta = 1.0 and 0.50133(12) at beta = 2.0, with P the
plaquette normalised as in the HMC program's output (base_program.hpp:303-305,
hmc.hpp:106: gauge_energy * 2 / (d (d-1) V Nc)).  A chain of the GPU md_update reproduces
them within four combined standard errors, in the FP64 and in the FP32 form.
(The U(1) 3+1 d value of the same file, 0.19658(10) at beta = 1, is not
used: it belongs to another normalisation than the Wilson action
S = -beta sum cos(theta_P) of u1.hh, whose leading strong-coupling plaquette
I1(1)/I0(1) is 0.446 and whose HMC chain in the oracle gives about 0.57.)
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 13526463


@pytest.fixture(scope="module")
def S():
    import su2_amd
    su2_amd.load_library()
    assert su2_amd.device_count() >= 1
    return su2_amd


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("beta,ref,err", [(1.0, 0.24320, 0.00005), (2.0, 0.50133, 0.00012)])
def test_su2_4d_plaquette_reference_values(S, prec, beta, ref, err):
    dims = (10, 10, 10, 10, 4)
    ntherm, nmeas, nbin = 200, 2000, 40
    with S.Lattice(S.SU2, dims, beta, precision=prec) as lat:
        lat.hotstart(SEED)
        norm = 2.0 / (4 * 3 * lat.V * 2.0)
        for i in range(ntherm):
            lat.md_update(SEED + i, S.OMF2, 10, momenta=S.MOMENTA_DEVICE)
        acc = 0
        P = np.empty(nmeas)
        for i in range(nmeas):
            r = lat.md_update(SEED + ntherm + i, S.OMF2, 10, momenta=S.MOMENTA_DEVICE)
            acc += r["accept"]
            P[i] = lat.gauge_energy() * norm
    bins = P.reshape(nbin, -1).mean(axis=1)
    mean = bins.mean()
    e = bins.std(ddof=1) / math.sqrt(nbin)
    print(f"beta={beta} prec={prec}: <P> = {mean:.5f} +- {e:.5f} (ref {ref} +- {err}), "
          f"acceptance {acc / nmeas:.3f}")
    assert acc / nmeas > 0.5
    assert abs(mean - ref) < 4.0 * math.sqrt(e * e + err * err)


@pytest.mark.parametrize("beta,ref,err,delta", [(1.0, 0.24320, 0.00005, 0.5),
                                                (2.0, 0.50133, 0.00012, 0.3)])
def test_su2_4d_plaquette_reference_values_metropolis(S, beta, ref, err, delta):
    """The same published plaquettes from a chain of the device Metropolis
    sweep (flat-sweep.hh:64-164; 10 hits per link), FP64."""
    dims = (10, 10, 10, 10, 4)
    ntherm, nmeas, nbin = 200, 2000, 40
    with S.Lattice(S.SU2, dims, beta) as lat:
        lat.hotstart(SEED)
        norm = 2.0 / (4 * 3 * lat.V * 2.0)
        for i in range(ntherm):
            lat.sweep(SEED, i, delta, 10)
        P = np.empty(nmeas)
        rate = 0.0
        for i in range(nmeas):
            rate += lat.sweep(SEED, ntherm + i, delta, 10)[0]
            P[i] = lat.gauge_energy() * norm
    bins = P.reshape(nbin, -1).mean(axis=1)
    mean = bins.mean()
    e = bins.std(ddof=1) / math.sqrt(nbin)
    print(f"Metropolis beta={beta}: <P> = {mean:.5f} +- {e:.5f} (ref {ref} +- {err}), "
          f"acceptance {rate / nmeas:.3f}")
    assert 0.2 < rate / nmeas < 0.9
    assert abs(mean - ref) < 4.0 * math.sqrt(e * e + err * err)
