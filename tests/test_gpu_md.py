"""GPU parity tests of the MD path (run on an MI355X: pytest -m gpu).

Every test calls the HIP library through the C ABI (su2_amd binding) and compares
with the CPU oracle (oracle/su2_oracle.c, itself pinned to the reference by
tests/test_oracle.py) on identical seeded inputs.

Tolerances (FP64 form): per-link and per-momentum 1e-12 for one force / update
sweep, 1e-10 after a full trajectory; dH to 1e-9 relative (both sides use
compensated sums); plaquette and energy density to 1e-12 relative.
"""
import math
import os

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


def fp32_dH_bound(dims):
    """Bound on |dH_FP32 - dH_FP64| for one trajectory: rounding the N stored
    FP32 degrees of freedom (4 per link, 3 per momentum) perturbs H by about
    sqrt(N) * 2^-24 * |dH/dx| per trajectory with |dH/dx| = O(1); the bound
    takes 20 times sqrt(N) 2^-24 (6.5e-3 at 32^4, 1.5e-2 at 48^4), a tenth of
    the earlier fixed 0.05 at 32^4."""
    V = dims[0] * dims[1] * dims[2] * dims[3]
    return 20.0 * math.sqrt(V * dims[4] * 7) * 2.0 ** -24


def _start(oracle, grp, dims, seed=SEED, delta=1.0):
    U = oracle.hotstart(grp, dims, seed, delta)
    return oracle.random_gauge_trafo(grp, dims, U, 654321)


GEOMS = [(1, (4, 4, 4, 4, 4)), (1, (3, 4, 5, 6, 4)), (1, (6, 1, 1, 5, 2)), (1, (4, 5, 1, 3, 3)),
         (0, (4, 4, 4, 4, 4)), (0, (8, 1, 1, 6, 2)), (0, (4, 5, 1, 3, 3))]


def test_start_configuration_bit_exact(S, oracle, golden):
    with S.Lattice(S.SU2, (4, 4, 4, 4, 4), 2.3) as lat:
        lat.hotstart(SEED)
        lat.random_gauge_trafo(654321)
        np.testing.assert_array_equal(lat.get_links(), golden["su2_4444_start"])
    with S.Lattice(S.U1, (8, 1, 1, 8, 2), 2.0) as lat:
        lat.hotstart(SEED)
        lat.random_gauge_trafo(654321)
        np.testing.assert_array_equal(lat.get_links(), golden["u1_88_start"])


@pytest.mark.parametrize("grp,dims", [(1, (4, 4, 4, 4, 4)), (1, (16, 16, 16, 8, 4)),
                                      (0, (8, 1, 1, 6, 2))])
def test_haarstart_bit_exact(S, oracle, grp, dims):
    """The Haar-uniform start (four normal_distribution draws per SU(2) link,
    normalised; a uniform phase per U(1) link) against the oracle's."""
    with S.Lattice(grp, dims, 2.3) as lat:
        lat.haarstart(SEED)
        np.testing.assert_array_equal(lat.get_links(), oracle.haarstart(grp, dims, SEED))


@pytest.mark.parametrize("dims", [(4, 4, 4, 4, 4), (16, 16, 16, 16, 4)])
def test_initnormal_bit_exact(S, oracle, dims):
    """The reference's momentum stream (mt19937 + libstdc++'s polar
    normal_distribution) bit for bit; 16^4 (786k normals) takes the
    multi-threaded second pass of the host draw."""
    with S.Lattice(S.SU2, dims, 2.3) as lat:
        lat.initnormal(SEED)
        np.testing.assert_array_equal(lat.get_momenta(), oracle.initnormal(1, dims, SEED))


@pytest.mark.parametrize("grp,dims", GEOMS)
@pytest.mark.parametrize("xi", [1.0, 1.3])
def test_derivative_parity(S, oracle, grp, dims, xi):
    U = _start(oracle, grp, dims)
    D0 = oracle.initnormal(grp, dims, 3)
    want = oracle.derivative(grp, dims, U, 2.3, xi, fac=0.7, deriv=D0)
    with S.Lattice(grp, dims, 2.3, xi=xi) as lat:
        lat.set_links(U)
        got = lat.derivative(0.7, D0)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("grp,dims", GEOMS)
def test_sums_parity(S, oracle, grp, dims):
    U = _start(oracle, grp, dims)
    P = oracle.initnormal(grp, dims, 4)
    with S.Lattice(grp, dims, 2.3) as lat:
        lat.set_links(U)
        lat.set_momenta(P)
        for sp in (False, True):
            if dims[4] == 2 and sp:
                continue
            assert lat.gauge_energy(sp) == pytest.approx(
                oracle.retr_sum_plaq(grp, dims, U, spatial_only=sp, compensated=True), rel=1e-13, abs=1e-12)
        assert lat.gauge_action() == pytest.approx(oracle.action(grp, dims, U, 2.3, compensated=True), rel=1e-13)
        assert lat.kinetic_energy() == pytest.approx(oracle.kinetic(grp, dims, P, compensated=True), rel=1e-14)
        assert lat.retr_sum_plaquettes(1.4, True) == pytest.approx(
            oracle.retr_sum_plaq(grp, dims, U, xi=1.4, compensated=True), rel=1e-13, abs=1e-12)


@pytest.mark.parametrize("grp,dims", GEOMS)
def test_update_momenta_and_gauge_parity(S, oracle, grp, dims):
    U = _start(oracle, grp, dims)
    P = oracle.initnormal(grp, dims, 4)
    P1 = oracle.update_momenta(grp, dims, U, P, 2.3, 0.05)
    U1_ = oracle.update_gauge(grp, dims, U, P1, 0.05)
    with S.Lattice(grp, dims, 2.3) as lat:
        lat.set_links(U)
        lat.set_momenta(P)
        lat.update_momenta(0.05)
        np.testing.assert_allclose(lat.get_momenta(), P1, atol=1e-12)
        lat.update_gauge(0.05)
        np.testing.assert_allclose(lat.get_links(), U1_, atol=1e-13)


@pytest.mark.parametrize("kind", range(7))
@pytest.mark.parametrize("grp,dims,xi", [(1, (4, 4, 4, 4, 4), 1.0), (1, (4, 3, 2, 5, 4), 1.2),
                                         (0, (8, 1, 1, 6, 2), 1.0), (1, (4, 5, 1, 3, 3), 1.0)])
def test_integrators_parity(S, oracle, kind, grp, dims, xi):
    U = _start(oracle, grp, dims)
    P = oracle.initnormal(grp, dims, 6)
    Uw, Pw = oracle.integrate(grp, dims, U, P, 2.3, kind, 5, xi=xi, exponent=9)
    with S.Lattice(grp, dims, 2.3, xi=xi) as lat:
        lat.set_links(U)
        lat.set_momenta(P)
        lat.integrate(kind, 5, exponent=9)
        Ug, Pg = lat.get_links(), lat.get_momenta()
    tol = 1e-6 if kind in (1, 3) else 1e-11  # lp_*: rounding at 1e-9 may flip a last digit
    np.testing.assert_allclose(Ug, Uw, atol=tol)
    np.testing.assert_allclose(Pg, Pw, atol=tol * 10)


@pytest.mark.parametrize("kind,n,beta", [(0, 10, 2.3), (6, 20, 2.4), (2, 10, 2.4)])
@pytest.mark.parametrize("dims", [(4, 4, 4, 4, 4), (8, 8, 8, 8, 4)])
def test_md_update_parity(S, oracle, kind, n, beta, dims):
    U = _start(oracle, 1, dims)
    Uw, info = oracle.md_update(1, dims, U, beta, kind, n, SEED + 3, revtest=True, compensated=True)
    with S.Lattice(S.SU2, dims, beta) as lat:
        lat.set_links(U)
        r = lat.md_update(SEED + 3, kind, n, revtest=True)
        Ug = lat.get_links()
    assert bool(r["accept"]) == info["accept"]
    assert r["deltaH"] == pytest.approx(info["dH"], rel=1e-9, abs=1e-11)
    assert r["deltadeltaH"] == pytest.approx(info["ddH"], abs=1e-9)
    assert r["kin_old"] == pytest.approx(info["K_old"], rel=1e-14)
    assert r["gauge_old"] == pytest.approx(info["S_old"], rel=1e-13)
    np.testing.assert_allclose(Ug, Uw, atol=1e-10)


def test_md_update_u1(S, oracle):
    dims = (16, 1, 1, 16, 2)
    U = _start(oracle, 0, dims)
    Uw, info = oracle.md_update(0, dims, U, 2.0, 0, 10, SEED, compensated=True)
    with S.Lattice(S.U1, dims, 2.0) as lat:
        lat.set_links(U)
        r = lat.md_update(SEED, 0, 10)
        np.testing.assert_allclose(lat.get_links(), Uw, atol=1e-10)
    assert r["deltaH"] == pytest.approx(info["dH"], rel=1e-9, abs=1e-11)
    assert bool(r["accept"]) == info["accept"]


def test_split_trajectory_reject_restores(S, oracle):
    dims = (4, 4, 4, 4, 4)
    U = _start(oracle, 1, dims)
    P = oracle.initnormal(1, dims, 9)
    with S.Lattice(S.SU2, dims, 2.4) as lat:
        lat.set_links(U)
        r = lat.md_trajectory(P, 6, 5)
        lat.md_finish(False)
        np.testing.assert_array_equal(lat.get_links(), U)
        r2 = lat.md_trajectory(P, 6, 5)
        lat.md_finish(True)
        assert r2["deltaH"] == r["deltaH"]
        Uw, _ = oracle.integrate(1, dims, U, P, 2.4, 6, 5)
        np.testing.assert_allclose(lat.get_links(), Uw, atol=1e-11)


def test_hmc_program_8888(S, oracle, golden):
    """The 8^4 beta=2.3 leapfrog(10), 20-trajectory run of the HMC program
    (hmc.hpp:86-129) driven through the C ABI, against the reference's rows."""
    dims = (8, 8, 8, 8, 4)
    ref = golden["hmc_8888_b2.3_lf10_rows"]
    with S.Lattice(S.SU2, dims, 2.3) as lat:
        lat.hotstart(SEED, 1.0)
        lat.random_gauge_trafo(654321)
        norm = 2.0 / 4 / 3 / lat.V / 2.0
        rate = 0.0
        for i in range(20):
            rev = i > 0 and i % 5 == 0
            r = lat.md_update(SEED + i, 0, 10, revtest=rev)
            plaq = lat.gauge_energy() * norm
            E, Q = lat.energy_density()
            rate += r["accept"]
            assert r["accept"] == ref[i, 1]
            assert plaq == pytest.approx(ref[i, 2], rel=1e-11)
            assert r["deltaH"] == pytest.approx(ref[i, 3], abs=5e-9)
            assert rate / (i + 1) == pytest.approx(ref[i, 4])
            if rev:
                assert r["deltadeltaH"] == pytest.approx(ref[i, 5], abs=1e-9)
            assert Q == pytest.approx(ref[i, 6], rel=1e-9, abs=1e-11)


@pytest.mark.parametrize("grp,dims", [(1, (4, 4, 4, 4, 4)), (1, (4, 5, 1, 3, 3)), (0, (4, 4, 4, 4, 4)),
                                      (0, (8, 1, 1, 6, 2)), (1, (6, 1, 1, 5, 2))])
def test_energy_density_parity(S, oracle, grp, dims):
    U = _start(oracle, grp, dims)
    with S.Lattice(grp, dims, 2.3) as lat:
        lat.set_links(U)
        for clover in (True, False):
            for ss in (False, True):
                e, q = lat.energy_density(clover, ss)
                ew, qw = oracle.energy_density(grp, dims, U, clover, ss)
                assert e == pytest.approx(ew, rel=1e-12, abs=1e-14)
                assert q == pytest.approx(qw, rel=1e-10, abs=1e-13)


def test_u1_charge_known_answer_on_gpu(S, oracle):
    dims = (4, 4, 4, 4, 4)
    U = np.zeros(4 ** 4 * 4)

    def idx(t, x, y, z, mu):
        return (((t * 4 + x) * 4 + y) * 4 + z) * 4 + mu
    pi = math.pi
    for (c, v) in [((1, 1, 1, 1, 1), pi / 2), ((1, 1, 1, 1, 2), pi / 2), ((0, 1, 1, 1, 0), pi),
                   ((1, 0, 1, 1, 1), pi / 2), ((1, 1, 0, 0, 2), pi), ((1, 1, 1, 0, 3), pi / 2)]:
        U[idx(*c)] = v
    with S.Lattice(S.U1, dims, 1.0) as lat:
        lat.set_links(U)
        _, Q = lat.energy_density()
    assert Q == pytest.approx(0.0126651, abs=5e-8)


@pytest.mark.parametrize("grp", [1, 0])
def test_runge_kutta_and_flow_parity(S, oracle, grp, golden):
    dims = (4, 4, 4, 4, 4)
    U = _start(oracle, grp, dims)
    with S.Lattice(grp, dims, 2.4) as lat:
        lat.set_links(U)
        rows = lat.gradient_flow(0.1, 0.01)
        np.testing.assert_array_equal(lat.get_links(), U)  # the flow works on a copy
        lat.runge_kutta(0.01)
        np.testing.assert_allclose(lat.get_links(), oracle.runge_kutta(grp, dims, U, 2.4, 0.01), atol=1e-13)
    _, want = oracle.gradient_flow(grp, dims, U, 2.4, 0.1, 0.01)
    assert rows.shape == want.shape
    np.testing.assert_allclose(rows, want, rtol=1e-10, atol=1e-12)
    if grp == 1:
        np.testing.assert_allclose(rows, golden["su2_4444_flow_rows"], rtol=1e-10, atol=1e-12)


def test_save_load_reference_format(S, oracle, tmp_path):
    dims = (4, 4, 4, 4, 4)
    U = _start(oracle, 1, dims)
    p = str(tmp_path / "conf.0")
    with S.Lattice(S.SU2, dims, 2.3) as lat:
        lat.set_links(U)
        lat.save(p)
    raw = open(p, "rb").read()
    assert len(raw) == 4 ** 4 * 4 * 40  # sizeof(_su2) = 40: size_t N_c, complex a, complex b
    rec = np.frombuffer(raw, dtype=np.float64).reshape(-1, 5)
    assert np.all(np.frombuffer(raw, dtype=np.uint64).reshape(-1, 5)[:, 0] == 2)
    np.testing.assert_array_equal(rec[:, 1:], U)
    with S.Lattice(S.SU2, dims, 2.3) as lat:
        lat.load(p)
        np.testing.assert_array_equal(lat.get_links(), U)
        with pytest.raises(S.SU2GPUError):
            lat.load(str(tmp_path / "missing"))


def test_fp32_form_tracks_fp64(S, oracle):
    dims = (8, 8, 8, 8, 4)
    U = _start(oracle, 1, dims)
    Uw, info = oracle.md_update(1, dims, U, 2.4, 6, 20, SEED, compensated=True)
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.md_update(SEED, 6, 20, revtest=True)
        Ug = lat.get_links()
    assert abs(r["deltaH"] - info["dH"]) < 2e-3
    assert abs(r["deltadeltaH"]) < 2e-3
    assert np.abs(Ug - Uw).max() < 1e-4


@pytest.mark.parametrize("grp,dims,tol", [(0, (16, 1, 1, 12, 2), 1e-5), (1, (6, 4, 1, 8, 3), 1e-4),
                                          (0, (4, 6, 4, 6, 4), 1e-5)])
def test_fp32_form_u1_and_lower_dimensions(S, oracle, grp, dims, tol):
    """The FP32 form on the generic kernels (U(1); SU(2) in d = 3; U(1) in
    d = 4) against the oracle's FP64 trajectory from the same start and host
    momenta."""
    U = oracle.hotstart(grp, dims, SEED + 9, 0.4)
    P = oracle.initnormal(grp, dims, SEED + 10)
    Uw, Pw = oracle.integrate(grp, dims, U, P, 2.0, 6, 8)
    h0 = oracle.action(grp, dims, U, 2.0, compensated=True) + oracle.kinetic(grp, dims, P, compensated=True)
    h1 = oracle.action(grp, dims, Uw, 2.0, compensated=True) + oracle.kinetic(grp, dims, Pw, compensated=True)
    with S.Lattice(grp, dims, 2.0, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.md_trajectory(P, 6, 8)
        lat.md_finish(True)
        Ug = lat.get_links()
    assert np.abs(Ug - Uw).max() < tol
    assert abs(r["deltaH"] - (h1 - h0)) < 1e-3


# specialised SU(2) d=4 layouts (Lx, Ly, Lz, Lt): the plaquette sum walks t
# with Tc = Lt / 8 slices per XCD = 1, 2, 4, with the +-z lane exchange
# (Lz = 8, 16) and without it (Lz = 12); the MD steps walk t from Lt = 32 on,
# in two t-chunks of 16 slices on (16, 8, 8, 32)
_PLAQ_GEOMS = [(8, 8, 8, 8, 4), (16, 16, 16, 16, 4), (8, 8, 8, 32, 4), (8, 8, 12, 32, 4),
               (16, 8, 8, 32, 4)]


@pytest.mark.parametrize("dims", _PLAQ_GEOMS)
@pytest.mark.parametrize("prec", [0, 1])
def test_specialised_plaquette_sums(S, oracle, dims, prec):
    """Plaquette sums on the specialised layouts (retr_sum_Wplaquettes,
    flat-gauge_energy.hpp:31-94) against the oracle's compensated sums of the
    same stored links: all planes, spatial planes only, and the anisotropic
    weights (xi = 1.3)."""
    U = _start(oracle, 1, dims)
    with S.Lattice(S.SU2, dims, 2.3, precision=prec) as lat:
        lat.set_links(U)
        Us = lat.get_links()
        for sp in (False, True):
            assert lat.gauge_energy(sp) == pytest.approx(
                oracle.retr_sum_plaq(1, dims, Us, spatial_only=sp, compensated=True), rel=1e-13)
        assert lat.retr_sum_plaquettes(1.3, True) == pytest.approx(
            oracle.retr_sum_plaq(1, dims, Us, xi=1.3, compensated=True), rel=1e-13)
        assert lat.gauge_action() == pytest.approx(
            oracle.action(1, dims, Us, 2.3, compensated=True), rel=1e-13)


@pytest.mark.parametrize("dims", _PLAQ_GEOMS)
@pytest.mark.parametrize("prec", [0, 1])
def test_trajectory_sums_on_specialised_layouts(S, oracle, dims, prec):
    """The sums of a trajectory on the specialised layouts, where the first
    step (kinetic-energy partials of the start momenta) and the last,
    momentum-only step (those of the final momenta) walk t with Lt >= 32: K_old
    and K_new equal the oracle's kinetic energies of the start and final
    momenta (monomial.hh:33-47).  FP64: the last kernel reunitarises the final
    links (integrator.hh:48,111) and sums their action (md_update.hh:47-51) in
    one launch; S_new equals the oracle's action of the links it leaves, S_old
    the action of the start links, and the final links are unit quaternions."""
    U = _start(oracle, 1, dims)
    P = oracle.initnormal(1, dims, SEED + 5)
    with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
        lat.set_links(U)
        Us = lat.get_links()
        lat.set_momenta(P)
        Ps = lat.get_momenta()
        r = lat.md_trajectory(P, S.OMF2, 6)
        lat.md_finish(True)
        Uend = lat.get_links()
        Pend = lat.get_momenta()
    assert r["kin_old"] == pytest.approx(oracle.kinetic(1, dims, Ps, compensated=True), rel=1e-13)
    assert r["kin_new"] == pytest.approx(oracle.kinetic(1, dims, Pend, compensated=True), rel=1e-13)
    assert r["gauge_old"] == pytest.approx(oracle.action(1, dims, Us, 2.4, compensated=True), rel=1e-13)
    assert r["gauge_new"] == pytest.approx(oracle.action(1, dims, Uend, 2.4, compensated=True), rel=1e-13)
    if prec == 0:
        q = Uend.reshape(-1, 4)
        assert np.abs((q ** 2).sum(1) - 1.0).max() < 1e-14


@pytest.mark.parametrize("dims", [(8, 8, 8, 8, 4), (16, 8, 8, 8, 4), (4, 4, 4, 4, 4)])
@pytest.mark.parametrize("kind", [6, 0])
def test_fp32_trajectory_action_sums(S, oracle, dims, kind):
    """S_old / S_new of an FP32 trajectory (summed inside the first and last
    fused launches where the specialised step applies, by the plaquette kernel
    elsewhere) equal the plaquette-kernel action of the start and final links."""
    U = _start(oracle, 1, dims)
    P = oracle.initnormal(1, dims, SEED + 3)
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP32) as lat:
        lat.set_links(U)
        s_start = lat.gauge_action()
        r = lat.md_trajectory(P, kind, 10)
        lat.md_finish(True)
        s_end = lat.gauge_action()
        Uend = lat.get_links()
    assert r["gauge_old"] == pytest.approx(s_start, rel=1e-12)
    assert r["gauge_new"] == pytest.approx(s_end, rel=1e-12)
    # the final links are unit quaternions (reunitarised in every link update)
    q = Uend.reshape(-1, 4)
    assert np.abs((q.astype(np.float64) ** 2).sum(1) - 1.0).max() < 5e-7


@pytest.mark.parametrize("grp,dims,nm", [(1, (8, 8, 8, 8, 4), 3), (1, (16, 8, 8, 32, 4), 3),
                                         (0, (64, 1, 1, 64, 2), 1)])
def test_device_momenta_statistics(S, grp, dims, nm):
    """K_old of md_update with device momenta is chi^2 / 2 with one degree of
    freedom per momentum component: drawn by their own launch (8^4, U(1)
    d = 2) or inside the first step's walk ((16, 8, 8, 32))."""
    with S.Lattice(grp, dims, 2.4) as lat:
        lat.hotstart(SEED)
        lat.md_update(SEED, 6, 4, momenta=S.MOMENTA_DEVICE)
        r = lat.md_update(SEED + 1, 6, 4, momenta=S.MOMENTA_DEVICE)
        assert math.isfinite(r["deltaH"])
        n = lat.nlinks * nm
        K = r["kin_old"]
        assert abs(2 * K / n - 1.0) < 5 * math.sqrt(2.0 / n)


def test_device_momenta_distribution(S):
    """Device Philox momenta (one Philox4x32-10 block per link, Box-Muller on
    32-bit uniforms): N(0,1) moments and a Kolmogorov-Smirnov test on 786k
    draws, independence of the components, and the FP32 form drawing the same
    momenta as the FP64 form to FP32 accuracy.  A leapfrog step of length
    1e-30 leaves the drawn momenta unchanged, so get_momenta returns them."""
    from scipy import stats
    dims = (16, 16, 16, 16, 4)
    draws = {}
    for prec in (S.FP64, S.FP32):
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.hotstart(SEED)
            lat.md_update(SEED + 5, S.LEAPFROG, 1, tau=1e-30, momenta=S.MOMENTA_DEVICE)
            draws[prec] = lat.get_momenta().reshape(-1, 3)  # (link, component)
    z = draws[S.FP64].ravel()
    n = z.size
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert abs((z ** 4).mean() - 3.0) < 5 * math.sqrt(96.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-4
    # the three components of one link (two Box-Muller pairs of one block)
    P = draws[S.FP64]
    nl = P.shape[0]
    assert np.abs(np.corrcoef(P.T) - np.eye(3)).max() < 5 / math.sqrt(nl)
    # neighbouring links (consecutive Philox counters) are independent
    assert abs(np.corrcoef(P[0::2, 0], P[1::2, 0])[0, 1]) < 5 / math.sqrt(nl / 2)
    d = np.abs(draws[S.FP32] - draws[S.FP64]).ravel()
    assert d.max() < 1e-3
    assert (d < 2e-6).mean() > 0.9999


def test_gauge_invariance_and_reversibility_32(S, oracle):
    """Size-independent properties at the benchmark size 32^4."""
    dims = (32, 32, 32, 32, 4)
    with S.Lattice(S.SU2, dims, 2.4) as lat:
        lat.hotstart(SEED)
        s0 = lat.gauge_action()
        e0, _ = lat.energy_density()
        lat.random_gauge_trafo(654321)
        assert lat.gauge_action() == pytest.approx(s0, rel=1e-12)
        assert lat.energy_density()[0] == pytest.approx(e0, rel=1e-11)
        U = lat.get_links()
        # one force sweep against the oracle at full size
        D = lat.derivative(1.0)
        want = oracle.derivative(1, dims, U, 2.4)
        np.testing.assert_allclose(D, want, atol=1e-12 * np.abs(want).max())
        # reversibility of OMF2 from the (far from equilibrium) hot start
        r = lat.md_update(SEED, 6, 20, revtest=True)
        assert abs(r["deltadeltaH"]) < 1e-7


def test_fp32_specialised_step_32(S, oracle):
    """The FP32 form at the benchmark size 32^4 (specialised step: t-blocked XCD
    layout, z-lane exchange, series exp, FP64 staple sums) against the FP64 form
    from the same links and host momenta, one OMF2(10) trajectory."""
    dims = (32, 32, 32, 32, 4)
    P = oracle.initnormal(1, dims, SEED + 7)
    with S.Lattice(S.SU2, dims, 2.4) as l64:
        l64.hotstart(SEED, 0.3)  # links within 0.3 * 2 pi of the identity
        U = l64.get_links()
        r64 = l64.md_trajectory(P, 6, 10)
        l64.md_finish(True)
        U64 = l64.get_links()
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP32) as l32:
        l32.set_links(U)
        s_start = l32.gauge_action()
        r32 = l32.md_trajectory(P, 6, 10)
        l32.md_finish(True)
        U32 = l32.get_links()
        s_end = l32.gauge_action()
    assert r32["gauge_old"] == pytest.approx(s_start, rel=1e-12)
    assert r32["gauge_new"] == pytest.approx(s_end, rel=1e-12)
    assert r32["kin_old"] == pytest.approx(r64["kin_old"], rel=1e-7)
    dU = np.abs(U32 - U64).max()
    ddh = abs(r32["deltaH"] - r64["deltaH"])
    print(f"32^4 FP32 vs FP64 OMF2(10): max|dU| = {dU:.3e}, |dH32 - dH64| = {ddh:.3e}")
    assert dU < 1e-4
    assert ddh < fp32_dH_bound(dims)


# dims = (Lx, Ly, Lz, Lt, nd)
@pytest.mark.parametrize("dims", [
    (4, 8, 24, 8, 4),    # t-blocked (Lt % 8 == 0), Lz = 24: no z-lane exchange
    (4, 4, 16, 6, 4),    # Lt % 8 != 0: XCD chunks without t-blocking, z-lane exchange
    (8, 4, 24, 12, 4),   # neither
    (2, 4, 32, 8, 4),    # Lx = 2, Ly = 4: x and y neighbours wrap within a block
    (16, 16, 1, 8, 4),   # Lz = 1: the z neighbours are the link itself (lane exchange with itself)
    (2, 64, 64, 8, 4),   # an x plane of 16 blocks: the x-fastest walk of the 3-volume blocks
])
def test_fp32_specialised_step_geometries(S, oracle, dims):
    """The specialised FP32 step on the launch layouts other than 32^4's (every
    3-volume here is a multiple of 256, so the FP32 form takes it) against the
    FP64 form from the same links and host momenta, one OMF2(5) trajectory."""
    assert (dims[0] * dims[1] * dims[2]) % 256 == 0
    P = oracle.initnormal(1, dims, SEED + 11)
    U = oracle.hotstart(1, dims, SEED + 11, 0.3)
    with S.Lattice(S.SU2, dims, 2.4) as l64:
        l64.set_links(U)
        r64 = l64.md_trajectory(P, 6, 5)
        l64.md_finish(True)
        U64 = l64.get_links()
    Uw, _ = oracle.integrate(1, dims, U, P, 2.4, 6, 5)
    np.testing.assert_allclose(U64, Uw, atol=1e-11)
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP32) as l32:
        l32.set_links(U)
        s_start = l32.gauge_action()
        r32 = l32.md_trajectory(P, 6, 5)
        l32.md_finish(True)
        U32 = l32.get_links()
        s_end = l32.gauge_action()
    assert r32["gauge_old"] == pytest.approx(s_start, rel=1e-12)
    assert r32["gauge_new"] == pytest.approx(s_end, rel=1e-12)
    assert np.abs(U32 - U64).max() < 2e-5
    assert abs(r32["deltaH"] - r64["deltaH"]) < 2e-3
    q = U32.reshape(-1, 4)
    assert np.abs((q ** 2).sum(1) - 1.0).max() < 5e-7


@pytest.mark.parametrize("xi", [1.3, 0.8])
def test_fp32_specialised_step_anisotropic(S, oracle, xi):
    """The specialised FP32 step with the anisotropic plane weights (w_t = 1/xi
    on planes touching t, w_s = xi on spatial ones, get_staples.hh:58-79)
    against the oracle's FP64 trajectory from the same links and momenta."""
    dims = (8, 8, 8, 8, 4)
    U = oracle.hotstart(1, dims, SEED + 5, 0.3)
    P = oracle.initnormal(1, dims, SEED + 6)
    Uw, Pw = oracle.integrate(1, dims, U, P, 2.4, 6, 5, xi=xi)
    h0 = oracle.action(1, dims, U, 2.4, xi=xi, compensated=True) + oracle.kinetic(1, dims, P, compensated=True)
    h1 = oracle.action(1, dims, Uw, 2.4, xi=xi, compensated=True) + oracle.kinetic(1, dims, Pw, compensated=True)
    with S.Lattice(S.SU2, dims, 2.4, xi=xi, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.md_trajectory(P, 6, 5)
        lat.md_finish(True)
        Ug = lat.get_links()
    assert np.abs(Ug - Uw).max() < 2e-5
    assert abs(r["deltaH"] - (h1 - h0)) < 2e-3


@pytest.mark.parametrize("dims,xi", [
    ((5, 4, 16, 2, 4), 1.0),   # 3-volume 320, V = 640: lane exchange, last block half full
    ((3, 5, 6, 7, 4), 1.0),    # V = 630: no lane exchange, ragged last block
    ((5, 4, 16, 2, 4), 1.3),   # anisotropic weights
])
def test_fp32_step_other_geometries(S, oracle, dims, xi):
    """The FP32 step on lattices whose 3-volume is not a multiple of 256 (the
    specialised body over the lattice's own geometry) against the oracle's FP64
    trajectory from the same links and host momenta."""
    assert (dims[0] * dims[1] * dims[2]) % 256 != 0
    U = oracle.hotstart(1, dims, SEED + 21, 0.3)
    P = oracle.initnormal(1, dims, SEED + 22)
    Uw, Pw = oracle.integrate(1, dims, U, P, 2.4, 6, 5, xi=xi)
    h0 = oracle.action(1, dims, U, 2.4, xi=xi, compensated=True) + oracle.kinetic(1, dims, P, compensated=True)
    h1 = oracle.action(1, dims, Uw, 2.4, xi=xi, compensated=True) + oracle.kinetic(1, dims, Pw, compensated=True)
    with S.Lattice(S.SU2, dims, 2.4, xi=xi, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.md_trajectory(P, 6, 5)
        lat.md_finish(True)
        Ug = lat.get_links()
    assert np.abs(Ug - Uw).max() < 2e-5
    assert abs(r["deltaH"] - (h1 - h0)) < 2e-3
    q = Ug.reshape(-1, 4)
    assert np.abs((q ** 2).sum(1) - 1.0).max() < 5e-7


def test_fp64_trajectory_parity_32(S, oracle):
    """The plain FP64 form at the benchmark size: one OMF2(20) trajectory on
    32^4 at beta = 2.4 from the seeded hot start with the reference's host
    momenta stream (initnormal, mt19937), against the oracle: max per-link
    deviation <= 1e-10, dH and the plaquette to 1e-9 relative."""
    dims = (32, 32, 32, 32, 4)
    beta = 2.4
    U = oracle.hotstart(1, dims, SEED, 1.0)
    P = oracle.initnormal(1, dims, SEED + 1)
    Uw, Pw = oracle.integrate(1, dims, U, P, beta, 6, 20)
    h_old = oracle.action(1, dims, U, beta, compensated=True) + oracle.kinetic(1, dims, P, compensated=True)
    h_new = oracle.action(1, dims, Uw, beta, compensated=True) + oracle.kinetic(1, dims, Pw, compensated=True)
    with S.Lattice(S.SU2, dims, beta) as lat:
        lat.set_links(U)
        r = lat.md_trajectory(P, 6, 20)
        lat.md_finish(True)
        Ug = lat.get_links()
        plaq = lat.gauge_energy()
    print(f"32^4 FP64 OMF2(20): max|dU| = {np.abs(Ug - Uw).max():.3e}, "
          f"dH gpu {r['deltaH']:.12f} oracle {h_new - h_old:.12f}")
    assert np.abs(Ug - Uw).max() <= 1e-10
    assert r["deltaH"] == pytest.approx(h_new - h_old, rel=1e-9, abs=1e-8)
    assert plaq == pytest.approx(oracle.retr_sum_plaq(1, dims, Uw, compensated=True), rel=1e-9)


def test_host_program_matches_oracle(S, oracle, tmp_path):
    """The C++ host program (hmc.hpp's flow over the adapter header) on a 4^4
    lattice: output lines against the oracle's HMC program, then a restart from
    nconf_counter.txt (base_program.hpp:284-297) continues the chain."""
    import subprocess
    from su2_amd import host
    exe = host.EXE
    assert os.path.exists(exe), "host program not built"
    conf = tmp_path / "confs"
    inp = tmp_path / "in.yaml"
    inp.write_text(f"""geometry:
  L : 4
  T : 4
  ndims : 4
hmc:
  seed: {SEED}
  n_save: 3
  n_meas: 6
  heat: true
  lenghty_conf_name: false
  conf_dir: "{conf}/"
integrator:
  name: omf2
  N_rev: 2
  n_steps: 10
  tau: 1.0
monomials:
  gauge:
    beta: 2.4
omeas:
  res_dir: "{tmp_path / 'omeas'}"
  nstep: 3
  gradient_flow:
    epsilon: 0.02
    tmax: 0.1
""")
    out = subprocess.run([exe, "-f", str(inp)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [ln.split() for ln in out.stdout.splitlines()
             if ln and ln[0].isdigit() and len(ln.split()) == 7]
    _, rows = oracle.hmc_run(1, (4, 4, 4, 4, 4), 2.4, SEED, 6, 10, 6, N_rev=2)
    assert len(lines) == 6
    for ln, r in zip(lines, rows):
        assert int(ln[0]) == r[0] and int(ln[1]) == r[1]
        assert float(ln[2]) == pytest.approx(r[2], rel=1e-12)
        assert float(ln[3]) == pytest.approx(r[3], abs=1e-9)
        if ln[5] != "NA":
            assert float(ln[5]) == pytest.approx(r[5], abs=1e-9)
        assert float(ln[6]) == pytest.approx(r[6], rel=1e-9, abs=1e-12)
    assert (conf / "conf.3").exists() and (conf / "conf.final").exists()
    # online gradient flow of trajectory 3 (if it was accepted) against the oracle
    gf = tmp_path / "omeas" / "gradient_flow.000003"
    if rows[3][1]:
        lines_gf = gf.read_text().splitlines()
        assert lines_gf[0] == "t xi P P_ss Ep Ep_ss Ec Ec_ss Q"
        table = np.array([[float(v) for v in ln.split()] for ln in lines_gf[1:]])
        raw = np.frombuffer((conf / "conf.3").read_bytes(), dtype=np.float64).reshape(-1, 5)[:, 1:]
        _, want = oracle.gradient_flow(1, (4, 4, 4, 4, 4), raw.copy(), 2.4, 0.1, 0.02)
        np.testing.assert_allclose(table, want, rtol=1e-10, atol=1e-12)
    else:
        assert not gf.exists()
    counter = (conf / "nconf_counter.txt").read_text().splitlines()
    assert counter[0] == "heat i path_conf" and counter[1].split()[1] == "3"
    # restart continues at i = 3
    inp.write_text(inp.read_text().replace("heat: true", "restart: true").replace("n_meas: 6", "n_meas: 2"))
    out2 = subprocess.run([exe, "-f", str(inp)], capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stderr
    lines2 = [ln.split() for ln in out2.stdout.splitlines()
              if ln and ln[0].isdigit() and len(ln.split()) == 7]
    assert [int(x[0]) for x in lines2] == [3, 4]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_large_lattice_64(S, prec):
    """64^4 (67 M links; 2.1 GB per FP64 link field): the 32-bit site and link
    indexing and the specialised step's launch layout at a size 16x the
    benchmark's, through size-independent properties of one OMF2(10)
    trajectory with the reversibility test: the action sums of the trajectory
    equal a separate action evaluation, the reversibility violation stays at
    the form's rounding floor, the links stay unitary, and a global gauge
    transformation leaves the action unchanged."""
    dims = (64, 64, 64, 64, 4)
    with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
        lat.hotstart(SEED, 0.3)
        s0 = lat.gauge_action()
        r = lat.md_update(SEED + 1, S.OMF2, 10, revtest=True, momenta=S.MOMENTA_DEVICE)
        assert math.isfinite(r["deltaH"])
        assert r["gauge_old"] == pytest.approx(s0, rel=1e-12)
        if r["accept"]:
            assert r["gauge_new"] == pytest.approx(lat.gauge_action(), rel=1e-12)
        assert abs(r["deltadeltaH"]) < (1e-6 if prec == 0 else 0.5)
        s1 = lat.gauge_action()
        lat.random_gauge_trafo(654321)
        assert lat.gauge_action() == pytest.approx(s1, rel=1e-9 if prec == 0 else 1e-5)
        q = lat.get_links().reshape(-1, 4)
        assert np.abs((q ** 2).sum(1) - 1.0).max() < (1e-14 if prec == 0 else 5e-7)


def test_fp64_and_fp32_trajectory_parity_48(S, oracle):
    """SU(2) 48^4, beta = 2.5: one OMF2(20) trajectory from the seeded Haar hot
    start with the reference's host momentum stream (initnormal, mt19937), in
    both forms against the oracle's FP64 trajectory (md_update.hh:18-81,
    integrator.hh:83-115).  The FP64 form: <= 1e-10 per link, dH and the
    plaquette to 1e-9 relative.  The FP32 form (at 48^4 the step runs with
    non-temporal stores and momentum loads and the general z lane exchange,
    64 % 48 != 0): <= 1e-4 per link, dH within the sqrt(N) rounding bound of
    the FP32 fields, the plaquette sum within 1e-8 relative."""
    dims = (48, 48, 48, 48, 4)
    beta = 2.5
    U = oracle.haarstart(1, dims, SEED)
    P = oracle.initnormal(1, dims, SEED + 1)
    Uw, Pw = oracle.integrate(1, dims, U, P, beta, 6, 20)
    h_old = oracle.action(1, dims, U, beta, compensated=True) + oracle.kinetic(1, dims, P, compensated=True)
    h_new = oracle.action(1, dims, Uw, beta, compensated=True) + oracle.kinetic(1, dims, Pw, compensated=True)
    dh_w = h_new - h_old
    plaq_w = oracle.retr_sum_plaq(1, dims, Uw, compensated=True)
    res = {}
    for prec in (S.FP64, S.FP32):
        with S.Lattice(S.SU2, dims, beta, precision=prec) as lat:
            lat.set_links(U)
            r = lat.md_trajectory(P, 6, 20)
            lat.md_finish(True)
            Ug = lat.get_links()
            plaq = lat.gauge_energy()
        res[prec] = (np.abs(Ug - Uw).max(), r["deltaH"], plaq)
        del Ug
    d64, dh64, pl64 = res[S.FP64]
    d32, dh32, pl32 = res[S.FP32]
    print(f"48^4 OMF2(20): dH oracle {dh_w:.10f}; FP64 max|dU| {d64:.3e}, dH {dh64:.10f}; "
          f"FP32 max|dU| {d32:.3e}, dH {dh32:.6f} (|diff| {abs(dh32 - dh_w):.3e}), "
          f"plaquette rel diff {abs(pl32 - plaq_w) / abs(plaq_w):.3e}")
    assert d64 <= 1e-10
    assert dh64 == pytest.approx(dh_w, rel=1e-9, abs=1e-8)
    assert pl64 == pytest.approx(plaq_w, rel=1e-9)
    assert d32 <= 1e-4
    assert abs(dh32 - dh_w) < fp32_dH_bound(dims)
    assert pl32 == pytest.approx(plaq_w, rel=1e-8)


_NT_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import su2_amd as S
from oracle.oracle import Oracle
orc = Oracle()
# (Lx, Ly, Lz, Lt) = (8, 16, 12, 8): Lz = 12 does not divide 64, the FP32
# step's general z exchange (neighbours in another wave loaded by the lanes
# that need them); (8, 8, 8, 32): 4 t-slices per XCD, the t-walk step;
# (8, 8, 12, 32): both (the FP32 general-exchange t-walk); (16, 8, 8, 32): the
# t-walk in two t-chunks of 16 slices over four 3-volume parts
for dims in [(8, 8, 8, 8, 4), (2, 64, 64, 8, 4), (8, 16, 12, 8, 4), (8, 8, 8, 32, 4),
             (8, 8, 12, 32, 4), (16, 8, 8, 32, 4)]:
    U = orc.hotstart(1, dims, 77, 0.3)
    P = orc.initnormal(1, dims, 78)
    Uw, _ = orc.integrate(1, dims, U, P, 2.4, 6, 5)
    for prec in (S.FP64, S.FP32):
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.set_links(U)
            lat.md_trajectory(P, 6, 5)
            lat.md_finish(True)
            d = float(np.abs(lat.get_links() - Uw).max())
        tol = 1e-11 if prec == S.FP64 else 2e-5
        print(dims, prec, d, flush=True)
        assert d < tol, (dims, prec, d)
print("NT_CHILD_OK", flush=True)
"""


_TW_GEOMS = (("z8", (8, 8, 8, 32, 4)), ("z12", (8, 8, 12, 32, 4)), ("z8p4", (16, 8, 8, 32, 4)),
             ("t40", (16, 8, 8, 40, 4)), ("z16", (8, 16, 16, 32, 4)))
_TW_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import su2_amd as S
from oracle.oracle import Oracle
orc = Oracle()
for tag, dims in %r:
    U = orc.hotstart(1, dims, 91, 0.5)
    P = orc.initnormal(1, dims, 92)
    for name, prec in (("fp32", S.FP32), ("fp64", S.FP64)):
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.set_links(U)
            r = lat.md_trajectory(P, 6, 6)
            lat.md_finish(True)
            np.save(sys.argv[2] + tag + name + ".npy", lat.get_links())
            print("TW_CHILD_OK", tag + name, repr(r["deltaH"]), flush=True)
""" % (_TW_GEOMS,)


def test_t_walk_step_parity(tmp_path, oracle):
    """The P-then-U step as a t-walk (k_md_su2_4d_tw: a thread walks the
    t-slices of its t-chunk, keeping the staple links that neighbouring slices
    share; on by default for lane-exchange lattices with >= 4 slices per XCD,
    FP32 within the Infinity Cache or with the general z exchange) against the
    one-launch-per-link step (SU2GPU_TW=0), and its default two t-chunks of
    Lt / 2 slices (four 3-volume parts, where the 3-volume blocks divide by
    four: the (Lx, Ly, Lz, Lt) = (16, 8, 8, 32) lattice) against one chunk per
    XCD (SU2GPU_TWC=8), and the FP64 walk with one direction per wave
    (k_md_su2_4d_tww, the default) against one per block (SU2GPU_TWW=0,
    k_md_su2_4d_tw): one OMF2(6) trajectory on (8, 8, 8, 32), (8, 8, 12, 32)
    (SU2GPU_ZG=1: the FP32 general exchange, k_md_su2_4d_twz), (16, 8, 8, 32),
    (16, 8, 8, 40) (chunks of 20 and of 5 slices) and (8, 16, 16, 32) in both
    forms, in child processes because the switches are read once
    per process.  FP32: links bit for bit (the same FMA chains), dH within
    1e-9 (the kinetic and action partial sums of the first and last steps, and
    with SU2GPU_TW=0 the plaquette sums, add the same terms in another order);
    FP64:
    the last bit of some values may differ (the compiler contracts the
    update's products per code path), so each against the oracle's trajectory,
    1e-11 per link."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    keys = [t + n for t, _ in _TW_GEOMS for n in ("fp32", "fp64")]
    cases = {"tw": {"SU2GPU_TW": "1"}, "tw8": {"SU2GPU_TW": "1", "SU2GPU_TWC": "8"},
             "blk": {"SU2GPU_TW": "1", "SU2GPU_TWW": "0"}, "one": {"SU2GPU_TW": "0"}}
    outs = {}
    for case, env in cases.items():
        base = str(tmp_path / f"links_{case}_")
        e = {k: v for k, v in os.environ.items() if k not in ("SU2GPU_TW", "SU2GPU_TWC", "SU2GPU_TWW")}
        e.update(SU2GPU_ZG="1", **env)
        out = subprocess.run([sys.executable, "-c", _TW_CHILD, root, base], env=e,
                             capture_output=True, text=True, timeout=300)
        print(case, out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
        dh = {ln.split()[1]: ln.split()[2] for ln in out.stdout.splitlines() if ln.startswith("TW_CHILD_OK")}
        outs[case] = {k: (np.load(base + k + ".npy"), dh[k]) for k in keys}
    for tag, dims in _TW_GEOMS:
        for case in ("tw8", "blk", "one"):
            np.testing.assert_array_equal(outs["tw"][tag + "fp32"][0], outs[case][tag + "fp32"][0])
        for case in ("tw8", "blk", "one"):
            assert abs(float(outs["tw"][tag + "fp32"][1]) - float(outs[case][tag + "fp32"][1])) < 1e-9
        U = oracle.hotstart(1, dims, 91, 0.5)
        P = oracle.initnormal(1, dims, 92)
        Uw, _ = oracle.integrate(1, dims, U, P, 2.4, 6, 6)
        for case in cases:
            d = float(np.abs(outs[case][tag + "fp64"][0] - Uw).max())
            print(dims, case, "FP64 max|dU| vs oracle", d)
            assert d < 1e-11


_GEN_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import su2_amd as S
from oracle.oracle import Oracle
orc = Oracle()
for tag, dims in %r:
    U = orc.hotstart(1, dims, 93, 0.5)
    for name, prec in (("fp32", S.FP32), ("fp64", S.FP64)):
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.set_links(U)
            r1 = lat.md_update(501, S.OMF2, 6, momenta=S.MOMENTA_DEVICE)
            r2 = lat.md_update(502, S.OMF2, 6, momenta=S.MOMENTA_DEVICE, revtest=True)
            np.save(sys.argv[2] + tag + name + "_U.npy", lat.get_links())
            np.save(sys.argv[2] + tag + name + "_P.npy", lat.get_momenta())
            vals = [r1[k] for k in ("deltaH", "kin_old", "kin_new")] + \
                   [r2[k] for k in ("deltaH", "kin_old", "kin_new", "deltadeltaH")]
            print("GEN_CHILD_OK", tag + name, " ".join(repr(v) for v in vals), flush=True)
""" % (_TW_GEOMS,)


def test_momenta_drawn_in_first_step_parity(tmp_path):
    """Device momenta of md_update (the Philox stream of k_momenta_philox)
    drawn inside the trajectory's first step where that walks t
    (k_md_su2_4d_tw_acc / _twz_acc with GEN: link (s, mu) draws counter
    (4 s + mu, 0, traj), the momenta kernel's) against their own launch before
    the first step (SU2GPU_GENFUSE=0): two OMF2(6) trajectories, the second
    with the reversibility test, on (Lx, Ly, Lz, Lt) = (8, 8, 8, 32),
    (8, 8, 12, 32) (SU2GPU_ZG=1: the general-exchange walks of both forms) and
    (16, 8, 8, 32), in child processes because the switch is read once per
    process.  FP32: links, final momenta, dH, K_old, K_new and ddH bit for bit
    (the same arithmetic on the same momenta); FP64: within the last bits (the
    compiler contracts the update's products per kernel), links and momenta
    1e-12, sums 1e-9."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = {}
    for case in ("1", "0"):
        base = str(tmp_path / f"gen{case}_")
        e = {k: v for k, v in os.environ.items() if not k.startswith("SU2GPU_")}
        e.update(SU2GPU_GENFUSE=case, SU2GPU_ZG="1")
        out = subprocess.run([sys.executable, "-c", _GEN_CHILD, root, base], env=e,
                             capture_output=True, text=True, timeout=300)
        print(case, out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
        vals = {ln.split()[1]: [float(x) for x in ln.split()[2:]] for ln in out.stdout.splitlines()
                if ln.startswith("GEN_CHILD_OK")}
        outs[case] = (base, vals)
    for tag, _ in _TW_GEOMS:
        for name in ("fp32", "fp64"):
            k = tag + name
            a = [np.load(outs[c][0] + k + f) for c in ("1", "0") for f in ("_U.npy", "_P.npy")]
            va, vb = outs["1"][1][k], outs["0"][1][k]
            if name == "fp32":
                np.testing.assert_array_equal(a[0], a[2])
                np.testing.assert_array_equal(a[1], a[3])
                assert va == vb
            else:
                assert float(np.abs(a[0] - a[2]).max()) < 1e-12
                assert float(np.abs(a[1] - a[3]).max()) < 1e-12
                np.testing.assert_allclose(va, vb, rtol=0, atol=1e-9)


@pytest.mark.parametrize("env", [{"SU2GPU_NT": "1", "SU2GPU_XWALK": "1"},
                                 {"SU2GPU_NT": "1", "SU2GPU_XWALK": "0"},
                                 {"SU2GPU_NT": "2"},
                                 {"SU2GPU_NT": "0", "SU2GPU_ZG": "1"}],
                         ids=["nt_xwalk", "nt", "nt2", "zg"])
def test_nontemporal_store_step_parity(env):
    """The specialised step with its non-temporal stores (quat<float> and
    quat<double> sto_nt, with the FP32 general z exchange where Lz does not
    divide 64), or with the general exchange alone, forced on small lattices,
    in a child process because the switches are read once per process: both
    forms against the oracle's FP64 trajectory from the same links and host
    momenta."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    e = dict(os.environ, **env)
    out = subprocess.run([sys.executable, "-c", _NT_CHILD, root], env=e, capture_output=True,
                         text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "NT_CHILD_OK" in out.stdout
