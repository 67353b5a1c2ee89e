"""CPU tests that pin the oracle (oracle/su2_oracle.c) before it is trusted.

Three kinds of pins:
  * the reference's own outputs: golden fixtures (tests/golden/golden.npz, written
    by tests/golden/make_golden.py from the reference build) and, where the
    reference build oracle/_ref is present, bit-for-bit agreement on random inputs;
  * the known answers printed by the reference's test program (test.cc:36-43,
    test.cc:80-105) and the standard's mt19937 known answer;
  * what the mathematics fixes: finite-difference force, gauge invariance,
    reversibility, integrator order, exact 2-d plaquettes, flow monotonicity and
    RK3 order.
"""
import math

import numpy as np
import pytest

from oracle.oracle import (EULER, LEAPFROG, LP_LEAPFROG, LP_OMF4, OMF2, OMF4, RUTH, SU2, U1)

D4 = (4, 4, 4, 4, 4)
SEED = 13526463


# ----------------------------------------------------------------- golden ---
def test_mt19937_known_answer(oracle, golden):
    # ISO C++ [rand.predef]: the 10000th invocation of a default-constructed mt19937
    assert oracle.mt_words(5489, 10000)[-1] == 4123659995
    assert golden["mt_seed5489_10000"][0] == 4123659995


def test_normals_match_reference_stream(oracle, golden):
    np.testing.assert_array_equal(oracle.normals(SEED, 1024), golden["normals_seed13526463_1024"])


def test_start_configurations_bit_exact(oracle, golden):
    U = oracle.hotstart(SU2, D4, SEED)
    np.testing.assert_array_equal(U, golden["su2_4444_hotstart"])
    U = oracle.random_gauge_trafo(SU2, D4, U, 654321)
    np.testing.assert_array_equal(U, golden["su2_4444_start"])
    d2 = (8, 1, 1, 8, 2)
    u = oracle.hotstart(U1, d2, SEED)
    np.testing.assert_array_equal(u, golden["u1_88_hotstart"])
    np.testing.assert_array_equal(oracle.random_gauge_trafo(U1, d2, u, 654321), golden["u1_88_start"])


def test_momenta_digest(oracle, golden):
    import hashlib
    P = oracle.initnormal(SU2, D4, SEED)
    dig = np.frombuffer(hashlib.sha256(P.tobytes()).digest(), dtype=np.uint8)
    np.testing.assert_array_equal(dig, golden["su2_4444_momenta_sha256"])


def test_force_action_kinetic_bit_exact(oracle, golden):
    U = golden["su2_4444_start"]
    np.testing.assert_array_equal(oracle.derivative(SU2, D4, U, 2.3), golden["su2_4444_deriv_b2.3"])
    assert oracle.retr_sum_plaq(SU2, D4, U) == golden["su2_4444_retr_plaq"][0]
    P = oracle.initnormal(SU2, D4, SEED)
    assert oracle.kinetic(SU2, D4, P) == golden["su2_4444_kinetic"][0]


@pytest.mark.parametrize("name,kind,n,beta", [("lf10_b2.3", LEAPFROG, 10, 2.3),
                                              ("omf2_20_b2.4", OMF2, 20, 2.4),
                                              ("omf4_10_b2.4", OMF4, 10, 2.4)])
def test_trajectory_bit_exact(oracle, golden, name, kind, n, beta):
    U, info = oracle.md_update(SU2, D4, golden["su2_4444_start"], beta, kind, n, SEED, revtest=True)
    np.testing.assert_array_equal(U, golden[f"traj_{name}_U"])
    ref = golden[f"traj_{name}_info"]
    # single-thread sums: the reference's order at one OpenMP thread
    assert info["dH"] == pytest.approx(ref[0], abs=1e-12)
    assert info["accept"] == bool(ref[1])
    assert info["ddH"] == pytest.approx(ref[2], abs=1e-12)


def test_u1_trajectory_bit_exact(oracle, golden):
    d2 = (8, 1, 1, 8, 2)
    u, info = oracle.md_update(U1, d2, golden["u1_88_start"], 2.0, LEAPFROG, 10, SEED)
    np.testing.assert_array_equal(u, golden["traj_u1_88_lf10_b2_U"])
    assert info["dH"] == pytest.approx(golden["traj_u1_88_lf10_b2_info"][0], abs=1e-12)


def test_energy_density_and_flow_bit_exact(oracle, golden):
    U = golden["su2_4444_start"]
    E, Q = oracle.energy_density(SU2, D4, U)
    assert (E, Q) == tuple(golden["su2_4444_EQ"])
    np.testing.assert_array_equal(oracle.runge_kutta(SU2, D4, U, 2.4, 0.01), golden["su2_4444_rk_eps0.01"])
    _, rows = oracle.gradient_flow(SU2, D4, U, 2.4, 0.1, 0.01)
    np.testing.assert_array_equal(rows, golden["su2_4444_flow_rows"])


def test_hmc_program_rows(oracle, golden):
    _, rows = oracle.hmc_run(SU2, (8, 8, 8, 8, 4), 2.3, SEED, LEAPFROG, 10, 20, N_rev=5)
    ref = golden["hmc_8888_b2.3_lf10_rows"]
    np.testing.assert_array_equal(rows[:, :2], ref[:, :2])  # i, accept
    np.testing.assert_allclose(rows[:, 2], ref[:, 2], rtol=1e-13)  # plaquette
    np.testing.assert_allclose(rows[:, 3], ref[:, 3], atol=1e-10)  # dH (sum order)
    np.testing.assert_allclose(rows[:, 6], ref[:, 6], rtol=1e-12, atol=1e-14)  # Q


# --------------------------------------------------- reference bit-exactness --
CASES = [(SU2, (4, 4, 4, 4, 4)), (SU2, (3, 4, 5, 6, 4)), (SU2, (6, 1, 1, 5, 2)),
         (SU2, (4, 5, 1, 3, 3)), (U1, (4, 4, 4, 4, 4)), (U1, (6, 1, 1, 5, 2)), (U1, (4, 5, 1, 3, 3))]


@pytest.mark.parametrize("grp,dims", CASES)
@pytest.mark.parametrize("xi", [1.0, 1.3])
def test_oracle_equals_reference(oracle, reference, grp, dims, xi):
    Uo = oracle.random_gauge_trafo(grp, dims, oracle.hotstart(grp, dims, 99), 7)
    Ur = reference.random_gauge_trafo(grp, dims, reference.hotstart(grp, dims, 99), 7)
    np.testing.assert_array_equal(Uo, Ur)
    np.testing.assert_array_equal(oracle.derivative(grp, dims, Uo, 2.3, xi),
                                  reference.derivative(grp, dims, Ur, 2.3, xi))
    P = oracle.initnormal(grp, dims, 5)
    for kind in (LEAPFROG, LP_LEAPFROG, OMF4, LP_OMF4, EULER, RUTH, OMF2):
        a = oracle.integrate(grp, dims, Uo, P, 2.3, kind, 4, xi=xi, exponent=7)
        b = reference.integrate(grp, dims, Ur, P, 2.3, kind, 4, xi=xi, exponent=7)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    assert oracle.energy_density(grp, dims, Uo) == reference.energy_density(grp, dims, Ur)
    assert oracle.energy_density(grp, dims, Uo, True, True) == reference.energy_density(grp, dims, Ur, True, True)
    np.testing.assert_array_equal(oracle.runge_kutta(grp, dims, Uo, 2.3, 0.02, xi=xi),
                                  reference.runge_kutta(grp, dims, Ur, 2.3, 0.02, xi=xi))


# ------------------------------------------------------------ known answers --
def _restore(q):
    return q / np.linalg.norm(q)


def test_su2_product_known_answer(oracle):
    # test.cc:36-43: U1 = (0.8+0.3i, 0.1+0.4i), U2 = (0.9, 0.3+0.2i), both restored;
    # "should be: 1, 1.3264 (0.6632,0.184826) (0.293548,0.6632)"
    u1 = oracle.su2_restore(np.array([0.8, 0.3, 0.1, 0.4]))
    u2 = oracle.su2_restore(np.array([0.9, 0.0, 0.3, 0.2]))
    u3 = oracle.su2_mul(u1, u2)
    assert 2 * u3[0] == pytest.approx(1.3264, abs=5e-5)
    np.testing.assert_allclose(u3, [0.6632, 0.184826, 0.293548, 0.6632], atol=5e-7)
    assert np.dot(u3, u3) == pytest.approx(1.0, abs=1e-15)
    # the same product as 2x2 complex matrices
    def mat(q):
        a, b = complex(q[0], q[1]), complex(q[2], q[3])
        return np.array([[a, b], [-b.conjugate(), a.conjugate()]])
    np.testing.assert_allclose(mat(u3), mat(u1) @ mat(u2), atol=1e-15)


def test_u1_topological_charge_known_answer(oracle):
    # test.cc:80-105: cold 4^4 U(1), a few links set by hand; "should be: 0.0126651"
    dims = (4, 4, 4, 4, 4)
    U = oracle.hotstart(U1, dims, 124665, 0.0)
    assert np.all(U == 0)
    pi = math.pi

    def idx(t, x, y, z, mu):
        return (((t * 4 + x) * 4 + y) * 4 + z) * 4 + mu
    U[idx(1, 1, 1, 1, 0)] = 0
    U[idx(1, 1, 1, 1, 1)] = pi / 2
    U[idx(1, 1, 1, 1, 2)] = pi / 2
    U[idx(1, 1, 1, 1, 3)] = 0
    U[idx(0, 1, 1, 1, 0)] = pi
    U[idx(1, 0, 1, 1, 1)] = pi / 2
    U[idx(1, 1, 0, 0, 2)] = pi
    U[idx(1, 1, 1, 0, 3)] = pi / 2
    _, Q = oracle.energy_density(U1, dims, U)
    assert Q == pytest.approx(0.0126651, abs=5e-8)
    _, Q2 = oracle.energy_density(U1, dims, oracle.random_gauge_trafo(U1, dims, U, 654321))
    assert Q2 == pytest.approx(Q, abs=1e-12)


# --------------------------------------------------------------- math pins --
def _su2_exp(p):
    a = np.linalg.norm(p)
    return np.array([math.cos(a), *(math.sin(a) / a * np.array([p[2], p[1], p[0]]))])


def _qmul(u, v):
    # (Re a, Im a, Re b, Im b) product, su2.hh:87-92 as complex arithmetic
    a1, b1 = complex(u[0], u[1]), complex(u[2], u[3])
    a2, b2 = complex(v[0], v[1]), complex(v[2], v[3])
    a = a1 * a2 - b1 * b2.conjugate()
    b = a1 * b2 + b1 * a2.conjugate()
    return np.array([a.real, a.imag, b.real, b.imag])


@pytest.mark.parametrize("grp", [SU2, U1])
@pytest.mark.parametrize("xi", [1.0, 1.7])
def test_force_is_gradient_of_action(oracle, grp, xi):
    """dS/de along U_mu(x) -> exp(e X) U_mu(x) equals X . F (the HMC force)."""
    dims = (3, 4, 2, 5, 4)
    beta = 2.1
    U = oracle.hotstart(grp, dims, 11)
    F = oracle.derivative(grp, dims, U, beta, xi)
    rng = np.random.default_rng(3)
    h = 1e-5
    for link in rng.choice(len(U), 6, replace=False):
        if grp == SU2:
            X = rng.normal(size=3)
            Up, Um = U.copy(), U.copy()
            Up[link] = _qmul(_su2_exp(h * X), U[link])
            Um[link] = _qmul(_su2_exp(-h * X), U[link])
            want = float(X @ F[link])
        else:
            Up, Um = U.copy(), U.copy()
            Up[link] += h
            Um[link] -= h
            want = float(F[link])
        dS = (oracle.action(grp, dims, Up, beta, xi) - oracle.action(grp, dims, Um, beta, xi)) / (2 * h)
        assert dS == pytest.approx(want, rel=1e-6, abs=1e-8)


@pytest.mark.parametrize("grp", [SU2, U1])
def test_gauge_invariance(oracle, grp):
    dims = (4, 4, 2, 6, 4)
    U = oracle.hotstart(grp, dims, 5)
    V = oracle.random_gauge_trafo(grp, dims, U, 77)
    assert oracle.retr_sum_plaq(grp, dims, V) == pytest.approx(oracle.retr_sum_plaq(grp, dims, U), rel=1e-13)
    e1, q1 = oracle.energy_density(grp, dims, U)
    e2, q2 = oracle.energy_density(grp, dims, V)
    assert e2 == pytest.approx(e1, rel=1e-12)
    assert q2 == pytest.approx(q1, abs=1e-12)


def test_cold_start_values(oracle):
    U = oracle.coldstart(SU2, D4)
    assert oracle.retr_sum_plaq(SU2, D4, U) == 6 * 256 * 2
    assert oracle.energy_density(SU2, D4, U) == (-0.0, -0.0) or oracle.energy_density(SU2, D4, U) == (0.0, 0.0)


@pytest.mark.parametrize("kind", [LEAPFROG, OMF2, OMF4])
def test_reversibility(oracle, kind):
    dims = (4, 4, 4, 4, 4)
    U = oracle.hotstart(SU2, dims, 3, 0.3)
    P = oracle.initnormal(SU2, dims, 4)
    U1_, P1 = oracle.integrate(SU2, dims, U, P, 2.4, kind, 8, restore=False)
    U2_, P2 = oracle.integrate(SU2, dims, U1_, -P1, 2.4, kind, 8, restore=False)
    np.testing.assert_allclose(U2_, U, atol=1e-11)
    np.testing.assert_allclose(-P2, P, atol=1e-10)


@pytest.mark.parametrize("kind,order", [(LEAPFROG, 2), (OMF2, 2)])
def test_integrator_order(oracle, kind, order):
    """|dH| ~ dtau^order: doubling the step count divides dH by ~2^order."""
    dims = (4, 4, 4, 4, 4)
    U = oracle.hotstart(SU2, dims, 8, 0.2)
    dh = []
    for n in (8, 16):
        _, info = oracle.md_update(SU2, dims, U, 2.4, kind, n, 21, tau=0.5, compensated=True)
        dh.append(abs(info["dH"]))
    ratio = dh[0] / dh[1]
    assert 2 ** order * 0.6 < ratio < 2 ** order * 1.6, ratio


def test_omf4_runs_n_minus_1_steps(oracle):
    """omf4 loops from i = 1 (integrator.hh:136): n steps integrate n-1 OMF4 blocks
    of size tau/n, i.e. the trajectory ends at tau*(n-1)/n; compare against a
    leapfrog with tiny steps to that time."""
    dims = (2, 2, 2, 2, 4)
    U = oracle.hotstart(SU2, dims, 2, 0.2)
    P = oracle.initnormal(SU2, dims, 3) * 0.3
    n = 10
    Uo, _ = oracle.integrate(SU2, dims, U, P, 1.0, OMF4, n, restore=False)
    Ul, _ = oracle.integrate(SU2, dims, U, P, 1.0, LEAPFROG, 4000, tau=(n - 1) / n, restore=False)
    Ul_full, _ = oracle.integrate(SU2, dims, U, P, 1.0, LEAPFROG, 4000, tau=1.0, restore=False)
    err = np.abs(Uo - Ul).max()
    assert err < 1e-5
    assert np.abs(Uo - Ul_full).max() > 100 * err


def _bessel_ratio(nu, beta):
    from scipy.special import iv
    return iv(nu + 1, beta) / iv(nu, beta)


@pytest.mark.parametrize("grp,beta,nu", [(SU2, 2.0, 1), (U1, 1.5, 0)])
def test_two_dimensional_plaquette_exact(oracle, grp, beta, nu):
    """d = 2: <P> = I_2(b)/I_1(b) for SU(2) (doc/reference-values.txt:3-4 quote
    0.43313(2) at b = 2), <cos> = I_1(b)/I_0(b) for U(1).  HMC chain on the oracle."""
    dims = (32, 1, 1, 32, 2)
    _, rows = oracle.hmc_run(grp, dims, beta, 1234, LEAPFROG, 8, 400)
    plaq = rows[60:, 2]
    exact = _bessel_ratio(nu, beta)
    # blocked error
    blocks = plaq[: len(plaq) // 20 * 20].reshape(20, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(plaq.mean() - exact) < 5 * err + 1e-3, (plaq.mean(), exact, err)
    assert rows[:, 4][-1] > 0.5  # acceptance


def test_flow_decreases_action_and_is_third_order(oracle):
    dims = (4, 4, 4, 4, 4)
    U = oracle.hotstart(SU2, dims, 17, 0.5)
    beta = 2.4
    s0 = oracle.retr_sum_plaq(SU2, dims, U)
    V = U
    prev = s0
    for _ in range(5):
        V = oracle.runge_kutta(SU2, dims, V, beta, 0.02)
        s = oracle.retr_sum_plaq(SU2, dims, V)
        assert s > prev  # sum ReTr P grows: Wilson action decreases
        prev = s

    def flow(eps, t=0.2):
        W = U
        for _ in range(int(round(t / eps))):
            W = oracle.runge_kutta(SU2, dims, W, beta, eps)
        return W
    ref = flow(0.0125)
    e1 = np.abs(flow(0.05) - ref).max()
    e2 = np.abs(flow(0.025) - ref).max()
    assert 5.0 < e1 / e2 < 11.0, e1 / e2


# ---------------------------------------------------------------- Metropolis --
PHILOX_KAT = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
              ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
              ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
               [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]


def test_oracle_philox_known_answers(oracle):
    """The oracle's Philox4x32-10 (the Metropolis random numbers) against the
    published test vectors of Salmon et al., SC11."""
    for ctr, key, want in PHILOX_KAT:
        assert list(oracle.philox(ctr, key)) == want


@pytest.mark.parametrize("grp,dims", [(SU2, (4, 4, 4, 4, 4)), (U1, (8, 1, 1, 6, 2)),
                                      (SU2, (4, 2, 1, 6, 3))])
def test_oracle_sweep_limits(oracle, grp, dims):
    """delta = 0 makes R the identity (every hit accepted, links unchanged up to
    restoreSU); beta = 0 accepts every hit; a sweep is a function of
    (seed, sweep number) and different sweep numbers give different chains."""
    U = oracle.hotstart(grp, dims, 11, 1.0)
    U0, r0 = oracle.sweep(grp, dims, U, 2.0, 5, 0, 0.0, 3)
    assert r0 == (1.0, 1.0)
    assert np.abs(U0 - U).max() < 1e-14
    _, rb = oracle.sweep(grp, dims, U, 0.0, 5, 0, 0.7, 3)
    assert rb == (1.0, 1.0)
    Ua, ra = oracle.sweep(grp, dims, U, 2.0, 5, 3, 0.4, 4)
    Ub, rb2 = oracle.sweep(grp, dims, U, 2.0, 5, 3, 0.4, 4)
    Uc, _ = oracle.sweep(grp, dims, U, 2.0, 5, 4, 0.4, 4)
    np.testing.assert_array_equal(Ua, Ub)
    assert ra == rb2 and 0.0 < ra[0] < 1.0
    assert np.abs(Uc - Ua).max() > 1e-3
    if grp == SU2:
        q = Ua.reshape(-1, 4)
        assert np.abs((q ** 2).sum(1) - 1.0).max() < 1e-14


def test_oracle_sweep_cold_start_frozen_at_large_beta(oracle):
    """At beta -> infinity every hit away from a cold start raises the action
    by a finite amount, so the Metropolis test rejects all of them."""
    dims = (4, 1, 1, 4, 2)
    U = np.zeros(16 * 2)  # U(1) cold start: ReTr P maximal
    Uc, r = oracle.sweep(U1, dims, U, 1e9, 3, 0, 0.5, 5)
    assert r[0] == 0.0
    np.testing.assert_array_equal(Uc, U)


@pytest.mark.parametrize("grp,beta,nu,delta", [(SU2, 2.0, 1, 0.35), (U1, 1.5, 0, 0.6)])
def test_two_dimensional_plaquette_exact_metropolis(oracle, grp, beta, nu, delta):
    """d = 2: the Metropolis chain of the oracle's sweep (the device order and
    Philox stream) reproduces <P> = I_2(b)/I_1(b) (SU(2),
    doc/reference-values.txt:3-4: 0.43313(2) at b = 2) and <cos> = I_1(b)/I_0(b)
    (U(1))."""
    dims = (16, 1, 1, 16, 2)
    U = oracle.hotstart(grp, dims, 77, 1.0)
    norm = 1.0 / (16 * 16) / (2.0 if grp == SU2 else 1.0)
    plaq = []
    for i in range(1200):
        U, rate = oracle.sweep(grp, dims, U, beta, 99, i, delta, 8)
        if i >= 200:
            plaq.append(oracle.retr_sum_plaq(grp, dims, U) * norm)
    plaq = np.array(plaq)
    exact = _bessel_ratio(nu, beta)
    blocks = plaq.reshape(20, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(plaq.mean() - exact) < 4 * err + 5e-4, (plaq.mean(), exact, err)
    assert 0.2 < rate[0] < 0.9


# ------------------------------------------------------------- Haar start ---
# orc_haarstart has no reference counterpart (the reference's hot start is
# random_element, random_element.hh:9-25, whose q0 = cos(alpha) is not Haar
# distributed).  It is pinned by (1) its draw order: link i is the i-th
# 4-tuple of the reference's normal_distribution stream over mt19937(seed)
# (the golden 1024 normals written from the reference build), normalised; and
# (2) what the Haar measure of SU(2) = the uniform measure on S^3 fixes:
# E[q_k^2] = 1/4 and E[q_k^4] = 1/8 for every component, the density
# (2/pi) sqrt(1 - x^2) of each component (Kolmogorov-Smirnov), invariance of
# that law under left multiplication by a fixed group element, and
# independence of neighbouring links.  A start with the reference's
# random_element recipe (E[q0^2] = 1/2), with uniform components normalised
# (E[q^4] = 0.139 instead of 0.125), or with three normals and a derived fourth
# component fails one of these.

def _semicircle_cdf(x):
    x = np.clip(x, -1.0, 1.0)
    return 0.5 + (x * np.sqrt(1.0 - x * x) + np.arcsin(x)) / np.pi


def test_haarstart_draw_order_is_the_reference_normal_stream(oracle, golden):
    z = golden["normals_seed13526463_1024"].reshape(-1, 4)
    U = oracle.haarstart(SU2, D4, SEED)[: z.shape[0]]
    want = z / np.sqrt((z * z).sum(1))[:, None]
    np.testing.assert_allclose(U, want, rtol=0, atol=1e-15)


@pytest.fixture(scope="module")
def haar16(oracle):
    return oracle.haarstart(SU2, (16, 16, 16, 16, 4), 20240601)


def test_haarstart_moments(haar16):
    q = haar16
    n = q.shape[0]
    np.testing.assert_allclose((q * q).sum(1), 1.0, rtol=0, atol=1e-15)
    # uniform on S^3 in R^4: E[x^2] = 1/4, E[x^4] = 3/(4*6) = 1/8, E[x^8] = 105/1920
    sd2 = math.sqrt(1.0 / 8 - 1.0 / 16) / math.sqrt(n)
    sd4 = math.sqrt(105.0 / 1920 - 1.0 / 64) / math.sqrt(n)
    for k in range(4):
        assert abs(q[:, k].mean()) < 6 * 0.5 / math.sqrt(n)
        assert abs((q[:, k] ** 2).mean() - 0.25) < 6 * sd2
        assert abs((q[:, k] ** 4).mean() - 0.125) < 6 * sd4
    # components of one link are uncorrelated
    c = np.corrcoef(q.T)
    assert np.abs(c - np.eye(4)).max() < 6 / math.sqrt(n)


def test_haarstart_component_law_ks(haar16):
    from scipy import stats
    for k in range(4):
        assert stats.kstest(haar16[:, k], _semicircle_cdf).pvalue > 1e-4


def test_haarstart_left_invariance(oracle, haar16):
    """g U has the law of U for a fixed g (the reference's own product,
    su2.hh:68-84, through the oracle's known-answer-pinned su2_mul)."""
    from scipy import stats
    g = np.array([0.3, -0.5, 0.7, 0.1])
    g /= np.linalg.norm(g)
    q = haar16[:65536]
    gq = np.array([oracle.su2_mul(g, x) for x in q])
    for k in range(4):
        assert stats.kstest(gq[:, k], _semicircle_cdf).pvalue > 1e-4
        assert stats.ks_2samp(gq[:, k], haar16[65536:, k]).pvalue > 1e-4


def test_haarstart_links_independent(haar16):
    q = haar16
    n = q.shape[0] // 2
    for k in range(4):
        assert abs(np.corrcoef(q[0::2, k], q[1::2, k])[0, 1]) < 6 / math.sqrt(n)


def test_haarstart_checks_reject_planted_mistakes(oracle):
    """The moment and KS checks above fail for the plausible wrong recipes:
    the reference's random_element start, and normalised uniform components."""
    from scipy import stats
    U = oracle.hotstart(SU2, (16, 16, 16, 16, 4), 20240601, 1.0)  # random_element recipe
    assert abs((U[:, 0] ** 2).mean() - 0.25) > 0.1
    assert stats.kstest(U[:, 0], _semicircle_cdf).pvalue < 1e-10
    rng = np.random.default_rng(1)
    c = rng.uniform(-1, 1, size=(262144, 4))
    c /= np.linalg.norm(c, axis=1)[:, None]
    n = c.shape[0]
    sd4 = math.sqrt(105.0 / 1920 - 1.0 / 64) / math.sqrt(n)
    assert abs((c[:, 0] ** 4).mean() - 0.125) > 6 * sd4


def test_haarstart_u1_uniform_phase(oracle):
    from scipy import stats
    u = oracle.haarstart(U1, (64, 1, 1, 64, 2), SEED)
    assert u.min() >= -math.pi and u.max() < math.pi
    assert stats.kstest(u, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-4


@pytest.mark.parametrize("dims", [(4, 4, 4, 3, 4), (3, 4, 5, 6, 4), (5, 5, 5, 5, 4), (4, 3, 1, 5, 3),
                                  (5, 1, 1, 7, 2), (2, 3, 2, 2, 4), (7, 1, 1, 1, 2)])
def test_sweep_colour_classes_are_independent_sets(oracle, dims):
    """The sweep's colour classes (orc_sweep_colour): for the links of direction
    mu, sites that differ by one step along any nu != mu never share a class,
    so every class is a set of links none of which is in another's staple; and
    with all extents even the classes are the site parities."""
    import ctypes as C
    lib = oracle.lib
    lib.orc_sweep_colour.restype = C.c_int
    lib.orc_sweep_colour.argtypes = [np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), C.c_int]
    Lx, Ly, Lz, Lt, nd = dims
    L = [Lt, Lx, Ly, Lz]
    d = np.asarray(dims, dtype=np.int32)
    import itertools
    cols = {}
    for c in itertools.product(*[range(n) for n in L]):
        for mu in range(nd):
            cols[(c, mu)] = lib.orc_sweep_colour(d, np.asarray(c, dtype=np.int32), mu)
    ncol = max(cols.values()) + 1
    assert ncol <= 3
    for (c, mu), k in cols.items():
        for nu in range(nd):
            if nu == mu or L[nu] == 1:
                continue
            for step in (1, -1):
                n = list(c)
                n[nu] = (n[nu] + step) % L[nu]
                assert cols[(tuple(n), mu)] != k, (c, mu, nu)
    if all(L[k] % 2 == 0 for k in range(nd)):
        for (c, mu), k in cols.items():
            assert k == sum(c) % 2
