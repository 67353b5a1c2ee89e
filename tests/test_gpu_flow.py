"""GPU tests of the gradient-flow adjoint (forward/backward C-ABI pair).

The backward pass has no reference counterpart; it is pinned by
  * central finite differences of the forward observable (FP64, h = 1e-4) along
    random su(2) direction fields X: (O(e^{hX}U) - O(e^{-hX}U)) / 2h = X . grad;
  * the plain force kernel: with zero flow steps the gradient of the Wilson
    action is gaugemonomial::derivative (flat-gaugemonomial.hh:105-131);
  * the forward flow itself, which is checked against the oracle's RK3 steps
    (tests/test_gpu_md.py::test_runge_kutta_and_flow_parity).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 13526463


@pytest.fixture(scope="module")
def S():
    import su2_amd
    su2_amd.load_library()
    return su2_amd


def _obs_at(lat, U, X, h, eps, nsteps, obs, ckpt=4):
    lat.set_links(U)
    if h != 0.0:
        lat.set_momenta(X)
        lat.update_gauge(h)
    return lat.flow_forward(eps, nsteps, ckpt, obs)


@pytest.mark.parametrize("obs", [0, 1])
@pytest.mark.parametrize("nsteps,ckpt", [(0, 1), (3, 1), (7, 3)])
def test_flow_gradient_finite_differences(S, oracle, obs, nsteps, ckpt):
    dims = (4, 4, 4, 4, 4)
    U = oracle.random_gauge_trafo(1, dims, oracle.hotstart(1, dims, SEED, 0.6), 654321)
    eps = 0.02
    rng = np.random.default_rng(5)
    with S.Lattice(S.SU2, dims, 2.4) as lat:
        lat.set_links(U)
        lat.flow_forward(eps, nsteps, ckpt, obs)
        g = lat.flow_backward()
        h = 1e-4
        for _ in range(4):
            # a random su(2) direction field of unit norm
            X = rng.normal(size=g.shape)
            X /= np.linalg.norm(X)
            fd = (_obs_at(lat, U, X, h, eps, nsteps, obs) - _obs_at(lat, U, X, -h, eps, nsteps, obs)) / (2 * h)
            an = float(np.sum(X * g))
            assert an == pytest.approx(fd, rel=1e-6, abs=1e-10), (an, fd)


def test_zero_step_action_gradient_is_the_force(S, oracle):
    dims = (4, 4, 4, 4, 4)
    U = oracle.random_gauge_trafo(1, dims, oracle.hotstart(1, dims, SEED), 654321)
    with S.Lattice(S.SU2, dims, 2.4) as lat:
        lat.set_links(U)
        S0 = lat.flow_forward(0.01, 0, 1, S.OBS_WILSON_ACTION)
        assert S0 == pytest.approx(lat.gauge_action(), rel=1e-14)
        g = lat.flow_backward()
        F = lat.derivative(1.0)
    np.testing.assert_allclose(g, F, rtol=0, atol=1e-13 * np.abs(F).max())
    np.testing.assert_allclose(F, oracle.derivative(1, dims, U, 2.4), atol=1e-12)


def test_checkpointing_is_exact_recomputation(S, oracle):
    dims = (6, 6, 6, 6, 4)
    U = oracle.hotstart(1, dims, SEED, 0.5)
    with S.Lattice(S.SU2, dims, 2.4, precision=S.FP32) as lat:
        lat.set_links(U)
        v1 = lat.flow_forward(0.01, 25, 1)
        g1 = lat.flow_backward()
        v2 = lat.flow_forward(0.01, 25, 20)
        g2 = lat.flow_backward()
    assert v1 == v2
    np.testing.assert_array_equal(g1, g2)


def test_fp32_gradient_close_to_fp64(S, oracle):
    dims = (8, 8, 8, 8, 4)
    U = oracle.hotstart(1, dims, SEED, 0.4)
    grads = []
    for prec in (S.FP64, S.FP32):
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.set_links(U)
            lat.flow_forward(0.01, 40, 20)
            grads.append(lat.flow_backward())
    rel = np.abs(grads[1] - grads[0]).max() / np.abs(grads[0]).max()
    assert rel < 1e-4, rel


@pytest.mark.parametrize("obs", [0, 1])
@pytest.mark.parametrize("dims,xi", [((4, 5, 1, 3, 3), 1.0), ((4, 4, 3, 4, 4), 1.3)])
def test_flow_gradient_finite_differences_d3_aniso(S, oracle, obs, dims, xi):
    """The same finite-difference check in d = 3 and, in d = 4, with the
    anisotropic weights (xi != 1) of the force and of its adjoint."""
    U = oracle.random_gauge_trafo(1, dims, oracle.hotstart(1, dims, SEED + 2, 0.6), 654321)
    eps = 0.02
    rng = np.random.default_rng(9)
    with S.Lattice(S.SU2, dims, 2.4, xi=xi) as lat:
        lat.set_links(U)
        lat.flow_forward(eps, 3, 2, obs)
        g = lat.flow_backward()
        h = 1e-4
        for _ in range(3):
            X = rng.normal(size=g.shape)
            X /= np.linalg.norm(X)
            fd = (_obs_at(lat, U, X, h, eps, 3, obs) - _obs_at(lat, U, X, -h, eps, 3, obs)) / (2 * h)
            assert float(np.sum(X * g)) == pytest.approx(fd, rel=1e-6, abs=1e-10)


@pytest.mark.parametrize("prec", [0, 1])
def test_kept_stage_states_match_recomputation(S, oracle, prec, monkeypatch):
    """The backward pass keeps the stage states of a segment from its forward
    recomputation when memory allows; forcing the per-step recomputation
    (SU2GPU_FLOW_RECOMPUTE=1) runs the same kernels on the same inputs, so the
    gradients are identical bit for bit."""
    dims = (8, 8, 8, 8, 4)  # 3-volume 512: the FP32 stages take the specialised step
    U = oracle.hotstart(1, dims, SEED, 0.5)
    grads = []
    for recompute in ("0", "1"):
        monkeypatch.setenv("SU2GPU_FLOW_RECOMPUTE", recompute)
        with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
            lat.set_links(U)
            lat.flow_forward(0.01, 23, 10)
            grads.append(lat.flow_backward())
    np.testing.assert_array_equal(grads[0], grads[1])


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_repeated_backward_is_identical(S, oracle, prec):
    """Two backward passes of one forward give the same gradient bit for bit.
    The forward keeps the last segment's stage states; the first backward
    then recomputes the earlier segments into the same buffers, so the second
    backward must recompute the last segment instead of reusing them."""
    dims = (8, 8, 8, 8, 4)
    U = oracle.hotstart(1, dims, SEED + 4, 0.5)
    with S.Lattice(S.SU2, dims, 2.4, precision=prec) as lat:
        lat.set_links(U)
        lat.flow_forward(0.01, 23, 10)
        g1 = lat.flow_backward()
        g2 = lat.flow_backward()
        g3 = lat.flow_backward()
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(g1, g3)


# ---------------------------------------------------------------------------
# The flow-adjoint configuration at full size: SU(2) 24^4, a start from 50 FP64
# HMC trajectories at beta = 2.4 after a Haar hot start, Luescher RK3 with
# eps = 0.01 to t = 2 (200 steps), checkpoints every 20 steps, observable
# t^2 <E(t)> (clover), 64 random unit su(2) directions for the central finite
# differences (h = 1e-4).  flat-gradient_flow.hh:17-50 (the RK3 step) and
# flat-energy_density.hh:57-164 (E) are the forward functions.

L24, BETA24, EPS24, NSTEP24, CKPT24 = 24, 2.4, 0.01, 200, 20


@pytest.fixture(scope="module")
def therm24(S):
    dims = (L24, L24, L24, L24, 4)
    with S.Lattice(S.SU2, dims, BETA24) as lat:
        lat.haarstart(4242)
        for i in range(50):
            lat.md_update(4242 + i, S.OMF2, 10, momenta=S.MOMENTA_DEVICE)
        return lat.get_links()


@pytest.fixture(scope="module")
def grad24_fp64(S, therm24):
    dims = (L24, L24, L24, L24, 4)
    with S.Lattice(S.SU2, dims, BETA24) as lat:
        lat.set_links(therm24)
        v = lat.flow_forward(EPS24, NSTEP24, CKPT24)
        return v, lat.flow_backward()


def test_flow_24_fp64_finite_differences_64_directions(S, therm24, grad24_fp64):
    dims = (L24, L24, L24, L24, 4)
    v0, g = grad24_fp64
    assert 0.0 < v0 < 1.0 and np.all(np.isfinite(g))
    rng = np.random.default_rng(2024)
    h = 1e-4
    worst = 0.0
    with S.Lattice(S.SU2, dims, BETA24) as lat:
        for _ in range(64):
            X = rng.normal(size=g.shape)
            X /= np.linalg.norm(X)
            fp = _obs_at(lat, therm24, X, h, EPS24, NSTEP24, 0, CKPT24)
            fm = _obs_at(lat, therm24, X, -h, EPS24, NSTEP24, 0, CKPT24)
            fd = (fp - fm) / (2 * h)
            an = float(np.sum(X * g))
            worst = max(worst, abs(an - fd) / abs(fd))
            assert an == pytest.approx(fd, rel=1e-6), (an, fd)
    print(f"24^4 FP64 t^2E = {v0:.10f}: worst relative |grad.X - FD| over 64 directions = {worst:.2e}")


def test_flow_24_fp32_backward_close_to_fp64(S, therm24, grad24_fp64):
    dims = (L24, L24, L24, L24, 4)
    v64, g64 = grad24_fp64
    with S.Lattice(S.SU2, dims, BETA24, precision=S.FP32) as lat:
        lat.set_links(therm24)
        v32 = lat.flow_forward(EPS24, NSTEP24, CKPT24)
        g32 = lat.flow_backward()
    rel_max = np.abs(g32 - g64).max() / np.abs(g64).max()
    rel_l2 = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"24^4 FP32 vs FP64 backward: max-norm relative {rel_max:.2e}, L2 relative {rel_l2:.2e}, "
          f"t^2E {v32:.8f} vs {v64:.8f}")
    assert v32 == pytest.approx(v64, rel=1e-5)
    assert rel_max <= 1e-4
    assert rel_l2 <= 1e-4


def test_flow_24_zero_step_action_gradient_is_the_force(S, oracle, therm24):
    dims = (L24, L24, L24, L24, 4)
    with S.Lattice(S.SU2, dims, BETA24) as lat:
        lat.set_links(therm24)
        S0 = lat.flow_forward(EPS24, 0, CKPT24, S.OBS_WILSON_ACTION)
        assert S0 == pytest.approx(lat.gauge_action(), rel=1e-14)
        g = lat.flow_backward()
        F = lat.derivative(1.0)
    np.testing.assert_allclose(g, F, rtol=0, atol=1e-13 * np.abs(F).max())
    np.testing.assert_allclose(F, oracle.derivative(1, dims, therm24, BETA24), rtol=0,
                               atol=1e-12 * np.abs(F).max())
