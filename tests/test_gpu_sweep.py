"""GPU tests of the Metropolis sweep (su2gpu_sweep, flat-sweep.hh:64-164).

The device sweep cannot reproduce the reference's draws (per-thread mt19937
engines walking t-slices in sequence); it is checked against the oracle's
restatement with the same visiting order and Philox stream (oracle/su2_oracle.c
orc_sweep, itself pinned in tests/test_oracle.py by the Philox known answers,
its limits and the exact 2-d plaquettes), and against the reference's
published 4-d plaquettes in tests/test_gpu_physics.py.
"""
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


# dims = (Lx, Ly, Lz, Lt, nd)
@pytest.mark.parametrize("grp,dims,xi", [(1, (4, 4, 4, 4, 4), 1.0), (1, (4, 4, 4, 4, 4), 1.3),
                                         (1, (4, 2, 1, 6, 3), 1.0), (0, (8, 1, 1, 6, 2), 1.0),
                                         (0, (4, 4, 2, 4, 4), 1.2)])
def test_sweep_parity_with_oracle(S, oracle, grp, dims, xi):
    U = oracle.hotstart(grp, dims, SEED, 1.0)
    with S.Lattice(grp, dims, 2.2, xi=xi) as lat:
        lat.set_links(U)
        Uw = U
        for w in range(2):
            r = lat.sweep(SEED + 1, w, 0.4, 5)
            Uw, rw = oracle.sweep(grp, dims, Uw, 2.2, SEED + 1, w, 0.4, 5, xi=xi)
            assert r == pytest.approx(rw, abs=1e-12)
        np.testing.assert_allclose(lat.get_links(), Uw, atol=1e-12)


def test_sweep_fp32_tracks_oracle(S, oracle):
    dims = (4, 4, 4, 4, 4)
    U = oracle.hotstart(1, dims, SEED, 1.0)
    with S.Lattice(S.SU2, dims, 2.2, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.sweep(SEED + 1, 0, 0.4, 5)
        Ug = lat.get_links()
    Uw, rw = oracle.sweep(1, dims, U, 2.2, SEED + 1, 0, 0.4, 5)
    assert abs(r[0] - rw[0]) < 0.01
    # a hit decided differently in FP32 changes its link completely: compare the
    # links whose hit decisions agreed (all but a few)
    d = np.abs(Ug - Uw).reshape(-1, 4).max(axis=1)
    assert (d < 1e-5).mean() > 0.99
    q = Ug.reshape(-1, 4).astype(np.float64)
    assert np.abs((q ** 2).sum(1) - 1.0).max() < 5e-7


# odd extents: three colour classes per direction over host-built site lists
@pytest.mark.parametrize("grp,dims,xi", [(1, (4, 4, 4, 3, 4), 1.0), (1, (3, 4, 5, 6, 4), 1.0),
                                         (1, (5, 5, 5, 5, 4), 1.3), (1, (4, 3, 1, 5, 3), 1.0),
                                         (0, (5, 1, 1, 7, 2), 1.0), (0, (3, 4, 2, 4, 4), 1.0)])
def test_sweep_odd_extents_parity_with_oracle(S, oracle, grp, dims, xi):
    U = oracle.hotstart(grp, dims, SEED + 3, 1.0)
    with S.Lattice(grp, dims, 2.2, xi=xi) as lat:
        lat.set_links(U)
        Uw = U
        for w in range(2):
            r = lat.sweep(SEED + 4, w, 0.4, 5)
            Uw, rw = oracle.sweep(grp, dims, Uw, 2.2, SEED + 4, w, 0.4, 5, xi=xi)
            assert r == pytest.approx(rw, abs=1e-12)
        np.testing.assert_allclose(lat.get_links(), Uw, atol=1e-12)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4, 4), (3, 4, 5, 6, 4)])
def test_sweep_fp32_deterministic_bound(S, oracle, dims):
    """At beta = 0 every hit is accepted in both forms (dS = 0), so no accept
    decision can differ and every FP32 link must stay within a rounding bound of
    the oracle's FP64 link: five hits, each one FP64 product and restoreSU
    rounded once to FP32 (|error| <= 5 * 2^-24 per component, plus the FP32
    start), bound 1e-6."""
    U = oracle.hotstart(1, dims, SEED + 5, 1.0)
    with S.Lattice(S.SU2, dims, 0.0, precision=S.FP32) as lat:
        lat.set_links(U)
        r = lat.sweep(SEED + 6, 0, 0.4, 5)
        Ug = lat.get_links()
    Uw, rw = oracle.sweep(1, dims, U, 0.0, SEED + 6, 0, 0.4, 5)
    assert r[0] == 1.0 and rw[0] == 1.0
    assert np.abs(Ug - Uw).max() < 1e-6


def test_host_program_metropolis(S, oracle, tmp_path):
    """The C++ host program in its Metropolis mode (metropolis.hpp's loop over
    the adapter's flat_spacetime::sweep, which keys the device stream with two
    draws of the caller's mt19937(seed + i)) against the same chain driven
    through the C ABI from Python; rows "i P E Q P_ss E_ss Q_ss"."""
    import os
    import subprocess
    from su2_amd import host
    exe = host.EXE
    assert os.path.exists(exe), "host program not built"
    seed, n_meas, n_hit, delta, beta = 4321, 5, 4, 0.4, 2.2
    inp = tmp_path / "in.yaml"
    inp.write_text(f"""geometry:
  L : 4
  T : 4
  ndims : 4
metropolis:
  seed: {seed}
  n_meas: {n_meas}
  N_hit: {n_hit}
  delta: {delta}
  heat: true
  lenghty_conf_name: false
  conf_dir: "{tmp_path / 'confs'}/"
monomials:
  gauge:
    beta: {beta}
""")
    out = subprocess.run([exe, "-f", str(inp)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [ln.split() for ln in out.stdout.splitlines()
            if ln and ln[0].isdigit() and len(ln.split()) == 7]
    assert len(rows) == n_meas
    with S.Lattice(S.SU2, (4, 4, 4, 4, 4), beta) as lat:
        lat.hotstart(seed, 1.0)
        lat.random_gauge_trafo(654321)
        norm = 2.0 / (4 * 3 * lat.V * 2.0)
        for i, row in enumerate(rows):
            w = oracle.mt_words(seed + i, 2)
            lat.sweep((int(w[0]) << 32) | int(w[1]), 0, delta, n_hit)
            assert int(row[0]) == i
            assert float(row[1]) == pytest.approx(lat.gauge_energy() * norm, rel=1e-13)
            E, Q = lat.energy_density(clover=False)
            assert float(row[2]) == pytest.approx(E, rel=1e-13)
    assert "## Acceptance rate" in out.stdout


def test_host_program_metropolis_restart(S, tmp_path):
    """Metropolis restart (init_gauge_conf_mcmc, base_program.hpp:276-318, and
    after_MCMC_step, :439-456): a run of 5 sweeps saving every 2 leaves
    nconf_counter.txt at sweep 4; a restart loads conf.4 (its plaquette is the
    saved configuration's), applies the gauge transformation and continues at
    sweep 4."""
    import os
    import subprocess
    from su2_amd import host
    exe = host.EXE
    assert os.path.exists(exe), "host program not built"
    conf = tmp_path / "confs"
    inp = tmp_path / "in.yaml"
    text = f"""geometry:
  L : 4
  T : 5
  ndims : 4
metropolis:
  seed: 77
  n_meas: 5
  N_hit: 3
  N_save: 2
  delta: 0.4
  heat: true
  lenghty_conf_name: false
  conf_dir: "{conf}/"
monomials:
  gauge:
    beta: 2.2
"""
    inp.write_text(text)
    out = subprocess.run([exe, "-f", str(inp)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [ln.split() for ln in out.stdout.splitlines() if ln and ln[0].isdigit()]
    assert [int(r[0]) for r in rows] == [0, 1, 2, 3, 4]
    counter = (conf / "nconf_counter.txt").read_text().splitlines()
    assert counter[0] == "heat i path_conf" and counter[1].split()[1] == "4"
    assert (conf / "conf.4").exists()
    inp.write_text(text.replace("heat: true", "restart: true").replace("n_meas: 5", "n_meas: 3"))
    out2 = subprocess.run([exe, "-f", str(inp)], capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stderr
    rows2 = [ln.split() for ln in out2.stdout.splitlines() if ln and ln[0].isdigit()]
    assert [int(r[0]) for r in rows2] == [4, 5, 6]
    # the loaded configuration is conf.4, i.e. the chain's state after sweep 4
    plaq = [float(ln.split()[-1]) for ln in out2.stdout.splitlines() if ln.startswith("## Initial Plaquette")]
    # printed with the default six significant digits, as the reference does
    assert plaq and plaq[0] == pytest.approx(float(rows[4][1]), rel=1e-5)
