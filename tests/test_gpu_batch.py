"""GPU tests of the batched ensemble (one launch per MD tick for many lattices).

Every member of a batch must follow exactly the trajectory that a single-lattice
handle (itself pinned to the oracle in test_gpu_md.py) gives for the same links
and momenta; graph replays must reproduce the eager launches; the autotuner
must leave the configurations untouched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 13526463

MIXED = [
    dict(group=0, dims=(8, 1, 1, 6, 2), beta=2.0, integrator="leapfrog", nsteps=7),
    dict(group=1, dims=(4, 5, 1, 3, 3), beta=3.5, integrator="omf2", nsteps=3),
    dict(group=1, dims=(4, 4, 4, 4, 4), beta=2.4, integrator="leapfrog", nsteps=10),
    dict(group=0, dims=(17, 1, 1, 19, 2), beta=4.0, integrator="omf2", nsteps=5),
    dict(group=1, dims=(6, 1, 1, 5, 2), beta=1.5, integrator="leapfrog", nsteps=4),
    dict(group=1, dims=(3, 4, 5, 6, 4), beta=2.3, integrator="omf2", nsteps=6),
    dict(group=1, dims=(8, 8, 1, 8, 3), beta=4.0, integrator="leapfrog", nsteps=5),
    dict(group=1, dims=(8, 8, 8, 8, 4), beta=2.5, integrator="omf2", nsteps=4),
]


@pytest.fixture(scope="module")
def S():
    import su2_amd
    su2_amd.load_library()
    return su2_amd


def _starts(oracle):
    out = []
    for i, m in enumerate(MIXED):
        U = oracle.hotstart(m["group"], m["dims"], SEED + i, 0.7)
        P = oracle.initnormal(m["group"], m["dims"], SEED + 50 + i)
        out.append((U, P))
    return out


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("block", [64, 256])
def test_batch_members_match_single_lattice(S, oracle, prec, block):
    starts = _starts(oracle)
    with S.Batch(MIXED, precision=prec) as b:
        b.set_launch(block, 0)
        for i, (U, P) in enumerate(starts):
            b.set_links(i, U)
            b.set_momenta(i, P)
        res = b.trajectory(0, momenta=S.MOMENTA_HOST, metropolis=False)
        got = [b.get_links(i) for i in range(len(MIXED))]
    for i, m in enumerate(MIXED):
        U, P = starts[i]
        with S.Lattice(m["group"], m["dims"], m["beta"], precision=prec) as lat:
            lat.set_links(U)
            r = lat.md_trajectory(P, S.INTEGRATORS[m["integrator"]], m["nsteps"])
            lat.md_finish(True)
            want = lat.get_links()
        tol = 1e-11 if prec == 0 else 2e-5
        np.testing.assert_allclose(got[i], want, atol=tol, err_msg=f"member {i}")
        assert res[i]["deltaH"] == pytest.approx(r["deltaH"], rel=1e-9 if prec == 0 else 1e-3,
                                                 abs=1e-10 if prec == 0 else 1e-4)
        if prec == 0:
            Uw, _ = oracle.integrate(m["group"], m["dims"], U, P, m["beta"],
                                     S.INTEGRATORS[m["integrator"]], m["nsteps"])
            np.testing.assert_allclose(got[i], Uw, atol=1e-10)


def test_batch_reject_keeps_start(S, oracle):
    starts = _starts(oracle)
    with S.Batch(MIXED, precision=0) as b:
        for i, (U, P) in enumerate(starts):
            b.set_links(i, U)
        # rough starts give large |dH|: the members whose trajectory is
        # rejected must keep their start links, the others must move
        res = b.trajectory(3, momenta=S.MOMENTA_DEVICE, metropolis=True)
        for i, r in enumerate(res):
            now = b.get_links(i)
            if not r["accept"]:
                np.testing.assert_array_equal(now, starts[i][0])
            else:
                assert not np.array_equal(now, starts[i][0])


def test_graph_replay_matches_eager(S, oracle):
    starts = _starts(oracle)
    runs = []
    for graph in (False, True):
        with S.Batch(MIXED, precision=1) as b:
            b.use_graph(graph)
            for i, (U, _) in enumerate(starts):
                b.set_links(i, U)
            hist = []
            for t in range(4):
                res = b.trajectory(100 + t, momenta=S.MOMENTA_DEVICE)
                # eager work and synchronisation between the replays
                hist.append(([r["deltaH"] for r in res], [r["accept"] for r in res],
                             [b.get_links(i).copy() for i in range(len(MIXED))]))
            runs.append(hist)
    for (dh0, acc0, U0), (dh1, acc1, U1) in zip(*runs):
        assert dh0 == dh1
        assert acc0 == acc1
        for a, c in zip(U0, U1):
            np.testing.assert_array_equal(a, c)


def test_profile_trajectory_matches_trajectory(S, oracle):
    """The profiled trajectory (eager, per-part launches between HIP events)
    leaves every member's links exactly as the ordinary trajectory of the same
    number does, and reports consistent part sizes and times."""
    starts = _starts(oracle)
    links = []
    prof = None
    for profiled in (False, True):
        with S.Batch(MIXED, precision=1) as b:
            for i, (U, _) in enumerate(starts):
                b.set_links(i, U)
            if profiled:
                prof = b.profile(77)
                total = b.link_updates()
            else:
                b.trajectory(77, momenta=S.MOMENTA_DEVICE)
            links.append([b.get_links(i).copy() for i in range(len(MIXED))])
    for a, c in zip(*links):
        np.testing.assert_array_equal(a, c)
    print(prof)
    assert sum(prof["link_updates"].values()) == total
    # MIXED: SU(2) d = 4 members 2, 5, 7; d = 3 members 1, 6; U(1) and the
    # SU(2) d = 2 member the generic part
    n4 = sum(np.prod(m["dims"][:4]) * 4 * m["nsteps"] for m in MIXED if m["group"] == 1 and m["dims"][4] == 4)
    assert prof["link_updates"]["su2_d4"] == n4
    for part in ("su2_d4", "su2_d3", "generic"):
        assert prof["ms_tick"][part] > 0
        assert 0 < prof["active_tile_fraction"][part] <= 1
    parts_ms = (prof["ms_momenta"] + prof["ms_plaq"] + sum(prof["ms_tick"].values()) +
                prof["ms_restore"] + prof["ms_reduce"])
    assert 0 < parts_ms <= prof["ms_total"] * 1.001


def test_autotune_cache_and_no_side_effects(S, oracle, tmp_path):
    starts = _starts(oracle)
    cache = str(tmp_path / "tune.json")
    with S.Batch(MIXED, precision=1) as b:
        for i, (U, _) in enumerate(starts):
            b.set_links(i, U)
        before = [b.get_links(i).copy() for i in range(len(MIXED))]
        blk, grd = b.autotune(cache)
        assert blk in (64, 128, 256) and grd > 0
        for i in range(len(MIXED)):
            np.testing.assert_array_equal(b.get_links(i), before[i])
        assert (tmp_path / "tune.json").exists()
        assert b.autotune(cache) == (blk, grd)
        r = b.trajectory(7)
        assert all(np.isfinite(x["deltaH"]) for x in r)


def test_mixed_ensemble_runs(S):
    """The 224-lattice mixed ensemble: one trajectory with finite dH everywhere."""
    members = S.ensemble_members()
    with S.Batch(members, precision=1) as b:
        b.use_graph(True)
        res = b.trajectory(1)
        assert len(res) == 224
        assert all(np.isfinite(r["deltaH"]) for r in res)
        assert b.link_updates() > 1e8


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_full_ensemble_member_parity(S, oracle, prec):
    """The 224-lattice mixed ensemble (128 U(1) d=2, 64 SU(2) d=3, 32 SU(2) d=4,
    leapfrog with 10..40 steps) after autotuning and with graph replay: the
    smallest, largest and a middle-sized member of every family follow the
    trajectory of a single-lattice handle from the same links and host momenta
    (md_update.hh:18-81): FP64 within 1e-10 per link (and the oracle's
    trajectory for the SU(2) d=4 members), FP32 within 2e-5."""
    members = S.ensemble_members()
    fams = {(0, 2): [], (1, 3): [], (1, 4): []}
    for i, m in enumerate(members):
        fams[(m["group"], m["dims"][4])].append(i)
    pick = []
    for idx in fams.values():
        by_size = sorted(idx, key=lambda i: (members[i]["dims"][0], i))
        pick += [by_size[0], by_size[len(by_size) // 2], by_size[-1]]
    starts = {}
    with S.Batch(members, precision=prec) as b:
        b.autotune(None)
        b.use_graph(True)
        for i, m in enumerate(members):
            P = oracle.initnormal(m["group"], m["dims"], SEED + 500 + i)
            b.set_momenta(i, P)
            if i in pick:
                U = oracle.hotstart(m["group"], m["dims"], SEED + 400 + i, 0.6)
                b.set_links(i, U)
                starts[i] = (U, P)
        res = b.trajectory(0, momenta=S.MOMENTA_HOST, metropolis=False)
        got = {i: b.get_links(i) for i in pick}
    for i in pick:
        m = members[i]
        U, P = starts[i]
        integ = S.INTEGRATORS[m["integrator"]] if isinstance(m["integrator"], str) else m["integrator"]
        with S.Lattice(m["group"], m["dims"], m["beta"], precision=prec) as lat:
            lat.set_links(U)
            r = lat.md_trajectory(P, integ, m["nsteps"])
            lat.md_finish(True)
            want = lat.get_links()
        tol = 1e-10 if prec == 0 else 2e-5
        np.testing.assert_allclose(got[i], want, atol=tol, err_msg=f"member {i} {m['dims']}")
        assert res[i]["deltaH"] == pytest.approx(r["deltaH"], rel=1e-9 if prec == 0 else 1e-3,
                                                 abs=1e-9 if prec == 0 else 5e-3)
        if prec == 0 and m["dims"][4] == 4:
            Uw, _ = oracle.integrate(m["group"], m["dims"], U, P, m["beta"], integ, m["nsteps"])
            np.testing.assert_allclose(got[i], Uw, atol=1e-10)
