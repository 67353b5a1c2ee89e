#!/usr/bin/env python3
"""Benchmark of the pure-gauge HMC trajectory on one MI355X.

Workload: SU(2), d = 4, 32^4 periodic, beta = 2.4, Omelyan OMF2 with 20 steps,
tau = 1, seeded hot start (Haar-uniform links), Gaussian momenta (device Philox
stream), Metropolis accept/reject.  A step is one HMC trajectory: momentum
refresh, H_old, 20 OMF2 steps (41 force evaluations, 40 link updates), final
restoreSU, H_new, accept/reject.  One trajectory performs 4 V * 20
link-updates.

`value` is the plain FP64 form's link-updates/s (the reference's precision).
The FP32 form (links, staples, forces and momenta in FP32, FP64 staple sums,
reunitarisation every step, FP64 action / kinetic sums and accept/reject) is
timed the same way and reported under "fp32", with its dH differences against
the FP64 form from identical starts and momenta at 32^4 and whether they meet
|delta dH| <= 1e-6.  The 48^4 leg ("cfg48") times both forms beyond the 256 MB
Infinity Cache and repeats the dH comparison at beta = 2.5.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def sweep_bytes_per_link(prec: int):
    """Algorithmic bytes per link of the two sweeps of an MD step, each field
    read or written once: F = force + momentum update (read U, read P, write
    P), G = link update (read P, read U, write U).  FP64: 80 and 88 B, FP32:
    40 and 44 B."""
    r = 8 if prec == 0 else 4
    return r * (4 + 3 + 3), r * (3 + 4 + 4)


def fused_compulsory_bytes_per_link(prec: int) -> int:
    # the fused F + G launch keeps the new momentum in registers: read U and
    # P, write P and U once each
    r = 8 if prec == 0 else 4
    return r * (4 + 3 + 3 + 4)


def run_form(S, prec, L, nsteps, steps, warmup, seed, beta):
    dims = (L, L, L, L, 4)
    lat = S.Lattice(S.SU2, dims, beta, precision=prec)
    lat.haarstart(seed)
    V = lat.V
    for i in range(warmup):
        lat.md_update(seed + 1000 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
    lat.synchronize()
    acc = 0
    dhs = []
    t0 = time.perf_counter()
    for i in range(steps):
        r = lat.md_update(seed + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
        acc += r["accept"]
        dhs.append(r["deltaH"])
    lat.synchronize()
    dt = time.perf_counter() - t0
    # kernel timing pass (HIP events on the handle's stream), outside the timed region
    lat.profile(True)
    lat.profile_reset()
    nprof = 2
    for i in range(nprof):
        lat.md_update(seed + 5000 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
    ms_fused, n_fused = lat.profile_read("md_fused")
    kinds = {}
    for k in ("md_fused", "md_p_only", "plaq", "sumsq", "reduce", "momenta", "update_gauge"):
        ms, n = lat.profile_read(k)
        if n:
            kinds[k] = {"ms_total": round(ms, 4), "launches": n, "us_avg": round(1000 * ms / n, 2)}
    lat.profile(False)
    lat.close()
    link_updates = 4 * V * nsteps * steps
    avg_us = 1000.0 * ms_fused / max(n_fused, 1)
    F, G = sweep_bytes_per_link(prec)
    # one fused launch performs an F and a G sweep over all 4 V links
    bytes_launch = (F + G) * 4 * V
    compulsory = fused_compulsory_bytes_per_link(prec) * 4 * V
    # OMF2 with n steps: 2 n + 1 F sweeps and 2 n G sweeps per trajectory
    traj_bytes = ((2 * nsteps + 1) * F + 2 * nsteps * G) * 4 * V
    ms_traj = 1000.0 * dt / steps
    return {
        "link_updates_per_s": link_updates / dt,
        "ms_per_traj": ms_traj,
        "acceptance": acc / steps,
        "mean_abs_dH": sum(abs(x) for x in dhs) / len(dhs),
        "fused_avg_us": avg_us,
        "fused_bytes_per_launch": bytes_launch,
        "fused_achieved_GBs": bytes_launch / (avg_us * 1e-6) / 1e9,
        "fused_compulsory_bytes_per_launch": compulsory,
        "fused_compulsory_GBs": compulsory / (avg_us * 1e-6) / 1e9,
        "traj_bytes": traj_bytes,
        "traj_achieved_GBs": traj_bytes / (ms_traj * 1e-3) / 1e9,
        "kernels": kinds,
    }


def delta_dH(S, L, nsteps, beta, seed, ntraj, ntherm=50, nchain=100):
    """FP32 form against the FP64 form at equilibrium.

    Start: Haar hot start + `ntherm` FP64 OMF2 trajectories (the relaxation
    trajectories lower H and are accepted).  Paired trajectories: from the same
    links and the same Philox momentum stream, both forms run one trajectory
    with the reversibility test; dH32 - dH64, both forms' ddH (|H after
    forward + flipped backward - H_start|) and the plaquette after the
    trajectory are compared.  Chains: both forms then continue `nchain`
    trajectories from the same start; acceptance and mean plaquette are
    compared with their binomial / standard errors."""
    import statistics
    dims = (L, L, L, L, 4)
    norm = 1.0 / (6 * L ** 4 * 2)
    l64 = S.Lattice(S.SU2, dims, beta, precision=S.FP64)
    l32 = S.Lattice(S.SU2, dims, beta, precision=S.FP32)
    l64.haarstart(seed)
    for i in range(ntherm):
        l64.md_update(seed + 100 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
    start = l64.get_links()
    plaq_start = l64.gauge_energy() * norm
    pairs = []
    for i in range(ntraj):
        x = l64.get_links()
        l32.set_links(x)
        r64 = l64.md_update(seed + 200 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE, revtest=True)
        r32 = l32.md_update(seed + 200 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE, revtest=True)
        # plaquette of each form's proposal from S_new = -(beta / 2) sum ReTr
        pairs.append({"dH64": r64["deltaH"], "dH32": r32["deltaH"],
                      "ddH64": r64["deltadeltaH"], "ddH32": r32["deltadeltaH"],
                      "same_accept": r32["accept"] == r64["accept"],
                      "plaq64": -r64["gauge_new"] / (beta / 2) * norm,
                      "plaq32": -r32["gauge_new"] / (beta / 2) * norm})
    # independent chains from the same equilibrium start
    l64.set_links(start)
    l32.set_links(start)
    chain = {}
    for name, lat in (("fp64", l64), ("fp32", l32)):
        acc, pl = [], []
        for i in range(nchain):
            r = lat.md_update(seed + 300 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
            acc.append(r["accept"])
            pl.append(lat.gauge_energy() * norm)
        a = sum(acc) / nchain
        chain[name] = {"acceptance": a, "acceptance_err": math.sqrt(max(a * (1 - a), 1e-12) / nchain),
                       "mean_plaquette": statistics.fmean(pl),
                       "plaquette_err": statistics.stdev(pl) / math.sqrt(nchain)}
    l64.close()
    l32.close()
    d = [abs(p["dH32"] - p["dH64"]) for p in pairs]
    return {
        "lattice": [L] * 4, "beta": beta,
        "start": f"Haar hot start + {ntherm} FP64 OMF2({nsteps}) trajectories, plaquette {plaq_start:.5f}",
        "paired_trajectories": ntraj,
        "mean_abs_dH64": statistics.fmean(abs(p["dH64"]) for p in pairs),
        "max_abs_dH32_minus_dH64": max(d),
        "mean_abs_dH32_minus_dH64": statistics.fmean(d),
        "max_abs_ddH_fp32": max(abs(p["ddH32"]) for p in pairs),
        "max_abs_ddH_fp64": max(abs(p["ddH64"]) for p in pairs),
        "max_abs_plaquette_diff": max(abs(p["plaq32"] - p["plaq64"]) for p in pairs),
        "same_accept_decisions": sum(p["same_accept"] for p in pairs),
        "chains": dict(chain, trajectories=nchain),
    }


def precision_floor(S, Ls, nsteps, beta, seed, ntraj, ntherm=30):
    """dH of the FP32 form and of the FP64 form with emulated FP32 storage
    (links, momenta, both rounded to FP32 after every update; FP64
    arithmetic) against the plain FP64 form, per lattice size: paired OMF2
    trajectories with the reversibility test from the same equilibrium links
    and the same Philox momentum stream.  Shows which stored field sets the
    FP32 form's dH floor and how the floor grows with the volume."""
    import statistics
    out = []
    forms = (("fp32", S.FP32, None), ("fp32_links", S.FP64, (True, False)),
             ("fp32_momenta", S.FP64, (False, True)), ("fp32_both", S.FP64, (True, True)))
    for L in Ls:
        dims = (L, L, L, L, 4)
        progress(f"precision floor at {L}^4")
        l64 = S.Lattice(S.SU2, dims, beta, precision=S.FP64)
        l64.haarstart(seed)
        for i in range(ntherm):
            l64.md_update(seed + 100 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE)
        lats = {}
        for name, prec, emu in forms:
            lats[name] = S.Lattice(S.SU2, dims, beta, precision=prec)
            if emu:
                lats[name].set_fp32_storage_emulation(*emu)
        rows = {name: [] for name, _, _ in forms}
        ref = []
        for i in range(ntraj):
            x = l64.get_links()
            for name, _, _ in forms:
                lats[name].set_links(x)
                r = lats[name].md_update(seed + 200 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE,
                                         revtest=True)
                rows[name].append(r)
            ref.append(l64.md_update(seed + 200 + i, S.OMF2, nsteps, momenta=S.MOMENTA_DEVICE,
                                     revtest=True))
        for lat in lats.values():
            lat.close()
        l64.close()
        e = {"lattice": [L] * 4, "paired_trajectories": ntraj,
             "mean_abs_dH64": statistics.fmean(abs(r["deltaH"]) for r in ref),
             "max_abs_ddH_fp64": max(abs(r["deltadeltaH"]) for r in ref)}
        for name, _, _ in forms:
            d = [abs(a["deltaH"] - b["deltaH"]) for a, b in zip(rows[name], ref)]
            e[name] = {"max_abs_dH_minus_dH64": max(d), "mean_abs_dH_minus_dH64": statistics.fmean(d),
                       "max_abs_ddH": max(abs(r["deltadeltaH"]) for r in rows[name])}
        out.append(e)
    return out


def batch_ensemble(S, steps, warmup):
    """The mixed ensemble of 224 lattices (128 U(1) d=2, 64 SU(2) d=3, 32 SU(2)
    d=4), FP32, one launch per part and MD tick (the three parts' tick chains
    on three streams, parallel branches of the replayed hipGraph), autotuned
    launch shape.  kernel_ms_per_traj: the kernel time of an eager trajectory
    with the parts one after another (HIP events around each launch group, so
    each includes a little launch latency); gpu_busy_fraction: that over the
    graph replay's wall time per trajectory, above 1 by the overlap of the
    parts (the wall time also holds the launch gaps, the host's descriptor
    upload, sums copy and Metropolis); launch_gap_ms_per_traj: what the wall
    time exceeds the kernel time by (0 when the overlap and the event
    overhead cover it); per part the ticks' link
    updates/s, algorithmic bandwidth (84 B per SU(2), 24 B per U(1) link
    update, FP32) and the active tile fraction (the share of a full launch per
    tick that lattices still running need; the launches cover only that)."""
    import statistics
    members = S.ensemble_members()
    b = S.Batch(members, precision=S.FP32)
    import tempfile
    # a fresh cache per run, so that the tuning time is measured every time
    cache = os.path.join(tempfile.mkdtemp(prefix="su2_tune_"), "batch_tune.json")
    t0 = time.perf_counter()
    blk, grd = b.autotune(cache)
    t_tune = time.perf_counter() - t0
    b.use_graph(True)
    for i in range(warmup):
        b.trajectory(10_000 + i)
    acc = 0
    t0 = time.perf_counter()
    for i in range(steps):
        res = b.trajectory(i)
        acc += sum(r["accept"] for r in res)
    dt = time.perf_counter() - t0
    lu = b.link_updates()
    # launch gaps and per-part efficiency: eager trajectories with HIP events
    # around every launch group (after the timed graph replays)
    prof = [b.profile(20_000 + i) for i in range(3)]
    b.close()
    ms_graph = 1000 * dt / steps
    kern = statistics.fmean(p["ms_momenta"] + p["ms_plaq"] + sum(p["ms_tick"].values()) +
                            p["ms_restore"] + p["ms_reduce"] for p in prof)
    parts = {}
    for part, nbytes in (("su2_d4", 84), ("su2_d3", 84), ("generic", 24)):
        ms = statistics.fmean(p["ms_tick"][part] for p in prof)
        n = prof[0]["link_updates"][part]
        parts[part] = {"tick_ms_per_traj": ms, "link_updates_per_traj": n,
                       "link_updates_per_s": n / (ms * 1e-3) if ms > 0 else None,
                       "algorithmic_GBs": n * nbytes / (ms * 1e-3) / 1e9 if ms > 0 else None,
                       "roofline_frac": n * nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS if ms > 0 else None,
                       "bytes_per_link_update": nbytes,
                       "active_tile_fraction": prof[0]["active_tile_fraction"][part]}
    return {"lattices": len(members), "link_updates_per_traj": lu,
            "link_updates_per_s": lu * steps / dt, "ms_per_traj": ms_graph,
            "acceptance": acc / (steps * len(members)), "block": blk, "grid": grd,
            "autotune_s": round(t_tune, 3), "graph": True, "precision": "fp32",
            "kernel_ms_per_traj": kern, "gpu_busy_fraction": kern / ms_graph,
            "launch_gap_ms_per_traj": max(0.0, ms_graph - kern),
            "eager_ms_per_traj": statistics.fmean(p["ms_total"] for p in prof),
            "parts": parts}


def flow_adjoint(S, L=24, beta=2.4, eps=0.01, tmax=2.0, ckpt=20, ntherm=50, seed=4242):
    """Forward RK3 Wilson flow to tmax and the hand-written backward pass giving
    d(t^2 <E(t)>)/dU, FP32 with FP64 sums, checkpoints every `ckpt` steps; start
    from `ntherm` FP64 OMF2 trajectories after a Haar hot start."""
    dims = (L, L, L, L, 4)
    lat64 = S.Lattice(S.SU2, dims, beta, precision=S.FP64)
    lat64.haarstart(seed)
    for i in range(ntherm):
        lat64.md_update(seed + i, S.OMF2, 10, momenta=S.MOMENTA_DEVICE)
    U = lat64.get_links()
    lat64.close()
    lat = S.Lattice(S.SU2, dims, beta, precision=S.FP32)
    lat.set_links(U)
    nsteps = int(round(tmax / eps))
    lat.flow_forward(eps, nsteps, ckpt)  # warm-up (allocations, code objects)
    lat.flow_backward()
    lat.set_links(U)
    lat.synchronize()
    t0 = time.perf_counter()
    val = lat.flow_forward(eps, nsteps, ckpt)
    lat.synchronize()
    t1 = time.perf_counter()
    g = lat.flow_backward()
    lat.synchronize()
    t2 = time.perf_counter()
    lat.close()
    V = L ** 4
    return {"lattice": [L] * 4, "eps": eps, "tmax": tmax, "steps": nsteps, "checkpoint_every": ckpt,
            "precision": "fp32", "t2E": val, "forward_ms": 1000 * (t1 - t0),
            "backward_ms": 1000 * (t2 - t1),
            "backward_link_steps_per_s": 4 * V * nsteps / (t2 - t1),
            "grad_norm": float((g ** 2).sum()) ** 0.5}


def metropolis(S, L=32, beta=2.4, n_hit=10, delta=0.3, sweeps=4, seed=4243):
    """Metropolis sweeps (flat-sweep.hh:64-164) at L^4 in both forms: link hits
    per second and the acceptance rate."""
    out = {"lattice": [L] * 4, "beta": beta, "n_hit": n_hit, "delta": delta, "sweeps": sweeps}
    for prec, name in ((S.FP64, "fp64"), (S.FP32, "fp32")):
        lat = S.Lattice(S.SU2, (L, L, L, L, 4), beta, precision=prec)
        lat.haarstart(seed)
        lat.sweep(seed, 0, delta, n_hit)
        lat.synchronize()
        t0 = time.perf_counter()
        rate = 0.0
        for i in range(sweeps):
            rate += lat.sweep(seed, 1 + i, delta, n_hit)[0]
        lat.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"ms_per_sweep": 1000 * dt / sweeps,
                     "link_hits_per_s": 4 * L ** 4 * n_hit * sweeps / dt,
                     "acceptance": rate / sweeps}
        lat.close()
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def affinity_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_sample(L, nsteps, beta, seed):
    """One OMF2(nsteps) trajectory of the reference's md_update (oracle/_ref
    build of its hot-path headers; the C oracle where that build is absent) on
    an L^4 Haar start, with the OpenMP thread count of this process's
    environment.  Runs in a child process (cpu_baseline) and prints one JSON
    line."""
    from oracle import oracle as orc
    dims = (L, L, L, L, 4)
    kind = "reference" if orc.have_reference() else "port"
    lib = orc.Reference() if kind == "reference" else orc.Oracle()
    U = orc.Oracle().haarstart(orc.SU2, dims, seed)
    # the reference's set_integrator prints the integrator's name on stdout;
    # keep this program's stdout to the one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        t0 = time.perf_counter()
        lib.md_update(orc.SU2, dims, U, beta, orc.OMF2, nsteps, seed)
        dt = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    print(json.dumps({"kind": kind, "seconds": dt, "link_updates_per_s": 4 * L ** 4 * nsteps / dt}),
          flush=True)


def cpu_baseline(L, nsteps, beta, seed):
    """The reference CPU path timed on this host: the reference's own md_update
    (its hot-path headers compiled by oracle/Makefile into oracle/_ref) on
    every core of this process's affinity mask for one trajectory at the
    benchmark size, and on one thread for one trajectory at (L/2)^4 (the same
    per-link work, 1/16 of the sites, so that the sample stays ~10 s).  Each
    sample runs in a child process with its own OMP_NUM_THREADS.  The
    reference's force loop is an OpenMP loop over t (flat-gaugemonomial.hh:
    109-110), i.e. L work items."""
    import subprocess
    ncores = affinity_cores()

    def sample(Ls, threads):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", str(Ls),
                              "--nsteps", str(nsteps), "--beta", str(beta), "--seed", str(seed)],
                             env=env, capture_output=True, text=True, timeout=600)
        if out.returncode != 0:
            raise RuntimeError("cpu sample failed: " + out.stderr[-2000:])
        return json.loads(out.stdout.strip().splitlines()[-1])

    full = sample(L, ncores)
    one = sample(L // 2, 1)
    return {"value": full["link_updates_per_s"], "unit": "link-updates/s", "cores": ncores,
            "kind": full["kind"],
            "sample": f"1 OMF2({nsteps}) trajectory of md_update at {L}^4 on {ncores} threads "
                      f"(incl. momenta, H, copy), {full['seconds']:.1f} s wall",
            "value_1thread": one["link_updates_per_s"],
            "sample_1thread": f"1 OMF2({nsteps}) trajectory at {L // 2}^4 on 1 thread, "
                              f"{one['seconds']:.1f} s wall",
            "cpu_model": cpu_model(),
            "note": "the reference parallelises its force sweep over t only (L work items)"}


def measured_traffic(L, form):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by the factor that
    the same counter reads for a copy of known size); None if no summary for
    this lattice size and form is committed, or if it was taken of another
    instantiation than the one this build launches."""
    path = os.path.join(ROOT, "profiles", "fused_step_traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    rec = tab.get(f"{L}^4/{form}")
    if rec is None or rec.get("kernel") != kernel_name(0 if form == "fp64" else 1, L):
        return None
    return rec.get("hbm_bytes_per_launch")


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def kernel_name(prec, L):
    """Instantiation of the specialised step that a P-then-U launch of the
    L^4 SU(2) lattice takes (su2_kernels.hip launch_fast), as rocprofv3 names
    it: <T, WRITE_U, ACCUM, SPEC, ZS, AN, NT, link in, link out>.  ZS: 1 the
    +-z lane exchange within a wave (64 % L == 0), 2 the general exchange
    (FP32 beyond 256 MiB otherwise), 0 none.  NT: non-temporal stores (1)
    and momentum loads (2: the FP32 general-exchange step, both forms with the
    lane exchange or in FP64 beyond 1 GiB) where the two link fields and the
    momenta exceed 256 MiB.
    The FP64 form steps on the half-field layout (QHalf in, QHalfW out); with
    >= 4 t-slices per XCD the step is a t-walk: FP64 with the lane exchange
    k_md_su2_4d_tww<T, NT, in, out> (one direction per wave; beyond 1 GiB of
    fields k_md_su2_4d_tw), FP32 with it (within the Infinity Cache)
    k_md_su2_4d_tw<T, ZS, NT, in, out>, and with the general exchange
    k_md_su2_4d_twz<T, NT, in, out>."""
    r = 8 if prec == 0 else 4
    big = 4 * L ** 4 * (2 * 4 * r + 3 * r) > 256 * 1024 * 1024
    if prec == 0:
        zs = 1 if 64 % L == 0 else 0
        nt = (2 if 4 * L ** 4 * (2 * 4 * r + 3 * r) > 1024 ** 3 else 1) if big else 0
        if zs == 1 and L % 8 == 0 and L // 8 >= 4:
            if nt < 2:
                return f"k_md_su2_4d_tww<double, {nt}, su2k::QHalf, su2k::QHalfW>"
            return f"k_md_su2_4d_tw<double, 1, {nt}, su2k::QHalf, su2k::QHalfW>"
        if zs == 0 and L % 8 == 0 and L // 8 >= 4:
            return f"k_md_su2_4d_twz<double, {nt}, su2k::QHalf, su2k::QHalfW>"
        return f"k_md_su2_4d<double, true, false, 1, {zs}, false, {nt}, su2k::QHalf, su2k::QHalfW>"
    zs = 1 if 64 % L == 0 else (2 if big else 0)
    huge = 4 * L ** 4 * (2 * 4 * r + 3 * r) > 1024 ** 3
    nt = (2 if zs == 2 or (zs == 1 and huge) else 1) if big else 0
    if zs == 1 and not big and L % 8 == 0 and L // 8 >= 4:
        # the t-walk: each thread walks the t-slices of its XCD's chunk
        return "k_md_su2_4d_tw<float, 1, 0, su2dev::quat<float> const*, su2dev::quat<float>*>"
    if zs == 2 and L % 8 == 0 and L // 8 >= 4:
        # the t-walk with the general exchange, one direction per wave
        return "k_md_su2_4d_twz<float, 2, su2dev::quat<float> const*, su2dev::quat<float>*>"
    return (f"k_md_su2_4d<float, true, false, 1, {zs}, false, {nt}, su2dev::quat<float> const*, "
            "su2dev::quat<float>*>")


def form_roofline(f, prec, L):
    return {"bound": "hbm", "achieved": f["fused_achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": f["fused_achieved_GBs"] / HBM_PEAK_GBS, "kernel": kernel_name(prec, L),
            "avg_us": f["fused_avg_us"], "bytes_per_launch": f["fused_bytes_per_launch"],
            "compulsory_frac": f["fused_compulsory_GBs"] / HBM_PEAK_GBS,
            "trajectory_frac": f["traj_achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(L, "fp64" if prec == 0 else "fp32")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed trajectories per form")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--nsteps", type=int, default=20)
    ap.add_argument("--beta", type=float, default=2.4)
    ap.add_argument("--seed", type=int, default=13526463)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ddh-traj", type=int, default=8, help="paired FP32/FP64 trajectories")
    ap.add_argument("--chain", type=int, default=100, help="chain length of the acceptance comparison")
    ap.add_argument("--prec-L", type=int, default=48, help="lattice extent of the 48^4 leg")
    ap.add_argument("--prec-beta", type=float, default=2.5, help="beta of the 48^4 leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-flow", action="store_true", help="skip the flow-adjoint leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-vs-FP64 studies")
    ap.add_argument("--no-48", action="store_true", help="skip the 48^4 leg")
    ap.add_argument("--no-metropolis", action="store_true", help="skip the Metropolis-sweep leg")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--precision-scan", type=str, default="",
                    help="lattice extents of the FP32 precision-floor study, e.g. 8,16,32,48")
    ap.add_argument("--precision-traj", type=int, default=6)
    ap.add_argument("--scan", type=str, default="",
                    help="comma-separated extents L: also time both forms on L^4 (scan entry)")
    ap.add_argument("--cpu-sample", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_sample:
        cpu_sample(args.cpu_sample, args.nsteps, args.beta, args.seed)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the lattice is not decomposed across GPUs", file=sys.stderr)
        sys.exit(2)

    import su2_amd as S
    S.load_library()
    progress("FP64 form")
    f64 = run_form(S, S.FP64, args.L, args.nsteps, args.steps, args.warmup, args.seed, args.beta)
    progress("FP32 form")
    f32 = run_form(S, S.FP32, args.L, args.nsteps, args.steps, args.warmup, args.seed, args.beta)
    ddh32 = ddh48 = None
    if not args.no_precision:
        progress(f"FP32 vs FP64 dH at {args.L}^4")
        ddh32 = delta_dH(S, args.L, args.nsteps, args.beta, args.seed, args.ddh_traj,
                         nchain=args.chain)
    leg48 = None
    if not args.no_48:
        Lp = args.prec_L
        leg48 = {"lattice": [Lp] * 4, "beta": args.prec_beta}
        for name, prec in (("fp64", S.FP64), ("fp32", S.FP32)):
            progress(f"{name} form at {Lp}^4")
            r = run_form(S, prec, Lp, args.nsteps, 3, 1, args.seed, args.prec_beta)
            leg48[name] = dict(r, roofline=form_roofline(r, prec, Lp))
        if not args.no_precision:
            progress(f"FP32 vs FP64 dH at {Lp}^4")
            ddh48 = delta_dH(S, Lp, args.nsteps, args.prec_beta, args.seed, args.ddh_traj,
                             nchain=args.chain)
            leg48["delta_dH"] = ddh48
    copy_ms = S.copy_probe(1 << 30, 20)
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    progress("batch, flow adjoint, Metropolis")
    batch = None if args.no_batch else batch_ensemble(S, args.batch_steps, 2)
    flow = None if args.no_flow else flow_adjoint(S)
    metro = None if args.no_metropolis else metropolis(S, args.L)
    pfloor = None
    if args.precision_scan:
        pfloor = precision_floor(S, [int(x) for x in args.precision_scan.split(",")], args.nsteps,
                                 args.prec_beta, args.seed, args.precision_traj)
    scan = []
    for Ls in filter(None, args.scan.split(",")):
        L = int(Ls)
        e = {"lattice": [L] * 4}
        for name, prec in (("fp32", S.FP32), ("fp64", S.FP64)):
            r = run_form(S, prec, L, args.nsteps, 3, 1, args.seed, args.beta)
            e[name] = {"link_updates_per_s": r["link_updates_per_s"], "ms_per_traj": r["ms_per_traj"],
                       "fused_avg_us": r["fused_avg_us"],
                       "fused_roofline_frac": r["fused_achieved_GBs"] / HBM_PEAK_GBS,
                       "fused_link_updates_per_s": 4 * L ** 4 / (r["fused_avg_us"] * 1e-6)}
        scan.append(e)
    cpu = None
    if not args.no_cpu:
        progress("CPU baseline")
        cpu = cpu_baseline(args.L, args.nsteps, args.beta, args.seed)
    n = args.nsteps
    roof = form_roofline(f64, 0, args.L)
    roof["traffic_source"] = "profiles/fused_step_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE)"
    fp32 = dict(f32, roofline=form_roofline(f32, 1, args.L))
    if ddh32 is not None:
        worst = max(ddh32["max_abs_dH32_minus_dH64"], ddh32["max_abs_ddH_fp32"])
        fp32["delta_dH"] = ddh32
        fp32["delta_dH_max"] = worst
        fp32["delta_dH_target"] = 1e-6
        fp32["delta_dH_target_met"] = bool(worst <= 1e-6)
    out = {
        "metric": "SU(2) 32^4 HMC link-updates/s, %HBM roofline; FP32 vs FP64 form at |dDH|<=1e-6",
        "value": f64["link_updates_per_s"],
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_traj"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SU(2) d=4 {args.L}^4 periodic, beta={args.beta}, OMF2 {n} steps, "
                               f"tau=1, Haar hot start, device Gaussian momenta, plain FP64 form "
                               f"(the reference's precision)",
                   "lattice": [args.L] * 4, "integrator": "omf2", "n_steps": n, "tau": 1.0},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": f64,
        "fp32": fp32,
        "fp32_over_fp64": f32["link_updates_per_s"] / f64["link_updates_per_s"],
        "cfg48": leg48,
        "measured_copy_GBs": copy_gbs,
        "fused_compulsory_frac_of_measured_copy_fp64": f64["fused_compulsory_GBs"] / copy_gbs,
        "batch": batch,
        "flow_adjoint": flow,
        "metropolis": metro,
    }
    if scan:
        out["scan"] = scan
    if pfloor:
        out["precision_floor"] = pfloor
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
