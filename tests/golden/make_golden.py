"""Writes the golden fixtures of tests/golden/ from the reference's own hot path.

Run in a tree where oracle/_ref/libsu2ref.so has been built from the reference
sources (make -C oracle ref).  The fixtures are data only: inputs and the
reference's outputs for them.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

# the reference's OpenMP reductions (flat-gauge_energy.hpp:40,65) are summed in
# loop order at one thread; fixtures of sums use that order
os.environ["OMP_NUM_THREADS"] = "1"

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import (LEAPFROG, OMF2, OMF4, SU2, U1, Reference)  # noqa: E402

SEED = 13526463


def main():
    R = Reference()
    out = {}
    # random streams
    out["mt_seed5489_10000"] = R.mt_words(5489, 10000)[-1:]
    out["normals_seed13526463_1024"] = R.normals(SEED, 1024)
    # start configurations: hotstart + random_gauge_trafo(654321) as the HMC program does
    d4 = (4, 4, 4, 4, 4)
    U = R.hotstart(SU2, d4, SEED, 1.0)
    out["su2_4444_hotstart"] = U
    U = R.random_gauge_trafo(SU2, d4, U, 654321)
    out["su2_4444_start"] = U
    d2 = (8, 1, 1, 8, 2)
    u = R.hotstart(U1, d2, SEED, 1.0)
    out["u1_88_hotstart"] = u
    out["u1_88_start"] = R.random_gauge_trafo(U1, d2, u, 654321)
    # momenta field digest
    P = R.initnormal(SU2, d4, SEED)
    out["su2_4444_momenta_sha256"] = np.frombuffer(hashlib.sha256(P.tobytes()).digest(), dtype=np.uint8)
    # force, action, kinetic
    out["su2_4444_deriv_b2.3"] = R.derivative(SU2, d4, out["su2_4444_start"], 2.3)
    out["su2_4444_retr_plaq"] = np.array([R.retr_sum_plaq(SU2, d4, out["su2_4444_start"])])
    out["su2_4444_kinetic"] = np.array([R.kinetic(SU2, d4, P)])
    # one trajectory each
    for name, kind, n, beta in (("lf10_b2.3", LEAPFROG, 10, 2.3), ("omf2_20_b2.4", OMF2, 20, 2.4),
                                ("omf4_10_b2.4", OMF4, 10, 2.4)):
        Ut, info = R.md_update(SU2, d4, out["su2_4444_start"], beta, kind, n, SEED, revtest=True)
        out[f"traj_{name}_U"] = Ut
        out[f"traj_{name}_info"] = np.array([info["dH"], float(info["accept"]), info["ddH"],
                                             info["K_old"], info["S_old"], info["K_new"], info["S_new"]])
    ut, info = R.md_update(U1, d2, out["u1_88_start"], 2.0, LEAPFROG, 10, SEED)
    out["traj_u1_88_lf10_b2_U"] = ut
    out["traj_u1_88_lf10_b2_info"] = np.array([info["dH"], float(info["accept"])])
    # energy density / charge, flow
    E, Q = R.energy_density(SU2, d4, out["su2_4444_start"])
    out["su2_4444_EQ"] = np.array([E, Q])
    out["su2_4444_rk_eps0.01"] = R.runge_kutta(SU2, d4, out["su2_4444_start"], 2.4, 0.01)
    out["su2_4444_flow_rows"] = R.gradient_flow_rows(SU2, d4, out["su2_4444_start"], 2.4, 0.1, 0.01,
                                                     "/tmp/_golden_flow.txt")
    # the 8^4 beta=2.3 leapfrog(10) run of 20 trajectories: i accept E*A dH rho ddH Q
    _, rows = R.hmc_run(SU2, (8, 8, 8, 8, 4), 2.3, SEED, LEAPFROG, 10, 20, N_rev=5)
    out["hmc_8888_b2.3_lf10_rows"] = rows
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
