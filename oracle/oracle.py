"""ctypes front end for the CPU oracle and the reference build.

TEST INFRASTRUCTURE, NOT PRODUCT: imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker.  The product package (su2_amd) never imports this module.

Two libraries sit behind it:
  * ``Oracle`` -> oracle/_build/libsu2oracle.so, the C restatement
    (su2_oracle.c) of the reference's algorithm;
  * ``Reference`` -> oracle/_ref/libsu2ref.so, the reference's own hot-path
    headers compiled from the reference tree (ref_driver.cc); present only where
    that tree was available when it was built.

Arrays are numpy float64 in the reference AoS order (gaugeconfig.hh:211-222):
SU(2) links (V*nd, 4) = (Re a, Im a, Re b, Im b); U(1) links (V*nd,) phases;
momenta (V*nd, 3) or (V*nd,).  ``dims`` is (Lx, Ly, Lz, Lt, nd).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libsu2oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsu2ref.so")

# integrator ids of the reference's enum (integrator.hh:15)
LEAPFROG, LP_LEAPFROG, OMF4, LP_OMF4, EULER, RUTH, OMF2 = range(7)
SU2, U1 = 1, 0

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and the reference driver when the tree exists)."""
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/include")
    if with_ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def dims_arr(dims) -> np.ndarray:
    return np.asarray(dims, dtype=np.int32)


def volume(dims) -> int:
    return int(dims[0] * dims[1] * dims[2] * dims[3])


def link_shape(group: int, dims):
    n = volume(dims) * dims[4]
    return (n, 4) if group == SU2 else (n,)


def mom_shape(group: int, dims):
    n = volume(dims) * dims[4]
    return (n, 3) if group == SU2 else (n,)


class _Lib:
    prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        p = self.prefix
        L = self.lib

        def sig(name, res, *args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = list(args)

        sig("hotstart", None, C.c_int, _ip, _dp, C.c_uint64, C.c_double)
        sig("random_gauge_trafo", None, C.c_int, _ip, _dp, C.c_uint64)
        sig("initnormal", None, C.c_int, _ip, _dp, C.c_uint64)
        sig("derivative", None, C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_double, _dp)
        sig("update_gauge", None, C.c_int, _ip, _dp, _dp, C.c_double)
        sig("integrate", None, C.c_int, _ip, _dp, _dp, C.c_double, C.c_double, C.c_int,
            C.c_long, C.c_double, C.c_int, C.c_long)
        sig("energy_density", None, C.c_int, _ip, _dp, C.c_int, C.c_int,
            C.POINTER(C.c_double), C.POINTER(C.c_double))
        sig("runge_kutta", None, *self._rk_args())
        sig("hmc_run", None, C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_uint64, C.c_int,
            C.c_int, C.c_long, C.c_double, C.c_long, C.c_long, C.c_long, _dp)
        sig("normals", None, C.c_uint64, C.c_size_t, _dp)
        sig("mt_words", None, C.c_uint64, C.c_size_t, _up)
        sig("su2_mul", None, _dp, _dp, _dp)

    def _rk_args(self):
        raise NotImplementedError

    # -- helpers ---------------------------------------------------------------
    def hotstart(self, group, dims, seed, delta=1.0):
        U = np.zeros(link_shape(group, dims))
        getattr(self.lib, self.prefix + "hotstart")(group, dims_arr(dims), U, seed, delta)
        return U

    def random_gauge_trafo(self, group, dims, U, seed=654321):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        getattr(self.lib, self.prefix + "random_gauge_trafo")(group, dims_arr(dims), U, seed)
        return U

    def initnormal(self, group, dims, seed):
        P = np.zeros(mom_shape(group, dims))
        getattr(self.lib, self.prefix + "initnormal")(group, dims_arr(dims), P, seed)
        return P

    def derivative(self, group, dims, U, beta, xi=1.0, fac=1.0, deriv=None):
        D = np.zeros(mom_shape(group, dims)) if deriv is None else deriv.copy()
        getattr(self.lib, self.prefix + "derivative")(
            group, dims_arr(dims), np.ascontiguousarray(U), beta, xi, fac, D)
        return D

    def update_gauge(self, group, dims, U, P, dtau):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        getattr(self.lib, self.prefix + "update_gauge")(
            group, dims_arr(dims), U, np.ascontiguousarray(P), dtau)
        return U

    def integrate(self, group, dims, U, P, beta, kind, nsteps, tau=1.0, xi=1.0, restore=True,
                  exponent=0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        P = np.ascontiguousarray(P, dtype=np.float64).copy()
        getattr(self.lib, self.prefix + "integrate")(
            group, dims_arr(dims), U, P, beta, xi, kind, nsteps, tau, int(restore), exponent)
        return U, P

    def energy_density(self, group, dims, U, clover=True, ss=False):
        E, Q = C.c_double(0), C.c_double(0)
        getattr(self.lib, self.prefix + "energy_density")(
            group, dims_arr(dims), np.ascontiguousarray(U), int(clover), int(ss),
            C.byref(E), C.byref(Q))
        return E.value, Q.value

    def hmc_run(self, group, dims, beta, seed, kind, nsteps, n_meas, tau=1.0, xi=1.0, heat=True,
                exponent=0, N_rev=0):
        U = np.zeros(link_shape(group, dims))
        rows = np.zeros((n_meas, 7))
        getattr(self.lib, self.prefix + "hmc_run")(
            group, dims_arr(dims), U, beta, xi, seed, int(heat), kind, nsteps, tau, exponent,
            n_meas, N_rev, rows)
        return U, rows

    def normals(self, seed, n):
        out = np.zeros(n)
        getattr(self.lib, self.prefix + "normals")(seed, n, out)
        return out

    def mt_words(self, seed, n):
        out = np.zeros(n, dtype=np.uint32)
        getattr(self.lib, self.prefix + "mt_words")(seed, n, out)
        return out

    def su2_mul(self, u, v):
        out = np.zeros(4)
        getattr(self.lib, self.prefix + "su2_mul")(
            np.ascontiguousarray(u, dtype=np.float64), np.ascontiguousarray(v, dtype=np.float64), out)
        return out


class Oracle(_Lib):
    """The C restatement (su2_oracle.c)."""

    prefix = "orc_"

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.orc_retr_sum_plaq.restype = C.c_double
        L.orc_retr_sum_plaq.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_int, C.c_int]
        L.orc_action.restype = C.c_double
        L.orc_action.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_int]
        L.orc_kinetic.restype = C.c_double
        L.orc_kinetic.argtypes = [C.c_int, _ip, _dp, C.c_int]
        L.orc_update_momenta.restype = None
        L.orc_update_momenta.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, _dp, C.c_double]
        L.orc_restoreSU.restype = None
        L.orc_restoreSU.argtypes = [C.c_int, _ip, _dp]
        L.orc_coldstart.restype = None
        L.orc_coldstart.argtypes = [C.c_int, _ip, _dp]
        L.orc_haarstart.restype = None
        L.orc_haarstart.argtypes = [C.c_int, _ip, _dp, C.c_uint64]
        L.orc_md_update.restype = None
        L.orc_md_update.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_int, C.c_long,
                                    C.c_double, C.c_long, C.c_uint64, C.c_int, C.c_int, _dp,
                                    C.c_void_p]
        L.orc_gradient_flow.restype = C.c_long
        L.orc_gradient_flow.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _dp, C.c_long]
        for n in ("orc_su2_restore", "orc_su2_exp"):
            getattr(L, n).restype = None
            getattr(L, n).argtypes = [_dp, _dp]
        L.orc_sweep.restype = None
        L.orc_sweep.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                C.c_double, C.c_long, _dp]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]

    def sweep(self, group, dims, U, beta, seed, sweep, delta, n_hit, xi=1.0):
        """Metropolis sweep (flat-sweep.hh:64-164) with the device's order and
        Philox stream; returns (U_after, (rate, rate_time))."""
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        rates = np.zeros(2)
        self.lib.orc_sweep(group, dims_arr(dims), U, beta, xi, seed, sweep, delta, n_hit, rates)
        return U, (rates[0], rates[1])

    def philox(self, ctr, key):
        out = np.zeros(4, dtype=np.uint32)
        self.lib.orc_philox4x32_10(np.ascontiguousarray(ctr, dtype=np.uint32),
                                   np.ascontiguousarray(key, dtype=np.uint32), out)
        return out

    def _rk_args(self):
        return (C.c_int, _ip, _dp, _dp, C.c_double, C.c_double, C.c_double)

    def runge_kutta(self, group, dims, U, beta, eps, xi=1.0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        Z = np.zeros(mom_shape(group, dims))
        self.lib.orc_runge_kutta(group, dims_arr(dims), U, Z, beta, xi, eps)
        return U

    def retr_sum_plaq(self, group, dims, U, xi=1.0, spatial_only=False, compensated=False):
        return self.lib.orc_retr_sum_plaq(group, dims_arr(dims), np.ascontiguousarray(U), xi,
                                          int(spatial_only), int(compensated))

    def action(self, group, dims, U, beta, xi=1.0, compensated=False):
        return self.lib.orc_action(group, dims_arr(dims), np.ascontiguousarray(U), beta, xi,
                                   int(compensated))

    def kinetic(self, group, dims, P, compensated=False):
        return self.lib.orc_kinetic(group, dims_arr(dims), np.ascontiguousarray(P), int(compensated))

    def update_momenta(self, group, dims, U, P, beta, dtau, xi=1.0):
        P = np.ascontiguousarray(P, dtype=np.float64).copy()
        self.lib.orc_update_momenta(group, dims_arr(dims), np.ascontiguousarray(U), beta, xi, P, dtau)
        return P

    def restoreSU(self, group, dims, U):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        self.lib.orc_restoreSU(group, dims_arr(dims), U)
        return U

    def coldstart(self, group, dims):
        U = np.zeros(link_shape(group, dims))
        self.lib.orc_coldstart(group, dims_arr(dims), U)
        return U

    def haarstart(self, group, dims, seed):
        U = np.zeros(link_shape(group, dims))
        self.lib.orc_haarstart(group, dims_arr(dims), U, seed)
        return U

    def md_update(self, group, dims, U, beta, kind, nsteps, seed, tau=1.0, xi=1.0, exponent=0,
                  revtest=False, compensated=False, want_P=False):
        """Returns (U_after, info dict[, P_final])."""
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        out = np.zeros(8)
        P = np.zeros(mom_shape(group, dims)) if want_P else None
        self.lib.orc_md_update(group, dims_arr(dims), U, beta, xi, kind, nsteps, tau, exponent,
                               seed, int(revtest), int(compensated), out,
                               P.ctypes.data if want_P else None)
        info = dict(dH=out[0], accept=bool(out[1]), ddH=out[2], K_old=out[3], S_old=out[4],
                    K_new=out[5], S_new=out[6], u=out[7])
        return (U, info, P) if want_P else (U, info)

    def gradient_flow(self, group, dims, U, beta, tmax, eps, xi=1.0, tstart=0.0, max_rows=100000):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        rows = np.zeros((max_rows, 9))
        n = self.lib.orc_gradient_flow(group, dims_arr(dims), U, beta, xi, tmax, eps, tstart,
                                       rows, max_rows)
        return U, rows[:n]

    def su2_restore(self, u):
        out = np.zeros(4)
        self.lib.orc_su2_restore(np.ascontiguousarray(u, dtype=np.float64), out)
        return out

    def su2_exp(self, p):
        out = np.zeros(4)
        self.lib.orc_su2_exp(np.ascontiguousarray(p, dtype=np.float64), out)
        return out


class Reference(_Lib):
    """The reference's own hot path (oracle/_ref/libsu2ref.so)."""

    prefix = "ref_"

    def __init__(self, path: str = REF_SO):
        super().__init__(path)
        L = self.lib
        L.ref_retr_sum_plaq.restype = C.c_double
        L.ref_retr_sum_plaq.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_int]
        L.ref_kinetic.restype = C.c_double
        L.ref_kinetic.argtypes = [C.c_int, _ip, _dp]
        L.ref_md_update.restype = None
        L.ref_md_update.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_int, C.c_long,
                                    C.c_double, C.c_long, C.c_uint64, C.c_int, _dp]
        L.ref_gradient_flow.restype = None
        L.ref_gradient_flow.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_char_p]
        L.ref_uniforms.restype = None
        L.ref_uniforms.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, _dp]
        L.ref_save.restype = None
        L.ref_save.argtypes = [C.c_int, _ip, _dp, C.c_char_p]
        L.ref_load.restype = C.c_int
        L.ref_load.argtypes = [C.c_int, _ip, _dp, C.c_char_p]

    def _rk_args(self):
        return (C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_double)

    def runge_kutta(self, group, dims, U, beta, eps, xi=1.0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        self.lib.ref_runge_kutta(group, dims_arr(dims), U, beta, xi, eps)
        return U

    def retr_sum_plaq(self, group, dims, U, xi=1.0, spatial_only=False):
        return self.lib.ref_retr_sum_plaq(group, dims_arr(dims), np.ascontiguousarray(U), xi,
                                          int(spatial_only))

    def kinetic(self, group, dims, P):
        return self.lib.ref_kinetic(group, dims_arr(dims), np.ascontiguousarray(P))

    def md_update(self, group, dims, U, beta, kind, nsteps, seed, tau=1.0, xi=1.0, exponent=0,
                  revtest=False):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        out = np.zeros(7)
        self.lib.ref_md_update(group, dims_arr(dims), U, beta, xi, kind, nsteps, tau, exponent,
                               seed, int(revtest), out)
        info = dict(dH=out[0], accept=bool(out[1]), ddH=out[2], K_old=out[3], S_old=out[4],
                    K_new=out[5], S_new=out[6])
        return U, info

    def gradient_flow_rows(self, group, dims, U, beta, tmax, eps, path, xi=1.0):
        self.lib.ref_gradient_flow(group, dims_arr(dims), np.ascontiguousarray(U), beta, xi, tmax,
                                   eps, path.encode())
        with open(path) as f:
            lines = f.read().splitlines()[1:]
        return np.array([[float(v) for v in ln.split()] for ln in lines if ln.strip()])

    def uniforms(self, seed, n, a, b):
        out = np.zeros(n)
        self.lib.ref_uniforms(seed, n, a, b, out)
        return out

    def save(self, group, dims, U, path):
        self.lib.ref_save(group, dims_arr(dims), np.ascontiguousarray(U), path.encode())

    def load(self, group, dims, path):
        U = np.zeros(link_shape(group, dims))
        r = self.lib.ref_load(group, dims_arr(dims), U, path.encode())
        return r, U


def have_reference() -> bool:
    return os.path.exists(REF_SO)
