"""su2_amd -- MI355X (gfx950) pure-gauge HMC molecular-dynamics path of HISKP-LQCD/su2.

Python binding of the C ABI in include/su2_mi355x.h (libsu2mi355x.so, built in
tree by ``su2_amd.build.build()``).  The binding is used by the tests and by
bench.py; C/C++ programs link the library directly (see the C++ adapter in
include/su2_mi355x/).  There is no CPU fallback: every call goes to the HIP
library and fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsu2mi355x.so")

U1, SU2 = 0, 1
FP64, FP32 = 0, 1
LAYOUT_PACKED, LAYOUT_OBJECT = 0, 1
LEAPFROG, LP_LEAPFROG, OMF4, LP_OMF4, EULER, RUTH, OMF2 = range(7)
INTEGRATORS = {"leapfrog": LEAPFROG, "lp_leapfrog": LP_LEAPFROG, "omf4": OMF4,
               "lp_omf4": LP_OMF4, "euler": EULER, "ruth": RUTH, "omf2": OMF2}
MOMENTA_HOST, MOMENTA_DEVICE = 0, 1
OBS_CLOVER_E, OBS_WILSON_ACTION = 0, 1

_lib = None


class SU2GPUError(RuntimeError):
    pass


class MDResult(C.Structure):
    _fields_ = [("deltaH", C.c_double), ("deltadeltaH", C.c_double), ("accept", C.c_int),
                ("revtest", C.c_int), ("kin_old", C.c_double), ("gauge_old", C.c_double),
                ("kin_new", C.c_double), ("gauge_new", C.c_double), ("uniform", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_h = C.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "su2gpu_last_error": (C.c_char_p, []),
    "su2gpu_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "su2gpu_create": (C.c_int, [C.POINTER(_h), C.c_int, C.c_size_t, C.c_size_t, C.c_size_t,
                                C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.c_int]),
    "su2gpu_destroy": (C.c_int, [_h]),
    "su2gpu_set_beta": (C.c_int, [_h, C.c_double]),
    "su2gpu_get_beta": (C.c_int, [_h, _dp]),
    "su2gpu_get_geometry": (C.c_int, [_h, C.POINTER(C.c_size_t)]),
    "su2gpu_set_links": (C.c_int, [_h, _vp, C.c_int]),
    "su2gpu_get_links": (C.c_int, [_h, _vp, C.c_int]),
    "su2gpu_set_momenta": (C.c_int, [_h, _vp]),
    "su2gpu_get_momenta": (C.c_int, [_h, _vp]),
    "su2gpu_flip_momenta": (C.c_int, [_h]),
    "su2gpu_set_fp32_storage_emulation": (C.c_int, [_h, C.c_int, C.c_int]),
    "su2gpu_initnormal": (C.c_int, [_h, C.c_uint64]),
    "su2gpu_hotstart": (C.c_int, [_h, C.c_uint64, C.c_double]),
    "su2gpu_coldstart": (C.c_int, [_h]),
    "su2gpu_random_gauge_trafo": (C.c_int, [_h, C.c_uint64]),
    "su2gpu_haarstart": (C.c_int, [_h, C.c_uint64]),
    "su2gpu_restoreSU": (C.c_int, [_h]),
    "su2gpu_retr_sum_plaquettes": (C.c_int, [_h, C.c_double, C.c_int, C.c_int, _dp]),
    "su2gpu_retr_sum_plaquettes_noalloc": (C.c_int, [_h, C.c_double, C.c_int, C.c_int, _vp]),
    "su2gpu_gauge_energy": (C.c_int, [_h, C.c_int, _dp]),
    "su2gpu_gauge_action": (C.c_int, [_h, _dp]),
    "su2gpu_kinetic_energy": (C.c_int, [_h, _dp]),
    "su2gpu_derivative": (C.c_int, [_h, C.c_double, _vp]),
    "su2gpu_zero_derivative": (C.c_int, [_h]),
    "su2gpu_derivative_device": (C.c_int, [_h, C.c_double]),
    "su2gpu_get_derivative": (C.c_int, [_h, _vp]),
    "su2gpu_update_momenta": (C.c_int, [_h, C.c_double]),
    "su2gpu_update_gauge": (C.c_int, [_h, C.c_double]),
    "su2gpu_integrate": (C.c_int, [_h, C.c_int, C.c_size_t, C.c_double, C.c_int, C.c_size_t]),
    "su2gpu_md_update": (C.c_int, [_h, C.c_uint64, C.c_int, C.c_size_t, C.c_double, C.c_size_t,
                                   C.c_int, C.c_int, C.POINTER(MDResult)]),
    "su2gpu_md_trajectory": (C.c_int, [_h, _vp, C.c_int, C.c_size_t, C.c_double, C.c_size_t,
                                       C.c_int, C.POINTER(MDResult)]),
    "su2gpu_md_finish": (C.c_int, [_h, C.c_int]),
    "su2gpu_energy_density": (C.c_int, [_h, C.c_int, C.c_int, _dp, _dp]),
    "su2gpu_sweep": (C.c_int, [_h, C.c_uint64, C.c_uint64, C.c_double, C.c_size_t, _dp, _dp]),
    "su2gpu_runge_kutta": (C.c_int, [_h, C.c_double]),
    "su2gpu_gradient_flow": (C.c_int, [_h, C.c_double, C.c_double, C.c_double, _vp, C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "su2gpu_flow_forward": (C.c_int, [_h, C.c_double, C.c_size_t, C.c_size_t, C.c_int, _dp]),
    "su2gpu_flow_backward": (C.c_int, [_h, _vp]),
    "su2gpu_save": (C.c_int, [_h, C.c_char_p]),
    "su2gpu_load": (C.c_int, [_h, C.c_char_p]),
    "su2gpu_synchronize": (C.c_int, [_h]),
    "su2gpu_copy_probe": (C.c_int, [C.c_size_t, C.c_int, _dp]),
    "su2gpu_profile_enable": (C.c_int, [_h, C.c_int]),
    "su2gpu_profile_reset": (C.c_int, [_h]),
    "su2gpu_profile_read": (C.c_int, [_h, C.c_char_p, _dp, C.POINTER(C.c_long)]),
}


class BatchMember(C.Structure):
    _fields_ = [("group", C.c_int), ("Lx", C.c_size_t), ("Ly", C.c_size_t), ("Lz", C.c_size_t),
                ("Lt", C.c_size_t), ("ndims", C.c_size_t), ("beta", C.c_double),
                ("integrator", C.c_int), ("nsteps", C.c_size_t), ("tau", C.c_double),
                ("seed", C.c_uint64)]


_SIGS.update({
    "su2gpu_batch_create": (C.c_int, [C.POINTER(_h), C.POINTER(BatchMember), C.c_size_t, C.c_int]),
    "su2gpu_batch_destroy": (C.c_int, [_h]),
    "su2gpu_batch_size": (C.c_int, [_h, C.POINTER(C.c_size_t)]),
    "su2gpu_batch_set_links": (C.c_int, [_h, C.c_size_t, _vp]),
    "su2gpu_batch_get_links": (C.c_int, [_h, C.c_size_t, _vp]),
    "su2gpu_batch_set_momenta": (C.c_int, [_h, C.c_size_t, _vp]),
    "su2gpu_batch_trajectory": (C.c_int, [_h, C.c_int, C.c_uint64, C.c_int, C.POINTER(MDResult)]),
    "su2gpu_batch_set_launch": (C.c_int, [_h, C.c_int, C.c_int]),
    "su2gpu_batch_get_launch": (C.c_int, [_h, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "su2gpu_batch_autotune": (C.c_int, [_h, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "su2gpu_batch_use_graph": (C.c_int, [_h, C.c_int]),
    "su2gpu_batch_link_updates": (C.c_int, [_h, _dp]),
    "su2gpu_batch_profile": (C.c_int, [_h, C.c_uint64, _dp, C.c_size_t]),
})


def exported_symbols():
    return sorted(_SIGS)


def load_library(path: str = LIB_PATH):
    """Load libsu2mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SU2GPUError(f"{path} not built: run __graft_entry__.build() or "
                          "python -m su2_amd.build")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ck(rc: int):
    if rc != 0:
        msg = _lib.su2gpu_last_error().decode(errors="replace")
        raise SU2GPUError(f"su2gpu error {rc}: {msg}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Lattice:
    """One lattice resident on the GPU (a handle of the C ABI).

    dims = (Lx, Ly, Lz, Lt, nd) as for gaugeconfig<G>(Lx, Ly, Lz, Lt, nd, beta).
    Host arrays follow the reference AoS order: links (V*nd, 4) / (V*nd,),
    momenta (V*nd, 3) / (V*nd,).
    """

    def __init__(self, group, dims, beta, xi=1.0, precision=FP64):
        lib = load_library()
        self.group, self.dims, self.precision = group, tuple(int(d) for d in dims), precision
        Lx, Ly, Lz, Lt, nd = self.dims
        self.V = Lx * Ly * Lz * Lt
        self.nd = nd
        self.nlinks = self.V * nd
        self.ncomp = 4 if group == SU2 else 1
        self.nmom = 3 if group == SU2 else 1
        h = C.c_void_p()
        _ck(lib.su2gpu_create(C.byref(h), group, Lx, Ly, Lz, Lt, nd, beta, xi, precision))
        self._h = h
        self.lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self.lib.su2gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- shapes
    def link_shape(self):
        return (self.nlinks, 4) if self.group == SU2 else (self.nlinks,)

    def mom_shape(self):
        return (self.nlinks, 3) if self.group == SU2 else (self.nlinks,)

    # -- fields
    def set_links(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.size == self.nlinks * self.ncomp
        _ck(self.lib.su2gpu_set_links(self._h, _ptr(U), LAYOUT_PACKED))

    def get_links(self):
        U = np.zeros(self.link_shape())
        _ck(self.lib.su2gpu_get_links(self._h, _ptr(U), LAYOUT_PACKED))
        return U

    def set_links_object(self, raw: bytes):
        buf = np.frombuffer(raw, dtype=np.uint8).copy()
        assert buf.size == self.nlinks * (self.ncomp + 1) * 8
        _ck(self.lib.su2gpu_set_links(self._h, _ptr(buf), LAYOUT_OBJECT))

    def get_links_object(self) -> bytes:
        buf = np.zeros(self.nlinks * (self.ncomp + 1) * 8, dtype=np.uint8)
        _ck(self.lib.su2gpu_get_links(self._h, _ptr(buf), LAYOUT_OBJECT))
        return buf.tobytes()

    def set_momenta(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.size == self.nlinks * self.nmom
        _ck(self.lib.su2gpu_set_momenta(self._h, _ptr(P)))

    def get_momenta(self):
        P = np.zeros(self.mom_shape())
        _ck(self.lib.su2gpu_get_momenta(self._h, _ptr(P)))
        return P

    def flip_momenta(self):
        _ck(self.lib.su2gpu_flip_momenta(self._h))

    def set_fp32_storage_emulation(self, links, momenta):
        """FP64 form with the updated links and / or momenta rounded to FP32
        (a precision study; generic MD kernels)."""
        _ck(self.lib.su2gpu_set_fp32_storage_emulation(self._h, int(bool(links)), int(bool(momenta))))

    def initnormal(self, seed):
        _ck(self.lib.su2gpu_initnormal(self._h, seed))

    # -- starts
    def hotstart(self, seed, delta=1.0):
        _ck(self.lib.su2gpu_hotstart(self._h, seed, delta))

    def coldstart(self):
        _ck(self.lib.su2gpu_coldstart(self._h))

    def haarstart(self, seed):
        _ck(self.lib.su2gpu_haarstart(self._h, seed))

    def random_gauge_trafo(self, seed=654321):
        _ck(self.lib.su2gpu_random_gauge_trafo(self._h, seed))

    def restoreSU(self):
        _ck(self.lib.su2gpu_restoreSU(self._h))

    # -- sums
    def retr_sum_plaquettes(self, xi=1.0, anisotropic=False, spatial_only=False):
        r = C.c_double()
        _ck(self.lib.su2gpu_retr_sum_plaquettes(self._h, xi, int(anisotropic), int(spatial_only),
                                                C.byref(r)))
        return r.value

    def retr_sum_plaquettes_noalloc(self, sum_dev_ptr, xi=1.0, anisotropic=False,
                                    spatial_only=False):
        """The plaquette sum written to device memory at sum_dev_ptr (one
        double) on the handle's stream, with no allocation; the caller
        synchronises (synchronize()) before reading it."""
        _ck(self.lib.su2gpu_retr_sum_plaquettes_noalloc(self._h, xi, int(anisotropic),
                                                        int(spatial_only), C.c_void_p(sum_dev_ptr)))

    def gauge_energy(self, spatial_only=False):
        r = C.c_double()
        _ck(self.lib.su2gpu_gauge_energy(self._h, int(spatial_only), C.byref(r)))
        return r.value

    def gauge_action(self):
        r = C.c_double()
        _ck(self.lib.su2gpu_gauge_action(self._h, C.byref(r)))
        return r.value

    def kinetic_energy(self):
        r = C.c_double()
        _ck(self.lib.su2gpu_kinetic_energy(self._h, C.byref(r)))
        return r.value

    # -- MD
    def derivative(self, fac=1.0, deriv=None):
        D = np.zeros(self.mom_shape()) if deriv is None else np.ascontiguousarray(deriv, dtype=np.float64).copy()
        _ck(self.lib.su2gpu_derivative(self._h, fac, _ptr(D)))
        return D

    def update_momenta(self, dtau):
        _ck(self.lib.su2gpu_update_momenta(self._h, dtau))

    def update_gauge(self, dtau):
        _ck(self.lib.su2gpu_update_gauge(self._h, dtau))

    def integrate(self, integrator, nsteps, tau=1.0, restore=True, exponent=0):
        if isinstance(integrator, str):
            integrator = INTEGRATORS[integrator]
        _ck(self.lib.su2gpu_integrate(self._h, integrator, nsteps, tau, int(restore), exponent))

    def md_update(self, seed, integrator, nsteps, tau=1.0, exponent=0, revtest=False,
                  momenta=MOMENTA_HOST):
        if isinstance(integrator, str):
            integrator = INTEGRATORS[integrator]
        r = MDResult()
        _ck(self.lib.su2gpu_md_update(self._h, seed, integrator, nsteps, tau, exponent,
                                      int(revtest), momenta, C.byref(r)))
        return r.as_dict()

    def md_trajectory(self, momenta, integrator, nsteps, tau=1.0, exponent=0, revtest=False):
        r = MDResult()
        P = None if momenta is None else np.ascontiguousarray(momenta, dtype=np.float64)
        _ck(self.lib.su2gpu_md_trajectory(self._h, None if P is None else _ptr(P), integrator,
                                          nsteps, tau, exponent, int(revtest), C.byref(r)))
        return r.as_dict()

    def md_finish(self, accept):
        _ck(self.lib.su2gpu_md_finish(self._h, int(accept)))

    # -- observables
    def sweep(self, seed, sweep, delta, n_hit):
        """n_hit Metropolis hits on every link (flat-sweep.hh:64-164); returns
        (rate, rate_time) as the reference's sweep does."""
        r, rt = C.c_double(), C.c_double()
        _ck(self.lib.su2gpu_sweep(self._h, seed, sweep, delta, n_hit, C.byref(r), C.byref(rt)))
        return r.value, rt.value

    def energy_density(self, clover=True, ss=False):
        E, Q = C.c_double(), C.c_double()
        _ck(self.lib.su2gpu_energy_density(self._h, int(clover), int(ss), C.byref(E), C.byref(Q)))
        return E.value, Q.value

    def runge_kutta(self, eps):
        _ck(self.lib.su2gpu_runge_kutta(self._h, eps))

    def gradient_flow(self, tmax, eps, tstart=0.0, max_rows=100000):
        rows = np.zeros((max_rows, 9))
        n = C.c_size_t()
        _ck(self.lib.su2gpu_gradient_flow(self._h, tmax, eps, tstart, _ptr(rows), max_rows,
                                          C.byref(n)))
        return rows[: n.value]

    # -- gradient-flow adjoint
    def flow_forward(self, eps, nsteps, ckpt_every=20, observable=OBS_CLOVER_E):
        v = C.c_double()
        _ck(self.lib.su2gpu_flow_forward(self._h, eps, nsteps, ckpt_every, observable, C.byref(v)))
        return v.value

    def flow_backward(self):
        g = np.zeros(self.mom_shape())
        _ck(self.lib.su2gpu_flow_backward(self._h, _ptr(g)))
        return g

    # -- files
    def save(self, path):
        _ck(self.lib.su2gpu_save(self._h, path.encode()))

    def load(self, path):
        _ck(self.lib.su2gpu_load(self._h, path.encode()))

    # -- timing
    def synchronize(self):
        _ck(self.lib.su2gpu_synchronize(self._h))

    def profile(self, on=True):
        _ck(self.lib.su2gpu_profile_enable(self._h, int(on)))

    def profile_reset(self):
        _ck(self.lib.su2gpu_profile_reset(self._h))

    def profile_read(self, name):
        ms, cnt = C.c_double(), C.c_long()
        _ck(self.lib.su2gpu_profile_read(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class Batch:
    """Many independent lattices advanced together (su2gpu_batch_* of the C ABI).

    members: sequence of dicts with keys group, dims=(Lx,Ly,Lz,Lt,nd), beta,
    integrator ('leapfrog'|'omf2' or id), nsteps, tau, seed.
    """

    def __init__(self, members, precision=FP32):
        lib = load_library()
        arr = (BatchMember * len(members))()
        self.members = []
        for i, m in enumerate(members):
            Lx, Ly, Lz, Lt, nd = m["dims"]
            integ = m.get("integrator", LEAPFROG)
            if isinstance(integ, str):
                integ = INTEGRATORS[integ]
            arr[i] = BatchMember(m["group"], Lx, Ly, Lz, Lt, nd, m["beta"], integ, m["nsteps"],
                                 m.get("tau", 1.0), m.get("seed", 13526463 + 1000 * i))
            self.members.append(dict(m, V=Lx * Ly * Lz * Lt, nd=nd))
        h = C.c_void_p()
        _ck(lib.su2gpu_batch_create(C.byref(h), arr, len(members), precision))
        self._h, self.lib, self.n = h, lib, len(members)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.su2gpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _shape(self, i, mom=False):
        m = self.members[i]
        n = m["V"] * m["nd"]
        if m["group"] == SU2:
            return (n, 3 if mom else 4)
        return (n,)

    def set_links(self, i, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == self._shape(i)
        _ck(self.lib.su2gpu_batch_set_links(self._h, i, _ptr(U)))

    def get_links(self, i):
        U = np.zeros(self._shape(i))
        _ck(self.lib.su2gpu_batch_get_links(self._h, i, _ptr(U)))
        return U

    def set_momenta(self, i, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == self._shape(i, True)
        _ck(self.lib.su2gpu_batch_set_momenta(self._h, i, _ptr(P)))

    def trajectory(self, traj, momenta=MOMENTA_DEVICE, metropolis=True):
        res = (MDResult * self.n)()
        _ck(self.lib.su2gpu_batch_trajectory(self._h, momenta, traj, int(metropolis), res))
        return [r.as_dict() for r in res]

    def set_launch(self, block, grid=0):
        _ck(self.lib.su2gpu_batch_set_launch(self._h, block, grid))

    def get_launch(self):
        b, g = C.c_int(), C.c_int()
        _ck(self.lib.su2gpu_batch_get_launch(self._h, C.byref(b), C.byref(g)))
        return b.value, g.value

    def autotune(self, cache_path=None):
        b, g = C.c_int(), C.c_int()
        _ck(self.lib.su2gpu_batch_autotune(self._h, cache_path.encode() if cache_path else None,
                                           C.byref(b), C.byref(g)))
        return b.value, g.value

    def use_graph(self, on=True):
        _ck(self.lib.su2gpu_batch_use_graph(self._h, int(on)))

    def link_updates(self):
        v = C.c_double()
        _ck(self.lib.su2gpu_batch_link_updates(self._h, C.byref(v)))
        return v.value

    def profile(self, traj):
        """One eager trajectory timed per launch group (su2gpu_batch_profile)."""
        out = (C.c_double * 14)()
        _ck(self.lib.su2gpu_batch_profile(self._h, traj, out, 14))
        v = list(out)
        parts = ("su2_d4", "su2_d3", "generic")
        return {"ms_total": v[0], "ms_momenta": v[1], "ms_plaq": v[2],
                "ms_tick": dict(zip(parts, v[3:6])), "ms_restore": v[6], "ms_reduce": v[7],
                "link_updates": dict(zip(parts, v[8:11])),
                "active_tile_fraction": dict(zip(parts, v[11:14]))}


def ensemble_members(seed=13526463):
    """The mixed ensemble: 128 U(1) d=2 lattices, L uniform in {16..128}, beta in [1,6];
    64 SU(2) d=3, L in {8..24}, beta in [3,8]; 32 SU(2) d=4, L in {8,12,16}, beta in
    [2.2,2.6]; leapfrog tau = 1 with 10..40 steps, seeded per lattice."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(128):
        L = int(rng.integers(16, 129))
        out.append(dict(group=U1, dims=(L, 1, 1, L, 2), beta=float(rng.uniform(1, 6))))
    for _ in range(64):
        L = int(rng.integers(8, 25))
        out.append(dict(group=SU2, dims=(L, L, 1, L, 3), beta=float(rng.uniform(3, 8))))
    for _ in range(32):
        L = int(rng.choice([8, 12, 16]))
        out.append(dict(group=SU2, dims=(L, L, L, L, 4), beta=float(rng.uniform(2.2, 2.6))))
    for i, m in enumerate(out):
        m.update(integrator=LEAPFROG, nsteps=int(rng.integers(10, 41)), tau=1.0,
                 seed=seed + 1000 * i)
    return out


def copy_probe(nbytes, reps=20):
    """Average ms of a plain device copy of nbytes (read + write nbytes each)."""
    lib = load_library()
    ms = C.c_double()
    _ck(lib.su2gpu_copy_probe(int(nbytes), reps, C.byref(ms)))
    return ms.value


def device_count() -> int:
    lib = load_library()
    n = C.c_int()
    _ck(lib.su2gpu_device_count(C.byref(n)))
    return n.value
