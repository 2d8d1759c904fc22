"""fmms_amd — MI355X (gfx950) FMM force evaluator, Python face of libfmmh.

Mirrors the reference solver API `crates/fmm` (Vedant9500/FMMS):

    Fmm(pos, mass, p, eps, n_leaf)          Fmm::new            fmm.rs:26-33
    Fmm.with_center(..., mode)              Fmm::with_center    fmm.rs:36-49
    fmm.refresh(pos, mass)                  Fmm::refresh        fmm.rs:54-56
    fmm.run(pos, mass) -> Result            Fmm::run            fmm.rs:60-62
    fmm.run_threaded(pos, mass, nthreads)   Fmm::run_threaded   fmm.rs:65-67
    Result(accel [n,3], pot [n], n_cells)   fmm::Result         fmm.rs:17-21
    plummer / hernquist                     models.rs:35-82

A particle list `&[Particle]` becomes two arrays: pos [n, 3] and mass [n]
(float64, particle-id order). Everything runs on the GPU through the C ABI
declared in include/fmmh.h; there is no CPU fallback. If libfmmh.so is
missing or cannot be loaded, the first call into it raises.
"""
import ctypes as ct
import math
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmmh.so")

FMMH_OK = 0
FMMH_E_EMPTY = 1
FMMH_E_PARAM = 2
FMMH_E_OVER_BUDGET = 3
FMMH_E_OVERFULL_LEAF = 4
FMMH_E_NONFINITE = 5
FMMH_E_HIP = 6
FMMH_E_STATE = 7
FMMH_FLAG_TIMING = 1
FMMH_PREC_FP64 = 0
FMMH_PREC_FP32BULK = 1
MAX_P = 40
MAX_SYSTEMS = 4096

_CODES = {1: "EMPTY", 2: "PARAM", 3: "OVER_BUDGET", 4: "OVERFULL_LEAF", 5: "NONFINITE", 6: "HIP",
          7: "STATE"}


class CenterMode:
    """tree.rs:19-25"""
    Com = 0
    Sphere = 1


class FmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmmh error {_CODES.get(code, code)}: {msg}")
        self.code = code


class _Params(ct.Structure):
    _fields_ = [("p", ct.c_int), ("eps_mac", ct.c_double), ("theta_max", ct.c_double),
                ("n_leaf", ct.c_int), ("softening", ct.c_double), ("center_mode", ct.c_int),
                ("precision", ct.c_int), ("flags", ct.c_uint), ("hbm_budget_bytes", ct.c_double),
                ("soft_mac", ct.c_double)]


class _LeapOpts(ct.Structure):
    _fields_ = [("dt", ct.c_double), ("steps", ct.c_int), ("refresh_every", ct.c_int),
                ("dt_eta", ct.c_double), ("dt_eps_len", ct.c_double), ("accel_cap", ct.c_double),
                ("drift_tol", ct.c_double), ("flags", ct.c_uint)]


class _Tuning(ct.Structure):
    _fields_ = [("m2l_chunk", ct.c_int), ("p2p_item_work", ct.c_uint64)]


FMMH_LF_ACC_VALID = 1
FMMH_LF_GRAPH = 2

_pd = ct.POINTER(ct.c_double)
_pu32 = ct.POINTER(ct.c_uint32)
_pi32 = ct.POINTER(ct.c_int32)
_pu64 = ct.POINTER(ct.c_uint64)
_pu8 = ct.POINTER(ct.c_uint8)
_vp = ct.c_void_p
_sz = ct.c_size_t

# Exported entry points and their signatures (must match include/fmmh.h).
SIGNATURES = {
    "fmmh_default_params": (None, [ct.POINTER(_Params)]),
    "fmmh_create": (ct.c_int, [ct.POINTER(_Params), ct.POINTER(_vp)]),
    "fmmh_destroy": (None, [_vp]),
    "fmmh_last_error": (ct.c_char_p, [_vp]),
    "fmmh_build": (ct.c_int, [_vp, _sz, _pd, _pd]),
    "fmmh_evaluate": (ct.c_int, [_vp, _sz, _pd, _pd, _pd, _pd]),
    "fmmh_accel": (ct.c_int, [_vp, _sz, _pd, _pd, _pd, _pd]),
    "fmmh_build_device": (ct.c_int, [_vp, _sz, _vp, _vp]),
    "fmmh_evaluate_device": (ct.c_int, [_vp, _sz, _vp, _vp, _vp, _vp]),
    "fmmh_accel_device": (ct.c_int, [_vp, _sz, _vp, _vp, _vp, _vp]),
    "fmmh_num_cells": (_sz, [_vp]),
    "fmmh_device_bytes": (_sz, [_vp]),
    "fmmh_default_tuning": (None, [ct.POINTER(_Tuning)]),
    "fmmh_set_tuning": (ct.c_int, [_vp, ct.POINTER(_Tuning)]),
    "fmmh_get_tuning": (ct.c_int, [_vp, ct.POINTER(_Tuning)]),
    "fmmh_get_tree": (ct.c_int, [_vp, _pu32, _pu64, _pu32, _pu32, _pi32, _pu8, _pu32, _pd, _pd, _pd]),
    "fmmh_get_list_sizes": (ct.c_int, [_vp, ct.POINTER(_sz), ct.POINTER(_sz)]),
    "fmmh_get_lists": (ct.c_int, [_vp, _pu32, _pu32]),
    "fmmh_get_moments": (ct.c_int, [_vp, _pd]),
    "fmmh_get_locals": (ct.c_int, [_vp, _pd]),
    "fmmh_get_pows": (ct.c_int, [_vp, _pd]),
    "fmmh_get_f": (ct.c_int, [_vp, _pd, _pd]),
    "fmmh_stage_times": (ct.c_int, [_vp, _pd, ct.c_int]),
    "fmmh_stage_name": (ct.c_char_p, [ct.c_int]),
    "fmmh_work_counts": (ct.c_int, [_vp, _pu64]),
    "fmmh_p2p_plan_counts": (ct.c_int, [_vp, _pu64]),
    "fmmh_direct_sampled": (ct.c_int, [_sz, _pd, _pd, ct.c_double, _sz, _pu32, _pd, _pd, _pd]),
    "fmmh_direct_sampled_noalloc": (ct.c_int, [_sz, _vp, _vp, ct.c_double, _sz, _vp, _vp, _vp,
                                               _vp, _vp]),
    "fmmh_plummer": (None, [_sz, _pd, ct.c_uint64, _pd, _pd]),
    "fmmh_hernquist": (None, [_sz, _pd, ct.c_uint64, _pd, _pd]),
    "fmmh_plummer_trunc": (None, [_sz, _pd, ct.c_double, ct.c_uint64, _pd, _pd]),
    "fmmh_set_systems": (ct.c_int, [_vp, _sz, _pu64]),
    "fmmh_evaluate_frozen": (ct.c_int, [_vp, _sz, _pd, _pd, _pd, _pd]),
    "fmmh_evaluate_frozen_device": (ct.c_int, [_vp, _sz, _vp, _vp, _vp, _vp]),
    "fmmh_leapfrog_default": (None, [ct.POINTER(_LeapOpts)]),
    "fmmh_leapfrog_stats": (ct.c_int, [_vp, _pu64]),
    "fmmh_leapfrog_device": (ct.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp,
                                        ct.POINTER(_LeapOpts), _pd]),
    "fmmh_ic_plummer": (ct.c_int, [_sz, ct.c_double, ct.c_uint64, _pd, _pd, _pd]),
    "fmmh_ic_hernquist": (ct.c_int, [_sz, ct.c_double, ct.c_double, ct.c_uint64, _pd, _pd, _pd]),
    "fmmh_ic_disk": (ct.c_int, [_sz, ct.c_double, ct.c_double, ct.c_double, ct.c_double,
                                ct.c_double, ct.c_uint64, _pd, _pd, _pd]),
    "fmmh_ic_halos_cube": (ct.c_int, [_sz, ct.c_int, ct.c_double, ct.c_double, ct.c_double,
                                      ct.c_double, ct.c_double, ct.c_uint64, _pd, _pd, _pd]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fmms_amd.build`")
    lib = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    """The loaded libfmmh.so (raises if it is missing or stale)."""
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _p(a, t=_pd):
    return a.ctypes.data_as(t)


Result = namedtuple("Result", ["accel", "pot", "n_cells"])
Tree = namedtuple("Tree", ["order", "keys", "first", "count", "children", "leaf", "level", "cen",
                           "mass", "rho"])


def dim(p):
    return (p + 1) * (p + 1)


class Fmm:
    """GPU FMM with the reference `Fmm` interface. Construction builds the
    tree (Fmm::new); `run` evaluates accelerations and potentials.

    Extensions beyond the reference (defaults reproduce it): theta_max
    (opening-angle bound, 1.0 = theta < 1), softening (Plummer softening of
    the kernel, near and far field; 0 = pure 1/r), soft_mac (softening guard
    of the MAC: M2L only where softening^2 (rhoA + rhoB)^3 <= soft_mac r^5;
    0 = off), eps = inf for the theta-only MAC."""

    def __init__(self, pos, mass, p, eps, n_leaf, center_mode=CenterMode.Com, theta_max=1.0,
                 softening=0.0, timing=False, hbm_budget_bytes=0.0, precision="fp64",
                 soft_mac=0.0, m2l_chunk=None, p2p_item_work=None):
        prm = _Params()
        lib().fmmh_default_params(ct.byref(prm))
        prm.p, prm.eps_mac, prm.theta_max, prm.n_leaf = int(p), float(eps), float(theta_max), int(n_leaf)
        prm.softening, prm.center_mode = float(softening), int(center_mode)
        prm.flags = FMMH_FLAG_TIMING if timing else 0
        prm.hbm_budget_bytes = float(hbm_budget_bytes)
        prm.soft_mac = float(soft_mac)
        prm.precision = {"fp64": FMMH_PREC_FP64, "fp32bulk": FMMH_PREC_FP32BULK}[precision]
        self.precision = precision
        self.p, self.eps, self.n_leaf = int(p), float(eps), int(n_leaf)
        self.theta_max, self.softening, self.center_mode = float(theta_max), float(softening), int(center_mode)
        h = _vp()
        rc = lib().fmmh_create(ct.byref(prm), ct.byref(h))
        if rc != FMMH_OK:
            raise FmmError(rc, "invalid parameters")
        self._h = h
        self.n = None
        if m2l_chunk is not None or p2p_item_work is not None:
            self.set_tuning(m2l_chunk=m2l_chunk, p2p_item_work=p2p_item_work)
        if pos is not None:
            self.refresh(pos, mass)

    def tuning(self):
        """Tile shapes of this context: {"m2l_chunk", "p2p_item_work"}."""
        t = _Tuning()
        self._check(lib().fmmh_get_tuning(self._h, ct.byref(t)))
        return {"m2l_chunk": int(t.m2l_chunk), "p2p_item_work": int(t.p2p_item_work)}

    def set_tuning(self, m2l_chunk=None, p2p_item_work=None):
        """Set tile shapes (None keeps the current value); they take effect at
        the next full evaluation."""
        cur = self.tuning()
        t = _Tuning()
        t.m2l_chunk = int(cur["m2l_chunk"] if m2l_chunk is None else m2l_chunk)
        t.p2p_item_work = int(cur["p2p_item_work"] if p2p_item_work is None else p2p_item_work)
        self._check(lib().fmmh_set_tuning(self._h, ct.byref(t)))

    @classmethod
    def with_center(cls, pos, mass, p, eps, n_leaf, mode, **kw):
        return cls(pos, mass, p, eps, n_leaf, center_mode=mode, **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib().fmmh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != FMMH_OK:
            raise FmmError(rc, lib().fmmh_last_error(self._h).decode())

    def refresh(self, pos, mass):
        mass = _f64(mass)
        n = mass.shape[0]
        pos = _f64(pos, (n, 3)) if n else _f64(pos)
        self._check(lib().fmmh_build(self._h, n, _p(pos), _p(mass)))
        self.n = n

    @property
    def n_cells(self):
        return int(lib().fmmh_num_cells(self._h))

    def device_bytes(self):
        """Device memory this context holds (bytes)."""
        return int(lib().fmmh_device_bytes(self._h))

    def run(self, pos, mass):
        mass = _f64(mass)
        n = mass.shape[0]
        pos = _f64(pos, (n, 3))
        acc = np.empty((n, 3))
        pot = np.empty(n)
        self._check(lib().fmmh_evaluate(self._h, n, _p(pos), _p(mass), _p(acc), _p(pot)))
        return Result(acc, pot, self.n_cells)

    def run_threaded(self, pos, mass, nthreads=1):
        """Same as `run`; the device decides its own parallelism."""
        return self.run(pos, mass)

    def accel(self, pos, mass):
        """Physics::accel (sim.rs:5-7): rebuild the tree and evaluate."""
        mass = _f64(mass)
        n = mass.shape[0]
        pos = _f64(pos, (n, 3))
        acc = np.empty((n, 3))
        pot = np.empty(n)
        self._check(lib().fmmh_accel(self._h, n, _p(pos), _p(mass), _p(acc), _p(pot)))
        self.n = n
        return Result(acc, pot, self.n_cells)

    # ---- device-resident entry points (raw device pointers, e.g. torch .data_ptr())
    def accel_device(self, n, pos_ptr, mass_ptr, acc_ptr, pot_ptr=0):
        self._check(lib().fmmh_accel_device(self._h, n, _vp(pos_ptr), _vp(mass_ptr), _vp(acc_ptr),
                                           _vp(pot_ptr) if pot_ptr else None))
        self.n = n

    def build_device(self, n, pos_ptr, mass_ptr):
        self._check(lib().fmmh_build_device(self._h, n, _vp(pos_ptr), _vp(mass_ptr)))
        self.n = n

    def evaluate_device(self, n, pos_ptr, mass_ptr, acc_ptr, pot_ptr=0):
        self._check(lib().fmmh_evaluate_device(self._h, n, _vp(pos_ptr), _vp(mass_ptr), _vp(acc_ptr),
                                              _vp(pot_ptr) if pot_ptr else None))

    # ---- batched ensembles
    def set_systems(self, sizes):
        """Treat the bodies as independent systems of the given sizes (in
        input order): one forest, one launch per stage, no cross-system
        interactions. `sizes` of length 1 restores single-system mode."""
        sizes = np.asarray(sizes, dtype=np.uint64)
        off = np.zeros(len(sizes) + 1, np.uint64)
        off[1:] = np.cumsum(sizes)
        self._check(lib().fmmh_set_systems(self._h, len(sizes), _p(off, _pu64)))
        self.n = None
        return off

    # ---- temporal reuse and time integration
    def run_frozen(self, pos, mass):
        """Evaluate on the tree and lists of the last full evaluation
        (gpufmm.rs:2710-2741 frozen step): no rebuild, no walk."""
        mass = _f64(mass)
        n = mass.shape[0]
        pos = _f64(pos, (n, 3))
        acc = np.empty((n, 3))
        pot = np.empty(n)
        self._check(lib().fmmh_evaluate_frozen(self._h, n, _p(pos), _p(mass), _p(acc), _p(pot)))
        return Result(acc, pot, self.n_cells)

    def evaluate_frozen_device(self, n, pos_ptr, mass_ptr, acc_ptr, pot_ptr):
        self._check(lib().fmmh_evaluate_frozen_device(self._h, n, _vp(pos_ptr), _vp(mass_ptr),
                                                     _vp(acc_ptr), _vp(pot_ptr)))

    def leapfrog_device(self, n, pos_ptr, vel_ptr, mass_ptr, acc_ptr, pot_ptr, dt, steps,
                        refresh_every=1, dt_eta=0.0, dt_eps_len=0.02, accel_cap=0.0,
                        acc_valid=False, graph=False, drift_tol=0.25):
        """Kick-drift-kick steps on device-resident id-order state
        (sim.rs:539-594); returns the time advanced."""
        o = _LeapOpts()
        lib().fmmh_leapfrog_default(ct.byref(o))
        o.dt, o.steps, o.refresh_every = float(dt), int(steps), int(refresh_every)
        o.dt_eta, o.dt_eps_len, o.accel_cap = float(dt_eta), float(dt_eps_len), float(accel_cap)
        o.drift_tol = float(drift_tol)
        o.flags = (FMMH_LF_ACC_VALID if acc_valid else 0) | (FMMH_LF_GRAPH if graph else 0)
        t = ct.c_double(0.0)
        self._check(lib().fmmh_leapfrog_device(self._h, n, _vp(pos_ptr), _vp(vel_ptr),
                                              _vp(mass_ptr), _vp(acc_ptr), _vp(pot_ptr),
                                              ct.byref(o), ct.byref(t)))
        self.n = n
        return t.value

    def leapfrog_stats(self):
        out = np.zeros(4, np.uint64)
        self._check(lib().fmmh_leapfrog_stats(self._h, _p(out, _pu64)))
        return dict(zip(["steps", "rebuilds", "frozen_steps", "drift_rebuilds"],
                        (int(v) for v in out)))

    # ---- parity views
    def tree(self):
        n, nc = self.n, self.n_cells
        t = Tree(np.empty(n, np.uint32), np.empty(n, np.uint64), np.empty(nc, np.uint32),
                 np.empty(nc, np.uint32), np.empty((nc, 8), np.int32), np.empty(nc, np.uint8),
                 np.empty(nc, np.uint32), np.empty((nc, 3)), np.empty(nc), np.empty(nc))
        self._check(lib().fmmh_get_tree(self._h, _p(t.order, _pu32), _p(t.keys, _pu64),
                                       _p(t.first, _pu32), _p(t.count, _pu32),
                                       _p(t.children, _pi32), _p(t.leaf, _pu8),
                                       _p(t.level, _pu32), _p(t.cen), _p(t.mass), _p(t.rho)))
        return t

    def lists(self):
        a, b = _sz(), _sz()
        self._check(lib().fmmh_get_list_sizes(self._h, ct.byref(a), ct.byref(b)))
        m2l = np.empty((a.value, 2), np.uint32)
        p2p = np.empty((b.value, 2), np.uint32)
        self._check(lib().fmmh_get_lists(self._h, _p(m2l, _pu32), _p(p2p, _pu32)))
        return m2l, p2p

    def moments(self):
        out = np.empty((self.n_cells, dim(self.p)), np.complex128)
        self._check(lib().fmmh_get_moments(self._h, out.ctypes.data_as(_pd)))
        return out

    def locals(self):
        out = np.empty((self.n_cells, dim(self.p)), np.complex128)
        self._check(lib().fmmh_get_locals(self._h, out.ctypes.data_as(_pd)))
        return out

    def pows(self):
        out = np.empty((self.n_cells, self.p + 1))
        self._check(lib().fmmh_get_pows(self._h, _p(out)))
        return out

    def f(self):
        f = np.empty(self.n)
        fmin = np.empty(self.n_cells)
        self._check(lib().fmmh_get_f(self._h, _p(f), _p(fmin)))
        return f, fmin

    def stage_times(self):
        buf = (ct.c_double * 32)()
        k = lib().fmmh_stage_times(self._h, buf, 32)
        return {lib().fmmh_stage_name(i).decode(): buf[i] for i in range(k)}

    def work_counts(self):
        out = np.zeros(8, np.uint64)
        self._check(lib().fmmh_work_counts(self._h, _p(out, _pu64)))
        keys = ["n", "ncells", "nleaves", "m2l_pairs", "p2p_pairs", "p2p_interactions",
                "max_level", "walk_rounds"]
        d = dict(zip(keys, (int(v) for v in out)))
        pl = np.zeros(2, np.uint64)
        self._check(lib().fmmh_p2p_plan_counts(self._h, _p(pl, _pu64)))
        d["p2p_items"], d["p2p_split_leaves"] = int(pl[0]), int(pl[1])
        return d


# ------------------------------------------------------------------ models.rs
def plummer(n, center=(0.0, 0.0, 0.0), seed=42):
    pos = np.empty((n, 3))
    mass = np.empty(n)
    c = _f64(center)
    lib().fmmh_plummer(n, _p(c), seed, _p(pos), _p(mass))
    return pos, mass


def plummer_truncated(n, rmax=20.0, center=(0.0, 0.0, 0.0), seed=42):
    """Plummer sphere with radii redrawn until r <= rmax (total mass 1)."""
    pos = np.empty((n, 3))
    mass = np.empty(n)
    c = _f64(center)
    lib().fmmh_plummer_trunc(n, _p(c), float(rmax), seed, _p(pos), _p(mass))
    return pos, mass


def hernquist(n, center=(0.0, 0.0, 0.0), seed=7):
    pos = np.empty((n, 3))
    mass = np.empty(n)
    c = _f64(center)
    lib().fmmh_hernquist(n, _p(c), seed, _p(pos), _p(mass))
    return pos, mass


# ------------------------------------------------- initial-condition library
class System:
    """Bodies of one initial condition: pos [n,3], vel [n,3], mass [n] (G = 1)."""

    def __init__(self, pos, vel, mass):
        self.pos, self.vel, self.mass = pos, vel, mass

    @property
    def n(self):
        return self.mass.shape[0]


def _ic(fn, n, *args):
    if n <= 0:
        raise FmmError(FMMH_E_PARAM, "n must be positive")
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    rc = fn(n, *args, _p(pos), _p(vel), _p(mass))
    if rc != FMMH_OK:
        raise FmmError(rc, f"{fn.__name__}: invalid arguments")
    return System(pos, vel, mass)


def ic_plummer(n, rmax=20.0, seed=42):
    """Plummer (a = 1, total mass 1) truncated at rmax with isotropic
    distribution-function velocities; positions equal plummer_truncated's."""
    return _ic(lib().fmmh_ic_plummer, n, float(rmax), seed)


def ic_hernquist(n, a=1.0, rmax=100.0, seed=7):
    """Hernquist (scale a, total mass 1) truncated at rmax, Jeans-Maxwellian
    velocities below 0.95 v_esc."""
    return _ic(lib().fmmh_ic_hernquist, n, float(a), float(rmax), seed)


def ic_disk(n, rd=1.0, zd=0.05, rmax=10.0, m_central=0.0, sigma=0.0, seed=11):
    """Exponential sech^2 disk on circular orbits (Keplerian + Freeman v_c)."""
    return _ic(lib().fmmh_ic_disk, n, float(rd), float(zd), float(rmax), float(m_central),
               float(sigma), seed)


def ic_halos_cube(n, nhalo=64, halo_frac=0.9, a_min=0.01, a_max=0.1, mass_decades=2.0,
                  rmax_over_a=20.0, seed=2024):
    """Hernquist halos in a uniform unit-cube background (halo bodies first)."""
    return _ic(lib().fmmh_ic_halos_cube, n, int(nhalo), float(halo_frac), float(a_min),
               float(a_max), float(mass_decades), float(rmax_over_a), seed)


class Integrator:
    """Device-resident N-body state advanced by kick-drift-kick leapfrog with
    this library's FMM forces (the energy-gate loop of
    crates/fmm/tests/energy_gate.rs:156-177 and Sim::step, sim.rs:539-594).
    pos/vel [n,3], mass [n] are copied to the device once; `step` runs on the
    device; `state()` copies pos, vel, acc, pot back."""

    def __init__(self, fmm, pos, vel, mass):
        from . import hip

        mass = _f64(mass)
        n = mass.shape[0]
        self.fmm, self.n = fmm, n
        self.pos = hip.DeviceArray.from_host(_f64(pos, (n, 3)))
        self.vel = hip.DeviceArray.from_host(_f64(vel, (n, 3)))
        self.mass = hip.DeviceArray.from_host(mass)
        self.acc = hip.DeviceArray((n, 3))
        self.pot = hip.DeviceArray((n,))
        self.acc_valid = False
        self.time = 0.0

    def step(self, dt, steps=1, refresh_every=1, graph=False, **kw):
        t = self.fmm.leapfrog_device(self.n, self.pos.ptr, self.vel.ptr, self.mass.ptr,
                                     self.acc.ptr, self.pot.ptr, dt, steps,
                                     refresh_every=refresh_every, acc_valid=self.acc_valid,
                                     graph=graph, **kw)
        self.acc_valid = True
        self.time += t
        return t

    def state(self):
        return (self.pos.download(), self.vel.download(), self.acc.download(),
                self.pot.download())


def sample_ids(n, samp, seed=7):
    """The bench's LCG target picker (crates/bench/src/main.rs:41-47)."""
    out = np.empty(samp, np.uint32)
    s = seed
    mask = (1 << 64) - 1
    for i in range(samp):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = (s >> 33) % n
    return out


def direct_sampled(pos, mass, ids, softening=0.0):
    """FP64 direct sum on the GPU at target ids: (acc [k,3], pot [k], f [k])."""
    mass = _f64(mass)
    n = mass.shape[0]
    pos = _f64(pos, (n, 3))
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    k = ids.shape[0]
    acc = np.empty((k, 3))
    pot = np.empty(k)
    f = np.empty(k)
    rc = lib().fmmh_direct_sampled(n, _p(pos), _p(mass), float(softening), k, _p(ids, _pu32),
                                  _p(acc), _p(pot), _p(f))
    if rc != FMMH_OK:
        raise FmmError(rc, "direct_sampled failed")
    return acc, pot, f


def rel_errors(acc, acc_ref):
    """Per-target relative force error |a - a_ref| / |a_ref|."""
    da = np.sqrt(np.sum((acc - acc_ref) ** 2, axis=1))
    ar = np.sqrt(np.sum(acc_ref ** 2, axis=1))
    return da / np.where(ar > 0, ar, 1.0)
