"""ctypes wrapper of the CPU oracle (liboracle.so, built from fmm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the GPU product is
compared against, and the CPU baseline. Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product never does.

Every function restates the reference solver `crates/fmm` (Vedant9500/FMMS);
the C++ source cites the reference file:line of each piece.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_d = ct.c_double
_pd = ct.POINTER(ct.c_double)
_pu32 = ct.POINTER(ct.c_uint32)
_pi32 = ct.POINTER(ct.c_int32)
_pu64 = ct.POINTER(ct.c_uint64)
_pu8 = ct.POINTER(ct.c_uint8)
_sz = ct.c_size_t


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


_LIB_TIMING = os.path.join(_HERE, "liboracle_timing.so")


def build_timing():
    """Compile liboracle_timing.so (the CPU-baseline build: host ISA,
    contraction on, 8-wide P2P, dynamic threads) on this machine."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "-B", "timing"], check=True)


def _load(path=None):
    path = path or _LIB
    if not os.path.exists(path):
        build_timing() if path == _LIB_TIMING else build()
    lib = ct.CDLL(path)
    sig = {
        "orc_plummer": (None, [_sz, _pd, ct.c_uint64, _pd, _pd]),
        "orc_hernquist": (None, [_sz, _pd, ct.c_uint64, _pd, _pd]),
        "orc_rng_unit_first": (_d, [ct.c_uint64, _sz, _pd]),
        "orc_dilate": (ct.c_uint64, [ct.c_uint64]),
        "orc_morton_key": (ct.c_uint64, [_d, _d, _d, _pd, _d]),
        "orc_theta_table": (None, [_pd, ct.c_int, _pd]),
        "orc_upsilon_table": (None, [_pd, ct.c_int, _pd]),
        "orc_factorials": (None, [ct.c_int, _pd]),
        "orc_binomials": (None, [ct.c_int, _pd]),
        "orc_m2l_mutual": (None, [ct.c_int, _pd, _pd, _pd, _pd, _pd, _pd]),
        "orc_m2l_rotation": (None, [ct.c_int, _pd, _pd, _pd, _pd, _pd, _pd]),
        "orc_m2l_mutual_soft": (None, [ct.c_int, _pd, _pd, _pd, _pd, _d, _pd, _pd]),
        "orc_soft_trace_pair": (None, [ct.c_int, _pd, _pd, _d, _d, _d, _pd, _pd]),
        "orc_l2p_trace": (None, [_pd, _pd, _pd, _pd, _pd]),
        "orc_rot_blocksum": (_sz, [ct.c_int]),
        "orc_rot_swap": (None, [ct.c_int, _pd]),
        "orc_l2l": (None, [ct.c_int, _pd, _pd, _pd, _pd]),
        "orc_l2p": (None, [ct.c_int, _pd, _pd, _pd, _pd, _pd]),
        "orc_direct": (None, [_sz, _pd, _pd, _d, _pd, _pd]),
        "orc_direct_targets": (None, [_sz, _pd, _pd, _d, _sz, _pu32, _pd, _pd, _pd, ct.c_int]),
        "orc_sample_ids": (None, [_sz, _sz, ct.c_uint64, _pu32]),
        "orc_new": (ct.c_void_p, []),
        "orc_free": (None, [ct.c_void_p]),
        "orc_build": (ct.c_int, [ct.c_void_p, _sz, _pd, _pd, ct.c_int, ct.c_int]),
        "orc_ncells": (_sz, [ct.c_void_p]),
        "orc_set_soft_mac": (None, [ct.c_void_p, _d]),
        "orc_max_level": (ct.c_uint32, [ct.c_void_p]),
        "orc_get_tree": (None, [ct.c_void_p, _pu32, _pu64, _pu32, _pu32, _pi32, _pu8, _pu32, _pd, _pd, _pd]),
        "orc_get_bbox": (None, [ct.c_void_p, _pd, _pd, _pd]),
        "orc_run": (ct.c_int, [ct.c_void_p, _pd, _pd, ct.c_int, _d, _d, _d, ct.c_int, _pd, _pd]),
        "orc_dim": (ct.c_int, [ct.c_void_p]),
        "orc_get_moments": (None, [ct.c_void_p, _pd]),
        "orc_get_pows": (None, [ct.c_void_p, _pd]),
        "orc_get_f": (None, [ct.c_void_p, _pd]),
        "orc_get_fmin": (None, [ct.c_void_p, _pd]),
        "orc_n_m2l": (_sz, [ct.c_void_p]),
        "orc_n_p2p": (_sz, [ct.c_void_p]),
        "orc_get_m2l": (None, [ct.c_void_p, _pu32]),
        "orc_get_p2p": (None, [ct.c_void_p, _pu32]),
        "orc_get_locals_walk": (None, [ct.c_void_p, _pd]),
        "orc_get_locals": (None, [ct.c_void_p, _pd]),
        "orc_get_near": (None, [ct.c_void_p, _pd]),
        "orc_get_times": (None, [ct.c_void_p, _pd]),
        "orc_ic_plummer": (ct.c_int, [_sz, _d, ct.c_uint64, _pd, _pd, _pd]),
        "orc_ic_hernquist": (ct.c_int, [_sz, _d, _d, ct.c_uint64, _pd, _pd, _pd]),
        "orc_ic_disk": (ct.c_int, [_sz, _d, _d, _d, _d, _d, ct.c_uint64, _pd, _pd, _pd]),
        "orc_ic_halos_cube": (ct.c_int, [_sz, ct.c_int, _d, _d, _d, _d, _d, ct.c_uint64, _pd,
                                         _pd, _pd]),
        "orc_hernquist_sigma2": (_d, [_d, _d]),
        "orc_warm_start": (None, [_sz, _pd, _pd, _pd, _pd]),
        "orc_kinetic": (_d, [_sz, _pd, _pd]),
        "orc_potential_energy": (_d, [_sz, _pd, _pd]),
        "orc_momentum_rel": (_d, [_sz, _pd, _pd]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None
_lib_timing = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def timing_lib(rebuild=False):
    """The timing build (never the checker). rebuild=True compiles it for
    this host first, since -march=native code may not run on another CPU."""
    global _lib_timing
    if _lib_timing is None:
        if rebuild:
            build_timing()
        _lib_timing = _load(_LIB_TIMING)
    return _lib_timing


def _p(a, ctype=_pd):
    return a.ctypes.data_as(ctype)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dim(p):
    return (p + 1) * (p + 1)


def idx(n, m):
    return n * (n + 1) + m


# ----------------------------------------------------------------- models.rs
def plummer(n, center=(0.0, 0.0, 0.0), seed=42):
    pos = np.empty((n, 3))
    mass = np.empty(n)
    c = _f64(center)
    lib().orc_plummer(n, _p(c), seed, _p(pos), _p(mass))
    return pos, mass


def hernquist(n, center=(0.0, 0.0, 0.0), seed=7):
    pos = np.empty((n, 3))
    mass = np.empty(n)
    c = _f64(center)
    lib().orc_hernquist(n, _p(c), seed, _p(pos), _p(mass))
    return pos, mass


# ------------------------------------------- initial conditions (ic_oracle.cpp)
def _ic(fn, n, *args):
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    rc = fn(n, *args, _p(pos), _p(vel), _p(mass))
    if rc != 0:
        raise ValueError(f"{fn.__name__}: bad arguments")
    return pos, vel, mass


def ic_plummer(n, rmax=20.0, seed=42):
    return _ic(lib().orc_ic_plummer, n, float(rmax), seed)


def ic_hernquist(n, a=1.0, rmax=100.0, seed=7):
    return _ic(lib().orc_ic_hernquist, n, float(a), float(rmax), seed)


def ic_disk(n, rd=1.0, zd=0.05, rmax=10.0, m_central=0.0, sigma=0.0, seed=11):
    return _ic(lib().orc_ic_disk, n, float(rd), float(zd), float(rmax), float(m_central),
               float(sigma), seed)


def ic_halos_cube(n, nhalo=64, halo_frac=0.9, a_min=0.01, a_max=0.1, mass_decades=2.0,
                  rmax_over_a=20.0, seed=2024):
    return _ic(lib().orc_ic_halos_cube, n, int(nhalo), float(halo_frac), float(a_min),
               float(a_max), float(mass_decades), float(rmax_over_a), seed)


def hernquist_sigma2(r, a=1.0):
    return lib().orc_hernquist_sigma2(float(r), float(a))


def warm_start(pos, mass, pot):
    """energy_gate.rs:39-79: virialised sigma(r) Maxwellian velocities."""
    pos, mass, pot = _f64(pos), _f64(mass), _f64(pot)
    vel = np.empty_like(pos)
    lib().orc_warm_start(mass.shape[0], _p(pos), _p(mass), _p(pot), _p(vel))
    return vel


def kinetic(mass, vel):
    mass, vel = _f64(mass), _f64(vel)
    return lib().orc_kinetic(mass.shape[0], _p(mass), _p(vel))


def potential_energy(mass, pot):
    mass, pot = _f64(mass), _f64(pot)
    return lib().orc_potential_energy(mass.shape[0], _p(mass), _p(pot))


def momentum_rel(mass, vel):
    mass, vel = _f64(mass), _f64(vel)
    return lib().orc_momentum_rel(mass.shape[0], _p(mass), _p(vel))


def rng_units(seed, k):
    out = np.empty(k)
    lib().orc_rng_unit_first(seed, k, _p(out))
    return out


# ------------------------------------------------------------ tree / harmonic
def dilate(x):
    return int(lib().orc_dilate(x))


def morton_key(x, y, z, lo, inv_side):
    lo = _f64(lo)
    return int(lib().orc_morton_key(x, y, z, _p(lo), inv_side))


def theta_table(r, p):
    out = np.empty(dim(p), dtype=np.complex128)
    rr = _f64(r)
    lib().orc_theta_table(_p(rr), p, out.ctypes.data_as(_pd))
    return out


def upsilon_table(r, p):
    out = np.empty(dim(p), dtype=np.complex128)
    rr = _f64(r)
    lib().orc_upsilon_table(_p(rr), p, out.ctypes.data_as(_pd))
    return out


def factorials(p):
    out = np.empty(2 * p + 1)
    lib().orc_factorials(p, _p(out))
    return out


def binomials(p):
    out = np.empty(p + 1)
    lib().orc_binomials(p, _p(out))
    return out


def _m2l(fn, p, cen_a, cen_b, ma, mb):
    ca, cb = _f64(cen_a), _f64(cen_b)
    ma = np.ascontiguousarray(ma, dtype=np.complex128)
    mb = np.ascontiguousarray(mb, dtype=np.complex128)
    la = np.zeros(dim(p), dtype=np.complex128)
    lb = np.zeros(dim(p), dtype=np.complex128)
    fn(p, _p(ca), _p(cb), ma.ctypes.data_as(_pd), mb.ctypes.data_as(_pd),
       la.ctypes.data_as(_pd), lb.ctypes.data_as(_pd))
    return la, lb


def m2l_mutual(p, cen_a, cen_b, ma, mb):
    """translate.rs:47-136 — returns (locals_a, locals_b)."""
    return _m2l(lib().orc_m2l_mutual, p, cen_a, cen_b, ma, mb)


def m2l_rotation(p, cen_a, cen_b, ma, mb):
    """rotation.rs:262-435 — returns (locals_a, locals_b)."""
    return _m2l(lib().orc_m2l_rotation, p, cen_a, cen_b, ma, mb)


def m2l_mutual_soft(p, cen_a, cen_b, ma, mb, eps2):
    """Mutual M2L with the harmonic derivatives of the Plummer kernel."""
    ma = np.ascontiguousarray(ma, dtype=np.complex128)
    mb = np.ascontiguousarray(mb, dtype=np.complex128)
    la = np.zeros(dim(p), np.complex128)
    lb = np.zeros(dim(p), np.complex128)
    a, b = _f64(cen_a), _f64(cen_b)
    lib().orc_m2l_mutual_soft(p, _p(a), _p(b), ma.ctypes.data_as(_pd), mb.ctypes.data_as(_pd),
                              float(eps2), la.ctypes.data_as(_pd), lb.ctypes.data_as(_pd))
    return la, lb


SOFT_LAP_WORDS = 10  # L, G (3), H (xx, yy, zz, xy, xz, yz)


def soft_trace_pair(p, cen_src, cen_dst, q, mass, eps2, f_dst):
    """Adds the source-side field (q/6) Lap K to second order into f_dst (in
    place); returns the target-side Laplacian field of the source at the
    target centre: [L, G (3), H (xx, yy, zz, xy, xz, yz)]."""
    assert f_dst.dtype == np.complex128 and f_dst.flags.c_contiguous and f_dst.size >= dim(p)
    lap = np.zeros(SOFT_LAP_WORDS)
    a, b = _f64(cen_src), _f64(cen_dst)
    lib().orc_soft_trace_pair(p, _p(a), _p(b), float(q), float(mass), float(eps2),
                              f_dst.ctypes.data_as(_pd), _p(lap))
    return lap


def l2p_trace(cen, pos, lap10):
    """Target-side trace polynomial of one cell's [L, G, H] at pos."""
    lap10 = _f64(lap10)
    assert lap10.shape == (SOFT_LAP_WORDS,)
    pot = np.zeros(1)
    acc = np.zeros(3)
    lib().orc_l2p_trace(_p(_f64(cen)), _p(_f64(pos)), _p(lap10), _p(pot), _p(acc))
    return pot[0], acc


def rot_swap(p):
    out = np.empty(lib().orc_rot_blocksum(p))
    lib().orc_rot_swap(p, _p(out))
    return out


def l2l(p, parent_cen, child_cen, fpar, fchild=None):
    fpar = np.ascontiguousarray(fpar, dtype=np.complex128)
    fch = np.zeros(dim(p), dtype=np.complex128) if fchild is None else np.array(fchild, dtype=np.complex128)
    a, b = _f64(parent_cen), _f64(child_cen)
    lib().orc_l2l(p, _p(a), _p(b), fpar.ctypes.data_as(_pd), fch.ctypes.data_as(_pd))
    return fch


def l2p(p, cen, pos, f):
    f = np.ascontiguousarray(f, dtype=np.complex128)
    pot = np.zeros(1)
    acc = np.zeros(3)
    a, b = _f64(cen), _f64(pos)
    lib().orc_l2p(p, _p(a), _p(b), f.ctypes.data_as(_pd), _p(pot), _p(acc))
    return pot[0], acc


# ------------------------------------------------------------------ direct.rs
def direct(pos, mass, softening=0.0):
    pos, mass = _f64(pos), _f64(mass)
    n = mass.shape[0]
    acc = np.empty((n, 3))
    pot = np.empty(n)
    lib().orc_direct(n, _p(pos), _p(mass), softening, _p(acc), _p(pot))
    return acc, pot


def direct_targets(pos, mass, ids, softening=0.0, nthreads=1):
    """Exact accel, potential and f = sum m/r^2 at target ids (sampled error)."""
    pos, mass = _f64(pos), _f64(mass)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    nt = ids.shape[0]
    acc = np.empty((nt, 3))
    pot = np.empty(nt)
    f = np.empty(nt)
    lib().orc_direct_targets(mass.shape[0], _p(pos), _p(mass), softening, nt, _p(ids, _pu32),
                             _p(acc), _p(pot), _p(f), nthreads)
    return acc, pot, f


def sample_ids(n, samp, seed=7):
    out = np.empty(samp, dtype=np.uint32)
    lib().orc_sample_ids(n, samp, seed, _p(out, _pu32))
    return out


def scaled_error(a_ref, a_fmm, pos, mass):
    """direct.rs:35-78: (max_scaled, max_rel, rms_scaled, rms_rel)."""
    n = mass.shape[0]
    ids = np.arange(n, dtype=np.uint32)
    _, _, f = direct_targets(pos, mass, ids)
    da = np.sqrt(np.sum((a_fmm - a_ref) ** 2, axis=1))
    aref = np.sqrt(np.sum(a_ref ** 2, axis=1))
    s = da / f
    rel = np.where(aref > 0, da / np.where(aref > 0, aref, 1), 0.0)
    return s.max(), rel.max(), np.sqrt(np.mean(s * s)), np.sqrt(np.mean(rel * rel))


# -------------------------------------------------------------------- Fmm
class Tree:
    pass


class OracleFmm:
    """Mirror of `Fmm` (fmm.rs:10-142) over the oracle: build on construction,
    `refresh` rebuilds, `run` evaluates and keeps every intermediate."""

    def __init__(self, pos, mass, p=8, eps=1e-5, n_leaf=16, center_mode=0, theta_max=1.0,
                 softening=0.0, soft_mac=0.0, timing_build=False):
        """timing_build=True runs the CPU-baseline build (timing_lib())
        instead of the parity build."""
        self.p, self.eps, self.n_leaf = p, eps, n_leaf
        self.center_mode, self.theta_max, self.softening = center_mode, theta_max, softening
        self._lib = timing_lib() if timing_build else lib()
        self._ctx = self._lib.orc_new()
        self._lib.orc_set_soft_mac(self._ctx, soft_mac)
        self.refresh(pos, mass)

    def __del__(self):
        if getattr(self, "_ctx", None):
            self._lib.orc_free(self._ctx)
            self._ctx = None

    def refresh(self, pos, mass):
        pos, mass = _f64(pos), _f64(mass)
        rc = self._lib.orc_build(self._ctx, mass.shape[0], _p(pos), _p(mass), self.n_leaf, self.center_mode)
        if rc != 0:
            raise ValueError("empty input")
        self.n = mass.shape[0]
        self.ncells = self._lib.orc_ncells(self._ctx)

    def tree(self):
        n, nc = self.n, self.ncells
        t = Tree()
        t.order = np.empty(n, np.uint32)
        t.keys = np.empty(n, np.uint64)
        t.first = np.empty(nc, np.uint32)
        t.count = np.empty(nc, np.uint32)
        t.children = np.empty((nc, 8), np.int32)
        t.leaf = np.empty(nc, np.uint8)
        t.level = np.empty(nc, np.uint32)
        t.cen = np.empty((nc, 3))
        t.mass = np.empty(nc)
        t.rho = np.empty(nc)
        self._lib.orc_get_tree(self._ctx, _p(t.order, _pu32), _p(t.keys, _pu64), _p(t.first, _pu32),
                           _p(t.count, _pu32), _p(t.children, _pi32), _p(t.leaf, _pu8),
                           _p(t.level, _pu32), _p(t.cen), _p(t.mass), _p(t.rho))
        lo, hi, inv = np.empty(3), np.empty(3), np.empty(1)
        self._lib.orc_get_bbox(self._ctx, _p(lo), _p(hi), _p(inv))
        t.lo, t.hi, t.inv_side = lo, hi, inv[0]
        return t

    def run(self, pos, mass, nthreads=1):
        pos, mass = _f64(pos), _f64(mass)
        n = mass.shape[0]
        acc = np.empty((n, 3))
        pot = np.empty(n)
        rc = self._lib.orc_run(self._ctx, _p(pos), _p(mass), self.p, self.eps, self.theta_max,
                           self.softening, nthreads, _p(acc), _p(pot))
        if rc != 0:
            raise ValueError("oracle run failed")
        return acc, pot

    def moments(self):
        out = np.empty(self.ncells * dim(self.p), dtype=np.complex128)
        self._lib.orc_get_moments(self._ctx, out.ctypes.data_as(_pd))
        return out.reshape(self.ncells, dim(self.p))

    def pows(self):
        out = np.empty(self.ncells * (self.p + 1))
        self._lib.orc_get_pows(self._ctx, _p(out))
        return out.reshape(self.ncells, self.p + 1)

    def f(self):
        out = np.empty(self.n)
        self._lib.orc_get_f(self._ctx, _p(out))
        return out

    def fmin(self):
        out = np.empty(self.ncells)
        self._lib.orc_get_fmin(self._ctx, _p(out))
        return out

    def m2l_pairs(self):
        k = self._lib.orc_n_m2l(self._ctx)
        out = np.empty((k, 2), np.uint32)
        if k:
            self._lib.orc_get_m2l(self._ctx, _p(out, _pu32))
        return out

    def p2p_pairs(self):
        k = self._lib.orc_n_p2p(self._ctx)
        out = np.empty((k, 2), np.uint32)
        if k:
            self._lib.orc_get_p2p(self._ctx, _p(out, _pu32))
        return out

    def locals_walk(self):
        out = np.empty(self.ncells * dim(self.p), dtype=np.complex128)
        self._lib.orc_get_locals_walk(self._ctx, out.ctypes.data_as(_pd))
        return out.reshape(self.ncells, dim(self.p))

    def locals(self):
        out = np.empty(self.ncells * dim(self.p), dtype=np.complex128)
        self._lib.orc_get_locals(self._ctx, out.ctypes.data_as(_pd))
        return out.reshape(self.ncells, dim(self.p))

    def near(self):
        out = np.empty((self.n, 4))
        self._lib.orc_get_near(self._ctx, _p(out))
        return out

    def times(self):
        out = np.empty(5)
        self._lib.orc_get_times(self._ctx, _p(out))
        return dict(zip(["moments", "prepass", "walk", "l2l", "l2p"], out))
