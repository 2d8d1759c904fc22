"""GPU tests of batched ensembles (fmmh_set_systems): many independent systems
in one forest, one launch per stage.

Bar: each system's accelerations and potentials in the ensemble equal its
standalone evaluation bit for bit (the forest's subtrees are the systems'
own octrees, the walk starts at the system roots, per-target summation
orders are preserved); leapfrog trajectories likewise. Systems overlap in
space on purpose, so any cross-system interaction would show.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fa():
    import fmms_amd

    return fmms_amd


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def make_systems(fa, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(300), np.log(6000))))
        kind = i % 3
        if kind == 0:
            s = fa.ic_plummer(n, rmax=20.0, seed=100 + i)
        elif kind == 1:
            s = fa.ic_hernquist(n, a=0.5, rmax=30.0, seed=200 + i)
        else:
            s = fa.ic_disk(n, m_central=0.1, sigma=0.02, seed=300 + i)
        out.append(s)
    return out


PARAMS = [(8, 0.5, 1e-3, 32, "fp64"), (6, 0.6, 0.0, 16, "fp64"), (16, 0.6, 1e-3, 32, "fp64"),
          (8, 0.5, 1e-3, 32, "fp32bulk"), (10, 0.6, 1e-2, 64, "fp32bulk")]


@pytest.mark.parametrize("p,theta,soft,n_leaf,precision", PARAMS)
@pytest.mark.parametrize("k", [1, 5, 13, 70])
def test_ensemble_equals_standalone(fa, k, p, theta, soft, n_leaf, precision):
    if k == 70 and p == 16:
        pytest.skip("covered at smaller k")
    systems = make_systems(fa, k, seed=k)
    say(f"[ensemble] k={k} p={p} theta={theta} soft={soft} {precision} "
        f"n={sum(s.n for s in systems)}")
    pos = np.concatenate([s.pos for s in systems])
    mass = np.concatenate([s.mass for s in systems])
    g = fa.Fmm(None, None, p=p, eps=math.inf, n_leaf=n_leaf, theta_max=theta, softening=soft,
               precision=precision)
    off = g.set_systems([s.n for s in systems])
    r = g.accel(pos, mass)
    for i, s in enumerate(systems):
        one = fa.Fmm(None, None, p=p, eps=math.inf, n_leaf=n_leaf, theta_max=theta,
                     softening=soft, precision=precision)
        ri = one.accel(s.pos, s.mass)
        a, b = int(off[i]), int(off[i + 1])
        assert np.array_equal(bits(r.accel[a:b]), bits(ri.accel)), f"system {i}"
        assert np.array_equal(bits(r.pot[a:b]), bits(ri.pot)), f"system {i}"


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_ensemble_leapfrog_equals_standalone(fa, precision):
    systems = make_systems(fa, 9, seed=5)
    say(f"[ensemble] leapfrog, 3 steps, rebuild every step, {precision}")
    pos = np.concatenate([s.pos for s in systems])
    vel = np.concatenate([s.vel for s in systems])
    mass = np.concatenate([s.mass for s in systems])
    g = fa.Fmm(None, None, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3,
               precision=precision)
    off = g.set_systems([s.n for s in systems])
    it = fa.Integrator(g, pos, vel, mass)
    it.step(1e-3, steps=3)
    x, v, a, _ = it.state()
    for i, s in enumerate(systems):
        one = fa.Fmm(None, None, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3,
                     precision=precision)
        it1 = fa.Integrator(one, s.pos, s.vel, s.mass)
        it1.step(1e-3, steps=3)
        x1, v1, a1, _ = it1.state()
        lo, hi = int(off[i]), int(off[i + 1])
        assert np.array_equal(bits(x[lo:hi]), bits(x1)), f"system {i}"
        assert np.array_equal(bits(v[lo:hi]), bits(v1)), f"system {i}"


def test_ensemble_with_split_p2p_leaves_equals_standalone(fa):
    """Systems whose large leaves are cut into several P2P work items: the cut
    depends on each leaf's own counts only, so every system still matches
    its standalone evaluation bit for bit."""
    systems = [fa.ic_plummer(20000, rmax=20.0, seed=71), fa.ic_hernquist(15000, a=0.5, rmax=30.0,
                                                                         seed=72)]
    say("[ensemble] split P2P leaves, n_leaf=256")
    pos = np.concatenate([s.pos for s in systems])
    mass = np.concatenate([s.mass for s in systems])
    kw = dict(p=6, eps=math.inf, n_leaf=256, theta_max=0.5, softening=1e-3)
    g = fa.Fmm(None, None, **kw)
    off = g.set_systems([s.n for s in systems])
    r = g.accel(pos, mass)
    assert g.work_counts()["p2p_split_leaves"] > 0
    for i, s in enumerate(systems):
        ri = fa.Fmm(None, None, **kw).accel(s.pos, s.mass)
        a, b = int(off[i]), int(off[i + 1])
        assert np.array_equal(bits(r.accel[a:b]), bits(ri.accel)), f"system {i}"
        assert np.array_equal(bits(r.pot[a:b]), bits(ri.pot)), f"system {i}"


@pytest.mark.parametrize("eps,theta", [(1e-6, 1.0), (1e-5, 0.6)])
def test_ensemble_error_mac_equals_standalone(fa, eps, theta):
    """Error-controlled MAC (mac.rs:23-95, the reference's default) on an
    ensemble: the f pre-pass walks each body's own system only, so f, fmin,
    the lists and the forces of every system equal its standalone run bit for
    bit (systems overlap in space, so a cross-system term would show)."""
    systems = make_systems(fa, 13, seed=29)
    say(f"[ensemble] error MAC eps={eps} theta={theta}")
    pos = np.concatenate([s.pos for s in systems])
    mass = np.concatenate([s.mass for s in systems])
    kw = dict(p=8, eps=eps, n_leaf=32, theta_max=theta, softening=1e-3)
    g = fa.Fmm(None, None, **kw)
    off = g.set_systems([s.n for s in systems])
    r = g.accel(pos, mass)
    f_all, _ = g.f()
    for i, s in enumerate(systems):
        one = fa.Fmm(None, None, **kw)
        ri = one.accel(s.pos, s.mass)
        fi, _ = one.f()
        a, b = int(off[i]), int(off[i + 1])
        assert np.array_equal(bits(r.accel[a:b]), bits(ri.accel)), f"system {i}"
        assert np.array_equal(bits(r.pot[a:b]), bits(ri.pot)), f"system {i}"
        assert np.array_equal(np.sort(f_all[a:b]), np.sort(fi)), f"system {i}"


def test_ensemble_argument_checks(fa):
    g = fa.Fmm(None, None, p=8, eps=math.inf, n_leaf=16, theta_max=0.5)
    g.set_systems([100, 200])
    pos = np.random.default_rng(1).random((250, 3))
    with pytest.raises(fa.FmmError):  # n != sum of the sizes
        g.accel(pos, np.ones(250))
    with pytest.raises(fa.FmmError):
        g.set_systems([1] * (fa.MAX_SYSTEMS + 1))
