"""Initial-condition library (fmmh_ic_*, CPU): the product draws bit for bit
what the oracle draws and what the committed golden fixture holds; the
distributions are pinned against closed forms and numerical integrals that
do not share the generators' code: the Plummer speed distribution of Aarseth,
Henon & Wielen (regularised incomplete beta) and the virial ratio from a
direct-sum potential; the Hernquist dispersion against the Jeans integral;
Freeman's disk curve against its Hankel-transform integral; sech^2 heights
and Gamma(2) radii. energy_gate.rs:39-79's warm start is pinned by its own
assertion (E0 ~ -0.147) and by exact virial balance.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fa():
    import fmms_amd

    return fmms_amd


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


GENS = [
    ("ic_plummer", 20000, {"rmax": 20.0, "seed": 5}),
    ("ic_hernquist", 20000, {"a": 0.3, "rmax": 30.0, "seed": 6}),
    ("ic_disk", 20000, {"m_central": 0.05, "sigma": 0.01, "seed": 8}),
    ("ic_disk", 5000, {"seed": 9}),
    ("ic_halos_cube", 60000, {"nhalo": 16, "seed": 10}),
]


@pytest.mark.parametrize("fn,n,kw", GENS, ids=[f"{g[0]}_{i}" for i, g in enumerate(GENS)])
def test_product_equals_oracle_bitwise(fa, fn, n, kw):
    s = getattr(fa, fn)(n, **kw)
    pos, vel, mass = getattr(O, fn)(n, **kw)
    assert np.array_equal(bits(s.pos), bits(pos))
    assert np.array_equal(bits(s.vel), bits(vel))
    assert np.array_equal(bits(s.mass), bits(mass))


def test_golden_first_samples(fa):
    with open(os.path.join(HERE, "golden", "ic_first8.json")) as f:
        gold = json.load(f)
    for name, g in gold.items():
        want = {k: np.array([float.fromhex(x) for x in g[k]]) for k in
                ("pos", "vel", "mass", "pos_last", "vel_last")}
        s = getattr(fa, g["fn"])(g["n"], **g["kwargs"])
        pos, vel, mass = getattr(O, g["fn"])(g["n"], **g["kwargs"])
        k = len(want["mass"])
        for p_, v_, m_ in ((s.pos, s.vel, s.mass), (pos, vel, mass)):
            assert np.array_equal(bits(p_[:k]).ravel(), bits(want["pos"])), name
            assert np.array_equal(bits(v_[:k]).ravel(), bits(want["vel"])), name
            assert np.array_equal(bits(m_[:k]), bits(want["mass"])), name
            assert np.array_equal(bits(p_[-k:]).ravel(), bits(want["pos_last"])), name
            assert np.array_equal(bits(v_[-k:]).ravel(), bits(want["vel_last"])), name


def test_plummer_positions_equal_truncated_sampler(fa):
    s = fa.ic_plummer(3000, rmax=20.0, seed=42)
    pos, mass = fa.plummer_truncated(3000, rmax=20.0, seed=42)
    assert np.array_equal(bits(s.pos), bits(pos)) and np.array_equal(s.mass, mass)
    # rmax = inf reproduces models.rs:35-58 (the reference sampler) exactly
    s = fa.ic_plummer(2000, rmax=math.inf, seed=3)
    pos, _ = O.plummer(2000, seed=3)
    assert np.array_equal(bits(s.pos), bits(pos))


def test_hernquist_positions_reduce_to_reference_sampler(fa):
    s = fa.ic_hernquist(2000, a=1.0, rmax=math.inf, seed=7)
    pos, _ = O.hernquist(2000, seed=7)
    assert np.array_equal(bits(s.pos), bits(pos))


def test_plummer_speed_distribution_and_virial_ratio(fa):
    from scipy import special, stats

    n = 60000
    s = fa.ic_plummer(n, rmax=1e9, seed=21)
    r2 = (s.pos ** 2).sum(1)
    vesc = np.sqrt(2.0) * (1.0 + r2) ** -0.25
    q = np.linalg.norm(s.vel, axis=1) / vesc
    # g(q) dq ~ q^2 (1 - q^2)^{7/2} dq  ->  CDF = I_{q^2}(3/2, 9/2)
    d = stats.kstest(q, lambda x: special.betainc(1.5, 4.5, np.clip(x, 0, 1) ** 2)).statistic
    assert d < 0.012, d
    # a wrong exponent (7/2 -> 5/2) fails the same test
    d_bad = stats.kstest(q, lambda x: special.betainc(1.5, 3.5, np.clip(x, 0, 1) ** 2)).statistic
    assert d_bad > 0.03
    # virial equilibrium: 2T / |W| = 1, W from the direct-sum potential
    m = 8000
    s = fa.ic_plummer(m, rmax=1e9, seed=22)
    acc = np.empty((m, 3))
    pot = np.empty(m)
    O.lib().orc_direct(m, O._p(s.pos), O._p(s.mass), 0.0, O._p(acc), O._p(pot))
    t = O.kinetic(s.mass, s.vel)
    w = O.potential_energy(s.mass, pot)
    assert abs(2 * t / abs(w) - 1.0) < 0.05, 2 * t / abs(w)
    # total mass 1 and zero total momentum
    assert abs(s.mass.sum() - 1.0) < 1e-12
    assert np.abs((s.mass[:, None] * s.vel).sum(0)).max() < 1e-15


@pytest.mark.parametrize("r", [1e-3, 0.05, 0.5, 2.0, 5.9, 6.0, 6.1, 10.0, 50.0, 300.0])
def test_hernquist_dispersion_matches_jeans_integral(r):
    from scipy import integrate

    # isotropic Jeans: rho sigma^2 = int_r^inf rho M / r'^2 dr' (G = M = a = 1)
    rho = lambda x: 1.0 / (2 * math.pi * x * (1 + x) ** 3)
    mass = lambda x: x * x / (1 + x) ** 2
    val, _ = integrate.quad(lambda x: rho(x) * mass(x) / (x * x), r, np.inf, epsabs=0,
                            epsrel=1e-13, limit=500)
    want = val / rho(r)
    got = O.hernquist_sigma2(r, 1.0)
    assert abs(got / want - 1.0) < 1e-9, (got, want)
    # scale radius a: sigma^2(r; a) = sigma^2(r/a; 1) / a
    assert abs(O.hernquist_sigma2(0.3 * r, 0.3) / (got / 0.3) - 1.0) < 1e-13


def test_hernquist_profile_and_dispersion(fa):
    n = 80000
    s = fa.ic_hernquist(n, a=1.0, rmax=100.0, seed=31)
    r = np.linalg.norm(s.pos, axis=1)
    assert r.max() <= 100.0
    # M(<r) = r^2 / (1 + r)^2, renormalised by the truncation at 100
    norm = (100.0 / 101.0) ** 2
    for rr in (0.5, 1.0, 3.0):
        assert abs((r <= rr).mean() - rr * rr / (1 + rr) ** 2 / norm) < 0.006
    from scipy import stats

    v2 = (s.vel ** 2).sum(1)
    sig2 = np.array([O.hernquist_sigma2(x, 1.0) for x in r])
    # Maxwellian truncated at 0.95 v_esc: E[v^2] = 3 s^2 F5(c) / F3(c), c = vmax^2 / s^2
    c = 0.95 ** 2 * 2.0 / (r + 1.0) / sig2
    ev2 = 3.0 * sig2 * stats.chi2.cdf(c, 5) / stats.chi2.cdf(c, 3)
    for lo, hi in ((0.1, 0.3), (0.3, 1.0), (1.0, 3.0), (3.0, 10.0)):
        sel = (r >= lo) & (r < hi)
        assert abs(v2[sel].mean() / ev2[sel].mean() - 1.0) < 0.02, (lo, hi)
    # speeds below 0.95 v_esc (up to the removed mean velocity, ~ sigma / sqrt(n))
    assert np.all(np.sqrt(v2) < 0.95 * np.sqrt(2.0 / (r + 1.0)) + 0.01)


def test_disk_structure_and_rotation_curve(fa):
    from scipy import special

    n = 60000
    s = fa.ic_disk(n, rd=1.0, zd=0.05, rmax=10.0, seed=41)
    R = np.hypot(s.pos[:, 0], s.pos[:, 1])
    # Gamma(2) radii truncated at 10: mean = (2 - e^-10 (10^2 + 2*10 + 2)) / (1 - 11 e^-10)
    e = math.exp(-10.0)
    assert abs(R.mean() - (2 - e * 122) / (1 - 11 * e)) < 0.015
    assert R.max() <= 10.0
    # sech^2 heights: median |z| = zd atanh(1/2)
    assert abs(np.median(np.abs(s.pos[:, 2])) / (0.05 * math.atanh(0.5)) - 1.0) < 0.02
    # circular speed: with a central mass m_c as body 0 and one disk body the
    # mean-velocity removal scales the disk body's velocity by m_c exactly
    mc = 0.25
    for seed in range(12):
        t = fa.ic_disk(2, m_central=mc, seed=100 + seed)
        Rb = math.hypot(t.pos[1, 0], t.pos[1, 1])
        vc = np.linalg.norm(t.vel[1]) / mc
        y = Rb / 2.0
        freeman = 2 * (1 - mc) * y * y * (special.i0(y) * special.k0(y)
                                         - special.i1(y) * special.k1(y))
        assert abs(vc * vc / (mc / Rb + freeman) - 1.0) < 1e-10
        # tangential, counter-clockwise
        assert abs(t.vel[1] @ np.array([t.pos[1, 0], t.pos[1, 1], 0.0])) < 1e-12 * vc * Rb
        assert (t.pos[1, 0] * t.vel[1, 1] - t.pos[1, 1] * t.vel[1, 0]) > 0


def test_freeman_term_matches_hankel_integral():
    import mpmath as mp
    from scipy import special

    # v_c^2 / R of a thin exponential disk (G = Md = rd = 1):
    # int_0^inf k J1(kR) (1 + k^2)^{-3/2} dk (Binney & Tremaine 2.165)
    mp.mp.dps = 20
    for R in (0.3, 2.0, 7.0):
        val = mp.quadosc(lambda k: k * mp.besselj(1, k * R) * (1 + k * k) ** -1.5, [0, mp.inf],
                         period=2 * mp.pi / R)
        y = R / 2
        fr = 2 * y * y * (special.i0(y) * special.k0(y) - special.i1(y) * special.k1(y)) / R
        assert abs(float(val) / fr - 1.0) < 1e-12


def test_halos_in_cube_layout(fa):
    n = 50000
    s = fa.ic_halos_cube(n, nhalo=64, halo_frac=0.9, seed=2024)
    nh = int(math.floor(0.9 * n + 0.5))
    assert np.all(s.mass == 1.0 / n)
    bg = s.pos[nh:]
    assert bg.shape[0] == n - nh
    assert np.all((bg >= 0.0) & (bg < 1.0)) and np.all(s.vel[nh:] == 0.0)
    # clustered: the halo bodies' nearest-grid occupancy is far from uniform
    h = np.floor(np.clip(s.pos[:nh], 0, 0.999) * 8).astype(int)
    occ = np.bincount(h[:, 0] * 64 + h[:, 1] * 8 + h[:, 2], minlength=512)
    assert occ.max() > 20 * nh / 512
    with pytest.raises(fa.FmmError):
        fa.ic_halos_cube(100, nhalo=0, halo_frac=0.5)


def test_invalid_arguments(fa):
    with pytest.raises(fa.FmmError):
        fa.ic_plummer(10, rmax=-1.0)
    with pytest.raises(fa.FmmError):
        fa.ic_hernquist(10, a=0.0)
    with pytest.raises(fa.FmmError):
        fa.ic_disk(10, m_central=1.0)
    with pytest.raises(fa.FmmError):
        fa.ic_disk(1, m_central=0.5)


def test_warm_start_energy_gate_initial_state():
    # energy_gate.rs:239-247: virialised Plummer has E0 = W/2 ~ -0.147
    n = 1024
    pos, mass = O.plummer(n, seed=42)
    acc = np.empty((n, 3))
    pot = np.empty(n)
    O.lib().orc_direct(n, O._p(pos), O._p(mass), 0.0, O._p(acc), O._p(pot))
    vel = O.warm_start(pos, mass, pot)
    t = O.kinetic(mass, vel)
    w = O.potential_energy(mass, pot)
    assert abs(2 * t + w) < 1e-13 * abs(w)          # exact virial rescale
    assert abs((t + w) + 0.147) < 0.03               # the reference's assertion
    assert O.momentum_rel(mass, vel) < 1e-14
