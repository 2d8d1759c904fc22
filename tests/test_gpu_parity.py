"""GPU parity of libfmmh against the oracle, through the C ABI.

Bars (stated per test):
  * tree, Morton order and keys, cell geometry (cen, mass, rho), moments,
    multipole powers, f pre-pass, fmin and the interaction lists: bit-exact;
  * accelerations / potentials against the oracle's serial FP64 run:
    |da| <= 1e-11 |a| + 1e-13 max|a| per body (summation order differs: the
    device groups directed interactions by target instead of mutual updates);
  * against the FP64 direct sum: the reference's integration thresholds.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fa():
    import fmms_amd

    return fmms_amd


def grid64():
    pos = np.array([[i % 4, (i // 4) % 4, i // 16] for i in range(64)], float)
    return pos, np.ones(64)


def dup_cloud():
    rng = np.random.default_rng(3)
    pos = rng.random((300, 3))
    pos[100:130] = pos[5]          # exact duplicates -> equal Morton keys (ties)
    pos[200:205] = pos[7] + 1e-12  # sub-quantum offsets -> equal keys too
    return pos, rng.random(300) + 0.5


def offset_cloud():
    pos, mass = O.plummer(1500, seed=17)
    return pos * np.array([1.0, 0.3, 0.2]) + np.array([0.0, 2.5, -4.0]), mass


TREE_CASES = {
    "grid64_leaf1": (grid64, 1, 0),
    "coincident5": (lambda: (np.ones((5, 3)), np.full(5, 2.0)), 2, 0),
    "single_body": (lambda: (np.array([[0.5, -0.25, 3.0]]), np.array([2.0])), 16, 0),
    "two_bodies": (lambda: (np.array([[0.0, 0, 0], [3.0, 0, 0]]), np.array([1.0, 2.0])), 1, 0),
    "plummer1000_leaf16": (lambda: O.plummer(1000, seed=1), 16, 0),
    "plummer4096_leaf8": (lambda: O.plummer(4096, seed=99), 8, 0),
    "plummer4096_leaf64": (lambda: O.plummer(4096, seed=99), 64, 0),
    "hernquist4096": (lambda: O.hernquist(4096, seed=7), 32, 0),
    "dup_ties": (dup_cloud, 4, 0),
    "offset_xmin_quirk": (offset_cloud, 16, 0),
    "sphere_plummer2048": (lambda: O.plummer(2048, seed=13), 16, 1),
    "sphere_grid64": (grid64, 1, 1),
    # leaves of more than 64 bodies: the leaf geometry runs over several
    # 64-body chunks (ordered sums, grow passes across chunks)
    "plummer4096_leaf200": (lambda: O.plummer(4096, seed=99), 200, 0),
    "sphere_plummer4096_leaf200": (lambda: O.plummer(4096, seed=99), 200, 1),
    "sphere_dup_ties_leaf40": (dup_cloud, 40, 1),
    "sphere_coincident130": (lambda: (np.full((130, 3), 0.25), np.linspace(1.0, 2.0, 130)), 2, 1),
}


def assert_tree_equal(g, o):
    for k in ("order", "first", "count", "children", "leaf", "level"):
        assert np.array_equal(getattr(g, k), getattr(o, k)), k
    assert np.array_equal(g.keys, o.keys), "keys"
    for k in ("cen", "mass", "rho"):
        a, b = getattr(g, k), getattr(o, k)
        assert a.shape == b.shape
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k


@pytest.mark.parametrize("case", list(TREE_CASES))
def test_tree_bit_exact(fa, case):
    gen, n_leaf, mode = TREE_CASES[case]
    pos, mass = gen()
    say(f"[tree] {case} n={len(mass)} n_leaf={n_leaf} mode={mode}")
    orc = O.OracleFmm(pos, mass, p=4, eps=1e-3, n_leaf=n_leaf, center_mode=mode)
    gpu = fa.Fmm(pos, mass, p=4, eps=1e-3, n_leaf=n_leaf, center_mode=mode)
    assert gpu.n_cells == orc.ncells
    assert_tree_equal(gpu.tree(), orc.tree())


UPWARD_CASES = [(4, 1e-2, 16, 0), (8, 1e-5, 16, 0), (6, 1e-4, 8, 1), (12, 1e-6, 32, 0),
                (16, 1e-8, 32, 0), (24, 1e-10, 32, 1)]


@pytest.mark.parametrize("p,eps,n_leaf,mode", UPWARD_CASES)
def test_moments_mac_inputs_and_lists_bit_exact(fa, p, eps, n_leaf, mode):
    pos, mass = O.plummer(2048, seed=13)
    say(f"[upward+lists] p={p} eps={eps} n_leaf={n_leaf} mode={mode}")
    orc = O.OracleFmm(pos, mass, p=p, eps=eps, n_leaf=n_leaf, center_mode=mode)
    orc.run(pos, mass)
    gpu = fa.Fmm(pos, mass, p=p, eps=eps, n_leaf=n_leaf, center_mode=mode)
    gpu.run(pos, mass)
    assert np.array_equal(gpu.moments(), orc.moments())  # == (signed zeros compare equal)
    assert np.array_equal(gpu.pows().view(np.uint64), orc.pows().view(np.uint64))
    f, fmin = gpu.f()
    assert np.array_equal(f.view(np.uint64), orc.f().view(np.uint64))
    assert np.array_equal(fmin, orc.fmin())
    m2l, p2p = gpu.lists()
    assert sorted(map(tuple, m2l)) == sorted(map(tuple, orc.m2l_pairs()))
    assert sorted(map(tuple, p2p)) == sorted(map(tuple, orc.p2p_pairs()))


def check_forces(a, pot, a_ref, pot_ref, rel=1e-11):
    an = np.linalg.norm(a_ref, axis=1)
    da = np.linalg.norm(a - a_ref, axis=1)
    assert np.all(da <= rel * an + 1e-13 * an.max()), (da / np.maximum(an, 1e-300)).max()
    dp = np.abs(pot - pot_ref)
    assert np.all(dp <= rel * np.abs(pot_ref) + 1e-13 * np.abs(pot_ref).max())


FORCE_CASES = [
    ("plummer2048_p8_eps1e-5", lambda: O.plummer(2048, seed=5), 8, 1e-5, 1.0, 0.0, 16),
    ("plummer4096_p4_eps1e-4", lambda: O.plummer(4096, seed=9), 4, 1e-4, 1.0, 0.0, 32),
    ("hernquist4096_p6", lambda: O.hernquist(4096, seed=7), 6, 1e-3, 1.0, 0.0, 32),
    ("theta0.5_p8_soft1e-3", lambda: O.plummer(3000, seed=4), 8, math.inf, 0.5, 1e-3, 32),
    ("theta0.7_p10_leaf64", lambda: O.plummer(3000, seed=8), 10, math.inf, 0.7, 0.0, 64),
    ("p1_eps1e-1", lambda: O.plummer(1000, seed=2), 1, 1e-1, 1.0, 0.0, 8),
    ("p12_eps1e-7", lambda: O.plummer(1500, seed=3), 12, 1e-7, 1.0, 0.0, 16),
    # runtime-p kernels: naive M2L for 13 <= p < 20, rotation M2L for p >= 20
    ("p13_naive_gen", lambda: O.plummer(1500, seed=23), 13, 1e-8, 1.0, 0.0, 16),
    ("p16_theta0.6_soft", lambda: O.plummer(1500, seed=24), 16, math.inf, 0.6, 1e-2, 16),
    ("p20_rotation", lambda: O.plummer(1500, seed=25), 20, 1e-10, 1.0, 0.0, 16),
    ("p24_rotation_soft", lambda: O.hernquist(1500, seed=26), 24, math.inf, 0.7, 1e-3, 32),
    ("p30_rotation", lambda: O.plummer(1000, seed=27), 30, 1e-12, 1.0, 0.0, 16),
    ("p40_rotation", lambda: O.plummer(600, seed=28), 40, math.inf, 0.8, 0.0, 8),
]


@pytest.mark.parametrize("case", FORCE_CASES, ids=[c[0] for c in FORCE_CASES])
def test_forces_match_oracle(fa, case):
    name, gen, p, eps, theta, soft, n_leaf = case
    pos, mass = gen()
    say(f"[forces] {name}")
    orc = O.OracleFmm(pos, mass, p=p, eps=eps, n_leaf=n_leaf, theta_max=theta, softening=soft)
    a_ref, pot_ref = orc.run(pos, mass)
    gpu = fa.Fmm(pos, mass, p=p, eps=eps, n_leaf=n_leaf, theta_max=theta, softening=soft)
    r = gpu.run(pos, mass)
    m2l, p2p = gpu.lists()
    assert sorted(map(tuple, m2l)) == sorted(map(tuple, orc.m2l_pairs()))
    assert sorted(map(tuple, p2p)) == sorted(map(tuple, orc.p2p_pairs()))
    # p > 10 runs the runtime-p kernels (rotation from p = 20); above 12 the bar is
    # the reference's own rotation-vs-naive contract (rotation.rs:505-556)
    rel = 1e-11 if p <= 12 else 1e-9
    check_forces(r.accel, r.pot, a_ref, pot_ref, rel=rel)
    # locals after L2L agree with the oracle's to the same order
    lg, lo = gpu.locals(), orc.locals()
    assert np.abs(lg - lo).max() <= (1e-10 if p <= 12 else 1e-8) * np.abs(lo).max()


@pytest.mark.parametrize("tau", [1e-4, 1e-5])
def test_softening_guard_matches_oracle(fa, tau):
    """Softening guard of the MAC (extension): lists bit-exact with the
    oracle's guarded walk, forces to the oracle bar; the guard rejects pairs
    (more P2P than without it) and no accepted M2L pair exceeds tau."""
    pos, mass = O.plummer(3000, seed=4)
    soft = 2e-2
    say(f"[soft-guard] tau={tau}")
    kw = dict(p=8, eps=math.inf, n_leaf=16, theta_max=0.5, softening=soft)
    orc = O.OracleFmm(pos, mass, soft_mac=tau, **kw)
    a_ref, pot_ref = orc.run(pos, mass)
    gpu = fa.Fmm(pos, mass, soft_mac=tau, **kw)
    r = gpu.run(pos, mass)
    m2l, p2p = gpu.lists()
    assert sorted(map(tuple, m2l)) == sorted(map(tuple, orc.m2l_pairs()))
    assert sorted(map(tuple, p2p)) == sorted(map(tuple, orc.p2p_pairs()))
    check_forces(r.accel, r.pot, a_ref, pot_ref)
    plain = fa.Fmm(pos, mass, **kw)
    plain.run(pos, mass)
    assert len(p2p) > len(plain.lists()[1])
    t = gpu.tree()
    m2l = np.asarray(m2l)
    s = t.rho[m2l[:, 0]] + t.rho[m2l[:, 1]]
    r2 = ((t.cen[m2l[:, 0]] - t.cen[m2l[:, 1]]) ** 2).sum(1)
    assert np.all(soft * soft * s ** 3 <= tau * r2 ** 2.5 * (1 + 1e-12))


@pytest.mark.parametrize("p", [4, 8, 10])
def test_fp32_m2l_locals_match_fp64(fa, p):
    """The mixed form's FP32 M2L (per-target scale lam = max(rho, octant edge
    / 4), moments pre-scaled by lam_s^-k, dimensionless offsets) against the
    FP64 M2L: after L2L every cell's local expansion within 5e-6 of the FP64
    one in the cell's scaled norm sum_n |F_n| lam^n, for 99 % of cells, and
    within 2e-5 for all (FP32 sums of up to 32 sources; a scaling slip would
    show as O(1) errors or overflow)."""
    pos, mass = O.plummer(8000, seed=81)
    say(f"[fp32-m2l] p={p}")
    kw = dict(p=p, eps=math.inf, n_leaf=16, theta_max=0.5, softening=1e-3)
    g64 = fa.Fmm(pos, mass, **kw)
    g64.run(pos, mass)
    g32 = fa.Fmm(pos, mass, precision="fp32bulk", **kw)
    g32.run(pos, mass)
    l64, l32 = g64.locals(), g32.locals()
    t = g64.tree()
    n_of = np.concatenate([[n] * (2 * n + 1) for n in range(p + 1)])
    w = np.maximum(t.rho, 1e-12)[:, None] ** n_of[None, :]  # lam^n with lam ~ rho
    num = (np.abs(l32 - l64) * w).sum(1)
    den = (np.abs(l64) * w).sum(1)
    live = den > 0
    rel = num[live] / den[live]
    assert np.percentile(rel, 99) < 5e-6 and rel.max() < 2e-5, (np.percentile(rel, 99), rel.max())


def test_fp32bulk_coincident_bodies(fa):
    """Cells of coincident bodies (radius 0) still get an FP32 M2L scale (a
    quarter of their octant edge): the mixed form stays finite and within
    FP32 precision of the FP64 path."""
    pos, mass = dup_cloud()
    say("[fp32bulk] coincident bodies")
    kw = dict(p=8, eps=math.inf, n_leaf=4, theta_max=0.5, softening=1e-3)
    r64 = fa.Fmm(None, None, **kw).accel(pos, mass)
    r32 = fa.Fmm(None, None, precision="fp32bulk", **kw).accel(pos, mass)
    assert np.all(np.isfinite(r32.accel)) and np.all(np.isfinite(r32.pot))
    d = np.linalg.norm(r32.accel - r64.accel, axis=1) / np.linalg.norm(r64.accel, axis=1)
    assert np.percentile(d, 99) < 1e-5 and d.max() < 1e-4, (np.percentile(d, 99), d.max())


def test_p2p_split_leaves_match_oracle(fa):
    """Large leaves (n_leaf = 256) beside dense regions carry P2P work above
    the per-item bound and are cut into several work items whose partial
    sums are added in item order: lists bit-exact, forces to the oracle bar
    in FP64; the mixed form agrees with FP64 to FP32 pair precision; both are
    deterministic run to run."""
    pos, mass = O.plummer(20000, seed=8)
    kw = dict(p=6, eps=math.inf, n_leaf=256, theta_max=0.5, softening=1e-3)
    say("[p2p-split] N=20000 n_leaf=256")
    orc = O.OracleFmm(pos, mass, **kw)
    a_ref, pot_ref = orc.run(pos, mass)
    gpu = fa.Fmm(pos, mass, **kw)
    r = gpu.run(pos, mass)
    wc = gpu.work_counts()
    assert wc["p2p_split_leaves"] > 0 and wc["p2p_items"] > wc["nleaves"], wc
    m2l, p2p = gpu.lists()
    assert sorted(map(tuple, p2p)) == sorted(map(tuple, orc.p2p_pairs()))
    check_forces(r.accel, r.pot, a_ref, pot_ref)
    r2 = gpu.run(pos, mass)
    assert np.array_equal(r.accel, r2.accel) and np.array_equal(r.pot, r2.pot)
    g32 = fa.Fmm(pos, mass, precision="fp32bulk", **kw)
    q1 = g32.run(pos, mass)
    q2 = g32.run(pos, mass)
    assert np.array_equal(q1.accel, q2.accel) and np.array_equal(q1.pot, q2.pot)
    rel = np.linalg.norm(q1.accel - r.accel, axis=1) / np.linalg.norm(r.accel, axis=1)
    assert np.percentile(rel, 99) < 1e-6 and rel.max() < 5e-5


def test_lists_with_long_segments(fa):
    """A cell with thousands of P2P neighbours (a large sparse leaf beside a
    dense clump): lists and forces match the oracle, reruns are
    bit-identical."""
    rng = np.random.default_rng(5)
    clump = rng.random((24000, 3)) * 0.02 + np.array([0.5, 0.5, 0.5])
    pos = np.vstack([clump, rng.random((400, 3))])
    mass = np.full(len(pos), 1.0 / len(pos))
    kw = dict(p=4, eps=math.inf, n_leaf=8, theta_max=0.5, softening=1e-3)
    say("[lists] long segments")
    orc = O.OracleFmm(pos, mass, **kw)
    a_ref, pot_ref = orc.run(pos, mass)
    deg = np.bincount(orc.p2p_pairs().ravel(), minlength=orc.ncells)
    assert deg.max() > 2048
    gpu = fa.Fmm(pos, mass, **kw)
    r = gpu.run(pos, mass)
    m2l, p2p = gpu.lists()
    assert sorted(map(tuple, m2l)) == sorted(map(tuple, orc.m2l_pairs()))
    assert sorted(map(tuple, p2p)) == sorted(map(tuple, orc.p2p_pairs()))
    check_forces(r.accel, r.pot, a_ref, pot_ref)
    for _ in range(2):
        r2 = fa.Fmm(pos, mass, **kw).run(pos, mass)
        assert np.array_equal(r.accel, r2.accel) and np.array_equal(r.pot, r2.pot)


def test_accel_entry_point_equals_build_plus_run(fa):
    pos, mass = O.plummer(3000, seed=21)
    say("[accel] build+evaluate in one call")
    g = fa.Fmm(None, None, p=8, eps=1e-5, n_leaf=16)
    r1 = g.accel(pos, mass)
    g2 = fa.Fmm(pos, mass, p=8, eps=1e-5, n_leaf=16)
    r2 = g2.run(pos, mass)
    assert np.array_equal(r1.accel, r2.accel) and np.array_equal(r1.pot, r2.pot)


def test_run_twice_is_deterministic(fa):
    pos, mass = O.plummer(5000, seed=31)
    say("[determinism] two runs")
    g = fa.Fmm(pos, mass, p=8, eps=1e-5, n_leaf=16)
    r1 = g.run(pos, mass)
    r2 = g.run(pos, mass)
    assert np.array_equal(r1.accel, r2.accel) and np.array_equal(r1.pot, r2.pot)


def test_run_with_moved_bodies_uses_built_tree(fa):
    # fmm.rs:69-88: run gathers the given bodies into the built tree's slots.
    pos, mass = O.plummer(2000, seed=41)
    moved = pos + 1e-3 * np.sin(np.arange(6000).reshape(2000, 3))
    say("[refresh] run with moved bodies on a stale tree")
    orc = O.OracleFmm(pos, mass, p=6, eps=1e-4, n_leaf=16)
    a_ref, pot_ref = orc.run(moved, mass)
    gpu = fa.Fmm(pos, mass, p=6, eps=1e-4, n_leaf=16)
    r = gpu.run(moved, mass)
    check_forces(r.accel, r.pot, a_ref, pot_ref)
    gpu.refresh(moved, mass)
    orc.refresh(moved, mass)
    assert_tree_equal(gpu.tree(), orc.tree())


def test_integration_thresholds_vs_direct(fa):
    # tests/integration.rs:11-42 on the device path
    n = 256
    pos, mass = O.plummer(n, seed=1234)
    a_ref, pot_ref = O.direct(pos, mass)
    say("[integration] N=256 vs direct")
    for p, eps in [(2, 1e-2), (3, 1e-3), (4, 1e-4), (5, 1e-5)]:
        r = fa.Fmm(pos, mass, p=p, eps=eps, n_leaf=16).run(pos, mass)
        max_s, _, rms_s, rms_rel = O.scaled_error(a_ref, r.accel, pos, mass)
        assert max_s < 10 * eps and rms_s < 10 * eps and rms_rel < 1e-2
        rel = np.sqrt(np.mean(((r.pot - pot_ref) / np.abs(pot_ref)) ** 2))
        assert rel < 1e-3


def test_momentum_conserved(fa):
    # tests/integration.rs:45-68. Target-grouped sums are antisymmetric only
    # up to rounding: stated bar 1e-13 (the reference's mutual kernel: 1e-14).
    pos, mass = O.plummer(512, seed=99)
    say("[momentum] N=512")
    r = fa.Fmm(pos, mass, p=4, eps=1e-4, n_leaf=16).run(pos, mass)
    fsum = (mass[:, None] * r.accel).sum(0)
    ftot = (mass * np.linalg.norm(r.accel, axis=1)).sum()
    assert np.linalg.norm(fsum) / ftot < 1e-13


def test_direct_sampled_matches_oracle(fa):
    pos, mass = O.plummer(20000, seed=6)
    ids = O.sample_ids(20000, 64, seed=7)
    say("[direct] sampled direct sum")
    for soft in (0.0, 1e-3):
        a, pot, f = fa.direct_sampled(pos, mass, ids, softening=soft)
        a2, pot2, f2 = O.direct_targets(pos, mass, ids, softening=soft, nthreads=4)
        assert np.abs(a - a2).max() <= 1e-12 * np.abs(a2).max()
        assert np.abs(pot - pot2).max() <= 1e-13 * np.abs(pot2).max()
        assert np.abs(f - f2).max() <= 1e-12 * np.abs(f2).max()


def test_error_paths(fa):
    say("[errors] empty / params / nonfinite / state / budget")
    with pytest.raises(fa.FmmError) as e:
        fa.Fmm(np.zeros((0, 3)), np.zeros(0), p=4, eps=1e-3, n_leaf=16)
    assert e.value.code == fa.FMMH_E_EMPTY
    for bad in (dict(p=0), dict(p=fa.MAX_P + 1), dict(n_leaf=0), dict(theta_max=0.0),
                dict(theta_max=1.5), dict(eps=0.0)):
        kw = dict(p=4, eps=1e-3, n_leaf=16)
        kw.update(bad)
        with pytest.raises(fa.FmmError) as e:
            fa.Fmm(None, None, **kw)
        assert e.value.code == fa.FMMH_E_PARAM
    pos, mass = O.plummer(100, seed=1)
    pos[17, 1] = np.nan
    with pytest.raises(fa.FmmError) as e:
        fa.Fmm(pos, mass, p=4, eps=1e-3, n_leaf=16)
    assert e.value.code == fa.FMMH_E_NONFINITE
    g = fa.Fmm(None, None, p=4, eps=1e-3, n_leaf=16)
    with pytest.raises(fa.FmmError) as e:
        g.run(*O.plummer(100, seed=1))
    assert e.value.code == fa.FMMH_E_STATE
    g.refresh(*O.plummer(100, seed=1))
    with pytest.raises(fa.FmmError) as e:
        g.run(*O.plummer(101, seed=1))
    assert e.value.code == fa.FMMH_E_STATE
    small = fa.Fmm(None, None, p=4, eps=1e-3, n_leaf=16, hbm_budget_bytes=1 << 20)
    with pytest.raises(fa.FmmError) as e:
        small.accel(*O.plummer(200000, seed=1))
    assert e.value.code == fa.FMMH_E_OVER_BUDGET


def test_large_theta_mac_properties(fa):
    """Size-independent checks at N=262,144: every body in exactly one leaf,
    every leaf in its own P2P self pair, M2L/P2P lists free of duplicates,
    and the sampled error against the FP64 direct sum."""
    n = 262144
    pos, mass = fa.plummer(n, seed=42)
    say(f"[large] N={n} p=8 theta=0.5 soft=1e-3")
    g = fa.Fmm(pos, mass, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    r = g.run(pos, mass)
    t = g.tree()
    leaves = np.flatnonzero(t.leaf)
    assert t.count[leaves].sum() == n
    m2l, p2p = g.lists()
    selfp = p2p[p2p[:, 0] == p2p[:, 1], 0]
    assert np.array_equal(np.sort(selfp), leaves)
    key = lambda a: np.minimum(a[:, 0], a[:, 1]).astype(np.uint64) << 32 | np.maximum(a[:, 0], a[:, 1])
    assert len(np.unique(key(m2l))) == len(m2l)
    assert len(np.unique(key(p2p))) == len(p2p)
    ids = fa.sample_ids(n, 256, 7)
    a_ref, _, _ = fa.direct_sampled(pos, mass, ids, softening=1e-3)
    rel = fa.rel_errors(r.accel[ids], a_ref)
    say(f"  median rel err {np.median(rel):.3e} p99 {np.percentile(rel, 99):.3e}")
    assert np.median(rel) < 1e-5


@pytest.mark.parametrize("soft,n_leaf", [(1e-3, 32), (0.0, 96), (1e-2, 64)])
def test_fp32bulk_form_accuracy(fa, soft, n_leaf):
    """FP32-bulk form (P2P pair terms in FP32 on leaf-centred positions; M2L,
    L2L and L2P in FP32 on radius-scaled operands; FP64 tile and chunk sums,
    compensated final sums): 99 % of bodies within 2e-6 of the FP64 path, every
    body within 5e-5 (an unsoftened close pair in a large leaf keeps only the
    FP32 resolution of leaf-centred coordinates), and no worse than the FP64
    path against the FP64 direct sum (the metric's bar)."""
    pos, mass = O.plummer(20000, seed=71)
    say(f"[fp32bulk] soft={soft} n_leaf={n_leaf}")
    kw = dict(p=8, eps=math.inf, n_leaf=n_leaf, theta_max=0.5, softening=soft)
    r64 = fa.Fmm(None, None, **kw).accel(pos, mass)
    r32 = fa.Fmm(None, None, precision="fp32bulk", **kw).accel(pos, mass)
    an = np.linalg.norm(r64.accel, axis=1)
    d = np.linalg.norm(r32.accel - r64.accel, axis=1) / an
    assert np.percentile(d, 99) < 2e-6 and np.median(d) < 5e-7, (np.percentile(d, 99), np.median(d))
    assert d.max() < 5e-5, d.max()
    assert np.all(np.abs(r32.pot - r64.pot) <= 5e-6 * np.abs(r64.pot))
    ids = fa.sample_ids(len(mass), 2000, 7)
    a_ref, _, _ = fa.direct_sampled(pos, mass, ids, softening=soft)
    e64 = fa.rel_errors(r64.accel[ids], a_ref)
    e32 = fa.rel_errors(r32.accel[ids], a_ref)
    assert np.percentile(e32, 99) <= 1.2 * np.percentile(e64, 99) + 1e-6
