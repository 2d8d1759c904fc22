"""Pins the CPU oracle (oracle/fmm_oracle.cpp) against the reference's own
known-answer tests, closed forms and invariants, chosen so that a plausible
restatement mistake (dropped term, wrong sign or index, transposed operand)
fails at least one of them. Every test cites the reference test it ports,
or states the mathematical fact it checks. CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def grid64():
    """tree.rs:452-464: 4x4x4 unit grid, unit masses."""
    pos = np.array([[i % 4, (i // 4) % 4, i // 16] for i in range(64)], float)
    return pos, np.ones(64)


# ------------------------------------------------------------------ models.rs
def test_splitmix_first_draws():
    # SplitMix64 (models.rs:20-26): known first output for seed 0 is
    # 0xE220A8397B1DCDAF (published reference value of the generator).
    u = O.rng_units(0, 1)[0]
    assert u == (0xE220A8397B1DCDAF >> 11) / float(1 << 53)


def test_plummer_shell_mass():
    # models.rs:89-106: M(<1)/M = 2^-3/2 = 0.3536 within 0.02
    pos, mass = O.plummer(100_000, seed=42)
    frac = mass[np.sum(pos * pos, axis=1) <= 1.0].sum() / mass.sum()
    assert abs(frac - 0.3536) < 0.02
    assert np.allclose(mass, 1e-5)


def test_hernquist_shell_mass():
    # models.rs:109-124: M(<1)/M = 1/4
    pos, mass = O.hernquist(100_000, seed=7)
    frac = mass[np.sum(pos * pos, axis=1) <= 1.0].sum() / mass.sum()
    assert abs(frac - 0.25) < 0.02


# ------------------------------------------------------------------ tree.rs
def test_dilate_kat():
    # tree.rs:549-552
    assert O.dilate(0b11) == 0b1001
    assert O.dilate(1) == 1


def test_dilate_brute_force():
    rng = np.random.default_rng(0)
    for v in rng.integers(0, 1 << 21, 200):
        v = int(v)
        want = 0
        for b in range(21):
            want |= ((v >> b) & 1) << (3 * b)
        assert O.dilate(v) == want
    assert O.dilate((1 << 21) | 5) == O.dilate(5)  # masked to 21 bits


def test_morton_key_interleave_and_x_min_quirk():
    # tree.rs:37-46: z in bit 0, y in bit 1, x in bit 2; every axis is
    # measured from lo[0] (the x minimum), then clamped to [0, 1].
    lo = np.array([0.0, 5.0, -3.0])
    inv = 1.0
    q = lambda c: min(max((c - lo[0]) * inv, 0.0), 1.0) * ((1 << 21) - 1)
    x, y, z = 0.25, 0.75, 0.5
    want = O.dilate(int(q(z))) | (O.dilate(int(q(y))) << 1) | (O.dilate(int(q(x))) << 2)
    assert O.morton_key(x, y, z, lo, inv) == want
    # y offset above lo[0] + side clamps to 1 even though y is inside its own range
    assert O.morton_key(0.0, 5.5, 0.0, lo, inv) == O.dilate((1 << 21) - 1) << 1


def test_grid_tree_kat():
    # tree.rs:467-480: 73 cells, root mass 64, centre 1.5, rho sqrt(6.75)
    pos, mass = grid64()
    f = O.OracleFmm(pos, mass, p=2, eps=1e-2, n_leaf=1)
    t = f.tree()
    assert f.ncells == 73
    assert abs(t.mass[0] - 64) < 1e-12
    assert np.allclose(t.cen[0], 1.5, atol=1e-9)
    assert abs(t.rho[0] - math.sqrt(6.75)) < 1e-9


def test_degenerate_points_single_leaf():
    # tree.rs:483-496 / 532-547
    pos = np.ones((5, 3))
    mass = np.full(5, 2.0)
    f = O.OracleFmm(pos, mass, p=2, eps=1e-2, n_leaf=5)
    t = f.tree()
    assert f.ncells == 1 and t.leaf[0] == 1 and abs(t.mass[0] - 10) < 1e-12
    g = O.OracleFmm(pos, mass, p=2, eps=1e-2, n_leaf=16)
    assert g.ncells == 1


def _tree_equal(a, b):
    for k in ("order", "first", "count", "children", "leaf", "level"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for k in ("cen", "mass", "rho"):
        assert np.array_equal(getattr(a, k).view(np.uint64), getattr(b, k).view(np.uint64)), k


@pytest.mark.parametrize("mode", [0, 1])
def test_refresh_matches_build(mode):
    # tree.rs:499-529
    pos, mass = O.plummer(4096, seed=99)
    fresh = O.OracleFmm(pos, mass, n_leaf=16, center_mode=mode)
    shifted = pos + np.array([5.0, -3.0, 1.0])
    re = O.OracleFmm(shifted, mass, n_leaf=16, center_mode=mode)
    re.refresh(pos, mass)
    _tree_equal(fresh.tree(), re.tree())


def test_tree_invariants():
    pos, mass = O.plummer(3000, seed=3)
    f = O.OracleFmm(pos, mass, n_leaf=8)
    t = f.tree()
    seen = np.zeros(3000, int)
    for c in range(f.ncells):
        ch = t.children[c]
        if t.leaf[c]:
            assert (ch < 0).all()
            seen[t.first[c]:t.first[c] + t.count[c]] += 1
            assert t.count[c] <= 8 or t.level[c] >= 21
        else:
            kids = ch[ch >= 0]
            assert (kids > c).all()  # DFS preorder: parents first
            assert t.count[kids].sum() == t.count[c]
            assert t.first[kids[0]] == t.first[c]
            assert np.isclose(t.mass[kids].sum(), t.mass[c], rtol=1e-13)
    assert (seen == 1).all()
    # keys sorted; order is a permutation
    assert (np.diff(t.keys.astype(np.int64)) >= 0).all()
    assert np.array_equal(np.sort(t.order), np.arange(3000))


def test_sphere_mode_encloses():
    # tree.rs:555-586
    pos, mass = grid64()
    com = O.OracleFmm(pos, mass, n_leaf=1).tree()
    sph = O.OracleFmm(pos, mass, n_leaf=1, center_mode=1).tree()
    assert sph.rho[0] <= com.rho[0] + 1e-9
    for c in range(len(sph.first)):
        if sph.leaf[c]:
            ids = sph.order[sph.first[c]:sph.first[c] + sph.count[c]]
            assert (np.linalg.norm(pos[ids] - sph.cen[c], axis=1) <= sph.rho[c] + 1e-9).all()
        else:
            for k in sph.children[c][sph.children[c] >= 0]:
                assert np.linalg.norm(sph.cen[k] - sph.cen[c]) + sph.rho[k] <= sph.rho[c] + 1e-9


# -------------------------------------------------------------- harmonic.rs
def test_table3_low_orders():
    # harmonic.rs:186-249 (Dehnen 2014 Table 3)
    x, y, z = 0.7, -0.4, 1.2
    p = 3
    r2 = x * x + y * y + z * z
    u = O.upsilon_table([x, y, z], p)
    t = O.theta_table([x, y, z], p)
    fact = math.factorial

    def val(tab, n, m):
        c = tab[O.idx(n, abs(m))]
        return c.real if m >= 0 else c.imag

    cases = [(0, 0, 1.0), (1, 0, z), (1, 1, x), (1, -1, y), (2, 0, 3 * z * z - r2),
             (2, 1, 3 * x * z), (2, -1, 3 * y * z), (2, 2, 3 * (x * x - y * y)), (2, -2, 6 * x * y),
             (3, 0, 3 * z * (5 * z * z - 3 * r2)), (3, 1, 3 * x * (5 * z * z - r2))]
    for n, m, poly in cases:
        c = fact(n - abs(m)) * fact(n + abs(m))
        assert abs(c * val(u, n, m) - poly) < 1e-12
        assert abs(r2 ** (n + 0.5) * val(t, n, m) - poly) < 1e-12


def test_theta_parity_upsilon_homogeneity_relation():
    # harmonic.rs:252-317
    r = np.array([1.3, -2.1, 0.4])
    p = 6
    t, tm = O.theta_table(r, p), O.theta_table(-r, p)
    u, us = O.upsilon_table(r, p), O.upsilon_table(1.7 * r, p)
    r2 = r @ r
    for n in range(p + 1):
        for m in range(-n, n + 1):
            i = O.idx(n, m)
            assert abs(tm[i] - (-1) ** n * t[i]) < 1e-12 * max(1, abs(t[i]))
            assert abs(us[i] - 1.7 ** n * u[i]) < 1e-12 * max(1, abs(us[i]))
            c = math.factorial(n - abs(m)) * math.factorial(n + abs(m))
            assert abs(u[i] - t[i] * r2 ** (n + 0.5) / c) < 1e-11 * max(1, abs(u[i]))
            # negative m: (-1)^m conj of positive m (harmonic.rs:230-236)
            j = O.idx(n, abs(m))
            assert u[i] == ((-1) ** abs(m) * np.conj(u[j]) if m < 0 else u[j])


def test_theta_is_gradient_of_inverse_r_closed_form():
    # Theta_0^0 = 1/r, Theta_1^0 = z/r^3, Theta_1^1 = (x+iy)/r^3 (Dehnen eq. 40)
    r = np.array([0.3, 0.9, -0.5])
    t = O.theta_table(r, 1)
    rn = np.linalg.norm(r)
    assert abs(t[0] - 1 / rn) < 1e-15
    assert abs(t[O.idx(1, 0)] - r[2] / rn ** 3) < 1e-14
    assert abs(t[O.idx(1, 1)] - (r[0] + 1j * r[1]) / rn ** 3) < 1e-14


def test_index_layout():
    assert [O.dim(k) for k in (0, 1, 3)] == [1, 4, 16]
    assert [O.idx(1, -1), O.idx(1, 0), O.idx(1, 1), O.idx(2, -2), O.idx(2, 0), O.idx(2, 2),
            O.idx(3, -3), O.idx(3, 3)] == [1, 2, 3, 4, 6, 8, 9, 15]


# ---------------------------------------------------------------- moments.rs
def test_factorials_binomials():
    # moments.rs:218-224
    assert list(O.factorials(4)[:6]) == [1, 1, 2, 6, 24, 120]
    assert list(O.binomials(4)) == [1, 4, 6, 4, 1]


def test_monopole_mass_dipole_zero():
    # moments.rs:190-216
    pos = np.array([[1.0, 0, 0], [0, 2.0, 0]])
    mass = np.array([2.0, 3.0])
    f = O.OracleFmm(pos, mass, p=2, eps=1e-2, n_leaf=1)
    f.run(pos, mass)
    m = f.moments()
    assert abs(m[0, 0] - 5.0) < 1e-12
    assert abs(m[0, O.idx(1, 0)]) < 1e-12 and abs(m[0, O.idx(1, 1)]) < 1e-12


def test_m2m_equals_direct_p2m():
    # M2M of the children's P2M equals P2M of all bodies about the parent
    # centre (exact shift of a degree-p multipole truncated at p).
    pos, mass = O.plummer(500, seed=21)
    p = 6
    f = O.OracleFmm(pos, mass, p=p, eps=1e-4, n_leaf=8)
    f.run(pos, mass)
    t = f.tree()
    mom = f.moments()
    for c in [0] + [int(k) for k in t.children[0] if k >= 0][:3]:
        ids = t.order[t.first[c]:t.first[c] + t.count[c]]
        direct = np.zeros(O.dim(p), complex)
        for i in ids:
            direct += mass[i] * O.upsilon_table(pos[i] - t.cen[c], p)
        assert np.allclose(mom[c], direct, rtol=1e-11, atol=1e-14 * np.abs(direct).max())


def test_multipole_power_formula():
    pos, mass = O.plummer(300, seed=2)
    p = 5
    f = O.OracleFmm(pos, mass, p=p, eps=1e-4, n_leaf=8)
    f.run(pos, mass)
    mom, pw = f.moments(), f.pows()
    for n in range(p + 1):
        s = sum(math.factorial(n - abs(m)) * math.factorial(n + abs(m)) * abs(mom[0, O.idx(n, m)]) ** 2
                for m in range(-n, n + 1))
        assert abs(pw[0, n] - math.sqrt(s)) <= 1e-12 * max(1.0, math.sqrt(s))


# -------------------------------------------------------------------- mac.rs
def test_f_prepass_two_body():
    # mac.rs:221-238
    pos = np.array([[0.0, 0, 0], [3.0, 0, 0]])
    mass = np.array([1.0, 2.0])
    f = O.OracleFmm(pos, mass, p=2, eps=1e-2, n_leaf=1)
    f.run(pos, mass)
    fv = f.f()
    t = f.tree()
    got = dict(zip(t.order, fv))
    assert abs(got[0] - 2 / 9) < 1e-12 and abs(got[1] - 1 / 9) < 1e-12


def test_f_prepass_brute_force_bound():
    # The p=0, theta<1 monopole walk approximates f_b = sum m/r^2; leaves
    # summed exactly, far cells by mass/d^2 with d > rho: within 50%.
    pos, mass = O.plummer(400, seed=8)
    f = O.OracleFmm(pos, mass, p=3, eps=1e-3, n_leaf=4)
    f.run(pos, mass)
    t = f.tree()
    fv = f.f()
    ids = t.order
    for k in range(0, 400, 37):
        i = ids[k]
        d2 = np.sum((pos - pos[i]) ** 2, axis=1)
        d2[i] = np.inf
        exact = np.sum(mass / d2)
        assert 0.5 < fv[k] / exact < 1.5
    _, fmin = fv, f.fmin()
    assert fmin[0] == fv.min()


# ------------------------------------------------------------------ walk.rs
def _brute_coverage(t, m2l, p2p, n):
    """Count how often every ordered body pair (i != j) is covered by a P2P
    leaf pair or by an M2L pair of ancestors (both directions count)."""
    nc = len(t.first)
    slots = [np.arange(t.first[c], t.first[c] + t.count[c]) for c in range(nc)]
    cov = np.zeros((n, n), int)
    for a, b in p2p:
        sa, sb = slots[a], slots[b]
        if a == b:
            cov[np.ix_(sa, sa)] += 1
        else:
            cov[np.ix_(sa, sb)] += 1
            cov[np.ix_(sb, sa)] += 1
    for a, b in m2l:
        sa, sb = slots[a], slots[b]
        cov[np.ix_(sa, sb)] += 1
        cov[np.ix_(sb, sa)] += 1
    return cov


@pytest.mark.parametrize("p,eps,theta", [(4, 1e-2, 1.0), (6, 1e-5, 1.0), (8, math.inf, 0.5)])
def test_lists_cover_every_pair_exactly_once(p, eps, theta):
    n = 300
    pos, mass = O.plummer(n, seed=11)
    f = O.OracleFmm(pos, mass, p=p, eps=eps, n_leaf=6, theta_max=theta)
    f.run(pos, mass)
    t = f.tree()
    m2l, p2p = f.m2l_pairs(), f.p2p_pairs()
    cov = _brute_coverage(t, m2l, p2p, n)
    off = ~np.eye(n, dtype=bool)
    assert (cov[off] == 1).all()
    assert (cov.diagonal() == 1).all()  # only the self pair (a,a) of i's leaf
    for a, b in p2p:
        assert t.leaf[a] and t.leaf[b]
    assert set(np.flatnonzero(t.leaf)) <= set(p2p.ravel())  # walk.rs:689-720


def test_threaded_walk_same_pair_set():
    # walk.rs:663-686: the pair set does not depend on the thread schedule
    pos, mass = O.plummer(4096, seed=77)
    f = O.OracleFmm(pos, mass, p=4, eps=1e-2, n_leaf=16)
    f.run(pos, mass, nthreads=1)
    s1 = {tuple(r) for r in f.m2l_pairs()}, {tuple(r) for r in f.p2p_pairs()}
    f.run(pos, mass, nthreads=4)
    s4 = {tuple(r) for r in f.m2l_pairs()}, {tuple(r) for r in f.p2p_pairs()}
    assert s1 == s4


# ------------------------------------------------------------- translate.rs
def test_rotation_matches_naive():
    # rotation.rs:505-556, extended over p in {8, 20, 28}
    rng = np.random.default_rng(11)
    for p in (8, 20, 28):
        d = O.dim(p)
        ma = rng.random(d) - 0.5 + 1j * (rng.random(d) - 0.5)
        mb = rng.random(d) - 0.5 + 1j * (rng.random(d) - 0.5)
        ca, cb = np.array([0.1, 0.2, 0.3]), np.array([1.1, 0.9, 1.6])
        na, nb = O.m2l_mutual(p, ca, cb, ma, mb)
        ra, rb = O.m2l_rotation(p, ca, cb, ma, mb)
        for x, y in ((na, ra), (nb, rb)):
            tol = 1e-9 * (np.abs(x.real) + np.abs(x.imag)).max()
            assert np.abs(x - y).max() < tol


def test_swap_involution_and_rotates_theta():
    # rotation.rs:446-502
    p = 8
    sw = O.rot_swap(p)
    x = np.array([0.6, -0.9, 1.3])
    t, tp = O.theta_table(x, p), O.theta_table(x[::-1].copy(), p)
    off = 0
    for n in range(p + 1):
        nb = 2 * n + 1
        B = sw[off:off + nb * nb].reshape(nb, nb)
        assert np.abs(B @ B - np.eye(nb)).max() < 1e-9
        tn = np.array([t[O.idx(n, l)] for l in range(-n, n + 1)])
        tpn = np.array([tp[O.idx(n, m)] for m in range(-n, n + 1)])
        assert np.abs(B @ tn - tpn).max() < 1e-9
        off += nb * nb


def test_m2l_matches_direct_potential():
    # Far-field potential/field of a cluster from M2L + L2P vs direct sum.
    rng = np.random.default_rng(5)
    p = 10
    src = rng.random((20, 3)) * 0.2
    ms = rng.random(20)
    tgt = np.array([1.5, 1.2, 1.7]) + rng.random((5, 3)) * 0.2
    za, zb = src.mean(0), tgt.mean(0)
    M = sum(m * O.upsilon_table(x - za, p) for x, m in zip(src, ms))
    zero = np.zeros(O.dim(p), complex)
    _, F = O.m2l_mutual(p, za, zb, M, zero)
    for x in tgt:
        pot, acc = O.l2p(p, zb, x, F)
        d = src - x
        r = np.linalg.norm(d, axis=1)
        assert abs(pot - np.sum(ms / r)) < 1e-7 * np.sum(ms / r)
        ad = np.sum(ms[:, None] * d / r[:, None] ** 3, axis=0)
        assert np.linalg.norm(acc - ad) < 1e-6 * np.linalg.norm(ad)


def test_l2l_is_exact_shift():
    # L2P(L2L(F), x) == L2P(F, x) for a degree-p local expansion
    rng = np.random.default_rng(3)
    p = 8
    d = O.dim(p)
    F = np.zeros(d, complex)
    # a valid local expansion: M2L of a random multipole
    M = rng.random(d) - 0.5 + 1j * (rng.random(d) - 0.5)
    for n in range(p + 1):  # impose the Delta^{-m} = (-1)^m conj symmetry
        for m in range(1, n + 1):
            M[O.idx(n, -m)] = (-1) ** m * np.conj(M[O.idx(n, m)])
        M[O.idx(n, 0)] = M[O.idx(n, 0)].real
    _, F = O.m2l_mutual(p, np.array([5.0, 4.0, 6.0]), np.zeros(3), M, np.zeros(d, complex))
    child = np.array([0.1, -0.05, 0.08])
    Fc = O.l2l(p, np.zeros(3), child, F)
    x = child + np.array([0.02, 0.03, -0.01])
    p1, a1 = O.l2p(p, np.zeros(3), x, F)
    p2, a2 = O.l2p(p, child, x, Fc)
    assert abs(p1 - p2) < 1e-13 * abs(p1)
    assert np.abs(a1 - a2).max() < 1e-12 * np.abs(a1).max()


# -------------------------------------------------------------------- fmm.rs
def test_integration_thresholds():
    # tests/integration.rs:11-42
    n = 256
    pos, mass = O.plummer(n, seed=1234)
    a_ref, pot_ref = O.direct(pos, mass)
    for p, eps in [(2, 1e-2), (3, 1e-3), (4, 1e-4), (5, 1e-5)]:
        f = O.OracleFmm(pos, mass, p=p, eps=eps, n_leaf=16)
        a, pot = f.run(pos, mass)
        max_s, _, rms_s, rms_rel = O.scaled_error(a_ref, a, pos, mass)
        assert max_s < 10 * eps and rms_s < 10 * eps and rms_rel < 1e-2
        rel = np.sqrt(np.mean(((pot - pot_ref) / np.maximum(np.abs(pot_ref), 1e-12)) ** 2))
        assert rel < 1e-3


def test_momentum_conserved():
    # tests/integration.rs:45-68
    pos, mass = O.plummer(512, seed=99)
    f = O.OracleFmm(pos, mass, p=4, eps=1e-4, n_leaf=16)
    a, _ = f.run(pos, mass)
    fsum = (mass[:, None] * a).sum(0)
    ftot = (mass * np.linalg.norm(a, axis=1)).sum()
    assert np.linalg.norm(fsum) / ftot < 1e-14


def test_hernquist_smoke():
    # tests/exit_gate.rs:27-36
    pos, mass = O.hernquist(4096, seed=7)
    f = O.OracleFmm(pos, mass, p=4, eps=1e-3, n_leaf=32)
    a, _ = f.run(pos, mass)
    a_ref, _ = O.direct(pos, mass)
    _, _, rms_s, _ = O.scaled_error(a_ref, a, pos, mass)
    assert rms_s < 1e-2


def test_threaded_matches_serial():
    # walk.rs:803-826
    pos, mass = O.plummer(4096, seed=5)
    for p in (8, 28):
        f = O.OracleFmm(pos, mass, p=p, eps=1e-5, n_leaf=16)
        a1, p1 = f.run(pos, mass, 1)
        a2, p2 = f.run(pos, mass, 8)
        assert np.all(np.abs(a1 - a2) < 1e-8 * (1 + np.abs(a1) + np.abs(a2)))
        assert np.all(np.abs(p1 - p2) < 1e-8 * (1 + np.abs(p1) + np.abs(p2)))


def test_softened_direct_and_fmm():
    # Extension: softening eps_s. The direct sum follows the closed form of
    # two bodies, and the softened FMM converges to the softened direct sum.
    pos = np.array([[0.0, 0, 0], [0.3, 0.4, 0]])
    mass = np.array([1.0, 2.0])
    a, pot = O.direct(pos, mass, softening=0.1)
    r2 = 0.25 + 0.01
    assert abs(pot[0] - 2 / math.sqrt(r2)) < 1e-15
    assert np.allclose(a[0], 2 * np.array([0.3, 0.4, 0]) / r2 ** 1.5, rtol=1e-14)
    pos, mass = O.plummer(2000, seed=4)
    a_ref, _ = O.direct(pos, mass, softening=1e-3)
    f = O.OracleFmm(pos, mass, p=8, eps=math.inf, n_leaf=16, theta_max=0.5, softening=1e-3)
    a, _ = f.run(pos, mass)
    rel = np.linalg.norm(a - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
    assert np.median(rel) < 1e-5


def test_direct_targets_matches_direct():
    pos, mass = O.plummer(500, seed=6)
    a, pot = O.direct(pos, mass, softening=1e-3)
    ids = O.sample_ids(500, 40, seed=7)
    at, pt, _ = O.direct_targets(pos, mass, ids, softening=1e-3, nthreads=2)
    assert np.allclose(at, a[ids], rtol=1e-12, atol=1e-12)
    assert np.allclose(pt, pot[ids], rtol=1e-12)


def _cluster(rng, nb, rad):
    d = rng.normal(size=(nb, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d *= rad * rng.random(nb)[:, None] ** (1 / 3)
    m = rng.random(nb) + 0.5
    d -= (m[:, None] * d).sum(0) / m.sum()
    return d, m


def _laplacian_field_exact(eps):
    """L = Lap K, G = grad L, H = Hess L of the Plummer kernel K, by computer
    algebra (independent of the oracle's closed forms)."""
    sp = pytest.importorskip("sympy")
    x, y, z = sp.symbols("x y z", real=True)
    K = 1 / sp.sqrt(x * x + y * y + z * z + eps * eps)
    L = sp.diff(K, x, 2) + sp.diff(K, y, 2) + sp.diff(K, z, 2)
    G = [sp.diff(L, v) for v in (x, y, z)]
    H = [sp.diff(L, x, 2), sp.diff(L, y, 2), sp.diff(L, z, 2), sp.diff(L, x, y),
         sp.diff(L, x, z), sp.diff(L, y, z)]
    f = sp.lambdify((x, y, z), [L] + G + H, "math")
    return lambda r: np.array(f(*r), float)


@pytest.mark.parametrize("eps", [0.05, 0.2])
def test_softened_laplacian_field_against_computer_algebra(eps):
    """The target-side Laplacian field of one M2L source, [L, G, H] at the
    target centre, against the exact sum over the source's bodies of Lap K,
    grad Lap K and Hess Lap K (computer algebra). A point mass matches to
    rounding; a small octahedral cluster (radius a, no dipole, no
    quadrupole, no third moments) matches L and G to O(a^4) because the
    source-side field (q/6) Lap^2 K is in them; leaving it out (a plausible
    omission) misses by O(a^2) and fails the same bound."""
    fld = _laplacian_field_exact(eps)
    zA = np.zeros(3)
    u = np.array([0.3, -0.5, 0.8])
    zB = u / np.linalg.norm(u)
    zero = np.zeros(O.dim(4), np.complex128)
    lap = O.soft_trace_pair(4, zA, zB, 0.0, 2.5, eps * eps, zero.copy())
    want = 2.5 * fld(zB - zA)
    assert np.allclose(lap, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    a = 0.02
    d = a * np.vstack([np.eye(3), -np.eye(3)])
    m = np.full(6, 0.5)
    q = float(np.sum(m * (d * d).sum(1)))
    lap = O.soft_trace_pair(4, zA, zB, q, m.sum(), eps * eps, zero.copy())
    want = sum(mi * fld(zB - (zA + di)) for mi, di in zip(m, d))
    lap0 = O.soft_trace_pair(4, zA, zB, 0.0, m.sum(), eps * eps, zero.copy())
    for sl, tol in ((slice(0, 1), 1e-5), (slice(1, 4), 1e-5), (slice(4, 10), 5e-3)):
        scale = np.abs(want[sl]).max()
        assert np.abs(lap[sl] - want[sl]).max() < tol * scale, sl
    for sl in (slice(0, 1), slice(1, 4)):
        scale = np.abs(want[sl]).max()
        assert np.abs(lap0[sl] - want[sl]).max() > 10 * 1e-5 * scale, sl


@pytest.mark.parametrize("eps", [0.02, 0.05])
def test_softened_far_field_against_exact_sum(eps):
    """One well-separated pair of clusters (radii 0.15, separation 1, p = 8):
    M2L with the harmonic derivatives of the Plummer kernel, the source-side
    field (q/6) Lap K to second order, L2P and the target-side trace
    polynomial to fourth order, against the exact softened sum. Each
    plausible omission -- no trace terms at all, no second-order (F_2) part
    of the source-side field, no Hessian H, no source-side part of L and G --
    fails the same bound; eps -> 0 reproduces the unsoftened M2L bit for
    bit."""
    p = 8
    rng = np.random.default_rng(11)
    dA, mA = _cluster(rng, 300, 0.15)
    dB, _ = _cluster(rng, 40, 0.15)
    zA = np.zeros(3)
    u = rng.normal(size=3)
    zB = u / np.linalg.norm(u)
    M = sum(mm * O.upsilon_table(dd, p) for mm, dd in zip(mA, dA))
    q = float(np.sum(mA * (dA ** 2).sum(1)))
    zero = np.zeros(O.dim(p), np.complex128)
    _, F = O.m2l_mutual_soft(p, zA, zB, M, zero, eps * eps)
    Fc = F.copy()
    lap = O.soft_trace_pair(p, zA, zB, q, mA.sum(), eps * eps, Fc)
    corr = Fc - F
    lap_noq = O.soft_trace_pair(p, zA, zB, 0.0, mA.sum(), eps * eps, zero.copy())

    def median_err(Fx, lx, with_trace=True):
        errs = []
        for db in dB:
            x = zB + db
            diff = x - (zA + dA)
            s = np.sqrt((diff ** 2).sum(1) + eps * eps)
            a_ex = -(mA[:, None] * diff / s[:, None] ** 3).sum(0)
            p_ex = np.sum(mA / s)
            p1, a1 = O.l2p(p, zB, x, Fx)
            if with_trace:
                pt, at = O.l2p_trace(zB, x, lx)
                p1, a1 = p1 + pt, a1 + at
            errs.append(max(np.linalg.norm(a1 - a_ex) / np.linalg.norm(a_ex),
                            abs(p1 - p_ex) / p_ex))
        return np.median(errs)

    bound = 1.4e-5 * (eps / 0.05) ** 2
    assert median_err(Fc, lap) < bound
    assert median_err(F, lap, with_trace=False) > 5 * bound  # no trace terms
    no_f2 = F + np.where(np.arange(O.dim(p)) >= 4, 0, corr)
    assert median_err(no_f2, lap) > bound  # source-side field to first order only
    no_h = lap.copy()
    no_h[4:] = 0.0
    assert median_err(Fc, no_h) > bound  # target-side polynomial to third order only
    no_ql = lap.copy()
    no_ql[:4] = lap_noq[:4]
    assert median_err(Fc, no_ql) > bound  # source-side field missing from L, G
    # eps -> 0: the unsoftened M2L exactly
    _, F0 = O.m2l_mutual_soft(p, zA, zB, M, zero, 0.0)
    _, Fu = O.m2l_mutual(p, zA, zB, M, zero)
    assert np.array_equal(F0, Fu)


def test_softened_core_error_floor():
    """Full oracle FMM where softening reaches into the far field: Plummer
    N = 262,144, eps = 4e-3, p = 10, theta = 0.5, n_leaf = 128 (eps relative
    to the core's leaf size as eps = 1e-3 at 16.7M bodies). Thousands of M2L
    pairs per core target add their softening residuals coherently; with
    the trace terms to the orders kept here the 99th-percentile error stays
    within 1.5x of the unsoftened run's. Moving the target-side polynomial
    through L2L as a Laplacian field, without the harmonic terms that a
    centre shift creates, leaves it 8x above; stopping the polynomial at
    third order, 4x."""
    n = 262144
    pos, _, mass = O.ic_plummer(n, seed=42)
    ids = O.sample_ids(n, 2048, 7)
    p99 = {}
    for soft in (0.0, 4e-3):
        acc = np.empty((len(ids), 3))
        pot = np.empty(len(ids))
        ff = np.empty(len(ids))
        O.lib().orc_direct_targets(n, O._p(pos), O._p(mass), soft, len(ids),
                                   O._p(ids, O._pu32), O._p(acc), O._p(pot), O._p(ff), 8)
        f = O.OracleFmm(pos, mass, p=10, eps=math.inf, n_leaf=128, theta_max=0.5, softening=soft)
        a, _ = f.run(pos, mass, nthreads=8)
        rel = np.linalg.norm(a[ids] - acc, axis=1) / np.linalg.norm(acc, axis=1)
        p99[soft] = np.percentile(rel, 99)
    assert p99[4e-3] < 1.5 * p99[0.0], p99


def test_softening_guard():
    """Softening guard of the MAC (extension): every accepted M2L pair keeps
    eps^2 (rhoA + rhoB)^3 <= tau r^5, rejected pairs end in P2P or deeper
    M2L (the forces still converge to the direct sum), tau = 0 leaves the
    lists unchanged, and the guard only acts with softening."""
    pos, mass = O.plummer(3000, seed=4)
    soft = 2e-2
    a_ref, _ = O.direct(pos, mass, softening=soft)
    kw = dict(p=8, eps=math.inf, n_leaf=16, theta_max=0.5, softening=soft)
    base = O.OracleFmm(pos, mass, **kw)
    a0, _ = base.run(pos, mass)
    lists0 = (sorted(map(tuple, base.m2l_pairs())), sorted(map(tuple, base.p2p_pairs())))
    off = O.OracleFmm(pos, mass, soft_mac=0.0, **kw)
    off.run(pos, mass)
    assert (sorted(map(tuple, off.m2l_pairs())), sorted(map(tuple, off.p2p_pairs()))) == lists0
    for tau in (1e-4, 1e-5):
        f = O.OracleFmm(pos, mass, soft_mac=tau, **kw)
        a, _ = f.run(pos, mass)
        t = f.tree()
        pr = f.m2l_pairs()
        s = t.rho[pr[:, 0]] + t.rho[pr[:, 1]]
        r2 = ((t.cen[pr[:, 0]] - t.cen[pr[:, 1]]) ** 2).sum(1)
        assert np.all(soft * soft * s ** 3 <= tau * r2 ** 2.5 * (1 + 1e-12))
        assert len(f.p2p_pairs()) > len(lists0[1])
        rel = np.linalg.norm(a - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
        rel0 = np.linalg.norm(a0 - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
        assert np.percentile(rel, 99) <= 1.05 * np.percentile(rel0, 99)
    g = O.OracleFmm(pos, mass, soft_mac=1e-5, **{**kw, "softening": 0.0})
    g.run(pos, mass)
    nosoft = O.OracleFmm(pos, mass, **{**kw, "softening": 0.0})
    nosoft.run(pos, mass)
    assert sorted(map(tuple, g.m2l_pairs())) == sorted(map(tuple, nosoft.m2l_pairs()))


def test_softened_fmm_error_matches_unsoftened_truncation():
    """Full oracle FMM, Plummer N = 30,000, p = 8, theta = 0.5: with Plummer
    softening 1e-3 the error against the softened direct sum stays at the
    unsoftened truncation level (the softened far field adds no error of its
    own at this eps)."""
    n = 30000
    pos, mass = O.plummer(n, seed=3)
    ids = O.sample_ids(n, 400, 7)
    meds = {}
    for soft in (0.0, 1e-3):
        f = O.OracleFmm(pos, mass, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=soft)
        a, _ = f.run(pos, mass, nthreads=4)
        acc = np.empty((len(ids), 3))
        pot = np.empty(len(ids))
        ff = np.empty(len(ids))
        O.lib().orc_direct_targets(n, O._p(pos), O._p(mass), soft, len(ids),
                                   O._p(ids, O._pu32), O._p(acc), O._p(pot), O._p(ff), 4)
        rel = np.linalg.norm(a[ids] - acc, axis=1) / np.linalg.norm(acc, axis=1)
        meds[soft] = np.median(rel)
    assert meds[1e-3] < 1.3 * meds[0.0], meds
