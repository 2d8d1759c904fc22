"""GPU tests of temporal reuse (frozen evaluations) and the device-resident
kick-drift-kick leapfrog, through the C ABI.

Bars:
  * frozen evaluation on the positions of the last full evaluation: bit-equal
    to it; on moved bodies under the theta MAC (decisions depend only on the
    frozen cell geometry, so the reference's `Fmm::run` on a stale tree walks
    the same lists): within 1e-11 of the oracle's run on the stale tree;
  * one leapfrog step: bit-equal to sim.rs:558-572's arithmetic done in numpy
    on the device forces;
  * graph replay: bit-equal to eager frozen steps, every step checked, with
    eager work between replays;
  * the energy-gate thresholds of crates/fmm/tests/energy_gate.rs:236-255.
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


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_frozen_on_same_positions_equals_full(fa):
    pos, mass = O.plummer(4000, seed=51)
    say("[frozen] same positions")
    g = fa.Fmm(pos, mass, p=8, eps=1e-5, n_leaf=16)
    r1 = g.run(pos, mass)
    r2 = g.run_frozen(pos, mass)
    assert np.array_equal(bits(r1.accel), bits(r2.accel))
    assert np.array_equal(bits(r1.pot), bits(r2.pot))


@pytest.mark.parametrize("p,theta,soft", [(6, 0.5, 1e-3), (8, 0.7, 0.0), (16, 0.6, 0.0)])
def test_frozen_moved_bodies_match_oracle_stale_tree(fa, p, theta, soft):
    pos, mass = O.plummer(3000, seed=52)
    rng = np.random.default_rng(5)
    moved = pos + 1e-3 * rng.standard_normal(pos.shape)
    say(f"[frozen] moved bodies p={p} theta={theta} soft={soft}")
    orc = O.OracleFmm(pos, mass, p=p, eps=math.inf, n_leaf=16, theta_max=theta, softening=soft)
    a_ref, pot_ref = orc.run(moved, mass)  # stale tree, lists walked on its geometry
    g = fa.Fmm(pos, mass, p=p, eps=math.inf, n_leaf=16, theta_max=theta, softening=soft)
    g.run(pos, mass)
    m2l0, p2p0 = g.lists()
    r = g.run_frozen(moved, mass)
    an = np.linalg.norm(a_ref, axis=1)
    da = np.linalg.norm(r.accel - a_ref, axis=1)
    rel = 1e-11 if p <= 12 else 1e-9
    assert np.all(da <= rel * an + 1e-13 * an.max()), (da / an).max()
    assert np.all(np.abs(r.pot - pot_ref) <= rel * np.abs(pot_ref) + 1e-13)
    # the full run on the stale tree walks the same lists -> the same bits
    r_full = g.run(moved, mass)
    m2l1, p2p1 = g.lists()
    assert sorted(map(tuple, m2l0)) == sorted(map(tuple, m2l1))
    assert sorted(map(tuple, p2p0)) == sorted(map(tuple, p2p1))
    assert np.array_equal(bits(r_full.accel), bits(r.accel))


def test_frozen_requires_full_evaluation(fa):
    pos, mass = O.plummer(500, seed=53)
    g = fa.Fmm(pos, mass, p=4, eps=1e-3, n_leaf=16)  # built, never evaluated
    with pytest.raises(fa.FmmError) as e:
        g.run_frozen(pos, mass)
    assert e.value.code == fa.FMMH_E_STATE


def test_leapfrog_step_is_sim_rs_arithmetic(fa):
    s = fa.ic_plummer(3000, rmax=20.0, seed=61)
    dt = 1e-3
    say("[leapfrog] one KDK step vs numpy arithmetic")
    g = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    it = fa.Integrator(g, s.pos, s.vel, s.mass)
    it.step(dt, steps=0)  # evaluates a(x0) only
    x0, v0, a0, _ = it.state()
    t = it.step(dt, steps=1)
    x1, v1, a1, p1 = it.state()
    assert t == dt
    vh = v0 + a0 * dt * 0.5
    xw = x0 + vh * dt
    assert np.array_equal(bits(x1), bits(xw))
    # forces at the drifted positions from an independent context
    ref = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    r = ref.accel(xw, s.mass)
    assert np.array_equal(bits(a1), bits(r.accel)) and np.array_equal(bits(p1), bits(r.pot))
    assert np.array_equal(bits(v1), bits(vh + r.accel * dt * 0.5))


def test_leapfrog_matches_oracle_trajectory(fa):
    s = fa.ic_plummer(2000, rmax=20.0, seed=62)
    dt, steps = 2e-3, 4
    say("[leapfrog] 4 steps vs oracle-force KDK")
    g = fa.Fmm(None, None, p=8, eps=1e-5, n_leaf=16)
    it = fa.Integrator(g, s.pos, s.vel, s.mass)
    it.step(dt, steps=steps)
    x, v, a, pot = it.state()
    xr, vr = s.pos.copy(), s.vel.copy()
    ar, _ = O.OracleFmm(xr, s.mass, p=8, eps=1e-5, n_leaf=16).run(xr, s.mass)
    for _ in range(steps):
        vr = vr + ar * dt * 0.5
        xr = xr + vr * dt
        ar, potr = O.OracleFmm(xr, s.mass, p=8, eps=1e-5, n_leaf=16).run(xr, s.mass)
        vr = vr + ar * dt * 0.5
    assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max()
    assert np.abs(v - vr).max() <= 1e-10 * np.abs(vr).max()
    assert np.abs(pot - potr).max() <= 1e-10 * np.abs(potr).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_graph_replay_equals_eager_steps(fa, precision):
    s = fa.ic_plummer(6000, rmax=20.0, seed=63)
    dt = 1e-3
    say(f"[leapfrog] graph replay vs eager, refresh every 3, {precision}")
    runs = []
    for graph in (False, True):
        g = fa.Fmm(None, None, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3,
                   precision=precision)
        it = fa.Integrator(g, s.pos, s.vel, s.mass)
        states = []
        for k in range(7):  # one call per step, host copies between replays
            it.step(dt, steps=1, refresh_every=3, graph=graph)
            states.append(it.state())
            if graph:  # eager work of this process between the replays
                g2 = fa.Fmm(None, None, p=4, eps=1e-3, n_leaf=16)
                g2.accel(s.pos[:500], s.mass[:500])
        runs.append(states)
    for k, (e, gph) in enumerate(zip(*runs)):
        for a, b in zip(e, gph):
            assert np.array_equal(bits(a), bits(b)), f"step {k}"
    # several steps in one call replay the captured graph back to back
    outs = []
    for graph in (False, True):
        g = fa.Fmm(None, None, p=8, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3,
                   precision=precision)
        it = fa.Integrator(g, s.pos, s.vel, s.mass)
        it.step(dt, steps=9, refresh_every=4, graph=graph)
        outs.append(it.state())
    for a, b in zip(*outs):
        assert np.array_equal(bits(a), bits(b))


def test_dt_clamp_and_accel_cap(fa):
    s = fa.ic_plummer(3000, rmax=20.0, seed=64)
    say("[leapfrog] dt clamp and accel cap")
    g = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5)
    it = fa.Integrator(g, s.pos, s.vel, s.mass)
    it.step(0.05, steps=0, accel_cap=0.5)
    for _ in range(3):
        _, _, a0, _ = it.state()
        amax = max(1e-9, np.abs(a0).max())
        want = min(0.05, 0.25 * math.sqrt(0.02 / amax))  # sim.rs:553-557
        t = it.step(0.05, steps=1, dt_eta=0.25, dt_eps_len=0.02, accel_cap=0.5)
        assert t == want
        _, _, a1, _ = it.state()
        assert np.linalg.norm(a1, axis=1).max() <= 0.5 * (1 + 1e-15)


def _energy_run(fa, refresh_every, graph):
    # energy_gate.rs:128-160 at the smoke scale: N=1024, p=4, eps=1e-3,
    # n_leaf=16, dt=1e-4, 100 steps, virialised warm start.
    n, steps, dt = 1024, 100, 1e-4
    pos, mass = O.plummer(n, seed=42)
    g = fa.Fmm(pos, mass, p=4, eps=1e-3, n_leaf=16)
    r0 = g.run(pos, mass)
    vel = O.warm_start(pos, mass, r0.pot)
    it = fa.Integrator(g, pos, vel, mass)
    it.step(dt, steps=0)
    _, v, _, pot = it.state()
    e0 = O.kinetic(mass, v) + O.potential_energy(mass, pot)
    es = []
    for _ in range(steps):
        it.step(dt, steps=1, refresh_every=refresh_every, graph=graph)
        _, v, _, pot = it.state()
        es.append(O.kinetic(mass, v) + O.potential_energy(mass, pot))
    es = np.array(es)
    drift = (es[-1] - e0) / abs(e0)
    max_dev = max(abs(es.min() - e0), abs(es.max() - e0)) / abs(e0)
    return e0, drift, max_dev, O.momentum_rel(mass, v)


@pytest.mark.parametrize("refresh_every,graph", [(1, False), (4, True)])
def test_energy_gate_smoke(fa, refresh_every, graph):
    say(f"[energy gate] refresh_every={refresh_every} graph={graph}")
    e0, drift, max_dev, mom = _energy_run(fa, refresh_every, graph)
    say(f"  E0={e0:.6e} drift={drift:.3e} max_dev={max_dev:.3e} momentum={mom:.3e}")
    assert abs(e0 + 0.147) < 0.03      # energy_gate.rs:241-245
    assert mom < 1e-12                 # :246
    assert abs(drift) < 1e-2           # :247
    assert max_dev < 1e-2              # :248


def test_drift_monitor_forces_rebuilds(fa):
    s = fa.ic_plummer(20000, rmax=20.0, seed=65)
    dt = 2e-2  # bodies cross core leaves within a few steps
    say("[leapfrog] drift monitor")
    g = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    it = fa.Integrator(g, s.pos, s.vel, s.mass)
    it.step(dt, steps=8, refresh_every=1000, drift_tol=0.25)
    st = g.leapfrog_stats()
    assert st["steps"] == 8 and st["drift_rebuilds"] >= 1
    assert st["rebuilds"] == 1 + st["drift_rebuilds"]
    assert st["frozen_steps"] == 8 - st["drift_rebuilds"]
    x, _, a, _ = it.state()
    # the forces of the last step stay close to a fresh evaluation's
    ref = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    r = ref.accel(x, s.mass)
    rel = np.linalg.norm(a - r.accel, axis=1) / np.linalg.norm(r.accel, axis=1)
    assert np.percentile(rel, 99) < 1e-3
    # monitor off: the tree is never rebuilt within refresh_every
    g2 = fa.Fmm(None, None, p=6, eps=math.inf, n_leaf=32, theta_max=0.5, softening=1e-3)
    it2 = fa.Integrator(g2, s.pos, s.vel, s.mass)
    it2.step(dt, steps=3, refresh_every=1000, drift_tol=0.0)
    st2 = g2.leapfrog_stats()
    assert st2["rebuilds"] == 1 and st2["frozen_steps"] == 3
