"""GPU tests of the continuous-batching server (fmms_amd.serve): systems that
arrive at different step boundaries, join and leave the batched ensemble,
and queue under the HBM budget end in exactly the state they reach when
stepped alone (bitwise: the ensemble evaluates every system as its own tree
does, and the leapfrog arithmetic is per body), in the plain FP64 form and
in the mixed FP32-bulk form; requests larger than the whole budget are
rejected instead of waiting forever."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _requests(serve):
    models = ["plummer", "hernquist", "disk"]
    sizes = [900, 3000, 1500, 4200, 700, 2500, 1800]
    steps = [3, 5, 2, 6, 4, 1, 3]
    return [serve.make_request(i, models[i % 3], sizes[i], steps[i], seed=40 + i)
            for i in range(len(sizes))]


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
@pytest.mark.parametrize("budget_bodies", [None, 5000])
def test_served_trajectories_equal_alone(budget_bodies, precision):
    from fmms_amd import serve

    reqs = _requests(serve)
    kw = dict(p=8, theta=0.5, softening=1e-3, n_leaf=32, dt=1e-3)
    per_body = serve.bytes_per_body(8) + serve.STATE_BYTES
    budget = 200e9 if budget_bodies is None else serve.OVERHEAD_BYTES + budget_bodies * per_body
    srv = serve.Server(hbm_budget=budget, precision=precision, soft_mac=1e-4, **kw)
    arrive_at = [0, 0, 1, 2, 2, 4, 5]  # step boundary at which each request arrives
    say(f"[serve] {precision}: {len(reqs)} requests, budget {budget_bodies} bodies")
    boundary, max_active = 0, 0
    while any(r.final is None for r in reqs):
        for r, b in zip(reqs, arrive_at):
            if b == boundary:
                srv.submit(r)
        srv.step(float(boundary))
        max_active = max(max_active, sum(s.req.n for s in srv.active))
        boundary += 1
        assert boundary < 100
    if budget_bodies is not None:
        assert max_active <= budget_bodies
    for r in reqs:
        (x, v, a, pot), _ = serve.run_alone(r, precision=precision, soft_mac=1e-4, **kw)
        assert r.steps_done == r.steps and not r.rejected
        assert np.array_equal(bits(r.final[0]), bits(x)), f"request {r.rid} pos"
        assert np.array_equal(bits(r.final[1]), bits(v)), f"request {r.rid} vel"
        assert np.array_equal(bits(r.final[2]), bits(a)), f"request {r.rid} acc"
        assert np.array_equal(bits(r.final[3]), bits(pot)), f"request {r.rid} pot"
    st = srv.stats()
    assert st["steps"] == boundary or st["steps"] <= boundary


def test_oversize_request_is_rejected():
    """A request whose footprint alone exceeds the budget is rejected at
    submission (FMMH_E_OVER_BUDGET) and the serving loop still ends with the
    other requests served."""
    from fmms_amd import FMMH_E_OVER_BUDGET, FmmError, serve

    per_body = serve.bytes_per_body(6) + serve.STATE_BYTES
    budget = serve.OVERHEAD_BYTES + 3000 * per_body
    say("[serve] oversize request")
    srv = serve.Server(p=6, theta=0.5, softening=1e-3, n_leaf=32, dt=1e-3, hbm_budget=budget)
    big = serve.make_request(0, "plummer", 4000, 2, seed=1)
    with pytest.raises(FmmError) as e:
        srv.submit(big)
    assert e.value.code == FMMH_E_OVER_BUDGET and big.rejected
    small = serve.make_request(1, "hernquist", 1200, 2, seed=2, arrival=0.0)
    big2 = serve.make_request(2, "disk", 5000, 2, seed=3, arrival=0.0)
    srv.run([small, big2], max_seconds=60.0)
    assert small.final is not None and small.steps_done == 2
    assert big2.rejected and big2.final is None


def test_footprint_estimate_grows_with_p():
    from fmms_amd import serve

    assert serve.bytes_per_body(8) == pytest.approx(serve.BYTES_PER_BODY)
    assert serve.bytes_per_body(4) < serve.bytes_per_body(8) < serve.bytes_per_body(12)
    assert math.isfinite(serve.bytes_per_body(40))
