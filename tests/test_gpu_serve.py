"""GPU tests of the continuous-batching server (fmms_amd.serve): systems that
arrive at different step boundaries, join and leave the batched ensemble,
and queue under the HBM budget end in exactly the state they reach when
stepped alone (bitwise: the ensemble evaluates every system as its own tree
does, and the leapfrog arithmetic is per body)."""
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


@pytest.mark.parametrize("budget_bodies", [None, 5000])
def test_served_trajectories_equal_alone(budget_bodies):
    from fmms_amd import serve

    reqs = _requests(serve)
    kw = dict(p=8, theta=0.5, softening=1e-3, n_leaf=32, dt=1e-3)
    budget = 200e9 if budget_bodies is None else budget_bodies * serve.BYTES_PER_BODY
    srv = serve.Server(hbm_budget=budget, **kw)
    arrive_at = [0, 0, 1, 2, 2, 4, 5]  # step boundary at which each request arrives
    say(f"[serve] {len(reqs)} requests, budget {budget_bodies} bodies")
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
        (x, v, a, pot), _ = serve.run_alone(r, **kw)
        assert r.steps_done == r.steps
        assert np.array_equal(bits(r.final[0]), bits(x)), f"request {r.rid} pos"
        assert np.array_equal(bits(r.final[1]), bits(v)), f"request {r.rid} vel"
        assert np.array_equal(bits(r.final[2]), bits(a)), f"request {r.rid} acc"
    st = srv.stats()
    assert st["steps"] == boundary or st["steps"] <= boundary
