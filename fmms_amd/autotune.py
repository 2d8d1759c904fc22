"""Measured choice of an evaluation's shape: expansion order p, leaf capacity
n_leaf and the tile shapes of fmmh_tuning (M2L chunk length, P2P work-item
bound), under a force-error bar.

The reference fixes these by hand: its GPU preset (p = 4, n_leaf = 64,
crates/visual/src/gpufmm.rs:55-58) and its P2P split constants (FAT_THR /
FAT_CHUNK, gpufmm.rs:686-689). Here they are measured on the workload itself:

  1. every (p, n_leaf) of the grid with the default tiles: the time of one
     full evaluation (tree build included; best of `repeats` after a warm-up)
     and the relative force error against the FP64 direct sum on sampled
     targets; the fastest pair whose 99th percentile meets `max_rel_err` wins
     (the most accurate one if none does);
  2. for that pair, every (m2l_chunk, p2p_item_work) of the tile grid (the
     tiles change the work decomposition, not the approximation), fastest
     wins.

The choice is returned and, with `cache`, stored in a JSON file under a key
of the workload's parameters, so later runs can `load` it instead of
measuring again.
"""
import json
import math
import os
import time

import numpy as np

from . import Fmm, direct_sampled, hip, rel_errors, sample_ids

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "fmms_amd", "tuning.json")


def workload_key(n, precision, theta, softening, soft_mac=0.0, eps=math.inf, p_grid=(8,),
                 leaf_grid=(), max_rel_err=1e-5):
    """Cache key of a workload: its size and every parameter the choice
    depends on."""
    return (f"{precision}|n{n}|theta{theta:g}|soft{softening:g}|tau{soft_mac:g}|eps{eps:g}|"
            f"p{','.join(map(str, p_grid))}|leaf{','.join(map(str, leaf_grid))}|bar{max_rel_err:g}")


def load(key, cache=DEFAULT_CACHE):
    """A stored choice, or None."""
    try:
        with open(cache) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def _store(key, result, cache):
    os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
    try:
        with open(cache) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[key] = result
    tmp = cache + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, cache)


def _time(ctx, n, dev, repeats):
    d_pos, d_mass, d_acc, d_pot = dev
    ctx.accel_device(n, d_pos.ptr, d_mass.ptr, d_acc.ptr, d_pot.ptr)  # warm-up (grows buffers)
    best = math.inf
    for _ in range(repeats):
        hip.synchronize()
        t = time.perf_counter()
        ctx.accel_device(n, d_pos.ptr, d_mass.ptr, d_acc.ptr, d_pot.ptr)
        hip.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def tune(pos=None, mass=None, *, p_grid=(8,), leaf_grid=(32, 64, 96, 128, 192),
         chunk_grid=(16, 32, 64), item_grid=(1 << 16, 1 << 18, 1 << 20), precision="fp32bulk",
         theta=0.5, softening=1e-3, soft_mac=0.0, eps=math.inf, max_rel_err=1e-5, samples=4096,
         repeats=2, cache=None, device=None, reference=None, log=None):
    """Fastest (p, n_leaf, m2l_chunk, p2p_item_work) meeting the error bar.

    pos [n, 3] and mass [n] are host arrays; `device` = (d_pos, d_mass,
    d_acc, d_pot) DeviceArrays of the same bodies skips the upload, and
    `reference` = (ids, a_ref) skips the direct sum. Returns a dict with the
    choice, its time and error, and every trial."""
    if device is None:
        mass = np.ascontiguousarray(mass, np.float64)
        n = mass.shape[0]
        device = (hip.DeviceArray.from_host(np.ascontiguousarray(pos, np.float64).reshape(n, 3)),
                  hip.DeviceArray.from_host(mass), hip.DeviceArray((n, 3)), hip.DeviceArray((n,)))
    n = device[1].shape[0]
    if reference is None:
        ids = sample_ids(n, samples, 7)
        a_ref, _, _ = direct_sampled(pos, mass, ids, softening=softening)
    else:
        ids, a_ref = reference
    say = log or (lambda *a: None)
    mk = lambda p, nl, **t: Fmm(None, None, p=p, eps=eps, n_leaf=nl, theta_max=theta,
                                softening=softening, precision=precision, soft_mac=soft_mac, **t)
    trials, best = [], None
    for p in p_grid:
        for nl in leaf_grid:
            ctx = mk(p, nl)
            el = _time(ctx, n, device, repeats)
            ctx.close()
            acc = device[2].download()
            p99 = float(np.percentile(rel_errors(acc[ids], a_ref), 99))
            ok = p99 <= max_rel_err
            trials.append({"p": p, "n_leaf": nl, "ms": 1e3 * el, "rel_p99": p99, "valid": ok})
            say(f"[autotune] {precision} p={p} n_leaf={nl}: {1e3 * el:.2f} ms, rel p99 {p99:.2e}"
                f"{'' if ok else ' (above the bar)'}")
            if ok and (best is None or el < best[2]):
                best = (p, nl, el, p99)
    valid = best is not None
    if best is None:  # nothing meets the bar: the most accurate candidate
        t = min(trials, key=lambda r: r["rel_p99"])
        best = (t["p"], t["n_leaf"], t["ms"] * 1e-3, t["rel_p99"])
    p, nl = best[0], best[1]
    tiles, tbest = [], None
    for ch in chunk_grid:
        for iw in item_grid:
            ctx = mk(p, nl, m2l_chunk=ch, p2p_item_work=iw)
            el = _time(ctx, n, device, repeats)
            ctx.close()
            tiles.append({"m2l_chunk": ch, "p2p_item_work": iw, "ms": 1e3 * el})
            say(f"[autotune] {precision} p={p} n_leaf={nl} m2l_chunk={ch} p2p_item_work={iw}: "
                f"{1e3 * el:.2f} ms")
            if tbest is None or el < tbest[2]:
                tbest = (ch, iw, el)
    result = {"p": p, "n_leaf": nl, "m2l_chunk": tbest[0], "p2p_item_work": tbest[1],
              "ms": 1e3 * tbest[2], "rel_p99": best[3], "valid": valid, "precision": precision,
              "trials": trials, "tile_trials": tiles}
    if cache:
        _store(workload_key(n, precision, theta, softening, soft_mac, eps, p_grid, leaf_grid,
                            max_rel_err), result, cache)
    return result


def make_fmm(choice, *, theta=0.5, softening=1e-3, soft_mac=0.0, eps=math.inf, **kw):
    """An Fmm context with a tuned choice."""
    return Fmm(None, None, p=choice["p"], eps=eps, n_leaf=choice["n_leaf"], theta_max=theta,
               softening=softening, precision=choice.get("precision", "fp64"), soft_mac=soft_mac,
               m2l_chunk=choice["m2l_chunk"], p2p_item_work=choice["p2p_item_work"], **kw)
