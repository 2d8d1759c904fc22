"""GPU test of the allocation-free entry point fmmh_direct_sampled_noalloc:
its results equal fmmh_direct_sampled's bit for bit (both add the same
per-range partial sums in the same order), on two calls in a row, the second
with fewer bodies and targets (the static scratch keeps the first call's
partials, which the second must overwrite); sizes beyond the scratch are
rejected with FMMH_E_PARAM."""
import ctypes as ct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _noalloc(fa, hip, pos, mass, ids, softening):
    n, nt = len(mass), len(ids)
    d_pos, d_mass = hip.DeviceArray.from_host(pos), hip.DeviceArray.from_host(mass)
    d_ids = hip.DeviceArray.from_host(np.ascontiguousarray(ids, np.uint32))
    d_acc, d_pot, d_f = hip.DeviceArray((nt, 3)), hip.DeviceArray((nt,)), hip.DeviceArray((nt,))
    rc = fa.lib().fmmh_direct_sampled_noalloc(n, d_pos.ptr, d_mass.ptr, float(softening), nt,
                                              d_ids.ptr, d_acc.ptr, d_pot.ptr, d_f.ptr, None)
    assert rc == fa.FMMH_OK, rc
    hip.synchronize()
    out = d_acc.download(), d_pot.download(), d_f.download()
    for a in (d_pos, d_mass, d_ids, d_acc, d_pot, d_f):
        a.free()
    return out


@pytest.mark.parametrize("softening", [0.0, 1e-3])
def test_noalloc_direct_equals_direct_sampled(softening):
    import fmms_amd as fa
    from fmms_amd import hip

    cases = [(300000, 3000, 11), (70000, 500, 12)]  # second call: smaller input
    for n, nt, seed in cases:
        pos, mass = fa.plummer(n, seed=seed)
        ids = fa.sample_ids(n, nt, seed)
        say(f"[noalloc] n={n} nt={nt} softening={softening}")
        a, pot, f = _noalloc(fa, hip, pos, mass, ids, softening)
        a2, pot2, f2 = fa.direct_sampled(pos, mass, ids, softening=softening)
        assert np.array_equal(bits(a), bits(a2))
        assert np.array_equal(bits(pot), bits(pot2))
        assert np.array_equal(bits(f), bits(f2))
        assert np.all(np.isfinite(a))


def test_noalloc_direct_rejects_sizes_beyond_scratch():
    import fmms_amd as fa

    lib = fa.lib()
    big = 209715 + 1  # FMMH_DIRECT_NOALLOC_MAX_TARGETS + 1
    dummy = ct.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.fmmh_direct_sampled_noalloc(1000, dummy, dummy, 0.0, big, dummy, dummy, None, None,
                                         None)
    assert rc == fa.FMMH_E_PARAM
    assert lib.fmmh_direct_sampled_noalloc(0, dummy, dummy, 0.0, 10, dummy, dummy, None, None,
                                           None) == fa.FMMH_E_EMPTY
