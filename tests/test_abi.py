"""CPU checks of the C-ABI library: libfmmh.so loads, exports every symbol
include/fmmh.h declares (and the Python mirror binds each), and the
host-only entry points (initial-condition samplers) agree with the oracle's
restatement of models.rs bit for bit. No device compute is called here."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fmmh.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fmmh_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_all_declared_symbols():
    import fmms_amd

    lib = ct.CDLL(fmms_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in fmmh.h but not exported"
    assert set(syms) == set(fmms_amd.SIGNATURES), "Python bindings differ from the header"


def test_exported_symbols_are_c_linkage():
    import fmms_amd

    out = subprocess.run(["nm", "-D", "--defined-only", fmms_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in declared_symbols():
        assert s in names


def test_library_is_gfx950_code_object():
    import fmms_amd

    data = open(fmms_amd.LIB_PATH, "rb").read()
    assert b"__CLANG_OFFLOAD_BUNDLE__" in data
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_plummer_generator_matches_oracle_bitwise():
    import fmms_amd

    for n, seed in [(1000, 42), (257, 7)]:
        p1, m1 = fmms_amd.plummer(n, seed=seed)
        p2, m2 = O.plummer(n, seed=seed)
        assert np.array_equal(p1.view(np.uint64), p2.view(np.uint64))
        assert np.array_equal(m1, m2)
    h1, _ = fmms_amd.hernquist(500, center=(1.0, -2.0, 0.5), seed=3)
    h2, _ = O.hernquist(500, center=(1.0, -2.0, 0.5), seed=3)
    assert np.array_equal(h1.view(np.uint64), h2.view(np.uint64))


def test_sample_ids_matches_oracle():
    import fmms_amd

    assert np.array_equal(fmms_amd.sample_ids(100_000, 300, 7), O.sample_ids(100_000, 300, 7))


def test_default_tuning_is_host_only():
    import fmms_amd

    t = fmms_amd._Tuning()
    fmms_amd.lib().fmmh_default_tuning(ct.byref(t))
    assert t.m2l_chunk == 32 and t.p2p_item_work == 1 << 18


def test_autotune_cache_roundtrip(tmp_path):
    from fmms_amd import autotune

    k1 = autotune.workload_key(1 << 20, "fp32bulk", 0.5, 1e-3, p_grid=(8,), leaf_grid=(64, 128))
    k2 = autotune.workload_key(1 << 20, "fp64", 0.5, 1e-3, p_grid=(8,), leaf_grid=(64, 128))
    k3 = autotune.workload_key(1 << 20, "fp32bulk", 0.5, 1e-4, p_grid=(8,), leaf_grid=(64, 128))
    assert len({k1, k2, k3}) == 3
    cache = str(tmp_path / "tune.json")
    assert autotune.load(k1, cache) is None
    choice = {"p": 8, "n_leaf": 128, "m2l_chunk": 32, "p2p_item_work": 1 << 18, "ms": 1.0}
    autotune._store(k1, choice, cache)
    autotune._store(k2, dict(choice, n_leaf=96), cache)
    assert autotune.load(k1, cache) == choice
    assert autotune.load(k2, cache)["n_leaf"] == 96
