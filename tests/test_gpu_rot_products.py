"""The rotation route's batched swap-matrix products (rotation.rs:125-245) on
the matrix cores agree with the same products on the vector ALUs.

tests/perf/rot_products (built by __graft_entry__.build()) runs both kernels
over the same pairs at expansion order 8 in FP32 and compares a weighted sum
of each pair's coefficients; both apply orthonormal matrices, so the sums
stay O(100) and FP32 rounding leaves them within 1e-4 relative of the
largest. This pins the MFMA operand layout the timing comparison relies on.
"""
import json
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "perf", "rot_products")


@pytest.mark.gpu
def test_mfma_products_match_valu():
    assert os.path.exists(EXE), "tests/perf/rot_products is not built (run __graft_entry__.build())"
    print("[rot products] 65536 pairs, 8 applications", flush=True)
    r = subprocess.run([EXE, "16", "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["pairs"] == 65536 and d["applications"] == 8
    assert d["max_rel_diff"] < 1e-4
    assert d["valu_ms"] > 0 and d["mfma_ms"] > 0
