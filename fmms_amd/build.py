"""Builds libfmmh.so (the gfx950 FMM evaluator) in-tree with hipcc.

Every translation unit is compiled for gfx950 only. Files whose results must
match the reference bit for bit (tree, upward pass, walk, host orchestration
with the initial-condition samplers) are compiled with -ffp-contract=off; the
force kernels (M2L, L2L/L2P, P2P) let the compiler fuse multiply-adds.
Packed FP32 instructions are disabled for all device code.

Usage: python -m fmms_amd.build  (or fmms_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libfmmh.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-command-line-argument",
]

# source -> contraction mode
SOURCES = {
    "prims.hip": "off",
    "tree.hip": "off",
    "upward.hip": "off",
    "upward32.hip": "off",
    "walk.hip": "off",
    "m2l.hip": "fast",
    "m2l_p1_6.hip": "fast",
    "m2l_p7_8.hip": "fast",
    "m2l_p9.hip": "fast",
    "m2l_p10.hip": "fast",
    "downward.hip": "fast",
    "near.hip": "fast",
    "direct.hip": "fast",
    "generic_up.hip": "off",
    "generic_far.hip": "fast",
    "fmmh.hip": "off",
    "ic.hip": "off",
    "integrate.hip": "off",
    "soft.hip": "fast",
}


def _compile(src, mode):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, f"-ffp-contract={mode}", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(SOURCES)))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


# Measurement programs under tests/perf (not part of the library): source ->
# executable next to it.
PROBES = {os.path.join(ROOT, "tests", "perf", "rot_products.hip"):
          os.path.join(ROOT, "tests", "perf", "rot_products")}


def build_probes(verbose=False):
    for src, exe in PROBES.items():
        cmd = [HIPCC, *COMMON, "-ffp-contract=fast", "-o", exe + ".tmp", src]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print("built", exe)


if __name__ == "__main__":
    build(verbose=True)
    build_probes(verbose=True)
    sys.exit(0)
