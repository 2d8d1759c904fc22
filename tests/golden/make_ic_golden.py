"""Writes ic_first8.json: the first 8 bodies (pos, vel, mass as float.hex)
of each initial-condition generator at fixed arguments, drawn by the oracle's
restatement (oracle/ic_oracle.cpp). Run from the repository root:

    python tests/golden/make_ic_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

# (generator, n, keyword arguments)
CASES = {
    "plummer": ("ic_plummer", 1000, {"rmax": 20.0, "seed": 42}),
    "hernquist": ("ic_hernquist", 1000, {"a": 0.5, "rmax": 50.0, "seed": 7}),
    "disk": ("ic_disk", 1000, {"rd": 1.0, "zd": 0.05, "rmax": 10.0, "m_central": 0.1,
                               "sigma": 0.02, "seed": 11}),
    "halos_cube": ("ic_halos_cube", 4096, {"nhalo": 64, "halo_frac": 0.9, "a_min": 0.01,
                                           "a_max": 0.1, "mass_decades": 2.0,
                                           "rmax_over_a": 20.0, "seed": 2024}),
}
K = 8


def hexes(a):
    return [float(x).hex() for x in a.reshape(-1)]


def main():
    out = {}
    for name, (fn, n, kw) in CASES.items():
        pos, vel, mass = getattr(O, fn)(n, **kw)
        # halo bodies come first; keep the last K (background) bodies as well
        out[name] = {"fn": fn, "n": n, "kwargs": kw,
                     "pos": hexes(pos[:K]), "vel": hexes(vel[:K]), "mass": hexes(mass[:K]),
                     "pos_last": hexes(pos[-K:]), "vel_last": hexes(vel[-K:])}
    with open(os.path.join(HERE, "ic_first8.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
