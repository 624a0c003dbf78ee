"""Writes tests/golden/v8_maths.json: V8's Math.log / Math.exp / Math.sin /
Math.cos (V8's ieee754 port of fdlibm 5.3, evaluated by node through
v8_maths.js) on deterministic input sets.  For each set it records the
SHA-256 of all output bits and the first outputs in full.

The inputs come from a SplitMix64 stream (v8_inputs below), so the test that
reads the fixture rebuilds them without any generator version dependence.
Run from the repository root:  python tests/golden/make_v8_maths.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N = 200_000
FULL = 512  # outputs kept in full per set


def splitmix64(seed, n):
    """n outputs of SplitMix64 from `seed` (uint64 array)"""
    out = np.empty(n, np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed)
        g = np.uint64(0x9E3779B97F4A7C15)
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = x + g * idx
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        out[:] = z ^ (z >> np.uint64(31))
    return out


def unit(seed, n):
    """U[0,1) = (x >> 11) * 2^-53"""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0**-53


# (name, V8 function, inputs): the ranges the Independent algorithm draws
# (Box-Muller u0 in (0,1], angles -TAU*u1 in (-2pi, 0], Poisson exp(-lambda*dt))
# and wide ranges that reach every branch of the reductions
def v8_inputs():
    return [
        ("log_unit", "log", unit(1, N) + 2.0**-53),
        ("log_wide", "log", np.exp(-700.0 + 1400.0 * unit(2, N))),
        ("exp", "exp", -745.0 + 1454.0 * unit(3, N)),
        ("exp_small", "exp", -40.0 + 80.0 * unit(4, N)),
        ("sin_angle", "sin", -6.283185307179586 * unit(5, N)),
        ("cos_angle", "cos", -6.283185307179586 * unit(6, N)),
        ("sin_wide", "sin", -1e5 + 2e5 * unit(7, N)),
        ("cos_wide", "cos", -1e5 + 2e5 * unit(8, N)),
    ]


def run_v8(fn, x):
    js = os.path.join(HERE, "v8_maths.js")
    with tempfile.TemporaryDirectory() as d:
        i, o = os.path.join(d, "in.f64"), os.path.join(d, "out.f64")
        np.ascontiguousarray(x, dtype="<f8").tofile(i)
        subprocess.run(["node", js, fn, i, o], check=True)
        return np.fromfile(o, "<f8")


def digest(y):
    return hashlib.sha256(np.ascontiguousarray(y, dtype="<f8").tobytes()).hexdigest()


def main():
    ver = subprocess.run(["node", "--version"], capture_output=True, text=True, check=True).stdout.strip()
    out = {"node": ver, "n": N, "sets": {}}
    for name, fn, x in v8_inputs():
        y = run_v8(fn, x)
        out["sets"][name] = {"fn": fn, "sha256": digest(y),
                             "first_bits": [int(v) for v in y[:FULL].view(np.uint64)]}
    with open(os.path.join(HERE, "v8_maths.json"), "w") as f:
        json.dump(out, f)
    print(f"wrote v8_maths.json ({ver}, {len(out['sets'])} sets of {N})", file=sys.stderr)


if __name__ == "__main__":
    main()
