"""Writes tests/golden/golden.json: digests of oracle outputs on fixed seeded
inputs (regression pins for the oracle and the HIP path).

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "necsim-rust_amd")]

import ctypes  # noqa: E402

import numpy as np  # noqa: E402

import necsim_hip as nh  # noqa: E402  (map generator only: pure numpy)
import oracle as O  # noqa: E402


def digest(run):
    ev = run["events"]
    return {"n_events": int(len(ev)), "steps": int(run["steps"]),
            "time_bits": struct.unpack("<Q", struct.pack("<d", run["time"]))[0],
            "paused": bool(run["paused"]), "biodiversity": O.biodiversity(ev),
            "events_sha256": hashlib.sha256(ev.tobytes()).hexdigest(),
            "n_remaining": int(len(run["remaining"])),
            "head": [[int(e["global_reference"]), struct.unpack("<Q", struct.pack("<d", e["prior_time"]))[0],
                      struct.unpack("<Q", struct.pack("<d", e["event_time"]))[0], int(e["x"]), int(e["y"]),
                      int(e["index"])] for e in ev[:4]]}


def cases():
    out = {}
    s1 = O.Scenario.non_spatial(100, 100, 1)
    l1 = s1.origins(1.0)
    out["config1_none"] = digest(O.simulate(s1, O.make_params(42, 1e-3, dedup=None), l1))
    out["config1_relative1"] = digest(O.simulate(s1, O.make_params(42, 1e-3), l1))
    s2 = O.Scenario.almost_infinite_normal(10.0)
    l2 = s2.origins(1.0, circle=(2**31 - 1, 2**31 - 1, 100))
    out["config2_pause40_none"] = digest(O.simulate(s2, O.make_params(7, 1e-4, dedup=None, pause_before=40.0), l2))
    out["config2_pause40_relative1"] = digest(O.simulate(s2, O.make_params(7, 1e-4, pause_before=40.0), l2))
    h, rp, c, w = nh.generate_map(8, 8, seed=5, sigma=1.5, cutoff=2.5)
    s3 = O.Scenario.spatially_explicit(8, 8, h, rp, c, w)
    out["map8_none"] = digest(O.simulate(s3, O.make_params(3, 0.01, dedup=None), s3.origins(1.0)))
    hab, rp32, c32, w32 = nh.generate_map(32, 32, seed=42)
    out["map32_sha256"] = hashlib.sha256(hab.tobytes() + rp32.tobytes() + c32.tobytes() + w32.tobytes()).hexdigest()
    L = O.lib()
    seed = L.orc_seed_from_u64(42)
    kat = []
    for loc, t in [(0, 0), (1, 0), (0, 1), (2**32 + 5, 123456), (2**64 - 1, 2**63)]:
        st = ctypes.c_uint64(L.orc_prime(seed, loc, t))
        u = L.orc_uniform_closed_open(ctypes.byref(st))
        kat.append([loc, t, L.orc_prime(seed, loc, t), struct.unpack("<Q", struct.pack("<d", u))[0]])
    out["wyhash_seed42"] = {"seed": seed, "prime_uniform": kat}
    out["maths"] = {name: [[struct.unpack("<Q", struct.pack("<d", x))[0],
                            struct.unpack("<Q", struct.pack("<d", getattr(L, name)(x)))[0]] for x in xs]
                    for name, xs in [("orc_ln", [0.3, 2.0**-53, 0.999999]), ("orc_exp", [-1.0, 0.5, -700.0]),
                                     ("orc_sin", [-2.0, -3.14159, -6.2]), ("orc_cos", [-2.0, -1.5707963, -4.71])]}
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as f:
        json.dump(cases(), f, indent=1, sort_keys=True)
    print("wrote", path)
