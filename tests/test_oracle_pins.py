"""Pins of the CPU oracle (the restatement of the reference's CPU Independent
algorithm) against published vectors, closed forms, brute force and
invariants.  CPU only."""
import ctypes
import math
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as O

L = O.lib()


def u64(state):
    return ctypes.c_uint64(state)


# ---------------------------------------------------------------------------
# RNG: published known answers
def test_seahash_diffuse_published_vectors():
    # seahash 4.1.0 helper::diffuse unit test vectors (the function restated at
    # necsim/impls/no-std/src/cogs/rng/wyhash.rs:99-119)
    for x, want in [(94203824938, 17289265692384716055), (0xDEADBEEF, 12110756357096144265), (0, 0),
                    (1, 15197155197312260123), (2, 1571904453004118546), (3, 16467633989910088880)]:
        assert L.orc_seahash_diffuse(x) == want


def test_seed_from_u64_value_breakage():
    # rand_core's SeedableRng::seed_from_u64 value-breakage test: seed 0 fills
    # the 8-byte seed to 5029875928683246316 (necsim/core/src/cogs/rng.rs:33-63)
    assert L.orc_seed_from_u64(0) == 5029875928683246316


def test_fnv1a_published_vectors():
    # FNV-1a 64 test vectors (crate fnv 1.0.7, necsim/impls/no-std/src/cache.rs:4)
    for s, want in [(b"", 0xcbf29ce484222325), (b"a", 0xaf63dc4c8601ec8c), (b"foobar", 0x85944171f73967e8)]:
        assert L.orc_fnv1a64(s, len(s)) == want


def test_speciation_sample_hash_byte_order():
    # Hash stream of SpeciationSample: x, y, index (u32 LE), time bits, sample bits (u64 LE)
    u, t, x, y, i = 0.25, 17.5, 3, 4, 5
    b = struct.pack("<IIIQQ", x, y, i, struct.unpack("<Q", struct.pack("<d", t))[0],
                    struct.unpack("<Q", struct.pack("<d", u))[0])
    assert L.orc_speciation_sample_hash(u, t, x, y, i) == L.orc_fnv1a64(b, len(b))


def test_wymum_is_xor_fold_of_128_bit_product():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b = (int(v) for v in rng.integers(0, 2**63, 2, dtype=np.uint64) * 2 + 1)
        r = a * b
        want = (a ^ (r & (2**64 - 1))) ^ (b ^ (r >> 64))
        assert L.orc_wymum(a, b) == want


def test_uniform_closed_forms():
    # U[0,1) = (x >> 11) * 2^-53 and U(0,1] = ((x >> 11) + 1) * 2^-53 (rng.rs:75-93)
    for s in [0, 1, 2**64 - 1, 0x123456789abcdef0]:
        st = u64(s)
        st2 = u64(s)
        x = L.orc_sample_u64(ctypes.byref(st2))
        assert L.orc_uniform_closed_open(ctypes.byref(st)) == (x >> 11) * 2.0**-53
        st = u64(s)
        assert L.orc_uniform_open_closed(ctypes.byref(st)) == ((x >> 11) + 1) * 2.0**-53


def test_uniform_statistics():
    # mean, variance and low-bit balance of U[0,1) over primed streams
    vals = []
    for k in range(20000):
        st = u64(L.orc_prime(12345, k, k * 7))
        vals.append(L.orc_uniform_closed_open(ctypes.byref(st)))
    v = np.array(vals)
    assert abs(v.mean() - 0.5) < 0.01 and abs(v.var() - 1 / 12) < 0.005
    assert v.min() >= 0.0 and v.max() < 1.0


# ---------------------------------------------------------------------------
# MathsCore (crate libm 0.2.8): within 1 ulp of the correctly rounded value
getcontext().prec = 60


def _ulps(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def _round_decimal(d):
    return float(d)  # Decimal -> nearest double


def _dec_sin_cos(x):
    x = Decimal(x)
    s, c = Decimal(0), Decimal(0)
    term = x
    n = 1
    while True:  # sin
        s += term
        term = -term * x * x / ((n + 1) * (n + 2))
        n += 2
        if abs(term) < Decimal(10) ** -55:
            break
    term = Decimal(1)
    n = 0
    while True:  # cos
        c += term
        term = -term * x * x / ((n + 1) * (n + 2))
        n += 2
        if abs(term) < Decimal(10) ** -55:
            break
    return s, c


def test_ln_exp_sin_cos_within_one_ulp_of_exact():
    rng = np.random.default_rng(7)
    xs = ((rng.integers(0, 2**53, 600, dtype=np.uint64) + 1).astype(np.float64) * 2.0**-53)
    for x in xs:
        assert _ulps(L.orc_ln(float(x)), _round_decimal(Decimal(float(x)).ln())) <= 1
    for x in rng.uniform(-30, 30, 300):
        assert _ulps(L.orc_exp(float(x)), _round_decimal(Decimal(float(x)).exp())) <= 1
    for x in -6.283185307179586 * rng.random(300):
        s, c = _dec_sin_cos(float(x))
        assert _ulps(L.orc_sin(float(x)), _round_decimal(s)) <= 1
        assert _ulps(L.orc_cos(float(x)), _round_decimal(c)) <= 1


def test_maths_special_values():
    assert L.orc_ln(1.0) == 0.0
    assert struct.pack("<d", L.orc_exp(-1.0)) == struct.pack("<Q", 0x3FD78B56362CEF38)  # e^-1 rounded
    assert L.orc_exp(0.0) == 1.0
    assert L.orc_sin(0.0) == 0.0 and L.orc_cos(0.0) == 1.0
    for k in range(-60, 60):
        # log(2^k) = k ln2_hi + k ln2_lo exactly as the e_log.c reduction evaluates it
        assert abs(L.orc_ln(2.0**k) - k * math.log(2)) <= abs(k) * 2e-16 + 1e-300
    assert math.isinf(L.orc_ln(0.0)) and L.orc_ln(0.0) < 0
    assert math.isnan(L.orc_ln(-1.0))
    assert L.orc_round(2.5) == 3.0 and L.orc_round(-2.5) == -3.0 and L.orc_round(0.49999999999999994) == 0.0


def test_maths_agree_with_glibc_almost_everywhere():
    rng = np.random.default_rng(9)
    for f, ref, xs in [(L.orc_ln, math.log, rng.random(20000) + 1e-9),
                       (L.orc_sin, math.sin, -6.283185307179586 * rng.random(20000)),
                       (L.orc_cos, math.cos, -6.283185307179586 * rng.random(20000))]:
        d = [_ulps(f(float(x)), ref(float(x))) for x in xs]
        assert max(d) <= 1 and sum(1 for v in d if v) < 0.15 * len(d)


# ---------------------------------------------------------------------------
# alias method (necsim/impls/no-std/src/alias/packed.rs)
def _alias(weights):
    n = len(weights)
    ev = np.arange(n, dtype=np.uint64) * 3 + 1
    w = np.asarray(weights, dtype=np.float64)
    u = np.zeros(n)
    e = np.zeros(n, np.uint64)
    k = np.zeros(n, np.uint64)
    L.orc_alias_create(ev.ctypes.data, w.ctypes.data, n, u.ctypes.data, e.ctypes.data, k.ctypes.data)
    return ev, w, u, e, k


@pytest.mark.parametrize("seed", range(12))
def test_alias_tables_reproduce_weights(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 9))
    weights = rng.random(n) * rng.choice([1.0, 1e-3, 7.0], n)
    ev, w, u, e, k = _alias(weights)
    assert list(e) == list(ev)  # stores exactly the input events (packed.rs:48-51)
    assert all((uu != 1.0) or (ee == kk) for uu, ee, kk in zip(u, e, k))  # full buckets (:52-58)
    # exact probability mass: bucket i contributes u_i/n to e_i and (1-u_i)/n to k_i
    mass = {int(x): 0.0 for x in ev}
    for uu, ee, kk in zip(u, e, k):
        mass[int(ee)] += uu / n
        mass[int(kk)] += (1 - uu) / n
    for x, ww in zip(ev, w):
        assert abs(mass[int(x)] - ww / w.sum()) < 1e-12
    # sampling on a fine grid of U realises the same masses
    grid = 4096 * n
    counts = {int(x): 0 for x in ev}
    for j in range(grid):
        counts[int(L.orc_alias_sample(u.ctypes.data, e.ctypes.data, k.ctypes.data, n, 1.0, j / grid))] += 1
    for x, ww in zip(ev, w):
        assert abs(counts[int(x)] / grid - ww / w.sum()) <= 1.0 / 4096 + 1e-12


def test_separable_alias_self_dispersal_branches():
    # an 8x8 map exercises both separable branches (self included / excluded)
    import necsim_hip as nh
    h, rp, c, w = nh.generate_map(8, 8, seed=5, sigma=1.5, cutoff=2.5)
    s = O.Scenario.spatially_explicit(8, 8, h, rp, c, w)
    t = s.tables()
    included = excluded = 0
    for cell in range(64):
        if h[cell] == 0:
            continue
        a, b = int(t["range_start"][cell]), int(t["range_end"][cell])
        if t["self_event"][cell] != cell:
            included += 1
            # self-dispersal sits in the last atom's k slot (mod.rs:150-168)
            assert t["atom_k"][b - 1] == cell and t["atom_e"][b - 1] != cell
        else:
            excluded += 1
            assert cell not in set(t["atom_e"][a:b]) | set(t["atom_k"][a:b])
        # empirical self-dispersal frequency matches self_p
        row = slice(int(rp[cell]), int(rp[cell + 1]))
        wt = w[row] * h[c[row]]
        p_self = wt[c[row] == cell].sum() / wt.sum()
        assert abs(t["self_p"][cell] - p_self) < 1e-12
        hits = 0
        trials = 4000
        for j in range(trials):
            st = u64(L.orc_prime(99, cell, j))
            tx, ty = ctypes.c_uint32(), ctypes.c_uint32()
            L.orc_sample_dispersal(s.handle, cell % 8, cell // 8, ctypes.byref(st), ctypes.byref(tx), ctypes.byref(ty))
            hits += (ty.value * 8 + tx.value) == cell
        assert abs(hits / trials - p_self) < 4 * math.sqrt(p_self * (1 - p_self) / trials) + 1e-9
    assert included > 0 and excluded > 0


# ---------------------------------------------------------------------------
# Poisson event-time sampler (event_time_sampler/poisson.rs:30-109)
def test_poisson_sampler_contract_and_rate():
    gaps = []
    for k in range(20000):
        t = (k % 97) * 0.731
        st = ctypes.c_uint64()
        nxt = L.orc_poisson_next_event_time(77, k * 13 + 5, t, 0.5, 2.0, ctypes.byref(st))
        assert nxt > t  # strictly after (sampler.rs:100 then requires > before)
        assert math.floor(nxt / 2.0) >= math.floor(t / 2.0)
        gaps.append(nxt - t)
    # memoryless Poisson process with rate lambda = 0.5: mean gap 1/lambda
    assert abs(np.mean(gaps) - 2.0) < 0.06


def test_poisson_sampler_is_a_pure_function_of_location_and_time():
    a = L.orc_poisson_next_event_time(5, 42, 10.0, 0.5, 2.0, None)
    b = L.orc_poisson_next_event_time(5, 42, 10.0, 0.5, 2.0, None)
    c = L.orc_poisson_next_event_time(5, 42, a - 1e-9, 0.5, 2.0, None)
    assert a == b == c  # the next event after any time in [10, a) is a


# ---------------------------------------------------------------------------
# origin samplers / landscape extent order (necsim/core/src/landscape/extent.rs:77-100)
def test_circle_origins_wrap_around_extent():
    s = O.Scenario.almost_infinite_normal(1.0)
    out = s.origins(1.0, circle=(0, 1, 2))
    # brute force: the 5x5 box around (0, 1) row-major with wrapping coords
    want = []
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if dx * dx + dy * dy <= 4:
                want.append(((0 + dx) % 2**32, (1 + dy) % 2**32))
    assert [(int(o["x"]), int(o["y"])) for o in out] == want
    assert all(int(o["global_reference"]) == (int(o["y"]) << 32) | int(o["x"]) for o in out)


def test_non_spatial_origins_and_partition():
    s = O.Scenario.non_spatial(7, 5, 3)
    full = s.origins(1.0)
    assert len(full) == 105
    assert [int(v) for v in full["global_reference"]] == list(range(105))
    parts = [s.origins(1.0, r, 4) for r in range(4)]
    merged = np.sort(np.concatenate(parts), order="global_reference")
    assert merged.tobytes() == full.tobytes()
    sub = s.origins(0.3)
    assert 15 < len(sub) < 50 and len(set(sub["global_reference"].tolist())) == len(sub)


# ---------------------------------------------------------------------------
# whole-algorithm invariants (config 1 at reduced size)
@pytest.fixture(scope="module")
def small_runs():
    s = O.Scenario.non_spatial(40, 40, 1)
    lin = s.origins(1.0)
    runs = {}
    for d in [None, ("relative", 1.0), ("relative", 0.1), ("absolute", 7)]:
        runs[d] = O.simulate(s, O.make_params(42, 1e-2, dedup=d), lin)
    return s, lin, runs


def test_dedup_preserves_species_and_unique_events(small_runs):
    s, lin, runs = small_runs
    base = runs[None]
    assert len(base["events"]) == len(lin)  # every lineage speciates exactly once
    keys = lambda ev: set(zip(ev["event_time"].tolist(), ev["x"].tolist(), ev["y"].tolist(), ev["index"].tolist()))
    full = set(e.tobytes() for e in base["events"])
    for d, r in runs.items():
        assert O.biodiversity(r["events"]) == O.biodiversity(base["events"])
        assert keys(r["events"]) == keys(base["events"])
        assert all(e.tobytes() in full for e in r["events"])
        assert r["steps"] <= base["steps"]


def test_stream_is_sorted_by_packed_event_order(small_runs):
    _, _, runs = small_runs
    ev = runs[None]["events"]
    keys = list(zip(ev["event_time"], ev["x"], ev["y"], ev["index"], ev["prior_time"], ev["global_reference"]))
    assert keys == sorted(keys)


def test_pause_resume_continuity_oracle(small_runs):
    s, lin, runs = small_runs
    p1 = O.simulate(s, O.make_params(42, 1e-2, dedup=None, pause_before=150.0), lin)
    assert p1["paused"] and all(p1["events"]["event_time"] < 150.0)
    p2 = O.simulate(s, O.make_params(42, 1e-2, dedup=None, min_event_time=150.0), p1["remaining"])
    both = np.sort(np.concatenate([p1["events"], p2["events"]]),
                   order=["event_time", "x", "y", "index", "prior_time", "global_reference"])
    assert both.tobytes() == runs[None]["events"].tobytes()
    assert p1["steps"] + p2["steps"] == runs[None]["steps"]


def test_water_level_slicing_does_not_change_dedup_none_output(small_runs):
    s, lin, runs = small_runs
    for step_slice, es in [(1, ("relative", 0.5)), (1000, ("absolute", 3)), (3, ("relative", 50.0))]:
        r = O.simulate(s, O.make_params(42, 1e-2, dedup=None, step_slice=step_slice, event_slice=es), lin)
        assert r["events"].tobytes() == runs[None]["events"].tobytes()
        assert r["steps"] == runs[None]["steps"] and r["time"] == runs[None]["time"]


def test_partitioned_baseline_union_matches_monolithic(small_runs):
    s, lin, runs = small_runs
    r = O.simulate_partitioned(s, O.make_params(42, 1e-2, dedup=None), 1.0, 4)
    assert r["steps"] == runs[None]["steps"] and r["events"] == len(lin)


# ---------------------------------------------------------------------------
# dispersal reporting (R::ReportDispersal): every event reported as a PackedEvent
def _sign(x):
    return (np.ascontiguousarray(x, dtype=np.float64).view(np.uint64) >> np.uint64(63)).astype(bool)


def _speciations_of(packed):
    """the speciation PackedEvents (prior(-), event(+), event.rs:103-115) as SpeciationEvents"""
    sp = packed[_sign(packed["prior_time"]) & ~_sign(packed["event_time"])]
    out = np.zeros(len(sp), O.EVENT_DTYPE)
    out["global_reference"] = sp["global_reference"]
    out["prior_time"] = np.abs(sp["prior_time"])
    out["event_time"] = sp["event_time"]
    out["x"], out["y"], out["index"] = sp["ox"], sp["oy"], sp["oindex"]
    return out


@pytest.fixture(scope="module")
def dispersal_runs():
    s = O.Scenario.spatially_explicit(*_small_map(12, 12, seed=7))
    lin = s.origins(1.0)
    p = dict(dedup=None)
    full = O.simulate_packed(s, O.make_params(3, 2e-2, report_dispersal=True, **p), lin)
    spec = O.simulate(s, O.make_params(3, 2e-2, **p), lin)
    return s, lin, full, spec


def _small_map(w, h, seed):
    import necsim_hip as nh
    return (w, h) + tuple(nh.generate_map(w, h, seed=seed, sigma=1.5, cutoff=3.0))


def test_dispersal_reporting_keeps_the_speciation_stream(dispersal_runs):
    """the reported events below each water level do not depend on how the
    level advances (slicing independence), so the speciations of a
    dispersal-reporting run are the speciation-only stream"""
    s, lin, full, spec = dispersal_runs
    assert _speciations_of(full["events"]).tobytes() == spec["events"].tobytes()
    assert full["steps"] == spec["steps"] and full["time"] == spec["time"]
    assert len(full["events"]) == full["steps"]  # one reported event per step


def test_dispersal_events_are_maybe_encoded_and_chain_per_lineage(dispersal_runs):
    s, lin, full, spec = dispersal_runs
    ev = full["events"]
    disp = ~_sign(ev["prior_time"]) & _sign(ev["event_time"])  # LineageInteraction::Maybe
    spec_mask = _sign(ev["prior_time"]) & ~_sign(ev["event_time"])
    assert np.all(disp | spec_mask) and np.all(ev["coalescence"] == ev["global_reference"])
    assert np.all(ev["ox"][spec_mask] == ev["tx"][spec_mask])
    start = {int(l["global_reference"]): (float(l["last_event_time"]), int(l["x"]), int(l["y"]),
                                          int(l["index"])) for l in lin}
    by_ref = {}
    for e in ev:  # the stream is time-sorted, so each lineage's events come in order
        by_ref.setdefault(int(e["global_reference"]), []).append(e)
    for ref, chain in by_ref.items():
        t, x, y, i = start[ref]
        for k, e in enumerate(chain):
            assert abs(float(e["prior_time"])) == t and (int(e["ox"]), int(e["oy"]), int(e["oindex"])) == (x, y, i)
            t, x, y, i = abs(float(e["event_time"])), int(e["tx"]), int(e["ty"]), int(e["tindex"])
            # a full run ends each chain with its only speciation
            assert (k == len(chain) - 1) == bool(_sign(e["prior_time"]))
    assert len(by_ref) == len(lin)


def test_packed_order_helper_matches_the_oracle_sort(dispersal_runs):
    """nh.packed_order (the reference of the device PackedEvent sort tests)
    reproduces the oracle's PackedEvent::cmp sort of a shuffled stream"""
    import necsim_hip as nh
    s, lin, full, spec = dispersal_runs
    ev = np.ascontiguousarray(full["events"]).view(nh.PACKED_EVENT_DTYPE)
    shuffled = ev[np.random.default_rng(0).permutation(len(ev))]
    assert shuffled[nh.packed_order(shuffled)].tobytes() == ev.tobytes()


def test_dispersal_stream_is_sorted_by_packed_event_order(dispersal_runs):
    s, lin, full, spec = dispersal_runs
    ev = full["events"]
    keys = list(zip(np.abs(ev["event_time"]), ev["ox"], ev["oy"], ev["oindex"], ev["tx"], ev["ty"], ev["tindex"],
                    np.abs(ev["prior_time"]), ev["global_reference"]))
    assert keys == sorted(keys)


def test_dispersal_reporting_pause_resume(dispersal_runs):
    s, lin, full, spec = dispersal_runs
    p1 = O.simulate_packed(s, O.make_params(3, 2e-2, dedup=None, report_dispersal=True, pause_before=40.0), lin)
    p2 = O.simulate_packed(s, O.make_params(3, 2e-2, dedup=None, report_dispersal=True, min_event_time=40.0),
                           p1["remaining"])
    both = np.concatenate([p1["events"], p2["events"]])
    assert both.tobytes() == full["events"].tobytes()
    assert p1["steps"] + p2["steps"] == full["steps"]


# ---------------------------------------------------------------------------
# MathsCore bits against an independent fdlibm port: V8's Math.* (node).
# The libm crate (0.2.8) ports musl, which took these functions from FreeBSD
# msun; V8 ports fdlibm 5.3, the same algorithms with the same constants and
# the same pi/2 reduction but, for log and the sin/cos kernels, an older
# evaluation order.  exp has one evaluation order in both, so the oracle must
# equal V8 bit for bit.  For log / sin / cos the oracle carries test-only
# twins in fdlibm's order built from the oracle's own constants and reduction
# (orc_fdlibm_*): they must equal V8 bit for bit, which pins every constant
# and the reduction, and the restatement must stay within 1 ulp of its twin.
import hashlib  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402
import shutil  # noqa: E402
import sys  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_v8_maths as V8  # noqa: E402

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v8_maths.json")) as _f:
    V8_FIXTURE = json.load(_f)
V8_SETS = {name: (fn, x) for name, fn, x in V8.v8_inputs()}
# oracle op, and the twin op that V8 must match (exp: the oracle itself)
OPS = {"log": (0, 4), "exp": (1, 1), "sin": (2, 5), "cos": (3, 6)}


@pytest.mark.parametrize("name", sorted(V8_SETS))
def test_maths_twin_equals_v8_fdlibm(name):
    fn, x = V8_SETS[name]
    fx = V8_FIXTURE["sets"][name]
    assert fx["fn"] == fn and V8_FIXTURE["n"] == len(x)
    op, twin = OPS[fn]
    y = O.maths_vec(twin, x)
    head = np.array(fx["first_bits"], dtype=np.uint64)
    bad = np.nonzero(y[:len(head)].view(np.uint64) != head)[0]
    assert len(bad) == 0, f"{name}: x={x[bad[:3]]}"
    assert hashlib.sha256(y.astype("<f8").tobytes()).hexdigest() == fx["sha256"], name
    if shutil.which("node"):  # the live engine, when present, agrees with the fixture
        assert V8.digest(V8.run_v8(fn, x[:20000])) == V8.digest(y[:20000])


@pytest.mark.parametrize("fn", ["log", "sin", "cos"])
def test_maths_restatement_within_one_ulp_of_pinned_twin(fn):
    op, twin = OPS[fn]
    for name, (f, x) in V8_SETS.items():
        if f != fn:
            continue
        a = O.maths_vec(op, x).view(np.int64)
        b = O.maths_vec(twin, x).view(np.int64)
        d = np.abs(a - b)
        assert d.max() <= 1, name
        # the orders differ on a few percent of inputs at most
        assert (d != 0).mean() < 0.05, name


# ---------------------------------------------------------------------------
# WyHash composition (wyhash.rs:53-80): closed forms that follow from the
# reference's composition and fail for its plausible mis-restatements (P0 and
# P2 swapped or paired with the wrong operand, a rotation dropped, the state
# transition or output operands changed).
M_SD = 0x6EED0E9DA4D94A4F
MASK = 2**64 - 1
P0, P1, P2, P5 = 0xA0761D6478BD642F, 0xE7037ED1A0B428DB, 0x8EBC6AF09C88C6E3, 0xEB44ACCAB455D165


def rot32(x):
    return ((x << 32) | (x >> 32)) & MASK


def seahash_inverse(y):
    # seahash_diffuse is x*M, x ^= (x >> 32) >> (x >> 60), x*M: each step inverts
    m_inv = pow(M_SD, -1, 2**64)
    x = (y * m_inv) & MASK
    x ^= (x >> 32) >> (x >> 60)  # the shifted part leaves bits 32..63 untouched
    return (x * m_inv) & MASK


def test_seahash_inverse_roundtrip():
    rng = np.random.default_rng(5)
    for v in rng.integers(0, 2**63, 2000, dtype=np.uint64):
        v = int(v) * 2 + 1
        assert seahash_inverse(L.orc_seahash_diffuse(v)) == v


def test_prime_with_both_operands_zero_gives_16_xor_p5():
    # seed = P0 ^ rot32(sd(loc)) zeroes the first wymum operand and
    # time = sd^-1(rot32(P2)) the second; wymum(0, 0) = 0, so the primed state
    # is wymum(0, 16 ^ P5) = 16 ^ P5, for every location
    rng = np.random.default_rng(6)
    t = seahash_inverse(rot32(P2))
    for loc in rng.integers(0, 2**63, 500, dtype=np.uint64):
        loc = int(loc)
        seed = P0 ^ rot32(L.orc_seahash_diffuse(loc))
        assert L.orc_prime(seed, loc, t) == 16 ^ P5


def test_prime_with_first_operand_zero_ignores_location():
    # with the first operand zero, wymum(0, b) = b: the state depends on the
    # time index alone
    rng = np.random.default_rng(7)
    for t in rng.integers(0, 2**63, 50, dtype=np.uint64):
        t = int(t)
        want = None
        for loc in rng.integers(0, 2**63, 20, dtype=np.uint64):
            loc = int(loc)
            s = L.orc_prime(P0 ^ rot32(L.orc_seahash_diffuse(loc)), loc, t)
            want = s if want is None else want
            assert s == want
        assert want == L.orc_wymum(rot32(L.orc_seahash_diffuse(t)) ^ P2, 16 ^ P5)


def test_prime_seed_enters_as_xor_of_rotated_location():
    # prime(seed ^ X, loc) == prime(seed, loc') where rot32(sd(loc')) =
    # rot32(sd(loc)) ^ X: the seed is XORed into the rotated diffused location
    rng = np.random.default_rng(8)
    for seed, loc, t, x in rng.integers(0, 2**63, (300, 4), dtype=np.uint64):
        seed, loc, t, x = int(seed), int(loc), int(t), int(x) | (1 << 40)
        loc2 = seahash_inverse(rot32(rot32(L.orc_seahash_diffuse(loc)) ^ x))
        assert L.orc_prime(seed ^ x, loc, t) == L.orc_prime(seed, loc2, t)


def test_sample_u64_with_zero_first_operand():
    # state = P1 - P0 steps to P1, so the output is sd(wymum(0, P1)) = sd(P1)
    st = u64((P1 - P0) & MASK)
    assert L.orc_sample_u64(ctypes.byref(st)) == L.orc_seahash_diffuse(P1)
    assert st.value == P1
    # and U[0,1) of it is (sd(P1) >> 11) * 2^-53
    st = u64((P1 - P0) & MASK)
    assert L.orc_uniform_closed_open(ctypes.byref(st)) == (L.orc_seahash_diffuse(P1) >> 11) * 2.0**-53
