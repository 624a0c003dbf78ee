"""GPU parity: the HIP Independent algorithm against the CPU oracle.

Every test runs the product through the C ABI (libnecsim_hip.so) on cuda:0 /
the first HIP device and compares with the oracle on the same seeded inputs.
"""
import ctypes
import struct

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from necsim_hip import _ffi
from parity_utils import (as_bytes, assert_exact, assert_invariants, hip_run, map_scenarios, oracle_run,
                          unique_keys)

pytestmark = pytest.mark.gpu
PLAIN, PERSISTENT, TASK_LIST = nh.PLAIN, nh.PERSISTENT, nh.TASK_LIST


def announce(msg):
    print(f"[case] {msg}", flush=True)


def device_eval(ctx, op, words, n):
    words = np.ascontiguousarray(words, dtype=np.uint64)
    out = np.zeros(n, np.uint64)
    _ffi.check(_ffi.lib().nh_debug_eval(ctx.handle, op, _ffi.ptr(words), _ffi.ptr(out), n))
    return out


# ---------------------------------------------------------------------------
# device MathsCore and RNG primitives, bit-exact against the oracle
def _inputs(rng, kind, n):
    if kind == "unit_open":  # Box-Muller u0 in (0, 1]
        return ((rng.integers(0, 2**53, n, dtype=np.uint64) + 1).astype(np.float64) * 2.0**-53)
    if kind == "angle":  # -TAU * u1
        return -6.283185307179586 * (rng.integers(0, 2**53, n, dtype=np.uint64).astype(np.float64) * 2.0**-53)
    if kind == "wide":
        return np.exp(rng.uniform(-700, 700, n))
    if kind == "angle_wide":  # every case of the pi/2 reduction, and the ulps around k pi/4
        k = np.arange(1, 41)
        near = ((k * (np.pi / 4)).view(np.int64)[:, None] + np.arange(-300, 301)).ravel().view(np.float64)
        return np.concatenate([rng.uniform(-1e5, 1e5, n // 4), rng.uniform(-40, 40, n // 2), near, -near])
    if kind == "exp":
        return rng.uniform(-740, 705, n)
    if kind == "sqrt":
        return -2.0 * np.log(((rng.integers(0, 2**53, n, dtype=np.uint64) + 1).astype(np.float64) * 2.0**-53))
    if kind == "round":
        return rng.normal(0, 50, n) * rng.choice([1.0, 1e-3, 1e3], n)
    raise ValueError(kind)


SPECIALS = [1.0, 2.0, 0.5, 2.0**-53, 1.0 - 2.0**-53, np.pi / 2, -np.pi / 2, -np.pi, -3 * np.pi / 2,
            -2 * np.pi, -np.pi / 4, -6.283185307179586, 1e-300, 5e-324, 0.0, -0.0]


@pytest.mark.parametrize("op,name,kind", [(0, "orc_ln", "unit_open"), (0, "orc_ln", "wide"),
                                          (1, "orc_exp", "exp"), (2, "orc_sin", "angle"),
                                          (3, "orc_cos", "angle"), (2, "orc_sin", "angle_wide"),
                                          (3, "orc_cos", "angle_wide"), (4, "orc_sqrt", "sqrt"),
                                          (5, "orc_floor", "round"), (6, "orc_round", "round")])
def test_device_maths_core(ctx, op, name, kind):
    announce(f"device maths op {op} {name} {kind}")
    rng = np.random.default_rng(op * 7 + len(kind))
    x = np.concatenate([_inputs(rng, kind, 200_000), np.array(SPECIALS)])
    if op == 0:
        x = x[x > 0]
    got = device_eval(ctx, op, x.view(np.uint64), len(x))
    f = getattr(O.lib(), name)
    want = np.array([f(float(v)) for v in x]).view(np.uint64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{name}: {len(bad)} mismatches, e.g. x={x[bad[:3]]}"


def test_device_division_correctly_rounded(ctx):
    announce("device f64 division")
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.uniform(-2, 2, 200_000), rng.normal(0, 1, 1000) * 1e300])
    b = np.concatenate([2.0 + rng.uniform(-0.3, 0.5, 200_000), rng.normal(0, 1, 1000) * 1e-300])
    pairs = np.stack([a, b], axis=1).ravel()
    got = device_eval(ctx, 7, pairs.view(np.uint64), len(a))
    want = (a / b).view(np.uint64)  # numpy f64 division is IEEE correctly rounded
    assert np.array_equal(got, want)


def test_device_rng(ctx):
    announce("device wymum / seahash / prime / uniform")
    rng = np.random.default_rng(11)
    a = rng.integers(0, 2**63, 50_000, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 50_000, dtype=np.uint64)
    b = rng.integers(0, 2**63, 50_000, dtype=np.uint64) * np.uint64(3)
    got = device_eval(ctx, 8, np.stack([a, b], 1).ravel(), len(a))
    want = np.array([O.lib().orc_wymum(int(x), int(y)) for x, y in zip(a, b)], dtype=np.uint64)
    assert np.array_equal(got, want)
    got = device_eval(ctx, 9, a, len(a))
    want = np.array([O.lib().orc_seahash_diffuse(int(x)) for x in a], dtype=np.uint64)
    assert np.array_equal(got, want)
    triples = np.stack([a, b, a ^ b], 1)
    got = device_eval(ctx, 10, triples.ravel(), len(a))
    want = np.array([O.lib().orc_prime(int(s), int(l), int(t)) for s, l, t in triples], dtype=np.uint64)
    assert np.array_equal(got, want)
    got = device_eval(ctx, 11, triples.ravel(), len(a))
    st = ctypes.c_uint64()
    want = []
    for s, l, t in triples:
        st.value = O.lib().orc_prime(int(s), int(l), int(t))
        want.append(O.lib().orc_uniform_closed_open(ctypes.byref(st)))
    assert np.array_equal(got, np.array(want).view(np.uint64))


@pytest.mark.parametrize("delta_t", [2.0, 3.0, 40.0, 2000.0])
def test_device_poisson_sampler(ctx, delta_t):
    """PoissonEventTimeSampler::next_event_time_at_indexed_location_weakly_after
    (poisson.rs:30-109) on the device against the oracle: lambda * delta_t of 1
    (table), 1.5 (division path), 20 (counts past the table) and 1000 (the
    normal approximation); the event time and the state primed for its draws"""
    announce(f"device Poisson sampler, delta_t {delta_t}")
    rng = np.random.default_rng(int(delta_t))
    m = 20_000 if delta_t < 100 else 4_000
    seeds = rng.integers(0, 2**63, m, dtype=np.uint64)
    locs = rng.integers(0, 2**63, m, dtype=np.uint64)
    times = rng.uniform(0.0, 1e4, m)
    times[:16] = [0.0, 2.0, 4.0, 1.9999999999999998, 3.0, 6.0, 1e-300, 5e-324, 2000.0, 1999.9999999999998,
                  4000.0, 40.0, 39.99999999999999, 80.0, 1.0, 0.5]
    words = np.concatenate([np.stack([seeds, locs, times.view(np.uint64)], 1).ravel(),
                            np.array([0.5, delta_t]).view(np.uint64)])
    got = device_eval(ctx, 12, words, 2 * m).reshape(m, 2)
    st = ctypes.c_uint64()
    for k in range(m):
        t = O.lib().orc_poisson_next_event_time(int(seeds[k]), int(locs[k]), float(times[k]), 0.5, delta_t,
                                                ctypes.byref(st))
        assert got[k, 0] == np.float64(t).view(np.uint64) and got[k, 1] == st.value, \
            f"item {k}: time {times[k]!r}, device {got[k, 0].view(np.float64)!r}, oracle {t!r}"


# ---------------------------------------------------------------------------
# config 1: NonSpatial 100x100, deme 1, nu 1e-3, seed 42, 100% sampled
@pytest.fixture(scope="module")
def config1():
    cogs = nh.NonSpatialScenario((100, 100), 1).new(1e-3)
    orc = O.Scenario.non_spatial(100, 100, 1)
    lin = orc.origins(1.0)
    return cogs, orc, lin, {d: oracle_run(orc, lin, 42, 1e-3, dedup=d) for d in ("none", 1.0)}


# persistent form: automatic, plain (1) and cooperative (2) next-event search
@pytest.mark.parametrize("form,coop", [(PLAIN, 0), (PERSISTENT, 0), (PERSISTENT, 1), (PERSISTENT, 2),
                                       (TASK_LIST, 0)])
def test_config1_dedup_none_bit_exact(config1, form, coop):
    announce(f"config1 dedup None form {form} coop_search {coop}")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, form, dedup="none", coop_search=coop)
    assert_exact(ref["none"], out, ev, "config1 none")
    assert O.biodiversity(ev) == 67


@pytest.mark.parametrize("plain_batch", [1, 3, 0])
def test_config1_dedup_plain_bit_exact(config1, plain_batch):
    """one host round trip per pass (1), and passes batched on the device"""
    announce(f"config1 dedup Relative(1.0) plain form, plain_batch {plain_batch}")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, PLAIN, dedup=1.0, plain_batch=plain_batch)
    assert_exact(ref[1.0], out, ev, "config1 dedup plain")


@pytest.mark.parametrize("coop", [1, 2])
def test_config1_dedup_persistent_invariants(config1, coop):
    announce(f"config1 dedup Relative(1.0) persistent form coop_search {coop}")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup=1.0, coop_search=coop)
    assert_invariants(ev, ref["none"]["events"], ref[1.0]["events"], "config1 dedup persistent")
    assert out.steps <= ref["none"]["steps"]


def test_config1_graph_replays_match_eager(config1):
    announce("config1 persistent form, rounds replayed from hipGraphs")
    cogs, orc, lin, ref = config1
    # a small event slice gives many water-level rounds (the dedup-None output
    # does not depend on it), so both captured graphs are replayed repeatedly
    few = nh.EventSlice.Relative(0.05)
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup="none", use_graph=True, event_slice=few)
    assert out.rounds >= 6
    assert_exact(ref["none"], out, ev, "config1 graph")
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup=1.0, use_graph=True, event_slice=few)
    assert_invariants(ev, ref["none"]["events"], ref[1.0]["events"], "config1 graph dedup")


# ---------------------------------------------------------------------------
# config 2: almost-infinite Gaussian sigma=10, circle r=100, nu 1e-4, seed 7
@pytest.fixture(scope="module")
def config2():
    cogs = nh.AlmostInfiniteNormalDispersalScenario(100, 10.0).new(1e-4)
    orc = O.Scenario.almost_infinite_normal(10.0)
    lin = orc.origins(1.0, circle=(2**31 - 1, 2**31 - 1, 100))
    assert len(lin) == 31417
    return cogs, orc, lin


@pytest.mark.parametrize("form,coop", [(PLAIN, 0), (PERSISTENT, 1), (PERSISTENT, 2), (TASK_LIST, 0)])
def test_config2_window_dedup_none(config2, form, coop):
    announce(f"config2 pause 40 dedup None form {form} coop_search {coop}")
    cogs, orc, lin = config2
    ref = oracle_run(orc, lin, 7, 1e-4, dedup="none", pause_before=40.0)
    out, ev = hip_run(cogs, lin, 7, form, dedup="none", pause_before=40.0, coop_search=coop)
    assert_exact(ref, out, ev, "config2 window")
    rem = np.sort(out.lineages, order=["global_reference"])
    orem = np.sort(ref["remaining"], order=["global_reference"])
    assert rem.tobytes() == orem.tobytes()


@pytest.mark.parametrize("plain_batch", [1, 0])
def test_config2_window_dedup_plain_exact(config2, plain_batch):
    announce(f"config2 pause 40 dedup plain, plain_batch {plain_batch}")
    cogs, orc, lin = config2
    ref = oracle_run(orc, lin, 7, 1e-4, dedup=1.0, pause_before=40.0)
    out, ev = hip_run(cogs, lin, 7, PLAIN, dedup=1.0, pause_before=40.0, plain_batch=plain_batch)
    assert_exact(ref, out, ev, "config2 dedup plain")
    # the plain form keeps the reference's queue order of remaining lineages
    assert out.lineages.tobytes() == ref["remaining"].tobytes()


# ---------------------------------------------------------------------------
# spatially explicit maps (separable and packed alias)
@pytest.mark.parametrize("packed", [False, True])
def test_small_map_full_run(packed):
    announce(f"8x8 map packed={packed}")
    hip, orc = map_scenarios(8, 8, seed=5, packed=packed, sigma=1.5, cutoff=2.5)
    cogs = hip.new(0.01)
    lin = orc.origins(1.0)
    for form in (PLAIN, PERSISTENT):
        ref = oracle_run(orc, lin, 3, 0.01, dedup="none")
        out, ev = hip_run(cogs, lin, 3, form, dedup="none")
        assert_exact(ref, out, ev, f"8x8 form {form}")
    ref = oracle_run(orc, lin, 3, 0.01, dedup=1.0)
    out, ev = hip_run(cogs, lin, 3, PLAIN, dedup=1.0)
    assert_exact(ref, out, ev, "8x8 dedup plain")


@pytest.fixture(scope="module")
def map64():
    hip, orc = map_scenarios(64, 64, seed=42)
    return hip.new(1e-3), orc, orc.origins(1.0)


@pytest.mark.parametrize("form,coop", [(PLAIN, 0), (PERSISTENT, 1), (PERSISTENT, 2), (TASK_LIST, 0)])
def test_map64_window(map64, form, coop):
    announce(f"64x64 map pause 30 form {form} coop_search {coop}")
    cogs, orc, lin = map64
    ref = oracle_run(orc, lin, 9, 1e-3, dedup="none", pause_before=30.0)
    out, ev = hip_run(cogs, lin, 9, form, dedup="none", pause_before=30.0, coop_search=coop)
    assert_exact(ref, out, ev, "map64 window")
    refd = oracle_run(orc, lin, 9, 1e-3, dedup=1.0, pause_before=30.0)
    outd, evd = hip_run(cogs, lin, 9, form, dedup=1.0, pause_before=30.0, coop_search=coop)
    if form == PLAIN:
        assert_exact(refd, outd, evd, "map64 dedup plain")
    elif form == TASK_LIST:  # the CUDA loop's queue order, not the CPU algorithm's
        assert_invariants(evd, ref["events"], refd["events"], "map64 dedup task list")
    else:
        assert_invariants(evd, ref["events"], refd["events"], "map64 dedup persistent")


# ---------------------------------------------------------------------------
# pause / resume continuity (docs/simulate.ron:45-54)
@pytest.mark.parametrize("form", [PLAIN, PERSISTENT])
def test_pause_resume_continuity(config1, form):
    announce(f"pause/resume config1 form {form}")
    cogs, orc, lin, ref = config1
    out1, ev1 = hip_run(cogs, lin, 42, form, dedup="none", pause_before=500.0)
    assert out1.paused and len(out1.lineages) > 0
    out2, ev2 = hip_run(cogs, out1.lineages, 42, form, dedup="none", resume_after=500.0)
    both = np.sort(np.concatenate([ev1, ev2]),
                   order=["event_time", "x", "y", "index", "prior_time", "global_reference"])
    want = ref["none"]["events"]
    assert both.tobytes() == want.tobytes()
    assert out1.steps + out2.steps == ref["none"]["steps"]


# ---------------------------------------------------------------------------
# edge cases
def test_edge_empty_and_single(config1):
    announce("edge: empty and single lineage")
    cogs, orc, lin, ref = config1
    for form in (PLAIN, PERSISTENT):
        out, ev = hip_run(cogs, lin[:0], 42, form)
        assert out.steps == 0 and len(ev) == 0 and not out.paused
        one = lin[17:18]
        r1 = oracle_run(orc, one, 42, 1e-3, dedup=1.0)
        out, ev = hip_run(cogs, one, 42, form, dedup=1.0)
        assert_exact(r1, out, ev, "single lineage")


def test_edge_certain_speciation():
    announce("edge: nu = 1")
    cogs = nh.NonSpatialScenario((10, 10), 4).new(1.0)
    orc = O.Scenario.non_spatial(10, 10, 4)
    lin = orc.origins(1.0)
    ref = oracle_run(orc, lin, 1, 1.0, dedup=1.0)
    for form in (PLAIN, PERSISTENT):
        out, ev = hip_run(cogs, lin, 1, form, dedup=1.0)
        assert_exact(ref, out, ev, "nu=1")


@pytest.mark.parametrize("plain_batch", [1, 0])
def test_edge_tiny_absolute_cache_long_bucket_lists(plain_batch):
    """an Absolute(3) cache puts hundreds of inserts into one bucket per pass:
    the long lists are replayed in pass order by repeated minimum searches"""
    announce(f"edge: Absolute(3) dedup cache (bucket lists beyond 64 inserts), plain_batch {plain_batch}")
    cogs = nh.NonSpatialScenario((30, 30), 2).new(1e-2)
    orc = O.Scenario.non_spatial(30, 30, 2)
    lin = orc.origins(1.0)
    ref = oracle_run(orc, lin, 5, 1e-2, dedup=("absolute", 3))
    out, ev = hip_run(cogs, lin, 5, PLAIN, dedup=("absolute", 3), plain_batch=plain_batch)
    assert_exact(ref, out, ev, "absolute(3)")


def test_edge_out_of_deme_rejected(config1):
    announce("edge: lineage outside its deme")
    cogs, orc, lin, ref = config1
    bad = lin[:4].copy()
    bad[2]["index"] = 5
    bad[3]["x"] = 100  # outside the 100x100 extent
    # resume: ResumeError::Sample with every exceptional lineage, in order
    # (initialiser/resume.rs:71-85)
    with pytest.raises(nh.ResumeSampleError) as e:
        hip_run(cogs, bad, 42, PLAIN)
    kinds = [k for k, _ in e.value.exceptional]
    assert kinds == ["OutOfDeme", "OutOfHabitat"]
    assert e.value.exceptional[0][1].tobytes() == bad[2].tobytes()
    assert e.value.exceptional[1][1].tobytes() == bad[3].tobytes()
    # the C ABI refuses them as well, before any kernel runs
    ctx = nh.Context.get(0)
    job = nh.Job(ctx, nh.Tables(ctx, cogs), nh.IndependentArguments(), nh.WyHash.seed_from_u64(42),
                 1e-3, 0.5)
    with pytest.raises(nh.NecsimHipError) as e2:
        job.add_lineages(bad)
    assert e2.value.code == -6
    job.close()


def test_edge_general_delta_t_and_step_slice(config1):
    announce("edge: delta_t 3.0 (division path), step_slice 1")
    cogs, orc, lin, ref = config1
    sub = lin[::7]
    r = oracle_run(orc, sub, 42, 1e-3, dedup=1.0, delta_t=3.0, step_slice=1)
    out, ev = hip_run(cogs, sub, 42, PLAIN, dedup=1.0, delta_t=3.0, step_slice=1)
    assert_exact(r, out, ev, "delta_t 3, step_slice 1")


# Windows of many events: lambda * delta_t = 1.5 (delta_t not a power of two),
# 20 (the Poisson count past its inversion table; a window's draws spread over
# many calls of the stepped search) and 1000 (exp(-1000) underflows: the count
# comes from the normal approximation, poisson.rs:66-74)
@pytest.mark.parametrize("delta_t,pause", [(3.0, None), (40.0, 400.0), (2000.0, 200.0)])
@pytest.mark.parametrize("form,coop", [(PLAIN, 0), (PERSISTENT, 1), (PERSISTENT, 2), (TASK_LIST, 0)])
def test_edge_windows_of_many_events(config1, delta_t, pause, form, coop):
    announce(f"edge: delta_t {delta_t} form {form} coop_search {coop}")
    cogs, orc, lin, _ = config1
    sub = lin[::50]
    r = oracle_run(orc, sub, 42, 1e-3, dedup="none", delta_t=delta_t, pause_before=pause)
    out, ev = hip_run(cogs, sub, 42, form, dedup="none", delta_t=delta_t, pause_before=pause, coop_search=coop)
    assert r["steps"] > 10 * len(sub)
    assert_exact(r, out, ev, f"delta_t {delta_t}")


def test_autotuned_persistent_form_keeps_parity(config2):
    announce("autotune on a config2 sample, then a tuned run")
    cogs, orc, lin = config2
    ctx = nh.Context.get(0)
    tables = nh.Tables(ctx, cogs)
    args, ms = nh.autotune(ctx, tables, nh.IndependentArguments(dedup_cache=nh.DedupCache.none()),
                           nh.WyHash.seed_from_u64(7), cogs, lin[::4], 10.0)
    assert args.block_size in (64, 128, 256) and args.step_slice in (5, 10, 20, 50, 150) and ms > 0
    ref = oracle_run(orc, lin, 7, 1e-4, dedup="none", pause_before=40.0)
    out, ev = hip_run(cogs, lin, 7, PERSISTENT, dedup="none", pause_before=40.0,
                      step_slice=args.step_slice)
    assert_exact(ref, out, ev, "tuned")


# ---------------------------------------------------------------------------
# the reference CUDA algorithm's launch structure (task lists, host dedup)
@pytest.mark.parametrize("task_list_size", [64, 1000, 0])
def test_task_list_form_config1(config1, task_list_size):
    """dedup None: bit-exact whatever the task list size; dedup on: the CUDA
    loop's queue order (tasks requeued after the skipped lineages of their
    fill), so the dedup-on invariants"""
    announce(f"task-list form config1, task_list_size {task_list_size}")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, TASK_LIST, dedup="none", task_list_size=task_list_size)
    assert_exact(ref["none"], out, ev, "task list none")
    out, ev = hip_run(cogs, lin, 42, TASK_LIST, dedup=1.0, task_list_size=task_list_size)
    assert_invariants(ev, ref["none"]["events"], ref[1.0]["events"], "task list dedup")


def test_task_list_form_pause_resume(config1):
    announce("task-list form pause at 500 and resume")
    cogs, orc, lin, ref = config1
    out1, ev1 = hip_run(cogs, lin, 42, TASK_LIST, pause_before=500.0, task_list_size=700)
    assert out1.paused and len(out1.lineages)
    out2, ev2 = hip_run(cogs, out1.lineages, 42, TASK_LIST, resume_after=500.0, task_list_size=700)
    both = np.concatenate([ev1, ev2])
    assert as_bytes(np.sort(both, order=["event_time", "x", "y", "index", "prior_time", "global_reference"])) == \
        as_bytes(ref["none"]["events"])
    assert out1.steps + out2.steps == ref["none"]["steps"]
