"""GPU parity of dispersal reporting (a reporter with ReportDispersal).

Every step of the HIP path is reported as a 56-byte PackedEvent
(necsim/core/src/event.rs:16-29): speciations with the sign bit of prior_time
set, dispersals (LineageInteraction::Maybe) with the sign bit of event_time
set, and the water level advances by event_slice / sum(lambda)
(parallelisation/independent/monolithic/mod.rs:139-148).  The oracle's
orc_simulate_packed is the checker.

Parity bar: the sorted PackedEvent stream (all 56 bytes of every event), the
step count, the outcome time and the remaining lineages are identical for
DedupCache::None in both forms and with dedup on in the plain form; the
persistent form with dedup on keeps the species and reports only events of
the dedup-None stream.
"""
import ctypes

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from necsim_hip import _ffi
from parity_utils import map_scenarios

pytestmark = pytest.mark.gpu
PLAIN, PERSISTENT = nh.PLAIN, nh.PERSISTENT
PK = nh.PACKED_EVENT_DTYPE


def announce(msg):
    print(f"[gpu case] {msg}", flush=True)


def _dedup(dedup):
    if dedup == "none":
        return nh.DedupCache.none(), None
    return nh.DedupCache.Relative(dedup), ("relative", dedup)


def hip_packed(cogs, lineages, seed, form, dedup="none", pause_before=None, step_slice=10,
               event_slice=None, use_graph=False, resume_after=None, pool=0):
    dc, _ = _dedup(dedup)
    es = nh.EventSlice.Relative(2.0) if event_slice is None else event_slice
    args = nh.IndependentArguments(step_slice=step_slice, dedup_cache=dc, event_slice=es, form=form,
                                   use_graph=use_graph, pool=pool)
    rep = nh.PackedEventCollector()
    out = nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(seed), cogs, lineages,
                                                         resume_after=resume_after, pause_before=pause_before,
                                                         reporter=rep)
    return out, rep.events


def oracle_packed(orc, lineages, seed, nu, dedup="none", pause_before=None, step_slice=10):
    _, d = _dedup(dedup)
    p = O.make_params(seed, nu, step_slice=step_slice, dedup=d, pause_before=pause_before,
                      report_dispersal=True)
    return O.simulate_packed(orc, p, lineages)


def packed_bytes(ev):
    return np.ascontiguousarray(ev).view(np.uint8).tobytes()


def assert_packed_exact(ref, out, ev, what):
    assert PK.itemsize == O.PACKED_DTYPE.itemsize == 56
    assert len(ev) == len(ref["events"]), f"{what}: {len(ev)} vs {len(ref['events'])} events"
    if packed_bytes(ev) != packed_bytes(ref["events"]):
        a = np.ascontiguousarray(ev).view(np.uint8).reshape(-1, 56)
        b = np.ascontiguousarray(ref["events"]).view(np.uint8).reshape(-1, 56)
        bad = np.nonzero(np.any(a != b, axis=1))[0]
        raise AssertionError(f"{what}: PackedEvent streams differ at {bad[:5].tolist()}: "
                             f"{ev[bad[:2]]} vs {ref['events'][bad[:2]]}")
    assert out.steps == ref["steps"], f"{what}: steps {out.steps} vs {ref['steps']}"
    assert np.float64(out.time).tobytes() == np.float64(ref["time"]).tobytes(), f"{what}: time"
    assert out.paused == ref["paused"], what
    if out.paused:
        rem = np.sort(out.lineages, order=["global_reference"])
        orem = np.sort(ref["remaining"], order=["global_reference"])
        assert rem.tobytes() == orem.tobytes(), f"{what}: remaining lineages differ"


def _species(packed):
    return O.biodiversity(nh.speciation_events(np.ascontiguousarray(packed).view(PK)))


# ---------------------------------------------------------------------------
# one small case per scenario kind: (name, hip cogs, oracle scenario, lineages, seed, nu, pause_before)
def _non_spatial():
    return ("non-spatial 30x30 deme 2", nh.NonSpatialScenario((30, 30), 2).new(1e-2),
            O.Scenario.non_spatial(30, 30, 2), None, 5, 1e-2, None)


def _almost_infinite():
    orc = O.Scenario.almost_infinite_normal(3.0)
    lin = orc.origins(1.0, circle=(2**31 - 1, 2**31 - 1, 20))
    return ("almost-infinite sigma 3 r 20 pause 60", nh.AlmostInfiniteNormalDispersalScenario(20, 3.0).new(1e-3),
            orc, lin, 7, 1e-3, 60.0)


def _map(packed):
    hip, orc = map_scenarios(24, 24, seed=11, packed=packed, sigma=1.5, cutoff=3.0)
    return (f"24x24 map packed={packed}", hip.new(2e-2), orc, None, 3, 2e-2, None)


CASES = {"non_spatial": _non_spatial, "almost_infinite": _almost_infinite,
         "separable_map": lambda: _map(False), "packed_map": lambda: _map(True)}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    name, cogs, orc, lin, seed, nu, pause = CASES[request.param]()
    if lin is None:
        lin = orc.origins(1.0)
    ref = oracle_packed(orc, lin, seed, nu, pause_before=pause)
    assert len(ref["events"]) == ref["steps"] > 1000
    return name, cogs, orc, lin, seed, nu, pause, ref


@pytest.mark.parametrize("form", [PLAIN, PERSISTENT])
def test_dedup_none_bit_exact(case, form):
    name, cogs, orc, lin, seed, nu, pause, ref = case
    announce(f"dispersal reporting {name} dedup None form {form}")
    out, ev = hip_packed(cogs, lin, seed, form, pause_before=pause)
    assert_packed_exact(ref, out, ev, f"{name} form {form}")


@pytest.mark.parametrize("pool", [128, 192])
def test_pooled_dedup_none_bit_exact(case, pool):
    """the pooled round kernel reporting every step: the same stream"""
    name, cogs, orc, lin, seed, nu, pause, ref = case
    announce(f"dispersal reporting {name} dedup None, pooled kernel {pool}")
    out, ev = hip_packed(cogs, lin, seed, PERSISTENT, pause_before=pause, pool=pool)
    assert_packed_exact(ref, out, ev, f"{name} pool {pool}")


def test_speciations_equal_the_speciation_only_run(case):
    name, cogs, orc, lin, seed, nu, pause, ref = case
    announce(f"dispersal reporting {name}: speciation subset vs speciation-only run")
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=PERSISTENT)
    col = nh.EventCollector()
    out = nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(seed), cogs, lin,
                                                         pause_before=pause, reporter=col)
    _, ev = hip_packed(cogs, lin, seed, PERSISTENT, pause_before=pause)
    want = np.ascontiguousarray(col.events, dtype=nh.EVENT_DTYPE)
    assert nh.speciation_events(ev).tobytes() == want.tobytes()
    assert out.steps == len(ev)


def test_dedup_plain_bit_exact_and_persistent_invariants(case):
    name, cogs, orc, lin, seed, nu, pause, ref = case
    announce(f"dispersal reporting {name} dedup Relative(1.0)")
    refd = oracle_packed(orc, lin, seed, nu, dedup=1.0, pause_before=pause)
    out, ev = hip_packed(cogs, lin, seed, PLAIN, dedup=1.0, pause_before=pause)
    assert_packed_exact(refd, out, ev, f"{name} dedup plain")
    out, ev = hip_packed(cogs, lin, seed, PERSISTENT, dedup=1.0, pause_before=pause)
    assert _species(ev) == _species(refd["events"]) == _species(ref["events"])
    full = set(np.ascontiguousarray(ref["events"]).view(np.uint8).reshape(-1, 56).view(f"V56").ravel().tolist())
    mine = np.ascontiguousarray(ev).view(np.uint8).reshape(-1, 56).view("V56").ravel().tolist()
    assert all(e in full for e in mine), f"{name}: an event outside the dedup-None stream"
    assert out.steps <= ref["steps"]


@pytest.mark.parametrize("pool", [0, 128])
def test_graph_replays_match_eager(pool):
    announce(f"dispersal reporting non-spatial 30x30, rounds replayed from hipGraphs, pool {pool}")
    name, cogs, orc, _, seed, nu, pause = _non_spatial()
    lin = orc.origins(1.0)
    ref = oracle_packed(orc, lin, seed, nu)
    few = nh.EventSlice.Relative(0.25)  # many rounds; the sorted stream does not depend on them
    out, ev = hip_packed(cogs, lin, seed, PERSISTENT, event_slice=few, use_graph=True, pool=pool)
    assert out.rounds >= 20
    assert_packed_exact(ref, out, ev, "graph")


@pytest.mark.parametrize("form,pool", [(PLAIN, 0), (PERSISTENT, 0), (PERSISTENT, 128)])
def test_pause_resume_continuity(form, pool):
    announce(f"dispersal reporting pause at 150 and resume, form {form} pool {pool}")
    name, cogs, orc, _, seed, nu, _p = _non_spatial()
    lin = orc.origins(1.0)
    ref = oracle_packed(orc, lin, seed, nu)
    out1, ev1 = hip_packed(cogs, lin, seed, form, pause_before=150.0, pool=pool)
    assert out1.paused and len(out1.lineages)
    out2, ev2 = hip_packed(cogs, out1.lineages, seed, form, resume_after=150.0, pool=pool)
    assert not out2.paused
    both = np.concatenate([ev1, ev2])
    assert packed_bytes(both) == packed_bytes(ref["events"])
    assert out1.steps + out2.steps == ref["steps"]


@pytest.mark.parametrize("form,pool", [(PLAIN, 0), (PERSISTENT, 0), (PERSISTENT, 128)])
def test_event_buffer_overflow_is_an_error(ctx, form, pool):
    """a round that reports more events than the buffer holds fails with
    NH_E_CAPACITY, writes nothing past the buffer and leaves the job failed"""
    announce(f"dispersal event buffer of 64 entries, form {form} pool {pool}")
    cogs = nh.NonSpatialScenario((30, 30), 2).new(1e-2)
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    job = nh.Job(ctx, tables, nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form, pool=pool),
                 nh.WyHash.seed_from_u64(5), 1e-2, 0.5, report_dispersal=True, event_buffer=64)
    job.add_lineages(lin)
    with pytest.raises(_ffi.NecsimHipError) as e:
        job.step_round()
    assert e.value.code == -7  # NH_E_CAPACITY
    with pytest.raises(_ffi.NecsimHipError) as e:
        job.step_round()
    assert e.value.code == -4  # NH_E_STATE
    job.close()


def test_drain_calls_match_the_job_kind(ctx):
    announce("drain calls on jobs of the other reporting kind")
    cogs = nh.NonSpatialScenario((10, 10), 1).new(1e-1)
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    for disp in (False, True):
        job = nh.Job(ctx, tables, args, nh.WyHash.seed_from_u64(1), 1e-1, 0.5, report_dispersal=disp)
        job.add_lineages(lin)
        job.step_round()
        with pytest.raises(_ffi.NecsimHipError) as e:
            job.drain_events() if disp else job.drain_packed_events()
        assert e.value.code == -4
        job.close()


@pytest.mark.parametrize("device_sized", [0, 1])
@pytest.mark.parametrize("shape", ["mixed", "runs", "long_runs", "runs_and_ghosts"])
def test_device_sort_of_packed_events(ctx, device_sized, shape):
    """the PackedEvent order on the device against a lexicographic sort, on
    records with many equal times and signed times: runs of equal times with
    several origins and targets (per-field passes), runs sharing one origin
    and target as coalesced lineages produce them (ranked in place), and runs
    longer than the in-place limit (per-field passes again)"""
    announce(f"device PackedEvent sort, {shape}, device_sized={device_sized}")
    rng = np.random.default_rng(17)
    n = 200_000
    ev = np.zeros(n, PK)
    ev["global_reference"] = rng.permutation(n).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    n_times = {"mixed": 3000, "runs": 60000, "long_runs": 12, "runs_and_ghosts": 60000}[shape]
    times = rng.uniform(1, 50, n_times)
    pick = rng.integers(0, n_times, n)
    t = times[pick]
    sign = rng.random(n) < 0.5
    ev["event_time"] = np.where(sign, -t, t)
    prior = t - rng.choice([0.0, 0.5, 1.0, 1.5], n)
    ev["prior_time"] = np.where(sign, prior, -prior)
    fields = (("origin_x", 4), ("origin_y", 3), ("origin_index", 2), ("target_x", 1 << 32), ("target_y", 3),
              ("target_index", 70000))
    for f, hi in fields:
        if shape == "mixed":
            ev[f] = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        else:  # one origin and target per time
            ev[f] = rng.integers(0, hi, n_times, dtype=np.uint64).astype(np.uint32)[pick]
    ev["coalescence"] = ev["global_reference"]
    if shape == "runs_and_ghosts":  # unused event slots of a persistent round: time +inf, zeros
        ghosts = np.zeros(5000, PK)
        ghosts["event_time"] = np.inf
        ev = np.concatenate([ev, ghosts])[rng.permutation(n + 5000)]
        n = len(ev)
    want = ev[nh.packed_order(ev)]
    got = np.zeros(n, PK)
    _ffi.check(_ffi.lib().nh_debug_sort_packed_events(ctx.handle, _ffi.ptr(ev), n, device_sized, _ffi.ptr(got)))
    assert packed_bytes(got) == packed_bytes(want)


def test_event_log_records_dispersal(tmp_path):
    announce("dispersal-reporting run recorded to an event log")
    import struct
    name, cogs, orc, _, seed, nu, _p = _non_spatial()
    lin = orc.origins(1.0)
    col = nh.PackedEventCollector()
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(seed), cogs, lin, reporter=col)
    log = nh.EventLogReporter(str(tmp_path / "log"), segment_capacity=50_000, report_dispersal=True)
    nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(seed), cogs, lin, reporter=log)
    log.close()
    import os
    names = sorted(os.listdir(tmp_path / "log"), key=int)
    assert len(names) == -(-len(col.events) // 50_000)
    got = []
    for nme in names:
        raw = open(tmp_path / "log" / nme, "rb").read()
        lo, hi, length, with_s, with_d = struct.unpack_from("<ddQ??", raw, 0)
        assert with_s and with_d and len(raw) == 26 + 56 * length
        seg = np.frombuffer(raw[26:], PK)
        assert lo == abs(seg["event_time"][0]) and hi == abs(seg["event_time"][-1])
        got.append(seg)
    assert packed_bytes(np.concatenate(got)) == packed_bytes(col.events)


# ---------------------------------------------------------------------------
# EventCounterReporter on the device
def _device_counts(ctx, ev):
    ev = np.ascontiguousarray(ev, dtype=PK)
    c = _ffi.EventCounts()
    _ffi.check(_ffi.lib().nh_debug_count_packed_events(ctx.handle, _ffi.ptr(ev), len(ev), ctypes.byref(c)))
    return {f: int(getattr(c, f)) for f, _ in _ffi.EventCounts._fields_}


def test_device_counter_on_hand_and_coalescing_streams(ctx):
    """the device counter against the sequential restatement of
    event_counter.rs:58-109 (tests/test_reporters.py)"""
    announce("device EventCounterReporter on the hand-made and a coalescing stream")
    from test_reporters import hand_stream, reference_counts
    ev = hand_stream()
    assert _device_counts(ctx, ev) == reference_counts([ev])
    s = O.Scenario.non_spatial(6, 6, 2)
    full = O.simulate_packed(s, O.make_params(11, 0.02, dedup=None, report_dispersal=True), s.origins(1.0))
    ev = np.ascontiguousarray(full["events"]).view(PK)
    assert _device_counts(ctx, ev) == reference_counts([ev])


def _device_sort_counts(ctx, ev):
    ev = np.ascontiguousarray(ev, dtype=PK)
    c = _ffi.EventCounts()
    _ffi.check(_ffi.lib().nh_debug_sort_count_packed_events(ctx.handle, _ffi.ptr(ev), len(ev), ctypes.byref(c)))
    return {f: int(getattr(c, f)) for f, _ in _ffi.EventCounts._fields_}


def _duplicate_runs(rng, n_runs, max_len, priors):
    """runs of coalesced duplicates (one time, origin and target each; Maybe
    dispersals), prior times from a small set, references in random order"""
    rows = []
    ref = 1
    for k in range(n_runs):
        t = 1.0 + k * 0.001 + float(rng.uniform(0.0, 1e-4))
        org = (int(rng.integers(0, 50)), int(rng.integers(0, 50)), 0)
        tgt = (int(rng.integers(0, 50)), int(rng.integers(0, 50)), 0)
        for _ in range(int(rng.integers(1, max_len + 1))):
            rows.append((ref, float(rng.choice(priors)), t, org, tgt))
            ref += 1
    ev = np.zeros(len(rows), PK)
    refs = rng.permutation(len(rows)).astype(np.uint64) + 1
    for i, (_, p, t, org, tgt) in enumerate(rows):
        ev[i]["global_reference"] = refs[i]
        ev[i]["prior_time"] = p
        ev[i]["event_time"] = -t  # LineageInteraction::Maybe
        ev[i]["origin_x"], ev[i]["origin_y"], ev[i]["origin_index"] = org
        ev[i]["target_x"], ev[i]["target_y"], ev[i]["target_index"] = tgt
        ev[i]["coalescence"] = refs[i]
    return ev


def test_device_sort_and_count_of_unsorted_streams(ctx):
    """A counted round's windows order each run by prior time alone (its
    counters read nothing else); runs whose equal prior times carry
    references in another order, runs that mix classes (counted through the
    permutation), runs longer than a window (field passes) and the hand-made
    and coalescing streams, each shuffled, must count as the sequential
    restatement of event_counter.rs:58-109 counts the sorted stream."""
    announce("device sort + count of shuffled streams (counted-round windows)")
    from test_reporters import hand_stream, reference_counts
    rng = np.random.default_rng(21)
    s = O.Scenario.non_spatial(6, 6, 2)
    full = O.simulate_packed(s, O.make_params(11, 0.02, dedup=None, report_dispersal=True), s.origins(1.0))
    cases = {
        "hand": hand_stream(),
        "coalescing": np.ascontiguousarray(full["events"]).view(PK),
        "equal_priors": _duplicate_runs(rng, 600, 80, [0.25, 0.5, 0.75]),
        "one_prior": _duplicate_runs(rng, 300, 200, [0.5]),
        "long_runs": _duplicate_runs(rng, 3, 2500, [0.25, 0.5]),
    }
    # a run of one time and location that mixes speciations with self-dispersals
    mixed = hand_stream()[:3].copy()
    extra = mixed.copy()
    extra["prior_time"] = 0.5
    extra["event_time"] = 1.0
    extra["global_reference"] = [40, 41, 42]
    cases["mixed_classes"] = np.concatenate([mixed, extra, _duplicate_runs(rng, 100, 30, [0.5, 0.75])])
    for name, ev in cases.items():
        want = reference_counts([ev[nh.packed_order(ev)]])
        got = _device_sort_counts(ctx, ev[rng.permutation(len(ev))])
        assert got == want, (name, got, want)


@pytest.mark.parametrize("form,graph,pool", [(PLAIN, False, 0), (PERSISTENT, False, 0), (PERSISTENT, True, 0),
                                             (PERSISTENT, False, 128), (PERSISTENT, True, 192)])
def test_device_counter_job_matches_host_counter(form, graph, pool):
    announce(f"device_event_counter job vs host EventCounterReporter, form {form} graph {graph} pool {pool}")
    cogs = nh.NonSpatialScenario((12, 12), 2).new(5e-3)
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form, use_graph=graph,
                                   event_slice=nh.EventSlice.Relative(0.5), pool=pool)
    host, dev = nh.EventCounterReporter(), nh.EventCounterReporter(on_device=True)
    col = nh.PackedEventCollector()
    outs = [nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(2), cogs, lin, reporter=r)
            for r in (host, dev, col)]
    assert outs[0].steps == outs[1].steps == outs[2].steps
    assert outs[1].rounds > 5
    assert dev.counts == host.counts
    from test_reporters import reference_counts
    assert dev.counts == reference_counts([col.events])
    assert dev.counts["late_coalescence"] > 0 and dev.counts["raw_total"] == outs[1].steps


def test_partial_drains_keep_the_order(ctx):
    """events left undrained on the device move to the host before the next
    round; drains hand out host-held events first, then device-held ones"""
    announce("partial drains of PackedEvents across rounds")
    name, cogs, orc, _, seed, nu, _p = _non_spatial()
    lin = orc.origins(1.0)
    ref = oracle_packed(orc, lin, seed, nu)
    tables = nh.Tables(ctx, cogs)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=PERSISTENT, use_graph=True,
                                   event_slice=nh.EventSlice.Relative(0.25))
    job = nh.Job(ctx, tables, args, nh.WyHash.seed_from_u64(seed), nu, 0.5, report_dispersal=True)
    job.add_lineages(lin)
    got, fin, k = [], False, 0
    lib = _ffi.lib()
    while not fin:
        fin = job.step_round()
        n = ctypes.c_size_t(0)
        _ffi.check(lib.nh_job_drain_packed_events(job.handle, None, 0, ctypes.byref(n)))
        take = n.value // 3 if k % 3 else (0 if k % 2 else n.value)
        buf = np.empty(take, PK)
        m = ctypes.c_size_t(0)
        _ffi.check(lib.nh_job_drain_packed_events(job.handle, _ffi.ptr(buf) if take else None, take,
                                                  ctypes.byref(m)))
        got.append(buf[:m.value] if take else buf[:0])
        k += 1
    got.append(job.drain_packed_events())
    job.close()
    assert k >= 10
    assert packed_bytes(np.concatenate(got)) == packed_bytes(ref["events"])
