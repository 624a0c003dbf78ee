"""Shared helpers: build the same workload for the oracle and the HIP path,
and compare their outputs.

Parity bar (integer/index and bit-identical f64 work):
 * DedupCache::None: the sorted speciation stream (all 40 bytes of every
   event), the step count and the outcome time are identical.
 * Dedup on, plain form: identical as above (reference queue order).
 * Dedup on, persistent form: which of two coalesced duplicates survives may
   differ, so the checks are the order-independent invariants: species count,
   the set of unique (event_time bits, origin), and every event being an event
   of the dedup-None stream.
"""
import numpy as np

import necsim_hip as nh
import oracle as O

EV = nh.EVENT_DTYPE


def map_scenarios(width, height, seed=42, packed=False, **kw):
    h, rp, c, w = nh.generate_map(width, height, seed=seed, **kw)
    hip = nh.SpatiallyExplicitUniformTurnoverScenario(width, height, h, rp, c, w, packed=packed)
    orc = O.Scenario.spatially_explicit(width, height, h, rp, c, w, packed=packed)
    return hip, orc


def oracle_run(orc_scn, lineages, seed, nu, dedup="none", pause_before=None, step_slice=10,
               delta_t=2.0, event_slice=("relative", 2.0), min_event_time=0.0):
    d = None if dedup == "none" else (("relative", dedup) if isinstance(dedup, float) else dedup)
    # orc_simulate applies seed_from_u64 to the raw Seed value itself
    p = O.make_params(seed, nu, delta_t=delta_t, step_slice=step_slice, dedup=d,
                      event_slice=event_slice, pause_before=pause_before,
                      min_event_time=min_event_time)
    return O.simulate(orc_scn, p, lineages)


def hip_run(cogs, lineages, seed, form, dedup="none", pause_before=None, step_slice=10,
            delta_t=2.0, event_slice=None, resume_after=None, use_graph=False, plain_batch=0, coop_search=0,
            task_list_size=0, pool=0):
    if dedup == "none":
        dc = nh.DedupCache.none()
    elif isinstance(dedup, float):
        dc = nh.DedupCache.Relative(dedup)
    else:
        dc = nh.DedupCache.Absolute(dedup[1])
    es = nh.EventSlice.Relative(2.0) if event_slice is None else event_slice
    args = nh.IndependentArguments(delta_t=delta_t, step_slice=step_slice, dedup_cache=dc,
                                   event_slice=es, form=form, use_graph=use_graph, plain_batch=plain_batch,
                                   coop_search=coop_search, task_list_size=task_list_size, pool=pool)
    rep = nh.EventCollector()
    out = nh.HipIndependentAlgorithm.resume_and_simulate(
        args, nh.WyHash.seed_from_u64(seed), cogs, lineages, resume_after=resume_after,
        pause_before=pause_before, reporter=rep)
    return out, rep.events


def as_bytes(events):
    return np.ascontiguousarray(events, dtype=EV).tobytes()


def assert_exact(o, h_out, h_events, what=""):
    assert len(h_events) == len(o["events"]), f"{what}: {len(h_events)} vs {len(o['events'])} events"
    if as_bytes(h_events) != as_bytes(o["events"]):
        a, b = np.asarray(h_events, dtype=EV), o["events"]
        diff = [i for i in range(len(a)) if a[i].tobytes() != b[i].tobytes()]
        raise AssertionError(f"{what}: event streams differ at {diff[:5]}: {a[diff[:2]]} vs {b[diff[:2]]}")
    assert h_out.steps == o["steps"], f"{what}: steps {h_out.steps} vs {o['steps']}"
    assert np.float64(h_out.time).tobytes() == np.float64(o["time"]).tobytes(), \
        f"{what}: time {h_out.time!r} vs {o['time']!r}"
    assert h_out.paused == o["paused"], what


def unique_keys(events):
    ev = np.asarray(events, dtype=EV)
    return set(zip(ev["event_time"].view(np.uint64).tolist(), ev["x"].tolist(), ev["y"].tolist(),
                   ev["index"].tolist()))


def assert_invariants(h_events, ref_none_events, o_dedup_events=None, what=""):
    """Dedup-on invariants against the dedup-None stream."""
    h_events = np.asarray(h_events, dtype=EV)
    assert unique_keys(h_events) == unique_keys(ref_none_events), f"{what}: unique events differ"
    assert O.biodiversity(h_events) == O.biodiversity(ref_none_events), f"{what}: species differ"
    full = set(e.tobytes() for e in np.asarray(ref_none_events, dtype=EV))
    for e in h_events:
        assert e.tobytes() in full, f"{what}: event {e} is not in the dedup-None stream"
    if o_dedup_events is not None:
        assert O.biodiversity(h_events) == O.biodiversity(o_dedup_events)
