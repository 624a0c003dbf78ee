"""Serving: continuous batching of simulations on resident tables.

Each request's result must be the same alone and in the mix: bit-exact
(stream, steps, time) with DedupCache::None; with dedup on, the species count
and unique-event set (the order-independent invariants) plus every event
belonging to the request's dedup-None stream.
"""
import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from parity_utils import assert_invariants, map_scenarios, oracle_run

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def world():
    hip, orc = map_scenarios(48, 48, seed=11)
    cogs = hip.new(1e-3)
    ctx = nh.Context.get(0)
    tables = nh.Tables(ctx, cogs)
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(10):
        nu = float(10 ** rng.uniform(-4, -2))
        pct = float(rng.uniform(0.05, 1.0))
        lin = cogs.sample_habitat(nh.OriginPreSampler.all().percentage(pct))
        reqs.append(dict(seed=1000 + k, nu=nu, lineages=lin, pause=float(rng.uniform(20, 80))))
    return ctx, tables, orc, reqs


def serve(ctx, tables, reqs, dedup, budget=1 << 34, tick_steps=0, stagger=0):
    srv = nh.Server(ctx, tables, max_lineages=1 << 22, hbm_budget_bytes=budget, tick_steps=tick_steps)
    args = nh.IndependentArguments(dedup_cache=dedup)
    ids = {}
    pending = list(range(len(reqs)))
    first = pending[:len(reqs) - stagger]
    for k in first:
        r = reqs[k]
        ids[srv.submit(r["lineages"], nh.WyHash.seed_from_u64(r["seed"]), r["nu"], args,
                       pause_before=r["pause"])] = k
    late = pending[len(reqs) - stagger:]
    results = {}
    ticks = 0
    while len(results) < len(reqs):
        if late and ticks >= 1:
            k = late.pop(0)  # arrives while the others run
            r = reqs[k]
            ids[srv.submit(r["lineages"], nh.WyHash.seed_from_u64(r["seed"]), r["nu"], args,
                           pause_before=r["pause"])] = k
        srv.tick()
        ticks += 1
        while True:
            res = srv.poll()
            if res is None:
                break
            results[ids[res.request_id]] = (res, srv.events(res.request_id))
            srv.release(res.request_id)
        assert ticks < 100000
    srv.close()
    return results, ticks


def test_serving_dedup_none_exact_alone_and_mixed(world):
    announce("serving: 10 requests, dedup None, mixed vs alone vs oracle")
    ctx, tables, orc, reqs = world
    mixed, _ = serve(ctx, tables, reqs, nh.DedupCache.none(), stagger=4)
    for k, r in enumerate(reqs):
        ref = oracle_run(orc, r["lineages"], r["seed"], r["nu"], dedup="none", pause_before=r["pause"])
        res, ev = mixed[k]
        assert ev.tobytes() == ref["events"].tobytes(), k
        assert res.steps == ref["steps"] and res.time == ref["time"] and bool(res.paused) == ref["paused"], k
    alone, _ = serve(ctx, tables, reqs[3:4], nh.DedupCache.none())
    assert alone[0][1].tobytes() == mixed[3][1].tobytes()


def test_serving_admission_by_footprint(world):
    announce("serving: tight HBM budget admits requests as memory frees")
    ctx, tables, orc, reqs = world
    biggest = max(len(r["lineages"]) for r in reqs)
    budget = biggest * 130 + 1  # room for about one request at a time
    res, ticks = serve(ctx, tables, reqs[:5], nh.DedupCache.none(), budget=budget)
    for k, r in enumerate(reqs[:5]):
        ref = oracle_run(orc, r["lineages"], r["seed"], r["nu"], dedup="none", pause_before=r["pause"])
        assert res[k][1].tobytes() == ref["events"].tobytes()
        assert res[k][0].admit_s >= res[k][0].submit_s and res[k][0].done_s >= res[k][0].admit_s


def test_serving_tick_limit_and_dedup_invariants(world):
    announce("serving: tick limit 7 steps, dedup Relative(1.0)")
    ctx, tables, orc, reqs = world
    res, ticks = serve(ctx, tables, reqs, nh.DedupCache.Relative(1.0), tick_steps=7, stagger=3)
    assert ticks > 3
    for k, r in enumerate(reqs):
        ref = oracle_run(orc, r["lineages"], r["seed"], r["nu"], dedup="none", pause_before=r["pause"])
        assert_invariants(res[k][1], ref["events"], None, f"request {k}")
