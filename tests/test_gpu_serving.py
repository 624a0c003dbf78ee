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
    budget = nh.Server.request_footprint(biggest) + 1  # room for about one request at a time
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


def test_serving_rejects_bad_lineages_before_any_kernel(world):
    announce("serving: lineages outside the habitat or deme are rejected at submission")
    ctx, tables, orc, reqs = world
    srv = nh.Server(ctx, tables, max_lineages=1 << 16, hbm_budget_bytes=1 << 26)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    good = reqs[0]["lineages"]
    cases = []
    bad = good[:4].copy()
    bad["x"][2] = 48  # outside the 48x48 extent
    cases.append(bad)
    bad = good[:4].copy()
    bad["index"][1] = 10_000  # above any cell's habitat
    cases.append(bad)
    h = np.asarray(nh.generate_map(48, 48, seed=11)[0], dtype=np.uint32).reshape(48, 48)
    ys, xs = np.nonzero(h == 0)
    if len(xs):
        bad = good[:4].copy()
        bad["x"][0], bad["y"][0], bad["index"][0] = xs[0], ys[0], 0  # a cell of habitat 0
        cases.append(bad)
    for bad in cases:
        with pytest.raises(nh.NecsimHipError) as e:
            srv.submit(bad, nh.WyHash.seed_from_u64(1), 1e-3, args, pause_before=10.0)
        assert e.value.code == -6  # NH_E_OUT_OF_DEME
    with pytest.raises(nh.NecsimHipError) as e:  # event slice capacity 0
        srv.submit(good, nh.WyHash.seed_from_u64(1), 1e-3,
                   nh.IndependentArguments(event_slice=nh.EventSlice.Absolute(0)), pause_before=10.0)
    assert e.value.code == -1
    # nothing was queued: no tick ever saw the bad requests
    assert srv.poll() is None
    srv.close()


def test_serving_admission_passes_a_blocked_head(world):
    announce("serving: a request that does not fit yet is passed by a smaller one")
    ctx, tables, orc, reqs = world
    by_size = sorted(range(len(reqs)), key=lambda k: len(reqs[k]["lineages"]))
    small, big1, big2 = by_size[0], by_size[-1], by_size[-2]
    fp = nh.Server.request_footprint
    budget = fp(len(reqs[big1]["lineages"])) + fp(len(reqs[small]["lineages"]))
    assert budget < fp(len(reqs[big1]["lineages"])) + fp(len(reqs[big2]["lineages"]))
    srv = nh.Server(ctx, tables, max_lineages=1 << 22, hbm_budget_bytes=budget, max_bypass_ticks=1000)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    order = (big1, big2, small)
    ids = {}
    for k in order:
        r = reqs[k]
        ids[srv.submit(r["lineages"], nh.WyHash.seed_from_u64(r["seed"]), r["nu"], args,
                       pause_before=r["pause"])] = k
    got = {}
    ticks = 0
    while len(got) < 3:
        srv.tick()
        ticks += 1
        while (res := srv.poll()) is not None:
            got[ids[res.request_id]] = (res, srv.events(res.request_id))
            srv.release(res.request_id)
        assert ticks < 100000
    srv.close()
    # the small request was admitted before the blocked big2
    assert got[small][0].admit_s < got[big2][0].admit_s
    for k in order:
        r = reqs[k]
        ref = oracle_run(orc, r["lineages"], r["seed"], r["nu"], dedup="none", pause_before=r["pause"])
        assert got[k][1].tobytes() == ref["events"].tobytes(), k
        assert got[k][0].steps == ref["steps"], k


@pytest.fixture(scope="module")
def config5():
    """config 5: the 256x256 generator map (seed 42) resident in one server;
    requests drawn like the serving benchmark (nu = 10^U[-7, -3], sample
    U[1 %, 100 %], own seeds), short pause windows so the oracle finishes"""
    hip, orc = map_scenarios(256, 256, seed=42)
    base = hip.new(1e-3)
    ctx = nh.Context.get(0)
    tables = nh.Tables(ctx, base)
    rng = np.random.default_rng(1)
    reqs = []
    for k in range(8):
        nu = float(10 ** rng.uniform(-7, -3))
        pct = float(rng.uniform(0.01, 1.0))
        lin = base.sample_habitat(nh.OriginPreSampler.all().percentage(pct))
        reqs.append(dict(seed=k, nu=nu, lineages=lin, pause=float(rng.uniform(16.0, 24.0))))
    return ctx, tables, orc, reqs


def test_config5_served_on_resident_256_map(config5):
    announce("config 5: 8 requests served on the resident 256x256 tables, staggered, vs oracle")
    ctx, tables, orc, reqs = config5
    mixed, ticks = serve(ctx, tables, reqs, nh.DedupCache.none(), tick_steps=4, stagger=3)
    assert ticks > 3
    for k, r in enumerate(reqs):
        ref = oracle_run(orc, r["lineages"], r["seed"], r["nu"], dedup="none", pause_before=r["pause"])
        res, ev = mixed[k]
        assert res.status == 0
        assert ev.tobytes() == ref["events"].tobytes(), k
        assert res.steps == ref["steps"] and res.time == ref["time"] and bool(res.paused) == ref["paused"], k
    k = max(range(len(reqs)), key=lambda j: len(reqs[j]["lineages"]))
    alone, _ = serve(ctx, tables, reqs[k:k + 1], nh.DedupCache.none())
    assert alone[0][1].tobytes() == mixed[k][1].tobytes()
    assert alone[0][0].steps == mixed[k][0].steps
