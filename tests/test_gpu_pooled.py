"""The pooled persistent kernel (nh_job_params.pool) against the oracle.

Each wave of the pooled kernel keeps 128 or 192 lineages in LDS and hands
search steps and executions to whichever lanes are free, so lineages move
between lanes; each lineage's own computation is unchanged.  Bar: with
DedupCache::None the sorted speciation stream, steps, time and remaining
lineages equal the oracle's bit for bit, for every scenario kind, windows of
many events per Poisson window, pause/resume and graph replays.  With a
dedup cache or dispersal reporting the job falls back to the one-lineage-per-
lane kernel, which must still hold its own bar.
"""
import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from parity_utils import assert_exact, assert_invariants, hip_run, map_scenarios, oracle_run

pytestmark = pytest.mark.gpu
PERSISTENT = nh.PERSISTENT
POOLS = [128, 192]


def announce(msg):
    print(f"[case] {msg}", flush=True)


def same_remaining(out, ref):
    rem = np.sort(out.lineages, order=["global_reference"])
    orem = np.sort(ref["remaining"], order=["global_reference"])
    return rem.tobytes() == orem.tobytes()


@pytest.fixture(scope="module")
def config1():
    cogs = nh.NonSpatialScenario((100, 100), 1).new(1e-3)
    orc = O.Scenario.non_spatial(100, 100, 1)
    lin = orc.origins(1.0)
    return cogs, orc, lin, oracle_run(orc, lin, 42, 1e-3, dedup="none")


@pytest.mark.parametrize("pool", POOLS)
def test_pooled_config1_full_run(config1, pool):
    announce(f"pooled {pool}: config1 full run, dedup None")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup="none", pool=pool)
    assert_exact(ref, out, ev, f"pool {pool}")
    assert O.biodiversity(ev) == 67


@pytest.mark.parametrize("pool", POOLS)
def test_pooled_config2_window(pool):
    announce(f"pooled {pool}: config2 window (almost-infinite, sigma 10), pause 40")
    cogs = nh.AlmostInfiniteNormalDispersalScenario(100, 10.0).new(1e-4)
    orc = O.Scenario.almost_infinite_normal(10.0)
    lin = orc.origins(1.0, circle=(2**31 - 1, 2**31 - 1, 100))
    ref = oracle_run(orc, lin, 7, 1e-4, dedup="none", pause_before=40.0)
    out, ev = hip_run(cogs, lin, 7, PERSISTENT, dedup="none", pause_before=40.0, pool=pool)
    assert_exact(ref, out, ev, f"config2 pool {pool}")
    assert same_remaining(out, ref)


@pytest.mark.parametrize("pool", POOLS)
@pytest.mark.parametrize("packed", [False, True])
def test_pooled_map_window(pool, packed):
    announce(f"pooled {pool}: 64x64 map packed={packed}, pause 30")
    hip, orc = map_scenarios(64, 64, seed=42, packed=packed)
    cogs = hip.new(1e-3)
    lin = orc.origins(1.0)
    ref = oracle_run(orc, lin, 9, 1e-3, dedup="none", pause_before=30.0)
    out, ev = hip_run(cogs, lin, 9, PERSISTENT, dedup="none", pause_before=30.0, pool=pool)
    assert_exact(ref, out, ev, f"map64 pool {pool}")
    assert same_remaining(out, ref)


@pytest.mark.parametrize("pool", POOLS)
@pytest.mark.parametrize("packed", [False, True])
def test_pooled_map_to_completion(pool, packed):
    # the alias kinds split the execution pass around the search pass; run
    # every lineage to its speciation (no water level) on a small map
    announce(f"pooled {pool}: 16x16 map packed={packed}, nu 0.02, to completion")
    hip, orc = map_scenarios(16, 16, seed=5, packed=packed)
    cogs = hip.new(0.02)
    lin = orc.origins(1.0)
    ref = oracle_run(orc, lin, 3, 0.02, dedup="none")
    out, ev = hip_run(cogs, lin, 3, PERSISTENT, dedup="none", pool=pool)
    assert_exact(ref, out, ev, f"map16 pool {pool}")
    assert not out.paused and len(ev) == len(lin)


@pytest.mark.parametrize("pool", POOLS)
def test_pooled_config4_sample_window(pool):
    announce(f"pooled {pool}: config4 0.1% sample, pause 400")
    cogs = nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0).new(1e-3)
    orc = O.Scenario.almost_infinite_normal(50.0)
    lin = orc.origins(0.001, circle=(2**31 - 1, 2**31 - 1, 2000))
    ref = oracle_run(orc, lin, 42, 1e-3, dedup="none", pause_before=400.0)
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup="none", pause_before=400.0, pool=pool)
    assert_exact(ref, out, ev, f"config4 sample pool {pool}")
    assert same_remaining(out, ref)


# lambda * delta_t of 1.5 (division path), 20 (counts past the inversion
# table, draws spread over many search steps) and 1000 (normal approximation)
@pytest.mark.parametrize("delta_t,pause", [(3.0, None), (40.0, 400.0), (2000.0, 200.0)])
@pytest.mark.parametrize("pool", POOLS)
def test_pooled_windows_of_many_events(config1, delta_t, pause, pool):
    announce(f"pooled {pool}: delta_t {delta_t}")
    cogs, orc, lin, _ = config1
    sub = lin[::50]
    r = oracle_run(orc, sub, 42, 1e-3, dedup="none", delta_t=delta_t, pause_before=pause)
    out, ev = hip_run(cogs, sub, 42, PERSISTENT, dedup="none", delta_t=delta_t, pause_before=pause, pool=pool)
    assert_exact(r, out, ev, f"delta_t {delta_t} pool {pool}")


@pytest.mark.parametrize("pool", POOLS)
def test_pooled_graph_replays_and_pause_resume(config1, pool):
    announce(f"pooled {pool}: rounds from hipGraphs, and pause at 300 then resume")
    cogs, orc, lin, ref = config1
    few = nh.EventSlice.Relative(0.05)
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup="none", use_graph=True, event_slice=few, pool=pool)
    assert out.rounds >= 6
    assert_exact(ref, out, ev, f"graph pool {pool}")
    first, ev1 = hip_run(cogs, lin, 42, PERSISTENT, dedup="none", pause_before=300.0, pool=pool)
    second, ev2 = hip_run(cogs, first.lineages, 42, PERSISTENT, dedup="none", resume_after=300.0, pool=pool)
    assert np.concatenate([ev1, ev2]).tobytes() == ref["events"].tobytes()
    assert first.steps + second.steps == ref["steps"]


def test_pool_with_dedup_uses_the_lane_kernel(config1):
    announce("pool requested with a dedup cache: the one-lineage-per-lane kernel, invariants hold")
    cogs, orc, lin, ref = config1
    out, ev = hip_run(cogs, lin, 42, PERSISTENT, dedup=1.0, pool=128)
    assert_invariants(ev, ref["events"], None, "pool + dedup")
