"""GPU parity at the shapes of the benchmark configurations.

Windows of config 3 (128x128 map), config 4 (almost-infinite Gaussian
sigma=50 over a circle of radius 2000, here a 0.1% sample of it) and config
5's 256x256 map, short enough for the oracle to finish in about a second.
Bar: with DedupCache::None both forms reproduce the oracle's sorted
speciation stream, step count, time and remaining lineages bit for bit; with
the reference's default cache the plain form does the same and the
persistent form keeps the order-independent invariants.
"""
import os

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from parity_utils import assert_exact, assert_invariants, hip_run, map_scenarios, oracle_run

pytestmark = pytest.mark.gpu

PLAIN, PERSISTENT = nh.PLAIN, nh.PERSISTENT
CENTRE = 2**31 - 1


def announce(what):
    print(f"[gpu case] {what}", flush=True)


def same_remaining(out, ref):
    rem = np.sort(out.lineages, order=["global_reference"])
    orem = np.sort(ref["remaining"], order=["global_reference"])
    return rem.tobytes() == orem.tobytes()


@pytest.fixture(scope="module")
def config4_sample():
    cogs_nu = {nu: nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0).new(nu) for nu in (1e-6, 1e-3)}
    orc = O.Scenario.almost_infinite_normal(50.0)
    lin = orc.origins(0.001, circle=(CENTRE, CENTRE, 2000))
    assert 12000 < len(lin) < 13200
    return cogs_nu, orc, lin


@pytest.mark.parametrize("nu", [1e-6, 1e-3])
@pytest.mark.parametrize("form", [PLAIN, PERSISTENT])
def test_config4_sample_window_dedup_none(config4_sample, form, nu):
    announce(f"config4 0.1% sample, pause 400, nu {nu}, dedup None, form {form}")
    cogs_nu, orc, lin = config4_sample
    ref = oracle_run(orc, lin, 42, nu, dedup="none", pause_before=400.0)
    out, ev = hip_run(cogs_nu[nu], lin, 42, form, dedup="none", pause_before=400.0)
    assert_exact(ref, out, ev, "config4 sample window")
    assert same_remaining(out, ref)


def test_config4_sample_window_dedup(config4_sample):
    announce("config4 0.1% sample, pause 400, nu 1e-3, dedup Relative(1.0), both forms")
    cogs_nu, orc, lin = config4_sample
    none = oracle_run(orc, lin, 42, 1e-3, dedup="none", pause_before=400.0)
    ref = oracle_run(orc, lin, 42, 1e-3, dedup=1.0, pause_before=400.0)
    out, ev = hip_run(cogs_nu[1e-3], lin, 42, PLAIN, dedup=1.0, pause_before=400.0)
    assert_exact(ref, out, ev, "config4 sample dedup plain")
    out, ev = hip_run(cogs_nu[1e-3], lin, 42, PERSISTENT, dedup=1.0, pause_before=400.0)
    assert_invariants(ev, none["events"], ref["events"], "config4 sample dedup persistent")


@pytest.fixture(scope="module")
def config3():
    hip, orc = map_scenarios(128, 128, seed=42)
    lin = orc.origins(1.0)
    assert len(lin) == 104294
    return hip.new(1e-5), orc, lin


@pytest.mark.parametrize("form", [PLAIN, PERSISTENT])
def test_config3_window(config3, form):
    announce(f"config3 128x128 map, pause 100, form {form}")
    cogs, orc, lin = config3
    ref = oracle_run(orc, lin, 42, 1e-5, dedup="none", pause_before=100.0)
    out, ev = hip_run(cogs, lin, 42, form, dedup="none", pause_before=100.0)
    assert_exact(ref, out, ev, "config3 window")
    assert same_remaining(out, ref)
    refd = oracle_run(orc, lin, 42, 1e-5, dedup=1.0, pause_before=100.0)
    outd, evd = hip_run(cogs, lin, 42, form, dedup=1.0, pause_before=100.0)
    if form == PLAIN:
        assert_exact(refd, outd, evd, "config3 dedup plain")
    else:
        assert_invariants(evd, ref["events"], refd["events"], "config3 dedup persistent")


@pytest.mark.parametrize("form", [PLAIN, PERSISTENT])
def test_map256_window(form):
    announce(f"256x256 map (config 5 generator), pause 20, form {form}")
    hip, orc = map_scenarios(256, 256, seed=42)
    lin = orc.origins(1.0)
    cogs = hip.new(1e-5)
    ref = oracle_run(orc, lin, 42, 1e-5, dedup="none", pause_before=20.0)
    out, ev = hip_run(cogs, lin, 42, form, dedup="none", pause_before=20.0)
    assert_exact(ref, out, ev, "map256 window")
    assert same_remaining(out, ref)


def _host_threads():
    n = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, share) if share > 0 else min(n, 16))


def oracle_none_partitioned(orc, lin, seed, nu, pause, threads):
    """The dedup-None oracle over interleaved slices of the lineages, one
    thread each (the C oracle runs without the GIL).  Lineages never interact
    and with DedupCache::None nothing depends on which others share a run, so
    the merged, PackedEvent-sorted stream, the step total and the remaining
    lineages equal the monolithic run's."""
    from concurrent.futures import ThreadPoolExecutor
    parts = [lin[r::threads] for r in range(threads)]
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda p: oracle_run(orc, p, seed, nu, dedup="none", pause_before=pause), parts))
    ev = np.concatenate([r["events"] for r in res])
    ev = ev[np.lexsort((ev["global_reference"], ev["prior_time"], ev["index"], ev["y"], ev["x"],
                        ev["event_time"]))]
    return dict(events=ev, steps=sum(r["steps"] for r in res), time=max(r["time"] for r in res),
                paused=any(r["paused"] for r in res),
                remaining=np.concatenate([r["remaining"] for r in res]))


def test_config4_sample_heavy_tail_window(config4_sample):
    """config 4's heavy tail: T = 2e5 (about 1e5 dependent steps per
    surviving lineage, 100x the headline window) on the 0.1 % sample at
    nu = 1e-6: the persistent form without dedup reproduces the oracle bit
    for bit, and with the reference's default cache keeps the species and
    the unique events"""
    threads = _host_threads()
    announce(f"config4 0.1% sample, pause 2e5, nu 1e-6: oracle on {threads} threads, persistent dedup None / 1.0")
    cogs_nu, orc, lin = config4_sample
    ref = oracle_none_partitioned(orc, lin, 42, 1e-6, 2e5, threads)
    out, ev = hip_run(cogs_nu[1e-6], lin, 42, PERSISTENT, dedup="none", pause_before=2e5)
    assert_exact(ref, out, ev, "config4 sample tail dedup None")
    assert same_remaining(out, ref)
    out, ev = hip_run(cogs_nu[1e-6], lin, 42, PERSISTENT, dedup=1.0, pause_before=2e5)
    assert_invariants(ev, ref["events"], None, "config4 sample tail dedup persistent")
