"""GPU parity at the shapes of the benchmark configurations.

Windows of config 3 (128x128 map), config 4 (almost-infinite Gaussian
sigma=50 over a circle of radius 2000, here a 0.1% sample of it) and config
5's 256x256 map, short enough for the oracle to finish in about a second.
Bar: with DedupCache::None both forms reproduce the oracle's sorted
speciation stream, step count, time and remaining lineages bit for bit; with
the reference's default cache the plain form does the same and the
persistent form keeps the order-independent invariants.
"""
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
