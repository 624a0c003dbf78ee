"""Alias tables built on the GPU (nh_tables_create_from_map) against the
host builder (nh_build_alias_tables + nh_tables_create), byte for byte.

The host builder is itself checked against the oracle's restatement of
InMemoryPackedSeparableAliasDispersalSampler::new / InMemoryPackedAliasDispersalSampler::new
in test_host_abi.py (and the oracle's alias tables against their weights in
test_oracle_pins.py).  The maps cover both separable self-dispersal branches
(self-dispersal below one bucket, and at or above it), rows that only
self-disperse, zero-habitat cells, and the contract errors.
"""
import numpy as np
import pytest

import necsim_hip as nh
from necsim_hip import _ffi

pytestmark = pytest.mark.gpu


def announce(what):
    print(f"[gpu case] {what}", flush=True)


def tables_pair(ctx, scn):
    host = nh.Tables(ctx, scn.new(1e-3, host_tables=True))
    dev = nh.Tables(ctx, scn.new(1e-3, host_tables=False))
    return host, dev


MAPS = [
    (8, 8, dict(seed=5, sigma=1.5, cutoff=2.5)),
    (16, 16, dict(seed=11, sigma=0.4, cutoff=1.5)),   # self-dispersal above one bucket
    (32, 32, dict(seed=3, zero_fraction=0.5)),
    (64, 64, dict(seed=42)),
    (128, 128, dict(seed=42)),
    (256, 256, dict(seed=42)),
]


@pytest.mark.parametrize("packed", [False, True], ids=["separable", "packed"])
@pytest.mark.parametrize("w,h,kw", MAPS, ids=[f"{m[0]}x{m[1]}" for m in MAPS])
def test_device_tables_equal_host_tables(ctx, w, h, kw, packed):
    announce(f"device alias build {w}x{h} {kw} packed={packed}")
    hab, rp, col, wt = nh.generate_map(w, h, **kw)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(w, h, hab, rp, col, wt, packed=packed)
    host, dev = tables_pair(ctx, scn)
    hc, ha = host.export()
    dc, da = dev.export()
    assert len(hc) == w * h and len(ha) > 0
    assert hc.tobytes() == dc.tobytes(), "cell records differ"
    assert ha.tobytes() == da.tobytes(), "atoms differ"


@pytest.mark.parametrize("w,kw,branch", [(8, dict(seed=5, sigma=1.5, cutoff=2.5), "below"),
                                          (16, dict(seed=11, sigma=0.4, cutoff=1.5), "above")])
def test_self_dispersal_branches_are_exercised(ctx, w, kw, branch):
    """below one bucket, self-dispersal stays an atom, flipped into the last
    atom's k slot (non_self = its other event); at or above, it is dropped
    from the atoms (non_self = the cell itself)"""
    hab, rp, col, wt = nh.generate_map(w, w, **kw)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(w, w, hab, rp, col, wt)
    cells, atoms = nh.Tables(ctx, scn.new(1e-3, host_tables=False)).export()
    idx = np.arange(len(cells))
    live = (cells["habitat"] > 0) & (cells["self_p"] > 0)
    if branch == "below":
        rows = idx[live & (cells["non_self"] != idx)]
        assert len(rows) > 0
        last = atoms[cells["atom_start"][rows] + cells["atom_len"][rows] - 1]
        assert np.array_equal(last["k"], rows) and np.array_equal(last["e"], cells["non_self"][rows])
    else:
        rows = idx[live & (cells["non_self"] == idx) & (cells["self_p"] < 1.0)]
        assert len(rows) > 0
        for r in rows[:50]:
            a = atoms[cells["atom_start"][r]:cells["atom_start"][r] + cells["atom_len"][r]]
            assert not np.any(a["e"] == r) and not np.any(a["k"] == r)


def test_device_tables_simulate_like_host_tables(ctx):
    announce("64x64 map window on device-built vs host-built tables")
    hab, rp, col, wt = nh.generate_map(64, 64, seed=42)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(64, 64, hab, rp, col, wt)
    out = []
    for host_tables in (True, False):
        cogs = scn.new(1e-3, host_tables=host_tables)
        lin = cogs.sample_habitat(nh.OriginPreSampler.all())
        tables = nh.Tables(ctx, cogs)
        job = nh.Job(ctx, tables, nh.IndependentArguments(dedup_cache=nh.DedupCache.none()),
                     nh.WyHash.seed_from_u64(9), 1e-3, 0.5, pause_before=30.0)
        job.add_lineages(lin)
        ev = []
        fin = False
        while not fin:
            fin = job.step_round()
            ev.append(job.drain_events())
        o = job.outcome()
        out.append((np.concatenate(ev).tobytes(), o.steps, o.time))
        job.close()
    assert out[0] == out[1]


def _map_with(hab, entries, w, h):
    rp = np.zeros(w * h + 1, np.uint64)
    cols, wts = [], []
    for r in range(w * h):
        row = sorted(entries.get(r, []))
        cols += [c for c, _ in row]
        wts += [p for _, p in row]
        rp[r + 1] = len(cols)
    return (np.asarray(hab, np.uint32), rp, np.asarray(cols, np.uint32), np.asarray(wts, np.float64))


@pytest.mark.parametrize("case,code", [("to_zero_habitat", -5), ("habitable_no_dispersal", -5),
                                       ("column_outside", -1), ("column_repeated", -1), ("empty_habitat", -1)])
def test_device_build_contract_errors(ctx, case, code):
    announce(f"device alias build contract error {case}")
    w, h = 2, 2
    if case == "to_zero_habitat":
        hab, rp, col, wt = _map_with([1, 1, 0, 1], {0: [(2, 1.0)], 1: [(1, 1.0)], 3: [(3, 1.0)]}, w, h)
    elif case == "habitable_no_dispersal":
        hab, rp, col, wt = _map_with([1, 1, 1, 1], {0: [(0, 1.0)], 1: [(1, 1.0)], 3: [(3, 1.0)]}, w, h)
    elif case == "column_outside":
        hab, rp, col, wt = _map_with([1, 1, 1, 1], {0: [(7, 1.0)], 1: [(1, 1.0)], 2: [(2, 1.0)], 3: [(3, 1.0)]},
                                     w, h)
    elif case == "column_repeated":
        hab, rp, col, wt = _map_with([1, 1, 1, 1], {0: [(0, 1.0), (0, 2.0)], 1: [(1, 1.0)], 2: [(2, 1.0)],
                                                     3: [(3, 1.0)]}, w, h)
    else:
        hab, rp, col, wt = _map_with([0, 0, 0, 0], {}, w, h)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(w, h, hab, rp, col, wt)
    with pytest.raises(_ffi.NecsimHipError) as e:
        nh.Tables(ctx, scn.new(1e-3, host_tables=False))
    assert e.value.code == code
    # the host builder reports the same failure
    if case != "empty_habitat":
        with pytest.raises(_ffi.NecsimHipError) as e2:
            nh.build_alias_tables(w, h, hab, rp, col, wt)
        assert e2.value.code == code
