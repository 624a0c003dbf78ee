"""The water-level event sort on the GPU against numpy's lexsort.

LiveWaterLevelReporterProxy::advance_water_level sorts each round's events
by the PackedEvent order (necsim/core/src/event.rs:256-280): event_time,
origin (x, y, index), target (= origin for speciation), prior_time, global
lineage reference.  numpy.lexsort over the same fields is an independent
implementation of that lexicographic order.  The cases stress the paths of
the device sort: distinct times (one radix sort), runs of equal times (the
full per-field radix sort that orders ties), runs of whole duplicates that
differ only in the reference (what coalesced lineages produce when dedup is
off), and the graph-capturable variant that gates the tie passes on device.
"""
import zlib

import numpy as np
import pytest

from necsim_hip import EVENT_DTYPE, _ffi

pytestmark = pytest.mark.gpu


def device_sort(ctx, ev, device_sized):
    ev = np.ascontiguousarray(ev, dtype=EVENT_DTYPE)
    out = np.zeros_like(ev)
    _ffi.check(_ffi.lib().nh_debug_sort_events(ctx.handle, _ffi.ptr(ev), len(ev), int(device_sized),
                                                _ffi.ptr(out)))
    return out


def expected(ev):
    order = np.lexsort((ev["global_reference"], ev["prior_time"], ev["index"], ev["y"], ev["x"],
                        ev["event_time"]))
    return ev[order]


def random_events(rng, n, n_times=None, n_xy=1 << 20, n_index=1, n_prior=None):
    ev = np.zeros(n, EVENT_DTYPE)
    ev["global_reference"] = rng.permutation(n).astype(np.uint64) + 1
    if n_times is None:
        ev["event_time"] = rng.uniform(0.0, 2000.0, n)
    else:
        ev["event_time"] = rng.choice(rng.uniform(0.0, 2000.0, n_times), n)
    centre = 2**31 - 1
    ev["x"] = centre + rng.integers(-n_xy, n_xy, n)
    ev["y"] = centre + rng.integers(-n_xy, n_xy, n)
    ev["index"] = rng.integers(0, n_index, n)
    if n_prior is None:
        ev["prior_time"] = ev["event_time"] * rng.uniform(0.0, 1.0, n)
    else:
        ev["prior_time"] = rng.choice(np.linspace(0.0, 1.0, n_prior), n)
    return ev


def duplicate_runs(rng, n_groups, size):
    """groups of coalesced lineages: identical events except the reference"""
    base = random_events(rng, n_groups)
    ev = np.repeat(base, size)
    ev["global_reference"] = rng.permutation(len(ev)).astype(np.uint64) * 7919 + 3
    return ev[rng.permutation(len(ev))]


CASES = {
    "single": lambda rng: random_events(rng, 1),
    "two_equal_times": lambda rng: random_events(rng, 2, n_times=1),
    "distinct": lambda rng: random_events(rng, 5000),
    "all_times_equal": lambda rng: random_events(rng, 4097, n_times=1, n_xy=3, n_index=3, n_prior=5),
    "few_times": lambda rng: random_events(rng, 20000, n_times=37, n_xy=2, n_index=2, n_prior=3),
    "duplicate_runs": lambda rng: duplicate_runs(rng, 60, 300),
    "zero_prior_times": lambda rng: random_events(rng, 3000, n_times=500, n_prior=1),
    "large_with_ties": lambda rng: np.concatenate([random_events(rng, 1 << 19),
                                                   duplicate_runs(rng, 1000, 64)]),
    # runs that share time and origin but not prior time (ranked in place)
    "coalesced_runs": lambda rng: coalesced_runs(rng, 3000, 40),
    # runs longer than the in-place ranking handles (per-field passes)
    "long_duplicate_runs": lambda rng: duplicate_runs(rng, 3, 5000),
    # the staged radix passes rank tiles of 4096 keys, and their workgroups
    # take tiles in an order that spreads the 8 XCDs over consecutive tiles:
    # one full tile, one key short of it, and a tile count that is not a
    # multiple of 8 (grid blocks past the last tile exit)
    "one_tile": lambda rng: random_events(rng, 4096, n_times=3000),
    "one_tile_short": lambda rng: random_events(rng, 4095, n_times=3000),
    "nine_tiles_and_a_bit": lambda rng: np.concatenate([random_events(rng, 9 * 4096 + 17 - 600, n_times=20000),
                                                        duplicate_runs(rng, 20, 30)]),
}


def coalesced_runs(rng, n_groups, size):
    ev = duplicate_runs(rng, n_groups, size)
    ev["prior_time"] = ev["event_time"] * rng.choice([0.0, 0.25, 0.5, 0.75], len(ev))
    return ev


@pytest.mark.parametrize("device_sized", [0, 1], ids=["eager", "capturable"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_event_sort_matches_lexsort(ctx, case, device_sized):
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    ev = CASES[case](rng)
    got = device_sort(ctx, ev, device_sized)
    assert got.tobytes() == expected(ev).tobytes()


def test_event_sort_is_a_permutation_of_the_input(ctx):
    rng = np.random.default_rng(5)
    ev = duplicate_runs(rng, 10, 50)
    got = device_sort(ctx, ev, 1)
    assert np.array_equal(np.sort(got["global_reference"]), np.sort(ev["global_reference"]))


def test_event_sort_empty(ctx):
    out = np.zeros(0, EVENT_DTYPE)
    _ffi.check(_ffi.lib().nh_debug_sort_events(ctx.handle, None, 0, 0, None))
    assert len(out) == 0
