"""Lineage files (nh_lineage_file_write / nh_lineage_file_read).

LineageFileSaver / LineageFileLoader (necsim/impls/std/src/lineage_file/)
serialise a Vec<Lineage> with bincode 1.3's `bincode::options()`: varint
integers (v < 251 in one byte; otherwise a tag 251 / 252 / 253 followed by v
as little-endian u16 / u32 / u64), f64 as 8 little-endian bytes, sequences
as their varint length then the elements, structs as their fields in order
(Lineage: global_reference, last_event_time, indexed_location {x, y, index};
necsim/core/src/lineage.rs:88-108).  The expected bytes are laid out here
from that format, independently of the writer, with values on every side
of the varint boundaries.
"""
import os
import struct

import numpy as np
import pytest

import necsim_hip as nh
from necsim_hip import _ffi


def varint(v):
    if v < 251:
        return bytes([v])
    if v < 1 << 16:
        return bytes([251]) + struct.pack("<H", v)
    if v < 1 << 32:
        return bytes([252]) + struct.pack("<I", v)
    return bytes([253]) + struct.pack("<Q", v)


def lineages(rows):
    out = np.zeros(len(rows), nh.LINEAGE_DTYPE)
    for i, (ref, t, x, y, idx) in enumerate(rows):
        out[i] = (ref, t, x, y, idx, 0)
    return out


EDGES = [0, 1, 250, 251, 255, 65535, 65536, 2**32 - 1, 2**32, 2**64 - 1]
SAMPLE = lineages([(r, 0.5 * k, EDGES[k % 8] % 2**32, EDGES[(k + 3) % 8] % 2**32, k % 3)
                   for k, r in enumerate(EDGES)])


def test_lineage_file_bytes_follow_bincode_varint(tmp_path):
    path = str(tmp_path / "lineages.bincode")
    nh.save_lineages(path, SAMPLE)
    want = varint(len(SAMPLE))
    for l in SAMPLE:
        want += varint(int(l["global_reference"])) + struct.pack("<d", float(l["last_event_time"]))
        want += varint(int(l["x"])) + varint(int(l["y"])) + varint(int(l["index"]))
    assert open(path, "rb").read() == want


def test_lineage_file_round_trip_and_rules(tmp_path):
    path = str(tmp_path / "l")
    nh.save_lineages(path, SAMPLE)
    assert nh.load_lineages(path).tobytes() == SAMPLE.tobytes()
    with pytest.raises(_ffi.NecsimHipError):  # create_new: an existing file is not overwritten
        nh.save_lineages(path, SAMPLE)
    raw = open(path, "rb").read()
    open(str(tmp_path / "short"), "wb").write(raw[:-3])
    open(str(tmp_path / "long"), "wb").write(raw + b"\0")
    for bad in ("short", "long"):
        with pytest.raises(_ffi.NecsimHipError):
            nh.load_lineages(str(tmp_path / bad))
    empty = str(tmp_path / "empty")
    nh.save_lineages(empty, SAMPLE[:0])
    assert open(empty, "rb").read() == b"\0" and len(nh.load_lineages(empty)) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("form", [nh.PLAIN, nh.PERSISTENT])
def test_pause_save_load_resume_continuity(tmp_path, form):
    print(f"[gpu case] pause -> lineage file -> resume, form {form}", flush=True)
    cogs = nh.NonSpatialScenario((40, 40), 1).new(2e-3)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form)
    full = nh.EventCollector()
    out = nh.HipIndependentAlgorithm.initialise_and_simulate(args, nh.WyHash.seed_from_u64(3), cogs,
                                                             nh.OriginPreSampler.all(), None, full)
    first, second = nh.EventCollector(), nh.EventCollector()
    out1 = nh.HipIndependentAlgorithm.initialise_and_simulate(args, nh.WyHash.seed_from_u64(3), cogs,
                                                              nh.OriginPreSampler.all(), 300.0, first)
    assert out1.paused and len(out1.lineages) > 0
    path = str(tmp_path / "paused.bincode")
    nh.save_lineages(path, out1.lineages)
    out2 = nh.HipIndependentAlgorithm.resume_and_simulate(args, nh.WyHash.seed_from_u64(3), cogs,
                                                          nh.load_lineages(path), resume_after=300.0,
                                                          reporter=second)
    both = np.sort(np.concatenate([first.events, second.events]),
                   order=["event_time", "x", "y", "index", "prior_time", "global_reference"])
    assert both.tobytes() == np.ascontiguousarray(full.events, dtype=nh.EVENT_DTYPE).tobytes()
    assert out1.steps + out2.steps == out.steps and not out2.paused
