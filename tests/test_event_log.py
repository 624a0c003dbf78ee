"""Event-log segments (nh_event_log_write_segment, EventLogReporter).

A segment is what EventLogRecorder::sort_and_write_segment writes
(necsim/impls/std/src/event_log/recorder.rs:174-208): bincode 1.3 with its
default options (little-endian, fixed-width integers, usize as u64, bool as
one byte, newtypes as their inner value) of EventLogHeader
(event_log/mod.rs:9-19) and then of each PackedEvent
(necsim/core/src/event.rs:16-29).  The expected bytes below are laid out from
that format with struct.pack, independently of the writer; a speciation is a
PackedEvent with the sign bit of prior_time set (event.rs:103-115), pinned at
-0.0 by the reference's own make_negative test (event.rs:345-350).
"""
import os
import struct

import numpy as np
import pytest

import necsim_hip as nh
from necsim_hip import _ffi

HEADER = struct.Struct("<ddQ??")
PACKED = struct.Struct("<QQdIIIIIIQ")


def events(rows):
    ev = np.zeros(len(rows), nh.EVENT_DTYPE)
    for i, (ref, prior, t, x, y, idx) in enumerate(rows):
        ev[i] = (ref, prior, t, x, y, idx, 0)
    return ev


def expected_segment(ev, with_speciation=True, with_dispersal=False):
    out = HEADER.pack(float(ev["event_time"][0]), float(ev["event_time"][-1]), len(ev), with_speciation,
                      with_dispersal)
    for e in ev:
        prior_bits = struct.unpack("<Q", struct.pack("<d", float(e["prior_time"])))[0] | (1 << 63)
        out += PACKED.pack(int(e["global_reference"]), prior_bits, float(e["event_time"]), int(e["x"]),
                           int(e["y"]), int(e["index"]), int(e["x"]), int(e["y"]), int(e["index"]),
                           int(e["global_reference"]))
    return out


def read_segment(path):
    """decode a segment into (header tuple, speciation events)"""
    raw = open(path, "rb").read()
    head = HEADER.unpack_from(raw, 0)
    n = head[2]
    assert len(raw) == HEADER.size + n * PACKED.size
    rows = []
    for i in range(n):
        ref, pbits, t, x, y, idx, tx, ty, tidx, coal = PACKED.unpack_from(raw, HEADER.size + i * PACKED.size)
        assert pbits >> 63 == 1 and struct.unpack("<d", struct.pack("<d", t))[0] > 0, "not a speciation"
        assert (tx, ty, tidx) == (x, y, idx) and coal == ref
        prior = struct.unpack("<d", struct.pack("<Q", pbits & ((1 << 63) - 1)))[0]
        rows.append((ref, prior, t, x, y, idx))
    return head, events(rows)


SAMPLE = events([(7, 0.0, 1.25, 3, 4, 0), (2, 0.5, 2.0, 1, 1, 2), (9, 1.75, 2.0, 1, 1, 2)])


def test_segment_bytes_follow_the_bincode_layout(tmp_path):
    path = str(tmp_path / "0")
    _ffi.check(_ffi.lib().nh_event_log_write_segment(path.encode(), _ffi.ptr(SAMPLE), len(SAMPLE)))
    raw = open(path, "rb").read()
    assert len(raw) == 26 + 3 * 56
    assert raw == expected_segment(SAMPLE)
    # prior_time 0.0 of the first event is stored as -0.0 (event.rs:345-350)
    assert raw[26 + 8:26 + 16] == struct.pack("<d", -0.0)


def test_segment_rejects_unsorted_or_empty(tmp_path):
    path = str(tmp_path / "0").encode()
    bad = SAMPLE[[1, 0, 2]]
    with pytest.raises(_ffi.NecsimHipError) as e:
        _ffi.check(_ffi.lib().nh_event_log_write_segment(path, _ffi.ptr(bad), len(bad)))
    assert e.value.code == -1  # NH_E_INVALID
    with pytest.raises(_ffi.NecsimHipError):
        _ffi.check(_ffi.lib().nh_event_log_write_segment(path, _ffi.ptr(SAMPLE), 0))


def test_reporter_segments_like_the_recorder(tmp_path):
    rng = np.random.default_rng(3)
    n = 10
    ev = events([(int(i + 1), float(rng.uniform(0, 1)), float(1 + i + rng.uniform(0, 0.5)),
                  int(rng.integers(0, 9)), int(rng.integers(0, 9)), 0) for i in range(n)])
    log = nh.EventLogReporter(str(tmp_path / "log"), segment_capacity=4)
    for chunk in (ev[:3], ev[3:8], ev[8:]):
        log.report_speciation(chunk)
    log.close()
    names = sorted(os.listdir(tmp_path / "log"), key=int)
    assert names == ["0", "1", "2"]
    got = []
    for name, size in zip(names, (4, 4, 2)):
        head, seg = read_segment(str(tmp_path / "log" / name))
        assert head[2] == size and head[3] is True and head[4] is False
        assert head[0] == seg["event_time"][0] and head[1] == seg["event_time"][-1]
        got.append(seg)
    assert np.concatenate(got).tobytes() == ev.tobytes()


def test_reporter_directory_rules(tmp_path):
    os.mkdir(tmp_path / "exists")
    with pytest.raises(FileExistsError):
        nh.EventLogReporter(str(tmp_path / "exists"))
    log = nh.EventLogReporter(str(tmp_path / "nested" / "empty"))
    log.close()
    assert not os.path.exists(tmp_path / "nested" / "empty")


@pytest.mark.gpu
def test_simulation_event_log_matches_reported_stream(tmp_path):
    print("[gpu case] config-1-like run recorded to an event log", flush=True)
    cogs = nh.NonSpatialScenario((30, 30), 2).new(1e-2)
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    col = nh.EventCollector()
    nh.HipIndependentAlgorithm.initialise_and_simulate(args, nh.WyHash.seed_from_u64(5), cogs,
                                                       nh.OriginPreSampler.all(), None, col)
    log = nh.EventLogReporter(str(tmp_path / "log"), segment_capacity=500)
    nh.HipIndependentAlgorithm.initialise_and_simulate(args, nh.WyHash.seed_from_u64(5), cogs,
                                                       nh.OriginPreSampler.all(), None, log)
    log.close()
    names = sorted(os.listdir(tmp_path / "log"), key=int)
    got = np.concatenate([read_segment(str(tmp_path / "log" / n))[1] for n in names])
    want = np.ascontiguousarray(col.events, dtype=nh.EVENT_DTYPE)
    assert len(names) == -(-len(want) // 500)
    assert got.tobytes() == want.tobytes()


PACKED_ROWS = [  # (ref, prior bits, event bits, origin, target)
    (7, 0x8000000000000000, 0x3FF4000000000000, (3, 4, 0), (3, 4, 0)),       # speciation, prior -0.0, t 1.25
    (2, 0x3FE0000000000000, 0xC000000000000000, (1, 1, 2), (5, 0, 1)),       # dispersal 0.5 -> -2.0
    (9, 0x3FFC000000000000, 0xC000000000000000, (1, 1, 2), (5, 0, 3)),       # dispersal 1.75 -> -2.0
]


def packed(rows):
    ev = np.zeros(len(rows), nh.PACKED_EVENT_DTYPE)
    raw = ev.view(np.uint64).reshape(len(rows), 7)
    for i, (ref, pb, tb, o, t) in enumerate(rows):
        ev[i]["global_reference"], ev[i]["coalescence"] = ref, ref
        raw[i, 1], raw[i, 2] = pb, tb
        ev[i]["origin_x"], ev[i]["origin_y"], ev[i]["origin_index"] = o
        ev[i]["target_x"], ev[i]["target_y"], ev[i]["target_index"] = t
    return ev


def test_packed_segment_bytes_follow_the_bincode_layout(tmp_path):
    ev = packed(PACKED_ROWS)
    path = str(tmp_path / "0")
    _ffi.check(_ffi.lib().nh_event_log_write_packed_segment(path.encode(), _ffi.ptr(ev), len(ev), 1, 1))
    raw = open(path, "rb").read()
    want = HEADER.pack(1.25, 2.0, 3, True, True)  # header times are magnitudes (event.rs:33-36)
    for ref, pb, tb, o, t in PACKED_ROWS:
        want += struct.pack("<QQQIIIIIIQ", ref, pb, tb, *o, *t, ref)
    assert raw == want
    # the dispersal events tie on time and origin: the target orders them
    bad = ev[[0, 2, 1]]
    with pytest.raises(_ffi.NecsimHipError):
        _ffi.check(_ffi.lib().nh_event_log_write_packed_segment(path.encode(), _ffi.ptr(bad), 3, 1, 1))


def test_dispersal_reporter_segments(tmp_path):
    ev = packed(PACKED_ROWS)
    log = nh.EventLogReporter(str(tmp_path / "log"), segment_capacity=2, report_dispersal=True)
    with pytest.raises(TypeError):
        log.report_speciation(SAMPLE)
    log.report_packed(ev[[1, 0]])  # one segment, sorted on write
    log.report_packed(ev[2:])
    log.close()
    raw0 = open(tmp_path / "log" / "0", "rb").read()
    raw1 = open(tmp_path / "log" / "1", "rb").read()
    assert HEADER.unpack_from(raw0, 0) == (1.25, 2.0, 2, True, True)
    assert raw0[26:] == ev[:2].tobytes() and raw1[26:] == ev[2:].tobytes()


def test_packed_helpers():
    ev = packed(PACKED_ROWS)
    assert nh.is_speciation(ev).tolist() == [True, False, False]
    sp = nh.speciation_events(ev)
    assert len(sp) == 1 and sp[0]["prior_time"] == 0.0 and sp[0]["event_time"] == 1.25
    assert np.signbit(sp["prior_time"]).tolist() == [False]
    assert nh.packed_order(ev[[2, 1, 0]]).tolist() == [2, 1, 0]
