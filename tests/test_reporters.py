"""Reporters over PackedEvent streams: EventCounterReporter and CsvReporter.

The checker of the counter is a sequential restatement of
EventCounterReporter::report_speciation / report_dispersal
(necsim/plugins/common/src/event_counter.rs:58-109), one TypedEvent at a time
(PackedEvent -> TypedEvent by the sign bits, necsim/core/src/event.rs:186-229).
Streams: the oracle's dispersal-reporting runs (coalesced lineages repeat
their events with other priors) and a hand-made stream that reaches every
counter, with speciations and dispersals interleaved inside one key run.
"""
import struct

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O

PK = nh.PACKED_EVENT_DTYPE


def typed(e):
    """(kind, key, prior, interaction, self_location) of one PackedEvent"""
    pb = struct.unpack("<Q", struct.pack("<d", float(e["prior_time"])))[0]
    eb = struct.unpack("<Q", struct.pack("<d", float(e["event_time"])))[0]
    prior_pos, event_pos = pb >> 63 == 0, eb >> 63 == 0
    time, prior = abs(float(e["event_time"])), abs(float(e["prior_time"]))
    origin = (int(e["origin_x"]), int(e["origin_y"]), int(e["origin_index"]))
    target = (int(e["target_x"]), int(e["target_y"]), int(e["target_index"]))
    if not prior_pos and event_pos:
        return "speciation", (origin, time), prior, None, None
    interaction = {(False, False): "none", (True, False): "maybe", (True, True): "coalescence"}[(prior_pos,
                                                                                                 event_pos)]
    return "dispersal", (origin, time, target), prior, interaction, origin == target


def reference_counts(streams):
    c = dict.fromkeys(nh.reporters.COUNTER_FIELDS, 0)
    last = {"speciation": None, "dispersal": None}
    last_parent_prior = None
    for stream in streams:
        for e in stream:
            kind, key, prior, interaction, self_loc = typed(e)
            c["raw_total"] += 1
            if key == last[kind]:
                if prior != last_parent_prior:
                    c["late_coalescence"] += 1
                last_parent_prior = prior
                continue
            last[kind] = key
            last_parent_prior = prior
            if kind == "speciation":
                c["speciation"] += 1
            elif interaction == "maybe":
                c["late_dispersal_coalescence"] += 1
            else:
                coal = interaction == "coalescence"
                name = ("self_" if self_loc else "out_") + ("coalescence" if coal else "dispersal")
                c[name] += 1
    return c


def host_counts(streams):
    rep = nh.EventCounterReporter()
    for s in streams:
        rep.report_packed(s)
    return rep.counts


def hand_stream():
    """every counter, in PackedEvent order; speciations and self-dispersals
    share a (time, origin) run"""
    rows = []  # (ref, prior, t, origin, target, (prior_neg, event_neg))
    SP, MAYBE, NONE, COAL = (1, 0), (0, 1), (1, 1), (0, 0)
    o, o2 = (1, 1, 0), (2, 0, 1)
    rows += [(1, 0.5, 1.0, o, o, SP), (2, 0.5, 1.0, o, o, SP), (3, 0.7, 1.0, o, o, SP)]  # dup, dup+late
    rows += [(4, 0.1, 2.0, o, (0, 0, 0), NONE), (5, 0.1, 2.0, o, o, NONE), (6, 0.2, 2.0, o, o, SP),
             (7, 0.3, 2.0, o, o, NONE), (8, 0.3, 2.0, o, o, SP), (9, 0.3, 2.0, o, (9, 9, 9), COAL)]
    rows += [(10, 0.4, 3.0, o2, o2, COAL), (11, 0.4, 3.0, o2, (5, 5, 5), MAYBE), (12, 0.9, 3.0, o2, (5, 5, 5), MAYBE),
             (13, 1.0, 3.5, o2, (5, 5, 5), COAL), (14, 1.0, 3.5, o2, (5, 5, 5), COAL)]
    ev = np.zeros(len(rows), PK)
    for i, (ref, p, t, org, tgt, (pn, en)) in enumerate(rows):
        ev[i]["global_reference"] = ref
        ev[i]["prior_time"] = -p if pn else p
        ev[i]["event_time"] = -t if en else t
        ev[i]["origin_x"], ev[i]["origin_y"], ev[i]["origin_index"] = org
        ev[i]["target_x"], ev[i]["target_y"], ev[i]["target_index"] = tgt
        ev[i]["coalescence"] = ref
    return ev[nh.packed_order(ev)]


def test_hand_stream_reaches_every_counter():
    ev = hand_stream()
    want = reference_counts([ev])
    assert all(want[f] > 0 for f in nh.reporters.COUNTER_FIELDS), want
    assert host_counts([ev]) == want
    assert host_counts([ev[:4], ev[4:9], ev[9:]]) == want  # the state carries across reports


@pytest.fixture(scope="module")
def coalescing_stream():
    """a dedup-None run where coalesced lineages repeat each other's events"""
    s = O.Scenario.non_spatial(6, 6, 2)
    lin = s.origins(1.0)
    full = O.simulate_packed(s, O.make_params(11, 0.02, dedup=None, report_dispersal=True), lin)
    return np.ascontiguousarray(full["events"]).view(PK)


def test_counter_on_a_coalescing_run(coalescing_stream):
    ev = coalescing_stream
    want = reference_counts([ev])
    duplicates = want["raw_total"] - want["speciation"] - want["late_dispersal_coalescence"]
    assert duplicates > 1000 and want["late_coalescence"] > 10 and want["speciation"] > 0
    assert want["late_dispersal_coalescence"] > 0  # every Independent dispersal is Maybe
    assert want["out_dispersal"] == want["self_dispersal"] == 0
    for cuts in ([], [1], [17, 1000, 1001]):
        parts = np.split(ev, cuts)
        assert host_counts(parts) == want


def test_counter_summary_text():
    rep = nh.EventCounterReporter()
    rep.report_packed(hand_stream())
    text = rep.summary()
    c = rep.counts
    assert text.startswith("Event Summary:\n - Total #individuals:\n")
    assert f"   - raw:\n     {c['raw_total']}\n" in text
    assert (f"   - detected late, no coalescence:\n     {c['late_dispersal_coalescence'] - c['late_coalescence']}"
            in text)


def test_rust_f64_display():
    # Rust's Display of f64: shortest round-trip digits, positional, no '.0'
    cases = {2.0: "2", 1.25: "1.25", 1e-7: "0.0000001", 1.5e20: "150000000000000000000", 0.1: "0.1",
             123456789.123: "123456789.123", 2.0**-10: "0.0009765625", 1e16: "10000000000000000"}
    for x, want in cases.items():
        assert nh.rust_f64_display(x) == want
    for x in np.random.default_rng(1).uniform(0, 1e6, 200):
        assert float(nh.rust_f64_display(x)) == x


def test_csv_reporter_lines(tmp_path):
    ev = hand_stream()
    path = tmp_path / "events.csv"
    rep = nh.CsvReporter(str(path))
    rep.report_packed(ev[:5])
    rep.report_packed(ev[5:])
    rep.close()
    lines = path.read_text().splitlines()
    assert lines[0] == "reference,time,x,y,index,type"
    assert len(lines) == len(ev) + 1
    for line, e in zip(lines[1:], ev):
        kind = typed(e)[0]
        t = abs(float(e["event_time"]))
        ts = str(int(t)) if t == int(t) else repr(t)
        want = f"{int(e['global_reference'])},{ts},{int(e['origin_x'])},{int(e['origin_y'])}," \
               f"{int(e['origin_index'])},{kind[0]}"
        assert line == want
