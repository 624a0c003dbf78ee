"""Reporters over the PackedEvent stream of dispersal-reporting runs.

Both mirror plugins of the reference that report speciation and dispersal
events: EventCounterReporter (necsim/plugins/common/src/event_counter.rs) and
CsvReporter (necsim/plugins/csv/src/lib.rs).  A reporter whose
REPORT_DISPERSAL is true receives each water-level round's events, sorted by
the PackedEvent order, through report_packed.
"""
import numpy as np

from . import _ffi

SIGN = np.uint64(1 << 63)

COUNTER_FIELDS = ("raw_total", "speciation", "out_dispersal", "self_dispersal", "out_coalescence",
                  "self_coalescence", "late_dispersal_coalescence", "late_coalescence")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


class EventCounterReporter:
    """EventCounterReporter (event_counter.rs:12-165).

    Each event is compared with the last reported event of its own kind
    (SpeciationEvent equality: origin and event time; DispersalEvent: origin,
    event time and target, necsim/core/src/event.rs:301-320).  An equal one
    is a late coalescence if its prior time differs from that of the event
    reported just before it; otherwise a speciation counts as such and a
    dispersal by its LineageInteraction (Maybe: late_dispersal_coalescence;
    else by self / out location and coalescence), event_counter.rs:58-109.

    With on_device=True, HipIndependentAlgorithm keeps the events on the
    device and counts them there (nh_job_event_counts); each simulate call
    then adds its counts, starting from a fresh comparison state."""

    REPORT_DISPERSAL = True

    def __init__(self, on_device=False):
        self.DEVICE_EVENT_COUNTER = bool(on_device)
        self.counts = dict.fromkeys(COUNTER_FIELDS, 0)
        self._last_prior = None  # last_parent_prior_time (magnitude bits)
        self._last = {True: None, False: None}  # last speciation / dispersal key

    def add_counts(self, counts):
        for f in COUNTER_FIELDS:
            self.counts[f] += int(counts[f])

    def report_packed(self, events):
        ev = np.ascontiguousarray(events, dtype=_ffi.PACKED_EVENT_DTYPE)
        n = len(ev)
        if n == 0:
            return
        pb, eb = _bits(ev["prior_time"]), _bits(ev["event_time"])
        prior_pos, event_pos = (pb & SIGN) == 0, (eb & SIGN) == 0
        speciation = ~prior_pos & event_pos
        pmag, tmag = pb & ~SIGN, eb & ~SIGN
        origin = [ev["origin_x"], ev["origin_y"], ev["origin_index"]]
        target = [ev["target_x"], ev["target_y"], ev["target_index"]]
        dup = np.zeros(n, bool)
        for kind in (True, False):
            idx = np.nonzero(speciation == kind)[0]
            if len(idx) == 0:
                continue
            cols = [tmag[idx]] + [c[idx] for c in origin] + ([] if kind else [c[idx] for c in target])
            same = np.ones(len(idx), bool)
            last = self._last[kind]
            for k, c in enumerate(cols):
                prev = np.empty_like(c)
                prev[1:] = c[:-1]
                if last is not None:
                    prev[0] = last[k]
                same &= c == prev
            if last is None:
                same[0] = False
            dup[idx] = same
            self._last[kind] = tuple(c[-1] for c in cols)
        prev_prior = np.empty(n, np.uint64)
        prev_prior[1:] = pmag[:-1]
        has_prev = np.ones(n, bool)
        if self._last_prior is None:
            has_prev[0] = False
            prev_prior[0] = 0
        else:
            prev_prior[0] = self._last_prior
        self._last_prior = pmag[-1]
        c = self.counts
        c["raw_total"] += n
        c["late_coalescence"] += int(np.count_nonzero(dup & (~has_prev | (prev_prior != pmag))))
        new = ~dup
        c["speciation"] += int(np.count_nonzero(new & speciation))
        disp = new & ~speciation
        maybe = disp & prior_pos & ~event_pos
        c["late_dispersal_coalescence"] += int(np.count_nonzero(maybe))
        rest = disp & ~maybe
        self_loc = (origin[0] == target[0]) & (origin[1] == target[1]) & (origin[2] == target[2])
        coal = prior_pos & event_pos
        c["self_coalescence"] += int(np.count_nonzero(rest & self_loc & coal))
        c["self_dispersal"] += int(np.count_nonzero(rest & self_loc & ~coal))
        c["out_coalescence"] += int(np.count_nonzero(rest & ~self_loc & coal))
        c["out_dispersal"] += int(np.count_nonzero(rest & ~self_loc & ~coal))

    def summary(self):
        """the text the reference's finalise logs (event_counter.rs:113-163)"""
        c = self.counts
        individuals = c["speciation"] + c["self_coalescence"] + c["out_coalescence"] + c["late_coalescence"]
        dedup = (c["speciation"] + c["self_coalescence"] + c["out_coalescence"] + c["self_dispersal"]
                 + c["out_dispersal"] + c["late_dispersal_coalescence"])
        return ("Event Summary:\n"
                f" - Total #individuals:\n   {individuals}\n"
                f" - Total #events:\n   - raw:\n     {c['raw_total']}\n   - deduplicated:\n     {dedup}\n"
                f" - Speciation:\n    {c['speciation']}\n"
                " - Dispersal:\n"
                f"   - same location, no coalescence:\n     {c['self_dispersal']}\n"
                f"   - same location, with coalescence:\n     {c['self_coalescence']}\n"
                f"   - new location, no coalescence:\n     {c['out_dispersal']}\n"
                f"   - new location, with coalescence:\n     {c['out_coalescence']}\n"
                f"   - detected late, no coalescence:\n     "
                f"{c['late_dispersal_coalescence'] - c['late_coalescence']}\n"
                f"   - detected late, with coalescence:\n     {c['late_coalescence']}\n")


def rust_f64_display(x):
    """f64 as Rust's Display prints it: the shortest digits that read back
    as x, positional (no exponent), no trailing '.0'"""
    return np.format_float_positional(float(x), unique=True, trim="-")


class CsvReporter:
    """CsvReporter (necsim/plugins/csv/src/lib.rs:71-163): one line
    `reference,time,x,y,index,type` per event, the origin's location, type
    's' (speciation) or 'd' (dispersal), the file created (truncated) when
    the reporter is made and flushed on close."""

    REPORT_DISPERSAL = True

    def __init__(self, output):
        self.output = output
        self._f = open(output, "w", newline="\n")
        self._f.write("reference,time,x,y,index,type\n")

    def report_packed(self, events):
        ev = np.ascontiguousarray(events, dtype=_ffi.PACKED_EVENT_DTYPE)
        pb, eb = _bits(ev["prior_time"]), _bits(ev["event_time"])
        speciation = ((pb & SIGN) != 0) & ((eb & SIGN) == 0)
        times = np.abs(ev["event_time"])
        lines = [f"{int(r)},{rust_f64_display(t)},{int(x)},{int(y)},{int(i)},{'s' if s else 'd'}\n"
                 for r, t, x, y, i, s in zip(ev["global_reference"], times, ev["origin_x"], ev["origin_y"],
                                             ev["origin_index"], speciation)]
        self._f.write("".join(lines))

    def close(self):
        if self._f:
            self._f.close()
            self._f = None
