"""Host-side mirror of rustcoalescence's algorithm plugin interface for the
HIP Independent algorithm.

The names follow the reference so that a caller reads the same flow as the
Rust CLI:

    rng = WyHash.seed_from_u64(42)                       # rng/mod.rs:60
    cogs = NonSpatialScenario((100, 100), 1).new(1e-3)   # Scenario::new
    outcome = HipIndependentAlgorithm.initialise_and_simulate(
        IndependentArguments(), rng, cogs, OriginPreSampler.all(), None, reporter)

rustcoalescence/algorithms/src/lib.rs:55-112 (Algorithm),
rustcoalescence/algorithms/independent/src/arguments.rs:79-146 (arguments),
rustcoalescence/algorithms/src/result.rs:11-26 (SimulationOutcome).
All simulation work runs in libnecsim_hip.so on the GPU.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _ffi
from ._ffi import check, lib, ptr


# ---------------------------------------------------------------------------
# RNG
@dataclass(frozen=True)
class WyHash:
    """WyHash PrimeableRng (necsim/impls/no-std/src/cogs/rng/wyhash.rs:19-24).

    Only the seed matters: every event re-primes the state."""
    seed: int

    @staticmethod
    def seed_from_u64(seed):
        return WyHash(int(lib().nh_seed_from_u64(seed)))


# ---------------------------------------------------------------------------
# arguments (parallelisation/independent/mod.rs:15-82)
@dataclass(frozen=True)
class DedupCache:
    mode: int
    factor: float = 0.0
    capacity: int = 0

    @staticmethod
    def Relative(factor=1.0):
        return DedupCache(_ffi.DEDUP_RELATIVE, factor=factor)

    @staticmethod
    def Absolute(capacity):
        return DedupCache(_ffi.DEDUP_ABSOLUTE, capacity=capacity)

    @staticmethod
    def none():
        return DedupCache(_ffi.DEDUP_NONE)


@dataclass(frozen=True)
class EventSlice:
    mode: int
    factor: float = 0.0
    capacity: int = 0

    @staticmethod
    def Relative(factor=2.0):
        return EventSlice(_ffi.SLICE_RELATIVE, factor=factor)

    @staticmethod
    def Absolute(capacity):
        return EventSlice(_ffi.SLICE_ABSOLUTE, capacity=capacity)


PLAIN = _ffi.FORM_PLAIN
PERSISTENT = _ffi.FORM_PERSISTENT
TASK_LIST = _ffi.FORM_TASK_LIST


@dataclass(frozen=True)
class IndependentArguments:
    """Independent(delta_t, step_slice, dedup_cache, parallelism_mode: Monolithic(event_slice))
    with the reference defaults, plus the device form and launch knobs."""
    delta_t: float = 2.0
    step_slice: int = 10
    dedup_cache: DedupCache = field(default_factory=lambda: DedupCache.Relative(1.0))
    event_slice: EventSlice = field(default_factory=lambda: EventSlice.Relative(2.0))
    form: int = PERSISTENT
    block_size: int = 0
    blocks_per_cu: int = 0
    use_graph: bool = False
    inner_steps: int = 0
    exec_fill: int = 0
    plain_batch: int = 0  # plain form: passes per host synchronisation (1: one per pass; 0: 16)
    coop_search: int = 0  # persistent form: cooperative next-event search (0: automatic, 1: off, 2: on)
    task_list_size: int = 0  # task-list form: lineages per launch (0: 4096)
    pool: int = 0  # persistent form without dedup: lineages per wave in the pooled kernel's LDS pool (128, 192)
    device: int = 0


# ---------------------------------------------------------------------------
# origin pre-sampler (origin_sampler/pre_sampler.rs:50-114)
@dataclass(frozen=True)
class OriginPreSampler:
    proportion: float = 1.0
    rank: int = 0
    size: int = 1

    @staticmethod
    def all():
        return OriginPreSampler()

    def percentage(self, p):
        return OriginPreSampler(p, self.rank, self.size)

    def partition(self, rank, size):
        return OriginPreSampler(self.proportion, rank, size)


# ---------------------------------------------------------------------------
# scenarios (rustcoalescence/scenarios/src)
class ScenarioCogs:
    """Habitat + dispersal sampler + uniform turnover / speciation."""

    def __init__(self, desc, speciation_probability, turnover_rate=0.5, keep=(), circle=None):
        self.desc = desc
        self.speciation_probability = float(speciation_probability)
        self.turnover_rate = float(turnover_rate)
        self._keep = list(keep)
        self.circle = circle

    def sample_habitat(self, pre_sampler: OriginPreSampler):
        cx, cy, r = self.circle if self.circle else (0, 0, 0)
        n = ctypes.c_size_t(0)
        args = (ctypes.byref(self.desc), pre_sampler.proportion, pre_sampler.rank, pre_sampler.size,
                cx, cy, r)
        check(lib().nh_sample_origins(*args, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, _ffi.LINEAGE_DTYPE)
        check(lib().nh_sample_origins(*args, ptr(out), n.value, ctypes.byref(n)))
        return out


@dataclass
class NonSpatialScenario:
    """NonSpatial(area, deme) (rustcoalescence/scenarios/src/non_spatial.rs:39-89)"""
    area: tuple
    deme: int

    def new(self, speciation_probability):
        d = _ffi.Scenario(kind=_ffi.SCENARIO_NON_SPATIAL, width=self.area[0], height=self.area[1],
                          deme=self.deme)
        return ScenarioCogs(d, speciation_probability)


CIRCLE_CENTRE = 2**32 // 2 - 1  # Sample::default_circle_sample_centre (u32::MAX / 2)


@dataclass
class AlmostInfiniteNormalDispersalScenario:
    """AlmostInfinite(sample: Circle(centre, radius), dispersal: Normal(sigma))
    (rustcoalescence/scenarios/src/almost_infinite/normal.rs:43-94, mod.rs:90-123)"""
    radius: int
    sigma: float
    centre: tuple = (CIRCLE_CENTRE, CIRCLE_CENTRE)

    def new(self, speciation_probability):
        d = _ffi.Scenario(kind=_ffi.SCENARIO_ALMOST_INFINITE_NORMAL, sigma=self.sigma)
        return ScenarioCogs(d, speciation_probability,
                            circle=(self.centre[0], self.centre[1], self.radius))


def build_alias_tables(width, height, habitat, row_ptr, col, weight, separable=True):
    """InMemoryPacked(Separable)AliasDispersalSampler::new from a sparse dispersal map."""
    habitat = np.ascontiguousarray(habitat, dtype=np.uint32).ravel()
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    weight = np.ascontiguousarray(weight, dtype=np.float64)
    cells = width * height
    ranges = np.zeros(cells, _ffi.RANGE_DTYPE)
    selfd = np.zeros(cells, _ffi.SELF_DTYPE)
    n = ctypes.c_size_t(0)
    args = (width, height, ptr(habitat), ptr(row_ptr), ptr(col), ptr(weight), 1 if separable else 0,
            ptr(ranges), ptr(selfd))
    check(lib().nh_build_alias_tables(*args, None, 0, ctypes.byref(n)))
    atoms = np.zeros(max(n.value, 1), _ffi.ATOM_DTYPE)
    check(lib().nh_build_alias_tables(*args, ptr(atoms), n.value, ctypes.byref(n)))
    return ranges, (selfd if separable else None), atoms[:n.value]


@dataclass
class SpatiallyExplicitUniformTurnoverScenario:
    """SpatiallyExplicit(turnover: Uniform, mode: Strict) on an in-memory habitat
    map with a sparse dispersal map (rows in ascending column order)
    (rustcoalescence/scenarios/src/spatially_explicit/turnover/uniform.rs:49-117)."""
    width: int
    height: int
    habitat: np.ndarray
    row_ptr: np.ndarray
    col: np.ndarray
    weight: np.ndarray
    turnover_rate: float = 0.5
    packed: bool = False  # InMemoryPackedAliasDispersalSampler instead of the separable one
    _tables: tuple = None

    def tables(self):
        if self._tables is None:
            self._tables = build_alias_tables(self.width, self.height, self.habitat, self.row_ptr,
                                              self.col, self.weight, separable=not self.packed)
        return self._tables

    def kind(self):
        return (_ffi.SCENARIO_SPATIALLY_EXPLICIT_PACKED if self.packed
                else _ffi.SCENARIO_SPATIALLY_EXPLICIT)

    def new(self, speciation_probability, host_tables=True):
        """Scenario cogs.  With host_tables=False the alias tables are not
        built on the host: Tables(ctx, cogs) then builds them on the GPU
        (nh_tables_create_from_map)."""
        habitat = np.ascontiguousarray(self.habitat, dtype=np.uint32).ravel()
        if host_tables:
            ranges, selfd, atoms = self.tables()
            d = _ffi.Scenario(kind=self.kind(), width=self.width, height=self.height, habitat=ptr(habitat),
                              alias_ranges=ptr(ranges), self_dispersal=ptr(selfd), atoms=ptr(atoms),
                              n_atoms=len(atoms))
            keep = (habitat, ranges, selfd, atoms)
        else:
            d = _ffi.Scenario(kind=self.kind(), width=self.width, height=self.height, habitat=ptr(habitat))
            keep = (habitat,)
        cogs = ScenarioCogs(d, speciation_probability, self.turnover_rate, keep=keep)
        cogs.map_scenario = self
        cogs.host_tables = host_tables
        return cogs

    def map_desc(self):
        """nh_map_scenario view of the maps (arrays kept alive by the caller)"""
        arrs = (np.ascontiguousarray(self.habitat, dtype=np.uint32).ravel(),
                np.ascontiguousarray(self.row_ptr, dtype=np.uint64),
                np.ascontiguousarray(self.col, dtype=np.uint32),
                np.ascontiguousarray(self.weight, dtype=np.float64))
        d = _ffi.MapScenario(kind=self.kind(), width=self.width, height=self.height, habitat=ptr(arrs[0]),
                             row_ptr=ptr(arrs[1]), col=ptr(arrs[2]), weight=ptr(arrs[3]))
        return d, arrs


# ---------------------------------------------------------------------------
# outcome and reporters
@dataclass
class SimulationOutcome:
    """SimulationOutcome::{Done, Paused} (rustcoalescence/algorithms/src/result.rs:13-26)"""
    paused: bool
    time: float
    steps: int
    lineages: Optional[np.ndarray] = None
    rng: Optional[WyHash] = None
    rounds: int = 0
    launches: int = 0
    kernel_ms: float = 0.0


class BiodiversityReporter:
    """necsim/plugins/common/src/biodiversity.rs:32-41: counts speciation
    events, skipping one equal (same origin and event time) to the previous."""

    def __init__(self):
        self.biodiversity = 0
        self._last = None

    def report_speciation(self, events):
        for e in events:
            key = (int(e["x"]), int(e["y"]), int(e["index"]), float(e["event_time"]))
            if key == self._last:
                continue
            self._last = key
            self.biodiversity += 1


class EventCollector:
    """Keeps the reported (sorted) speciation stream."""

    def __init__(self):
        self.chunks = []

    def report_speciation(self, events):
        self.chunks.append(np.array(events, copy=True))

    @property
    def events(self):
        if not self.chunks:
            return np.zeros(0, _ffi.EVENT_DTYPE)
        return np.concatenate(self.chunks)


SIGN_BIT = np.uint64(1 << 63)


def _magnitude_bits(f64):
    return np.ascontiguousarray(f64, dtype=np.float64).view(np.uint64) & ~SIGN_BIT


def is_speciation(packed):
    """PackedEvent -> TypedEvent::Speciation: prior_time negative and
    event_time positive (necsim/core/src/event.rs:103-115, 186-229)"""
    prior = np.ascontiguousarray(packed["prior_time"]).view(np.uint64)
    event = np.ascontiguousarray(packed["event_time"]).view(np.uint64)
    return ((prior & SIGN_BIT) != 0) & ((event & SIGN_BIT) == 0)


def speciation_events(packed):
    """the speciation events of a PackedEvent stream, as SpeciationEvents"""
    sp = packed[is_speciation(packed)]
    out = np.zeros(len(sp), _ffi.EVENT_DTYPE)
    out["global_reference"] = sp["global_reference"]
    out["prior_time"] = _magnitude_bits(sp["prior_time"]).view(np.float64)
    out["event_time"] = sp["event_time"]
    out["x"], out["y"], out["index"] = sp["origin_x"], sp["origin_y"], sp["origin_index"]
    return out


def packed_order(packed):
    """indices that sort PackedEvents by PackedEvent::cmp (event.rs:256-280)"""
    return np.lexsort((packed["global_reference"], _magnitude_bits(packed["prior_time"]),
                       packed["target_index"], packed["target_y"], packed["target_x"],
                       packed["origin_index"], packed["origin_y"], packed["origin_x"],
                       _magnitude_bits(packed["event_time"])))


class PackedEventCollector:
    """A reporter with ReportDispersal (the reference's
    Reporter::ReportDispersal = True): keeps every reported event, speciations
    and dispersals, as the sorted PackedEvent stream."""

    REPORT_DISPERSAL = True

    def __init__(self):
        self.chunks = []

    def report_packed(self, events):
        self.chunks.append(np.array(events, copy=True))

    @property
    def events(self):
        if not self.chunks:
            return np.zeros(0, _ffi.PACKED_EVENT_DTYPE)
        return np.concatenate(self.chunks)


def save_lineages(path, lineages):
    """LineageFileSaver: paused lineages to a new lineage file (bincode varint)"""
    lin = np.ascontiguousarray(lineages, dtype=_ffi.LINEAGE_DTYPE)
    check(lib().nh_lineage_file_write(os.fsencode(path), ptr(lin), len(lin)))


def load_lineages(path):
    """LineageFileLoader: the lineages of a lineage file"""
    n = ctypes.c_size_t(0)
    check(lib().nh_lineage_file_read(os.fsencode(path), None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, _ffi.LINEAGE_DTYPE)
    check(lib().nh_lineage_file_read(os.fsencode(path), ptr(out), len(out), ctypes.byref(n)))
    return out


class EventLogReporter:
    """EventLogRecorder (necsim/impls/std/src/event_log/recorder.rs:20-210) as a
    reporter: events are buffered in reporting order and written as numbered
    segment files of `segment_capacity` events (the last one shorter), each
    sorted by the PackedEvent order; a log without events leaves no directory
    behind (recorder.rs:46-55).  The directory must not exist yet.  With
    report_dispersal the log records every event (speciations and
    dispersals, header with_dispersal = true) instead of speciations only."""

    def __init__(self, directory, segment_capacity=1 << 20, report_dispersal=False):
        if segment_capacity < 1:
            raise ValueError("segment_capacity must be positive")
        directory = os.path.abspath(directory)
        os.makedirs(os.path.dirname(directory), exist_ok=True)
        os.mkdir(directory)
        self.directory = directory
        self.capacity = int(segment_capacity)
        self.segment_index = 0
        self.REPORT_DISPERSAL = bool(report_dispersal)
        self._buffer = []
        self._buffered = 0

    def _append(self, events, dtype):
        if len(events) == 0:
            return
        self._buffer.append(np.ascontiguousarray(events, dtype=dtype))
        self._buffered += len(events)
        while self._buffered >= self.capacity:
            ev = np.concatenate(self._buffer)
            self._write(ev[:self.capacity])
            rest = ev[self.capacity:]
            self._buffer = [rest] if len(rest) else []
            self._buffered = len(rest)

    def report_speciation(self, events):
        if self.REPORT_DISPERSAL:
            raise TypeError("this log records PackedEvents (report_packed)")
        self._append(events, _ffi.EVENT_DTYPE)

    def report_packed(self, events):
        if not self.REPORT_DISPERSAL:
            raise TypeError("this log records speciation events only (report_speciation)")
        self._append(events, _ffi.PACKED_EVENT_DTYPE)

    def _write(self, ev):
        path = os.path.join(self.directory, str(self.segment_index)).encode()
        if self.REPORT_DISPERSAL:
            ev = np.ascontiguousarray(ev[packed_order(ev)])
            check(lib().nh_event_log_write_packed_segment(path, ptr(ev), len(ev), 1, 1))
        else:
            order = np.lexsort((ev["global_reference"], ev["prior_time"], ev["index"], ev["y"], ev["x"],
                                ev["event_time"]))
            ev = np.ascontiguousarray(ev[order])
            check(lib().nh_event_log_write_segment(path, ptr(ev), len(ev)))
        self.segment_index += 1

    def close(self):
        if self._buffered:
            self._write(np.concatenate(self._buffer))
            self._buffer, self._buffered = [], 0
        if self.segment_index == 0 and os.path.isdir(self.directory) and not os.listdir(self.directory):
            os.rmdir(self.directory)


class Context:
    """One device context (nh_init / nh_ctx_destroy)."""
    _cache = {}

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().nh_init(device, ctypes.byref(h)))
        self.handle = h

    @classmethod
    def get(cls, device=0):
        if device not in cls._cache:
            cls._cache[device] = Context(device)
        return cls._cache[device]

    def info(self):
        cus = ctypes.c_uint32(0)
        name = ctypes.create_string_buffer(128)
        check(lib().nh_device_info(self.handle, ctypes.byref(cus), name, 128))
        return cus.value, name.value.decode()


class Tables:
    """Resident device tables of one scenario: nh_tables_create from host-built
    tables, or nh_tables_create_from_map (alias tables built on the GPU) for
    map scenarios whose cogs were made with host_tables=False."""

    def __init__(self, ctx, cogs, on_device=None):
        h = ctypes.c_void_p()
        if on_device is None:
            on_device = not getattr(cogs, "host_tables", True)
        if on_device:
            d, keep = cogs.map_scenario.map_desc()
            check(lib().nh_tables_create_from_map(ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        else:
            check(lib().nh_tables_create(ctx.handle, ctypes.byref(cogs.desc), ctypes.byref(h)))
        self.handle = h
        self.cogs = cogs

    def classify(self, lineages):
        """NH_LINEAGE_* class of every lineage (nh_classify_lineages): 0 OK,
        1 OutOfDeme, 2 OutOfHabitat."""
        lineages = np.ascontiguousarray(lineages, dtype=_ffi.LINEAGE_DTYPE)
        kinds = np.zeros(len(lineages), np.int32)
        bad = ctypes.c_size_t(0)
        check(lib().nh_classify_lineages(self.handle, ptr(lineages), len(lineages), ptr(kinds),
                                         ctypes.byref(bad)))
        return kinds

    def export(self):
        """(cell records, atoms) of in-memory tables, as resident on the device"""
        nc, na = ctypes.c_size_t(0), ctypes.c_size_t(0)
        check(lib().nh_tables_export(self.handle, None, 0, None, 0, ctypes.byref(nc), ctypes.byref(na)))
        cells = np.zeros(nc.value, _ffi.CELL_RECORD_DTYPE)
        atoms = np.zeros(na.value, _ffi.DEVICE_ATOM_DTYPE)
        check(lib().nh_tables_export(self.handle, ptr(cells), len(cells), ptr(atoms), len(atoms),
                                     ctypes.byref(nc), ctypes.byref(na)))
        return cells, atoms

    def __del__(self):
        if getattr(self, "handle", None):
            lib().nh_tables_destroy(self.handle)
            self.handle = None


class Job:
    """One simulation on the device (nh_job_*)."""

    def __init__(self, ctx, tables, args: IndependentArguments, rng: WyHash, speciation_probability,
                 turnover_rate, pause_before=None, resume_after=0.0, report_dispersal=False, event_buffer=0,
                 device_event_counter=False):
        p = _job_params(args, rng, speciation_probability, turnover_rate, pause_before, resume_after)
        p.report_dispersal = 1 if report_dispersal else 0
        p.event_buffer = int(event_buffer)
        p.device_event_counter = 1 if device_event_counter else 0
        h = ctypes.c_void_p()
        check(lib().nh_job_create(ctx.handle, tables.handle, ctypes.byref(p), ctypes.byref(h)))
        self.handle = h
        self.rng = rng
        # the job reads the tables until it is destroyed: keep them alive
        self.tables = tables

    def add_lineages(self, lineages):
        lineages = np.ascontiguousarray(lineages, dtype=_ffi.LINEAGE_DTYPE)
        check(lib().nh_job_add_lineages(self.handle, ptr(lineages), len(lineages)))

    def add_origins(self, pre_sampler, circle=None):
        """the scenario's origin sampler run on the device into the job
        (nh_job_add_origins); `circle` = (centre x, centre y, radius) of the
        almost-infinite scenarios (ScenarioCogs.circle)"""
        cx, cy, r = circle if circle else (0, 0, 0)
        check(lib().nh_job_add_origins(self.handle, float(pre_sampler.proportion), int(pre_sampler.rank),
                                       int(pre_sampler.size), cx, cy, r))

    def start(self):
        check(lib().nh_job_start(self.handle))

    def step_round(self):
        fin = ctypes.c_int32(0)
        check(lib().nh_job_step_round(self.handle, ctypes.byref(fin)))
        return bool(fin.value)

    def drain_events(self):
        n = ctypes.c_size_t(0)
        check(lib().nh_job_drain_events(self.handle, None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, _ffi.EVENT_DTYPE)
        if n.value:
            check(lib().nh_job_drain_events(self.handle, ptr(buf), n.value, ctypes.byref(n)))
        return buf[:n.value]

    def drain_packed_events(self):
        """every event since the last drain (report_dispersal jobs), sorted"""
        n = ctypes.c_size_t(0)
        check(lib().nh_job_drain_packed_events(self.handle, None, 0, ctypes.byref(n)))
        buf = np.empty(n.value, _ffi.PACKED_EVENT_DTYPE)  # every byte is written
        if n.value:
            check(lib().nh_job_drain_packed_events(self.handle, ptr(buf), n.value, ctypes.byref(n)))
        return buf[:n.value]

    def event_counts(self):
        """the device EventCounterReporter's counters (device_event_counter jobs)"""
        c = _ffi.EventCounts()
        check(lib().nh_job_event_counts(self.handle, ctypes.byref(c)))
        return {f: int(getattr(c, f)) for f, _ in _ffi.EventCounts._fields_}

    def outcome(self):
        o = _ffi.Outcome()
        check(lib().nh_job_outcome(self.handle, ctypes.byref(o)))
        return o

    def take_lineages(self):
        n = ctypes.c_size_t(0)
        check(lib().nh_job_take_lineages(self.handle, None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, _ffi.LINEAGE_DTYPE)
        if n.value:
            check(lib().nh_job_take_lineages(self.handle, ptr(buf), n.value, ctypes.byref(n)))
        return buf

    def kernel_time(self):
        ms = ctypes.c_double(0)
        k = ctypes.c_uint64(0)
        check(lib().nh_job_kernel_time(self.handle, ctypes.byref(ms), ctypes.byref(k)))
        return ms.value, k.value

    def close(self):
        if getattr(self, "handle", None):
            lib().nh_job_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def _job_params(args, rng, speciation_probability, turnover_rate, pause_before=None, resume_after=0.0):
    return _ffi.JobParams(rng_seed=rng.seed, speciation_probability=speciation_probability,
                          turnover_rate=turnover_rate, delta_t=args.delta_t,
                          step_slice=args.step_slice, dedup_mode=args.dedup_cache.mode,
                          dedup_factor=args.dedup_cache.factor,
                          dedup_capacity=args.dedup_cache.capacity,
                          event_slice_mode=args.event_slice.mode,
                          event_slice_factor=args.event_slice.factor,
                          event_slice_capacity=args.event_slice.capacity,
                          has_pause_before=0 if pause_before is None else 1,
                          pause_before=0.0 if pause_before is None else pause_before,
                          resume_after=resume_after, form=args.form, block_size=args.block_size,
                          blocks_per_cu=args.blocks_per_cu, use_graph=1 if args.use_graph else 0,
                          inner_steps=args.inner_steps, exec_fill=args.exec_fill,
                          plain_batch=args.plain_batch, coop_search=args.coop_search,
                          task_list_size=args.task_list_size, pool=args.pool)


def autotune(ctx, tables, args: IndependentArguments, rng: WyHash, cogs, sample, window,
             tune_step_slice=True):
    """nh_autotune: the fastest persistent launch shape (and step slice) for a
    window of the workload on `sample`; returns (tuned IndependentArguments, ms)."""
    import dataclasses
    base = _job_params(dataclasses.replace(args, step_slice=0 if tune_step_slice else args.step_slice),
                       rng, cogs.speciation_probability, cogs.turnover_rate)
    sample = np.ascontiguousarray(sample, dtype=_ffi.LINEAGE_DTYPE)
    out = _ffi.JobParams()
    ms = ctypes.c_double(0)
    check(lib().nh_autotune(ctx.handle, tables.handle, ctypes.byref(base), ptr(sample), len(sample),
                            window, ctypes.byref(out), ctypes.byref(ms)))
    return dataclasses.replace(args, form=PERSISTENT, step_slice=int(out.step_slice),
                               block_size=int(out.block_size),
                               blocks_per_cu=int(out.blocks_per_cu),
                               inner_steps=int(out.inner_steps),
                               exec_fill=int(out.exec_fill), pool=int(out.pool)), ms.value


def copy_calibrate(ctx, nbytes, reps=10):
    """nh_copy_calibrate: mean duration (ms) of a plain copy kernel over
    `nbytes` and the copy bandwidth in GB/s (bytes read + written)."""
    ms = ctypes.c_double(0)
    check(lib().nh_copy_calibrate(ctx.handle, int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9


class ResumeSampleError(Exception):
    """ResumeError::Sample (rustcoalescence/algorithms/src/result.rs:29-32):
    the resumed lineages that IndependentActiveLineageSampler rejects
    (active_lineage_sampler/independent/mod.rs:98-117).  `exceptional` is a
    list of (kind, lineage) with kind "OutOfDeme" or "OutOfHabitat", in input
    order."""

    def __init__(self, exceptional):
        super().__init__(f"{len(exceptional)} exceptional lineage(s) on resume")
        self.exceptional = exceptional


class HipIndependentAlgorithm:
    """Algorithm::initialise_and_simulate / resume_and_simulate of the HIP
    Independent algorithm (Monolithic parallelism mode).

    A reporter whose REPORT_DISPERSAL is true (the reference's
    Reporter::ReportDispersal) receives every event of each water-level round,
    speciations and dispersals, as sorted PackedEvents through report_packed
    (or, with DEVICE_EVENT_COUNTER, counts them on the device and hands the
    counters over once through add_counts); any other reporter receives the
    speciation events through report_speciation."""

    @staticmethod
    def _run(args, rng, cogs, lineages, pause_before, resume_after, reporter, check_resume=False,
             pre_sampler=None):
        ctx = Context.get(args.device)
        tables = Tables(ctx, cogs)
        if check_resume:
            # IndependentLineageStoreSampleInitialiser's resume
            # (rustcoalescence/algorithms/independent/src/initialiser/resume.rs:71-85)
            kinds = tables.classify(lineages)
            bad = np.nonzero(kinds)[0]
            if len(bad):
                names = {1: "OutOfDeme", 2: "OutOfHabitat"}
                raise ResumeSampleError([(names[int(kinds[k])], lineages[k]) for k in bad])
        disp = bool(getattr(reporter, "REPORT_DISPERSAL", False))
        on_device = disp and bool(getattr(reporter, "DEVICE_EVENT_COUNTER", False))
        job = Job(ctx, tables, args, rng, cogs.speciation_probability, cogs.turnover_rate,
                  pause_before, resume_after, report_dispersal=disp, device_event_counter=on_device)
        try:
            if pre_sampler is not None:
                job.add_origins(pre_sampler, cogs.circle)
            else:
                job.add_lineages(lineages)
            finished = False
            while not finished:
                finished = job.step_round()
                if on_device:
                    continue
                if disp:
                    ev = job.drain_packed_events()
                    if len(ev):
                        reporter.report_packed(ev)
                else:
                    ev = job.drain_events()
                    if len(ev) and reporter is not None:
                        reporter.report_speciation(ev)
            if on_device:
                reporter.add_counts(job.event_counts())
            o = job.outcome()
            ms, _ = job.kernel_time()
            rem = job.take_lineages() if o.paused else None
            return SimulationOutcome(paused=bool(o.paused), time=o.time, steps=int(o.steps),
                                     lineages=rem, rng=rng if o.paused else None,
                                     rounds=int(o.rounds), launches=int(o.launches), kernel_ms=ms)
        finally:
            job.close()

    @classmethod
    def initialise_and_simulate(cls, args: IndependentArguments, rng: WyHash, cogs: ScenarioCogs,
                                pre_sampler: OriginPreSampler, pause_before=None, reporter=None):
        # the genesis initialiser's origin walk, on the device (nh_job_add_origins)
        return cls._run(args, rng, cogs, None, pause_before, 0.0, reporter, pre_sampler=pre_sampler)

    @classmethod
    def resume_and_simulate(cls, args: IndependentArguments, rng: WyHash, cogs: ScenarioCogs,
                            lineages, resume_after=None, pause_before=None, reporter=None):
        return cls._run(args, rng, cogs, lineages, pause_before,
                        0.0 if resume_after is None else resume_after, reporter, check_resume=True)


class Server:
    """Continuous-batching server on resident tables (nh_server_*).

    Requests are admitted by device footprint as memory frees up; every
    tick() runs one shared persistent kernel over all admitted requests.
    block_size / blocks_per_cu of 0 launch one block of 256 threads per CU,
    the latency-optimal shape of the serving stream."""

    def __init__(self, ctx, tables, max_lineages, hbm_budget_bytes, tick_steps=0, block_size=0,
                 blocks_per_cu=0, max_bypass_ticks=0):
        p = _ffi.ServerParams(max_lineages=max_lineages, hbm_budget_bytes=hbm_budget_bytes,
                              tick_steps=tick_steps, block_size=block_size,
                              blocks_per_cu=blocks_per_cu, max_bypass_ticks=max_bypass_ticks)
        h = ctypes.c_void_p()
        check(lib().nh_server_create(ctx.handle, tables.handle, ctypes.byref(p), ctypes.byref(h)))
        self.handle = h
        self.tables = tables

    def submit(self, lineages, rng: WyHash, speciation_probability, args: IndependentArguments,
               turnover_rate=0.5, pause_before=None):
        lineages = np.ascontiguousarray(lineages, dtype=_ffi.LINEAGE_DTYPE)
        r = _ffi.Request(rng_seed=rng.seed, speciation_probability=speciation_probability,
                         turnover_rate=turnover_rate, delta_t=args.delta_t,
                         step_slice=args.step_slice, dedup_mode=args.dedup_cache.mode,
                         dedup_factor=args.dedup_cache.factor,
                         dedup_capacity=args.dedup_cache.capacity,
                         event_slice_mode=args.event_slice.mode,
                         event_slice_factor=args.event_slice.factor,
                         event_slice_capacity=args.event_slice.capacity,
                         has_pause_before=0 if pause_before is None else 1,
                         pause_before=0.0 if pause_before is None else pause_before,
                         lineages=ptr(lineages), n_lineages=len(lineages))
        rid = ctypes.c_uint64(0)
        check(lib().nh_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    @staticmethod
    def request_footprint(n_lineages, cache_buckets=0):
        """Device bytes a request takes from the server's pool."""
        b = ctypes.c_uint64(0)
        check(lib().nh_server_request_footprint(int(n_lineages), int(cache_buckets), ctypes.byref(b)))
        return b.value

    def tick(self):
        a = ctypes.c_uint64(0)
        q = ctypes.c_uint64(0)
        check(lib().nh_server_tick(self.handle, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self):
        res = _ffi.RequestResult()
        rc = lib().nh_server_poll(self.handle, ctypes.byref(res))
        if rc == -4:  # NH_E_STATE: nothing finished yet
            return None
        check(rc)
        return res

    def events(self, request_id):
        n = ctypes.c_size_t(0)
        check(lib().nh_server_request_events(self.handle, request_id, None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, _ffi.EVENT_DTYPE)
        if n.value:
            check(lib().nh_server_request_events(self.handle, request_id, ptr(buf), n.value,
                                                 ctypes.byref(n)))
        return buf

    def release(self, request_id):
        check(lib().nh_server_release(self.handle, request_id))

    def close(self):
        if getattr(self, "handle", None):
            lib().nh_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
