"""ctypes wrapper of the CPU oracle (test infrastructure and CPU baseline only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/_build/libnecsim_oracle.so, the C restatement of
necsim-rust's CPU Independent algorithm (see necsim_oracle.h).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NECSIM_ORACLE_LIBRARY selects another build (the sanitizer build of the
# Makefile's `sanitize` target)
LIB_PATH = os.environ.get("NECSIM_ORACLE_LIBRARY", os.path.join(HERE, "_build", "libnecsim_oracle.so"))

LINEAGE_DTYPE = np.dtype([("global_reference", "<u8"), ("last_event_time", "<f8"),
                          ("x", "<u4"), ("y", "<u4"), ("index", "<u4"), ("_pad", "<u4")])
EVENT_DTYPE = np.dtype([("global_reference", "<u8"), ("prior_time", "<f8"), ("event_time", "<f8"),
                        ("x", "<u4"), ("y", "<u4"), ("index", "<u4"), ("_pad", "<u4")])
# PackedEvent (necsim/core/src/event.rs:16-29): sign bits of the times encode the event type
PACKED_DTYPE = np.dtype([("global_reference", "<u8"), ("prior_time", "<f8"), ("event_time", "<f8"),
                         ("ox", "<u4"), ("oy", "<u4"), ("oindex", "<u4"),
                         ("tx", "<u4"), ("ty", "<u4"), ("tindex", "<u4"), ("coalescence", "<u8")])

NON_SPATIAL, ALMOST_INFINITE_NORMAL, SPATIALLY_EXPLICIT, SPATIALLY_EXPLICIT_PACKED = 0, 1, 2, 3
DEDUP_NONE, DEDUP_RELATIVE, DEDUP_ABSOLUTE = 0, 1, 2
SLICE_RELATIVE, SLICE_ABSOLUTE = 1, 2


class ScenarioDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("deme", ctypes.c_uint32), ("sigma", ctypes.c_double),
                ("habitat", ctypes.c_void_p), ("row_ptr", ctypes.c_void_p),
                ("col", ctypes.c_void_p), ("weight", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("speciation_probability", ctypes.c_double),
                ("turnover_rate", ctypes.c_double), ("delta_t", ctypes.c_double),
                ("step_slice", ctypes.c_uint64), ("dedup_mode", ctypes.c_int),
                ("dedup_factor", ctypes.c_double), ("dedup_capacity", ctypes.c_uint64),
                ("event_slice_mode", ctypes.c_int), ("event_slice_factor", ctypes.c_double),
                ("event_slice_capacity", ctypes.c_uint64), ("has_pause", ctypes.c_int),
                ("pause_before", ctypes.c_double), ("min_event_time", ctypes.c_double),
                ("report_dispersal", ctypes.c_int)]


class Outcome(ctypes.Structure):
    _fields_ = [("paused", ctypes.c_int), ("time", ctypes.c_double), ("steps", ctypes.c_uint64),
                ("n_events", ctypes.c_size_t), ("n_remaining", ctypes.c_size_t)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, u64, u32, vp = ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
        for name in ("orc_ln", "orc_exp", "orc_sin", "orc_cos", "orc_sqrt", "orc_floor", "orc_round"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d]
        for name in ("orc_fdlibm_log", "orc_fdlibm_sin", "orc_fdlibm_cos"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d]
        L.orc_maths_vec.restype = ctypes.c_int
        L.orc_maths_vec.argtypes = [ctypes.c_int, vp, vp, ctypes.c_size_t]
        L.orc_seed_from_u64.restype = u64
        L.orc_seed_from_u64.argtypes = [u64]
        L.orc_seahash_diffuse.restype = u64
        L.orc_seahash_diffuse.argtypes = [u64]
        L.orc_wymum.restype = u64
        L.orc_wymum.argtypes = [u64, u64]
        L.orc_prime.restype = u64
        L.orc_prime.argtypes = [u64, u64, u64]
        L.orc_sample_u64.restype = u64
        L.orc_sample_u64.argtypes = [ctypes.POINTER(u64)]
        L.orc_uniform_closed_open.restype = d
        L.orc_uniform_closed_open.argtypes = [ctypes.POINTER(u64)]
        L.orc_uniform_open_closed.restype = d
        L.orc_uniform_open_closed.argtypes = [ctypes.POINTER(u64)]
        L.orc_fnv1a64.restype = u64
        L.orc_fnv1a64.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.orc_speciation_sample_hash.restype = u64
        L.orc_speciation_sample_hash.argtypes = [d, d, u32, u32, u32]
        L.orc_poisson_next_event_time.restype = d
        L.orc_poisson_next_event_time.argtypes = [u64, u64, d, d, d, ctypes.POINTER(u64)]
        L.orc_alias_create.restype = None
        L.orc_alias_create.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp]
        L.orc_alias_sample.restype = u64
        L.orc_alias_sample.argtypes = [vp, vp, vp, ctypes.c_size_t, d, d]
        L.orc_scenario_new.restype = vp
        L.orc_scenario_new.argtypes = [ctypes.POINTER(ScenarioDesc)]
        L.orc_scenario_free.restype = None
        L.orc_scenario_free.argtypes = [vp]
        L.orc_scenario_tables.restype = ctypes.c_size_t
        L.orc_scenario_tables.argtypes = [vp] + [vp] * 7
        L.orc_location_id.restype = u64
        L.orc_location_id.argtypes = [vp, u32, u32, u32]
        L.orc_deme_at.restype = u32
        L.orc_deme_at.argtypes = [vp, u32, u32]
        L.orc_sample_dispersal.restype = None
        L.orc_sample_dispersal.argtypes = [vp, u32, u32, ctypes.POINTER(u64),
                                           ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.orc_sample_origins.restype = ctypes.c_size_t
        L.orc_sample_origins.argtypes = [vp, d, u32, u32, u32, u32, u32, vp, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t)]
        L.orc_simulate_packed.restype = ctypes.c_int
        L.orc_simulate_packed.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.c_size_t,
                                          ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t), vp,
                                          ctypes.POINTER(Outcome)]
        L.orc_free.restype = None
        L.orc_free.argtypes = [vp]
        L.orc_simulate.restype = ctypes.c_int
        L.orc_simulate.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.c_size_t, vp, vp,
                                   ctypes.POINTER(Outcome)]
        L.orc_biodiversity.restype = u64
        L.orc_biodiversity.argtypes = [vp, ctypes.c_size_t]
        L.orc_simulate_partitioned.restype = u64
        L.orc_simulate_partitioned.argtypes = [vp, ctypes.POINTER(Params), d, u32, u32, u32, u32, u32,
                                               ctypes.POINTER(d), ctypes.POINTER(u64)]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class Scenario:
    """Oracle scenario (habitat + dispersal sampler tables)."""

    def __init__(self, kind, width=0, height=0, deme=1, sigma=0.0, habitat=None, row_ptr=None,
                 col=None, weight=None):
        self.kind = kind
        self._keep = []
        desc = ScenarioDesc(kind=kind, width=width, height=height, deme=deme, sigma=sigma)
        if habitat is not None:
            habitat = np.ascontiguousarray(habitat, dtype=np.uint32).ravel()
            row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
            col = np.ascontiguousarray(col, dtype=np.uint32)
            weight = np.ascontiguousarray(weight, dtype=np.float64)
            self._keep += [habitat, row_ptr, col, weight]
            desc.habitat, desc.row_ptr = _ptr(habitat), _ptr(row_ptr)
            desc.col, desc.weight = _ptr(col), _ptr(weight)
        self.width, self.height = width, height
        self.handle = lib().orc_scenario_new(ctypes.byref(desc))
        if not self.handle:
            raise ValueError("oracle rejected the scenario (in-memory dispersal contract)")

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_scenario_free(self.handle)
            self.handle = None

    @staticmethod
    def non_spatial(width, height, deme):
        return Scenario(NON_SPATIAL, width, height, deme)

    @staticmethod
    def almost_infinite_normal(sigma):
        return Scenario(ALMOST_INFINITE_NORMAL, sigma=sigma)

    @staticmethod
    def spatially_explicit(width, height, habitat, row_ptr, col, weight, packed=False):
        return Scenario(SPATIALLY_EXPLICIT_PACKED if packed else SPATIALLY_EXPLICIT, width, height,
                        habitat=habitat, row_ptr=row_ptr, col=col, weight=weight)

    def tables(self):
        cells = self.width * self.height
        n = lib().orc_scenario_tables(self.handle, None, None, None, None, None, None, None)
        rs = np.zeros(cells, np.uint64)
        re = np.zeros(cells, np.uint64)
        sp = np.zeros(cells, np.float64)
        se = np.zeros(cells, np.uint64)
        au = np.zeros(max(n, 1), np.float64)
        ae = np.zeros(max(n, 1), np.uint64)
        ak = np.zeros(max(n, 1), np.uint64)
        lib().orc_scenario_tables(self.handle, _ptr(rs), _ptr(re), _ptr(sp), _ptr(se), _ptr(au),
                                  _ptr(ae), _ptr(ak))
        return dict(range_start=rs, range_end=re, self_p=sp, self_event=se, atom_u=au[:n],
                    atom_e=ae[:n], atom_k=ak[:n])

    def origins(self, percentage=1.0, rank=0, size=1, circle=(2**31 - 1, 2**31 - 1, 0)):
        total = ctypes.c_size_t(0)
        cx, cy, r = circle
        lib().orc_sample_origins(self.handle, percentage, rank, size, cx, cy, r, None, 0,
                                 ctypes.byref(total))
        out = np.zeros(total.value, LINEAGE_DTYPE)
        lib().orc_sample_origins(self.handle, percentage, rank, size, cx, cy, r, _ptr(out),
                                 total.value, ctypes.byref(total))
        return out


def make_params(seed, speciation_probability, delta_t=2.0, step_slice=10, dedup=("relative", 1.0),
                event_slice=("relative", 2.0), pause_before=None, min_event_time=0.0,
                turnover_rate=0.5, report_dispersal=False):
    p = Params(seed=seed, speciation_probability=speciation_probability,
               turnover_rate=turnover_rate, delta_t=delta_t, step_slice=step_slice,
               min_event_time=min_event_time, report_dispersal=1 if report_dispersal else 0)
    if dedup is None or dedup == "none":
        p.dedup_mode = DEDUP_NONE
    elif dedup[0] == "relative":
        p.dedup_mode, p.dedup_factor = DEDUP_RELATIVE, dedup[1]
    else:
        p.dedup_mode, p.dedup_capacity = DEDUP_ABSOLUTE, dedup[1]
    if event_slice[0] == "relative":
        p.event_slice_mode, p.event_slice_factor = SLICE_RELATIVE, event_slice[1]
    else:
        p.event_slice_mode, p.event_slice_capacity = SLICE_ABSOLUTE, event_slice[1]
    if pause_before is not None:
        p.has_pause, p.pause_before = 1, pause_before
    return p


def simulate(scenario, params, lineages):
    lineages = np.ascontiguousarray(lineages, dtype=LINEAGE_DTYPE)
    n = len(lineages)
    events = np.zeros(max(n, 1), EVENT_DTYPE)
    remaining = np.zeros(max(n, 1), LINEAGE_DTYPE)
    out = Outcome()
    rc = lib().orc_simulate(scenario.handle, ctypes.byref(params), _ptr(lineages), n,
                            _ptr(events), _ptr(remaining), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_simulate failed: {rc}")
    return dict(paused=bool(out.paused), time=out.time, steps=out.steps,
                events=events[:out.n_events].copy(), remaining=remaining[:out.n_remaining].copy())


def simulate_packed(scenario, params, lineages):
    """orc_simulate_packed: every reported event as a 56-byte PackedEvent"""
    lineages = np.ascontiguousarray(lineages, dtype=LINEAGE_DTYPE)
    n = len(lineages)
    remaining = np.zeros(max(n, 1), LINEAGE_DTYPE)
    out = Outcome()
    buf = ctypes.c_void_p()
    nev = ctypes.c_size_t(0)
    rc = lib().orc_simulate_packed(scenario.handle, ctypes.byref(params), _ptr(lineages), n,
                                   ctypes.byref(buf), ctypes.byref(nev), _ptr(remaining), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_simulate_packed failed: {rc}")
    events = np.zeros(nev.value, PACKED_DTYPE)
    if nev.value:
        ctypes.memmove(events.ctypes.data, buf, nev.value * PACKED_DTYPE.itemsize)
    lib().orc_free(buf)
    return dict(paused=bool(out.paused), time=out.time, steps=out.steps, events=events,
                remaining=remaining[:out.n_remaining].copy())


def biodiversity(events):
    events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    return int(lib().orc_biodiversity(_ptr(events), len(events)))


def simulate_partitioned(scenario, params, percentage, threads, partitions=None,
                         circle=(2**31 - 1, 2**31 - 1, 0)):
    secs = ctypes.c_double(0.0)
    nev = ctypes.c_uint64(0)
    cx, cy, r = circle
    steps = lib().orc_simulate_partitioned(scenario.handle, ctypes.byref(params), percentage, cx,
                                           cy, r, threads, partitions or threads, ctypes.byref(secs),
                                           ctypes.byref(nev))
    return dict(steps=int(steps), seconds=secs.value, events=int(nev.value))


def maths_vec(op, x):
    """orc_maths_vec: op 0 ln, 1 exp, 2 sin, 3 cos, 4/5/6 the fdlibm-order
    twins of ln / sin / cos; x is a float64 array."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    if lib().orc_maths_vec(op, x.ctypes.data, y.ctypes.data, len(x)) != 0:
        raise ValueError(f"unknown maths op {op}")
    return y
