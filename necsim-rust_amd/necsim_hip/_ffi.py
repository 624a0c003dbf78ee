"""ctypes binding of include/necsim_hip.h (libnecsim_hip.so, built in-tree).

The library is loaded from the package directory; a missing library raises
instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# NECSIM_HIP_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("NECSIM_HIP_LIBRARY", os.path.join(PKG_DIR, "libnecsim_hip.so"))

NH_OK = 0
ERRORS = {-1: "NH_E_INVALID", -2: "NH_E_HIP", -3: "NH_E_NOMEM", -4: "NH_E_STATE",
          -5: "NH_E_CONTRACT", -6: "NH_E_OUT_OF_DEME", -7: "NH_E_CAPACITY"}

SCENARIO_NON_SPATIAL = 0
SCENARIO_ALMOST_INFINITE_NORMAL = 1
SCENARIO_SPATIALLY_EXPLICIT = 2
SCENARIO_SPATIALLY_EXPLICIT_PACKED = 3
DEDUP_NONE, DEDUP_RELATIVE, DEDUP_ABSOLUTE = 0, 1, 2
SLICE_RELATIVE, SLICE_ABSOLUTE = 1, 2
FORM_PLAIN, FORM_PERSISTENT, FORM_TASK_LIST = 0, 1, 2

LINEAGE_DTYPE = np.dtype([("global_reference", "<u8"), ("last_event_time", "<f8"),
                          ("x", "<u4"), ("y", "<u4"), ("index", "<u4"), ("_pad", "<u4")])
EVENT_DTYPE = np.dtype([("global_reference", "<u8"), ("prior_time", "<f8"), ("event_time", "<f8"),
                        ("x", "<u4"), ("y", "<u4"), ("index", "<u4"), ("_pad", "<u4")])
# necsim/core/src/event.rs:16-29 PackedEvent (nh_packed_event)
PACKED_EVENT_DTYPE = np.dtype([("global_reference", "<u8"), ("prior_time", "<f8"), ("event_time", "<f8"),
                               ("origin_x", "<u4"), ("origin_y", "<u4"), ("origin_index", "<u4"),
                               ("target_x", "<u4"), ("target_y", "<u4"), ("target_index", "<u4"),
                               ("coalescence", "<u8")])
ATOM_DTYPE = np.dtype([("u", "<f8"), ("e", "<u8"), ("k", "<u8")])
RANGE_DTYPE = np.dtype([("start", "<u8"), ("end", "<u8")])
SELF_DTYPE = np.dtype([("self_dispersal", "<f8"), ("non_self_dispersal_event", "<u8")])
# resident device tables (nh_tables_export)
CELL_RECORD_DTYPE = np.dtype([("prefix", "<u8"), ("habitat", "<u4"), ("atom_start", "<u4"),
                              ("atom_len", "<u4"), ("non_self", "<u4"), ("self_p", "<f8")])
DEVICE_ATOM_DTYPE = np.dtype([("u", "<f8"), ("e", "<u4"), ("k", "<u4")])

c_u64, c_u32, c_i32, c_f64, c_vp, c_sz = (ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32,
                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t)


class Scenario(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("width", c_u32), ("height", c_u32), ("origin_x", c_u32),
                ("origin_y", c_u32), ("deme", c_u32), ("sigma", c_f64), ("habitat", c_vp),
                ("alias_ranges", c_vp), ("self_dispersal", c_vp), ("atoms", c_vp),
                ("n_atoms", c_sz)]


class MapScenario(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("width", c_u32), ("height", c_u32), ("origin_x", c_u32),
                ("origin_y", c_u32), ("habitat", c_vp), ("row_ptr", c_vp), ("col", c_vp), ("weight", c_vp)]


class JobParams(ctypes.Structure):
    _fields_ = [("rng_seed", c_u64), ("speciation_probability", c_f64), ("turnover_rate", c_f64),
                ("delta_t", c_f64), ("step_slice", c_u64), ("dedup_mode", c_i32),
                ("dedup_factor", c_f64), ("dedup_capacity", c_u64), ("event_slice_mode", c_i32),
                ("event_slice_factor", c_f64), ("event_slice_capacity", c_u64),
                ("has_pause_before", c_i32), ("pause_before", c_f64), ("resume_after", c_f64),
                ("form", c_i32), ("block_size", c_u32), ("blocks_per_cu", c_u32),
                ("use_graph", c_u32), ("inner_steps", c_u32), ("exec_fill", c_u32),
                ("report_dispersal", c_i32), ("event_buffer", c_u64), ("device_event_counter", c_i32),
                ("plain_batch", c_u32), ("coop_search", c_u32), ("task_list_size", c_u32), ("pool", c_u32)]


class EventCounts(ctypes.Structure):
    _fields_ = [("raw_total", c_u64), ("speciation", c_u64), ("out_dispersal", c_u64),
                ("self_dispersal", c_u64), ("out_coalescence", c_u64), ("self_coalescence", c_u64),
                ("late_dispersal_coalescence", c_u64), ("late_coalescence", c_u64)]


class Outcome(ctypes.Structure):
    _fields_ = [("paused", c_i32), ("time", c_f64), ("steps", c_u64), ("n_events", c_u64),
                ("n_remaining", c_u64), ("rounds", c_u64), ("launches", c_u64)]


class ServerParams(ctypes.Structure):
    _fields_ = [("max_lineages", c_u64), ("hbm_budget_bytes", c_u64), ("tick_steps", c_u64),
                ("block_size", c_u32), ("blocks_per_cu", c_u32), ("max_bypass_ticks", c_u64)]


class Request(ctypes.Structure):
    _fields_ = [("rng_seed", c_u64), ("speciation_probability", c_f64), ("turnover_rate", c_f64),
                ("delta_t", c_f64), ("step_slice", c_u64), ("dedup_mode", c_i32),
                ("dedup_factor", c_f64), ("dedup_capacity", c_u64), ("event_slice_mode", c_i32),
                ("event_slice_factor", c_f64), ("event_slice_capacity", c_u64),
                ("has_pause_before", c_i32), ("pause_before", c_f64), ("lineages", c_vp),
                ("n_lineages", c_sz)]


class RequestResult(ctypes.Structure):
    _fields_ = [("request_id", c_u64), ("status", c_i32), ("paused", c_i32), ("time", c_f64),
                ("steps", c_u64), ("n_events", c_u64), ("submit_s", c_f64), ("admit_s", c_f64),
                ("done_s", c_f64)]


# every symbol declared in include/necsim_hip.h, with its signature
SIGNATURES = {
    "nh_abi_version": (ctypes.c_int, []),
    "nh_last_error": (ctypes.c_int, [ctypes.c_char_p, c_sz]),
    "nh_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    "nh_ctx_destroy": (ctypes.c_int, [c_vp]),
    "nh_device_info": (ctypes.c_int, [c_vp, ctypes.POINTER(c_u32), ctypes.c_char_p, c_sz]),
    "nh_tables_create": (ctypes.c_int, [c_vp, ctypes.POINTER(Scenario), ctypes.POINTER(c_vp)]),
    "nh_tables_destroy": (ctypes.c_int, [c_vp]),
    "nh_tables_footprint": (ctypes.c_int, [c_vp, ctypes.POINTER(c_u64)]),
    "nh_tables_create_from_map": (ctypes.c_int, [c_vp, ctypes.POINTER(MapScenario), ctypes.POINTER(c_vp)]),
    "nh_tables_export": (ctypes.c_int, [c_vp, c_vp, c_sz, c_vp, c_sz, ctypes.POINTER(c_sz),
                                        ctypes.POINTER(c_sz)]),
    "nh_job_create": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(JobParams), ctypes.POINTER(c_vp)]),
    "nh_classify_lineages": (ctypes.c_int, [c_vp, c_vp, c_sz, c_vp, ctypes.POINTER(c_sz)]),
    "nh_job_add_lineages": (ctypes.c_int, [c_vp, c_vp, c_sz]),
    "nh_job_add_origins": (ctypes.c_int, [c_vp, c_f64, c_u32, c_u32, c_u32, c_u32, c_u32]),
    "nh_job_start": (ctypes.c_int, [c_vp]),
    "nh_job_step_round": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i32)]),
    "nh_job_simulate": (ctypes.c_int, [c_vp]),
    "nh_job_drain_events": (ctypes.c_int, [c_vp, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_job_drain_packed_events": (ctypes.c_int, [c_vp, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_job_event_counts": (ctypes.c_int, [c_vp, ctypes.POINTER(EventCounts)]),
    "nh_job_outcome": (ctypes.c_int, [c_vp, ctypes.POINTER(Outcome)]),
    "nh_job_take_lineages": (ctypes.c_int, [c_vp, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_job_destroy": (ctypes.c_int, [c_vp]),
    "nh_job_kernel_time": (ctypes.c_int, [c_vp, ctypes.POINTER(c_f64), ctypes.POINTER(c_u64)]),
    "nh_autotune": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(JobParams), c_vp, c_sz, c_f64,
                                   ctypes.POINTER(JobParams), ctypes.POINTER(c_f64)]),
    "nh_server_create": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(ServerParams), ctypes.POINTER(c_vp)]),
    "nh_server_request_footprint": (ctypes.c_int, [c_u64, c_u64, ctypes.POINTER(c_u64)]),
    "nh_server_submit": (ctypes.c_int, [c_vp, ctypes.POINTER(Request), ctypes.POINTER(c_u64)]),
    "nh_server_tick": (ctypes.c_int, [c_vp, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "nh_server_poll": (ctypes.c_int, [c_vp, ctypes.POINTER(RequestResult)]),
    "nh_server_request_events": (ctypes.c_int, [c_vp, c_u64, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_server_release": (ctypes.c_int, [c_vp, c_u64]),
    "nh_server_destroy": (ctypes.c_int, [c_vp]),
    "nh_seed_from_u64": (c_u64, [c_u64]),
    "nh_build_alias_tables": (ctypes.c_int, [c_u32, c_u32, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                             c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_debug_eval": (ctypes.c_int, [c_vp, c_i32, c_vp, c_vp, c_sz]),
    "nh_event_log_write_segment": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_sz]),
    "nh_event_log_write_packed_segment": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_sz, c_i32, c_i32]),
    "nh_lineage_file_write": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_sz]),
    "nh_lineage_file_read": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_copy_calibrate": (ctypes.c_int, [c_vp, c_sz, c_i32, ctypes.POINTER(ctypes.c_double)]),
    "nh_debug_sort_events": (ctypes.c_int, [c_vp, c_vp, c_sz, c_i32, c_vp]),
    "nh_sort_events_noalloc": (ctypes.c_int, [c_vp, c_vp, c_sz, c_i32, c_vp]),
    "nh_debug_sort_packed_events": (ctypes.c_int, [c_vp, c_vp, c_sz, c_i32, c_vp]),
    "nh_debug_count_packed_events": (ctypes.c_int, [c_vp, c_vp, c_sz, ctypes.POINTER(EventCounts)]),
    "nh_debug_sort_count_packed_events": (ctypes.c_int, [c_vp, c_vp, c_sz, ctypes.POINTER(EventCounts)]),
    "nh_sample_origins": (ctypes.c_int, [ctypes.POINTER(Scenario), c_f64, c_u32, c_u32, c_u32, c_u32,
                                         c_u32, c_vp, c_sz, ctypes.POINTER(c_sz)]),
    "nh_sample_origins_device": (ctypes.c_int, [c_vp, c_vp, c_f64, c_u32, c_u32, c_u32, c_u32, c_u32, c_vp,
                                                c_sz, ctypes.POINTER(c_sz)]),
    "nh_debug_presample": (ctypes.c_int, [c_vp, c_f64, c_u64, c_i32, c_vp]),
    "nh_debug_repeated_sum": (ctypes.c_int, [c_f64, c_u64, ctypes.POINTER(c_f64)]),
}

_lib = None


class NecsimHipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


def lib():
    """Load libnecsim_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with necsim-rust_amd/build.py "
                              "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().nh_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != NH_OK:
        raise NecsimHipError(rc, last_error())
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data
