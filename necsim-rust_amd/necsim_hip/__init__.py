"""MI355X (gfx950) Independent coalescence algorithm for necsim-rust scenarios.

The simulation runs in libnecsim_hip.so (hand-written HIP kernels behind the C
ABI of include/necsim_hip.h); this package is the Python host mirror of
rustcoalescence's algorithm plugin interface and the synthetic workload
generators used by the tests and the benchmark.
"""
from . import _ffi
from ._ffi import EVENT_DTYPE, LINEAGE_DTYPE, PACKED_EVENT_DTYPE, NecsimHipError, lib
from .algorithm import (PERSISTENT, PLAIN, TASK_LIST, AlmostInfiniteNormalDispersalScenario,
                        BiodiversityReporter, Context, DedupCache, EventCollector, EventLogReporter,
                        EventSlice,
                        HipIndependentAlgorithm, IndependentArguments, Job, NonSpatialScenario,
                        OriginPreSampler, PackedEventCollector, ResumeSampleError, ScenarioCogs, Server,
                        SimulationOutcome,
                        SpatiallyExplicitUniformTurnoverScenario, Tables, WyHash,
                        autotune, build_alias_tables, copy_calibrate, is_speciation, load_lineages,
                        packed_order, save_lineages, speciation_events)
from .maps import generate_map
from .reporters import CsvReporter, EventCounterReporter, rust_f64_display

__all__ = [n for n in dir() if not n.startswith("_")]
