/*
 * necsim_hip.h -- C ABI of the MI355X (gfx950) Independent coalescence
 * algorithm.
 *
 * The library is a drop-in for necsim-rust's CPU `IndependentAlgorithm`
 * (rustcoalescence/algorithms/independent/src/lib.rs:39-173) and the existing
 * GPU `CudaAlgorithm` (rustcoalescence/algorithms/cuda/src/lib.rs:66-593):
 * a Rust glue crate implements the `Algorithm` trait
 * (rustcoalescence/algorithms/src/lib.rs:55-112) on top of these entry points.
 *
 * Conventions
 *  - Every function returns 0 (NH_OK) on success and a negative NH_E* code on
 *    failure; nh_last_error() then describes the failure of the calling thread.
 *  - All pointers are plain host pointers; the caller owns every host buffer.
 *    Device memory is owned by the library and freed by the nh_*_destroy calls.
 *  - A context owns one HIP stream on one device.  Handles are not
 *    thread-safe: use each context (and the objects created from it) from one
 *    thread at a time.  Concurrency comes from batching jobs inside one
 *    persistent kernel (the server API), not from concurrent calls.
 *  - Struct layouts of nh_lineage, nh_speciation_event, nh_packed_event,
 *    nh_alias_range, nh_self_dispersal and nh_alias_atom equal the reference's #[repr(C)]
 *    types on 64-bit targets, so the glue passes its own buffers unchanged.
 */
#ifndef NECSIM_HIP_H
#define NECSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NH_ABI_VERSION 1

/* ---- return codes ------------------------------------------------------- */
enum {
    NH_OK = 0,
    NH_E_INVALID = -1,     /* invalid argument (null pointer, bad enum, bad size) */
    NH_E_HIP = -2,         /* a HIP runtime call failed (no device, fault, ...) */
    NH_E_NOMEM = -3,       /* device or host allocation failed */
    NH_E_STATE = -4,       /* call not valid in the object's current state */
    NH_E_CONTRACT = -5,    /* scenario violates the in-memory dispersal contract */
    NH_E_OUT_OF_DEME = -6, /* a lineage lies outside the habitat or its deme */
    NH_E_CAPACITY = -7,    /* a caller buffer or a server limit is too small */
};

/* ---- plain data types (reference #[repr(C)] layouts) --------------------- */

/* necsim/core/src/lineage.rs:97-109 `Lineage` (32 B) */
typedef struct {
    uint64_t global_reference;
    double last_event_time;
    uint32_t x, y, index;
    uint32_t _pad;
} nh_lineage;

/* necsim/core/src/event.rs:58-71 `SpeciationEvent` (40 B) */
typedef struct {
    uint64_t global_reference;
    double prior_time;
    double event_time;
    uint32_t x, y, index;
    uint32_t _pad;
} nh_speciation_event;

/* necsim/core/src/event.rs:16-29 `PackedEvent` (56 B): origin (x, y, index),
 * target (x, y, index) and the interaction's coalescence reference.  A
 * speciation has the sign bit of prior_time set and target = origin; a
 * dispersal of the Independent algorithm (LineageInteraction::Maybe) has the
 * sign bit of event_time set; coalescence = the lineage's own reference for
 * both (event.rs:103-155). */
typedef struct {
    uint64_t global_reference;
    double prior_time;
    double event_time;
    uint32_t origin_x, origin_y, origin_index;
    uint32_t target_x, target_y, target_index;
    uint64_t coalescence;
} nh_packed_event;

/* necsim/impls/no-std/src/alias/packed.rs:11-15 `AliasMethodSamplerAtom<usize>` (24 B) */
typedef struct {
    double u;
    uint64_t e;
    uint64_t k;
} nh_alias_atom;

/* .../dispersal_sampler/in_memory/packed_alias/mod.rs `AliasSamplerRange` (16 B) */
typedef struct {
    uint64_t start;
    uint64_t end;
} nh_alias_range;

/* .../in_memory/packed_separable_alias/mod.rs:35-44 `SeparableAliasSelfDispersal` (16 B) */
typedef struct {
    double self_dispersal;
    uint64_t non_self_dispersal_event;
} nh_self_dispersal;

/* ---- scenarios ------------------------------------------------------------ */
enum {
    /* NonSpatialHabitat + NonSpatialDispersalSampler
     * (rustcoalescence/scenarios/src/non_spatial.rs:39-89) */
    NH_SCENARIO_NON_SPATIAL = 0,
    /* AlmostInfiniteHabitat + AlmostInfiniteNormalDispersalSampler
     * (rustcoalescence/scenarios/src/almost_infinite/normal.rs:43-94) */
    NH_SCENARIO_ALMOST_INFINITE_NORMAL = 1,
    /* InMemoryHabitat + InMemoryPackedSeparableAliasDispersalSampler
     * (rustcoalescence/scenarios/src/spatially_explicit/turnover/uniform.rs:49-117) */
    NH_SCENARIO_SPATIALLY_EXPLICIT = 2,
    /* InMemoryHabitat + InMemoryPackedAliasDispersalSampler
     * (necsim/impls/no-std/src/cogs/dispersal_sampler/in_memory/packed_alias/mod.rs:49-119) */
    NH_SCENARIO_SPATIALLY_EXPLICIT_PACKED = 3,
};

/* The scenario cogs of rustcoalescence_scenarios::ScenarioCogs
 * (rustcoalescence/scenarios/src/lib.rs:44-95) with uniform turnover and
 * speciation.  For the in-memory kinds the caller hands over the habitat and
 * the already built alias tables (row-major over the habitat extent). */
typedef struct {
    int32_t kind;
    uint32_t width, height;       /* extent (non-spatial, in-memory) */
    uint32_t origin_x, origin_y;  /* extent origin (non-spatial, in-memory) */
    uint32_t deme;                /* non-spatial deme size */
    double sigma;                 /* almost-infinite Gaussian sigma */
    /* in-memory only */
    const uint32_t *habitat;                 /* width*height */
    const nh_alias_range *alias_ranges;      /* width*height */
    const nh_self_dispersal *self_dispersal; /* width*height, separable kind only */
    const nh_alias_atom *atoms;
    size_t n_atoms;
} nh_scenario;

/* ---- algorithm arguments -------------------------------------------------- */
enum { NH_DEDUP_NONE = 0, NH_DEDUP_RELATIVE = 1, NH_DEDUP_ABSOLUTE = 2 };
enum { NH_SLICE_RELATIVE = 1, NH_SLICE_ABSOLUTE = 2 };
enum {
    /* one kernel launch per step slice, lineages processed in the reference's
     * water-level queue order (bit-exact with DedupCache::Relative/Absolute) */
    NH_FORM_PLAIN = 0,
    /* persistent lane-refilling kernel per water-level round (bit-exact with
     * DedupCache::None; with dedup on, which of two coalesced duplicates
     * survives may differ, while species and unique events do not) */
    NH_FORM_PERSISTENT = 1,
    /* the reference CUDA algorithm's launch structure
     * (rustcoalescence/algorithms/cuda/src/parallelisation/monolithic.rs:222-318):
     * a task list of task_list_size lineages per launch, uploaded before and
     * downloaded after it, with the dedup cache and the slow/fast queues on the
     * host (bit-exact with DedupCache::None; with dedup on it keeps that loop's
     * queue order, which is not the CPU algorithm's) */
    NH_FORM_TASK_LIST = 2,
};

/* IndependentArguments (rustcoalescence/algorithms/independent/src/arguments.rs:79-146)
 * plus the simulate call's rng / pause_before / resume_after inputs and the
 * device launch knobs. */
typedef struct {
    uint64_t rng_seed;             /* WyHash seed (= seed_from_u64 of the RON Seed) */
    double speciation_probability; /* UniformSpeciationProbability nu */
    double turnover_rate;          /* UniformTurnoverRate lambda (0.5 by default) */
    double delta_t;                /* default 2.0 */
    uint64_t step_slice;           /* default 10 */
    int32_t dedup_mode;            /* default NH_DEDUP_RELATIVE, factor 1.0 */
    double dedup_factor;
    uint64_t dedup_capacity;
    int32_t event_slice_mode;      /* default NH_SLICE_RELATIVE, factor 2.0 */
    double event_slice_factor;
    uint64_t event_slice_capacity;
    int32_t has_pause_before;
    double pause_before;
    double resume_after;           /* min event time; 0.0 for a fresh simulation */
    int32_t form;                  /* NH_FORM_* */
    uint32_t block_size;           /* threads per block, 0 = tuned default */
    uint32_t blocks_per_cu;        /* persistent form: resident blocks per CU, 0 = tuned */
    uint32_t use_graph;            /* persistent form: capture each round in a hipGraph */
    uint32_t inner_steps;          /* persistent form: iterations per lane between refills, 0 = default 16 */
    uint32_t exec_fill;            /* persistent form: a wave executes its lanes' next events once at least
                                      exec_fill eighths of its active lanes have found theirs (1..8);
                                      0 = default 5, or 1 for the spatially explicit (alias) scenarios */
    int32_t report_dispersal;      /* the reporter's ReportDispersal: every event is reported as a
                                      PackedEvent (nh_job_drain_packed_events) and the water level
                                      advances by event_slice / sum(lambda) (monolithic/mod.rs:139-148);
                                      0 = speciation only (nh_job_drain_events) */
    uint64_t event_buffer;         /* report_dispersal: PackedEvents one round may report;
                                      0 = default 2 * event_slice + lineages + 65536 */
    int32_t device_event_counter;  /* report_dispersal: the events stay on the device, where an
                                      EventCounterReporter counts them (nh_job_event_counts);
                                      nothing is left to drain */
    uint32_t plain_batch;          /* plain form: step-slice passes run back to back on the device per
                                      host synchronisation (pass sizes stay on the device); 1 = one
                                      host round trip per pass, the reference's launch structure;
                                      0 = default 16.  Results do not depend on it. */
    uint32_t coop_search;          /* persistent form: lanes that are not searching evaluate the next
                                      Poisson windows of those that are (shortens the serial chains of
                                      sparse waves); 0 = automatic (with dedup, or fewer lineages
                                      than lanes, and not with report_dispersal), 1 = off,
                                      2 = on.  Results do not depend on it. */
    uint32_t task_list_size;       /* NH_FORM_TASK_LIST: lineages per launch; 0 = 4096, the reference
                                      CUDA algorithm's default grid x block of 64 x 64 */
    uint32_t pool;                 /* persistent form without dedup cache: lineages per wave kept
                                      in LDS by the pooled kernel, whose lanes
                                      take whichever of them needs a search step or an execution
                                      (128 or 192; one-wave blocks, blocks_per_cu counts waves);
                                      0 = the one-lineage-per-lane kernel.  Results do not depend on it. */
} nh_job_params;

typedef struct {
    int32_t paused;      /* Status::Paused (lineages remain) */
    double time;         /* SimulationOutcome time: max event time */
    uint64_t steps;      /* SimulationOutcome steps: completed events */
    uint64_t n_events;   /* events reported so far (speciations, plus dispersals with report_dispersal) */
    uint64_t n_remaining;/* lineages left (paused) */
    uint64_t rounds;     /* water-level rounds run */
    uint64_t launches;   /* step-kernel launches */
} nh_outcome;

/* The counters of the EventCounterReporter
 * (necsim/plugins/common/src/event_counter.rs:12-26, report rules :58-109),
 * in field order. */
typedef struct {
    uint64_t raw_total;
    uint64_t speciation;
    uint64_t out_dispersal;
    uint64_t self_dispersal;
    uint64_t out_coalescence;
    uint64_t self_coalescence;
    uint64_t late_dispersal_coalescence;
    uint64_t late_coalescence;
} nh_event_counts;

typedef struct nh_ctx nh_ctx;
typedef struct nh_tables nh_tables;
typedef struct nh_job nh_job;
typedef struct nh_server nh_server;

/* ---- library / context ---------------------------------------------------- */
int nh_abi_version(void);
/* Describe the last failure of the calling thread (NUL-terminated, truncated). */
int nh_last_error(char *buf, size_t cap);
/* rustcoalescence/algorithms/cuda/src/cuda.rs:11-37 (`with_initialised_cuda`):
 * select the device and create the context's stream. */
int nh_init(int device, nh_ctx **out);
int nh_ctx_destroy(nh_ctx *ctx);
/* number of compute units and device name */
int nh_device_info(nh_ctx *ctx, uint32_t *compute_units, char *name, size_t name_cap);

/* ---- scenario tables ------------------------------------------------------ */
/* Upload the scenario cogs (the `LendRustToCuda` borrow of habitat and
 * dispersal sampler in the CUDA algorithm, .../packed_separable_alias/mod.rs:47-48)
 * into resident device tables. */
int nh_tables_create(nh_ctx *ctx, const nh_scenario *scenario, nh_tables **out);
int nh_tables_destroy(nh_tables *tables);

/* An in-memory map scenario as the reference's SpatiallyExplicit map scenario
 * reads it (rustcoalescence/scenarios/src/spatially_explicit/maps/mod.rs): a
 * habitat map and a dispersal map, the latter row-major per origin cell with
 * the row's entries in ascending column order (a dense Array2D row filtered
 * to its non-zero weights yields exactly these entries). */
typedef struct {
    int32_t kind;                 /* NH_SCENARIO_SPATIALLY_EXPLICIT (separable) or _PACKED */
    uint32_t width, height;
    uint32_t origin_x, origin_y;
    const uint32_t *habitat;      /* width*height */
    const uint64_t *row_ptr;      /* width*height + 1, row_ptr[0] = 0, non-decreasing */
    const uint32_t *col;          /* row_ptr[width*height] target cells */
    const double *weight;         /* row_ptr[width*height] dispersal weights */
} nh_map_scenario;

/* InMemoryPackedSeparableAliasDispersalSampler::new
 * (.../in_memory/packed_separable_alias/mod.rs:64-203) or
 * InMemoryPackedAliasDispersalSampler::new (.../in_memory/packed_alias/mod.rs:58-118)
 * followed by nh_tables_create, built on the device (one origin row per
 * thread): the same resident tables as nh_build_alias_tables +
 * nh_tables_create, bit for bit, with the same contract errors
 * (.../in_memory/contract.rs:13-69). */
int nh_tables_create_from_map(nh_ctx *ctx, const nh_map_scenario *map, nh_tables **out);

/* Diagnostics, no counterpart in the reference: copy resident in-memory
 * tables back, 32-byte cell records {prefix u64, habitat u32, atom_start u32,
 * atom_len u32, non_self u32, self_p f64} and 16-byte atoms {u f64, e u32,
 * k u32}.  With null buffers only the counts are returned. */
int nh_tables_export(const nh_tables *tables, void *cells, size_t cells_cap, void *atoms, size_t atoms_cap,
                     size_t *n_cells, size_t *n_atoms);
/* device bytes held by the tables */
int nh_tables_footprint(const nh_tables *tables, uint64_t *bytes);

/* ---- one simulation (Algorithm::initialise_and_simulate / resume_and_simulate) ---- */
int nh_job_create(nh_ctx *ctx, const nh_tables *tables, const nh_job_params *params, nh_job **out);

/* Lineage classes of IndependentActiveLineageSampler::resume_with_store_and_lineages
 * (necsim/impls/no-std/src/cogs/active_lineage_sampler/independent/mod.rs:98-117),
 * i.e. the variants of ExceptionalLineage it produces
 * (.../active_lineage_sampler/resuming/lineage.rs:7-14). */
enum {
    NH_LINEAGE_OK = 0,
    NH_LINEAGE_OUT_OF_DEME = 1,     /* index >= the habitat at its location */
    NH_LINEAGE_OUT_OF_HABITAT = 2,  /* location outside the extent or of habitat 0 */
};
/* Classify resumed lineages against the scenario (kinds[n] receives
 * NH_LINEAGE_*; *n_exceptional, if not NULL, the number not OK).  The resume
 * path of the Independent algorithm (rustcoalescence/algorithms/independent/
 * src/initialiser/resume.rs:71-85) fails with ResumeError::Sample(every
 * exceptional lineage) when any is not OK; the glue builds that list from
 * kinds before it adds the lineages to a job. */
int nh_classify_lineages(const nh_tables *tables, const nh_lineage *lineages, size_t n, int32_t *kinds,
                         size_t *n_exceptional);
/* Append lineages in origin-sampler order (the order defines the water-level
 * queue, parallelisation/independent/monolithic/mod.rs:94-103).  Lineages
 * that nh_classify_lineages does not class OK fail with NH_E_OUT_OF_DEME
 * (active_lineage_sampler/independent/mod.rs:98-117) and nothing is added. */
int nh_job_add_lineages(nh_job *job, const nh_lineage *lineages, size_t n);
/* The job's lineages from the scenario's origin sampler, run on the device
 * straight into the job's lineage store (nh_sample_origins_device's
 * lineages, never copied through the host; the task-list form, whose queues
 * live on the host, downloads them).  Replaces the genesis initialiser's
 * origin walk (rustcoalescence/algorithms/independent/src/initialiser/
 * genesis.rs:39-72 with OriginPreSampler::all().percentage(p).partition(rank,
 * size), pre_sampler.rs:71-114).  A job takes either nh_job_add_lineages or
 * one nh_job_add_origins (NH_E_STATE otherwise). */
int nh_job_add_origins(nh_job *job, double percentage, uint32_t rank, uint32_t size,
                       uint32_t circle_centre_x, uint32_t circle_centre_y, uint32_t circle_radius);
/* Allocate the job's device state and upload its lineages (done implicitly by
 * the first nh_job_step_round; callable earlier so that the upload stays out
 * of a timed region).  No counterpart in the reference: the CUDA algorithm
 * uploads inside its launch loop (cuda/src/parallelisation/monolithic.rs:222-246). */
int nh_job_start(nh_job *job);
/* Run one water-level round (monolithic/mod.rs:135-255).  *finished is set to
 * 1 when no lineage is left or pause_before has been reached. */
int nh_job_step_round(nh_job *job, int32_t *finished);
/* Run rounds until finished (parallelisation::independent::monolithic::simulate). */
int nh_job_simulate(nh_job *job);
/* Speciation events produced since the last drain, sorted by the PackedEvent
 * order (necsim/core/src/event.rs:256-280), i.e. what the live water-level
 * reporter (monolithic/reporter/live.rs:88-115) hands to the reporter.  Call
 * with buf == NULL to query the count in *n_out. */
int nh_job_drain_events(nh_job *job, nh_speciation_event *buf, size_t cap, size_t *n_out);
/* The same for a job with report_dispersal: every speciation and dispersal
 * event since the last drain as PackedEvents, sorted by the PackedEvent
 * order.  Each round's events are kept in a device buffer of event_buffer
 * entries (by default 2 * event_slice + lineages + 65536, about twice what a
 * round reports on average); a round that reports more fails with
 * NH_E_CAPACITY and leaves the job failed (later calls: NH_E_STATE).  Either drain call on a job of the other kind
 * fails with NH_E_STATE. */
int nh_job_drain_packed_events(nh_job *job, nh_packed_event *buf, size_t cap, size_t *n_out);
/* A device_event_counter job: the EventCounterReporter's counters over every
 * event reported so far, counted on the device from each round's sorted
 * events (the reporter's report_speciation / report_dispersal,
 * event_counter.rs:58-109).  Other jobs fail with NH_E_STATE. */
int nh_job_event_counts(nh_job *job, nh_event_counts *out);
int nh_job_outcome(const nh_job *job, nh_outcome *out);
/* Lineages remaining after a paused run (SimulationOutcome::Paused.lineages) */
int nh_job_take_lineages(nh_job *job, nh_lineage *buf, size_t cap, size_t *n_out);
int nh_job_destroy(nh_job *job);
/* Mean device time per step-kernel launch (ms) over the run, from HIP events
 * recorded on the job's stream, and the number of timed launches. */
int nh_job_kernel_time(const nh_job *job, double *total_ms, uint64_t *launches);

/* Autotuner of the persistent form (no counterpart in the reference, whose
 * CUDA launch shape is fixed by arguments, cuda/src/arguments.rs:124-138).
 * Times short pause windows ([0, window)) on `sample` for every candidate
 * block size (128/256), resident blocks per CU (occupancy limit and half of
 * it), iterations per lane between refills (4, 16), execution fill (1, 5, 8
 * eighths) and, when base->step_slice == 0, step slice (5, 10, 20, 50, 150);
 * without dedup cache also the pooled kernel with 128 and 192 lineages per
 * wave; and writes the fastest combination (least time
 * for the window) to *tuned (a copy of *base with block_size, blocks_per_cu,
 * inner_steps, exec_fill, step_slice and pool set).
 * *best_ms receives the winner's window time. */
int nh_autotune(nh_ctx *ctx, const nh_tables *tables, const nh_job_params *base,
                const nh_lineage *sample, size_t n, double window, nh_job_params *tuned,
                double *best_ms);

/* ---- serving: continuous batching of simulations on resident tables ------- */
/* No counterpart in the reference (which runs one simulation per process).
 * Request memory comes from one device allocation of hbm_budget_bytes made by
 * nh_server_create: admission and retirement never allocate or free device
 * memory.  A request's block is nh_server_request_footprint() bytes. */
typedef struct {
    uint64_t max_lineages;      /* lane-state capacity shared by requests in flight */
    uint64_t hbm_budget_bytes;  /* device memory of the request pool (allocated at creation) */
    uint64_t tick_steps;        /* max steps per lineage per kernel tick (latency knob); a lane
                                   stops only at a step-slice boundary, so a tick may run up to
                                   step_slice - 1 steps more */
    uint32_t block_size;        /* 0 = default (256) */
    uint32_t blocks_per_cu;     /* 0 = default (1): a tick's latency is set by the
                                   slowest requests' serial lineage chains, which
                                   run faster with few resident waves per CU */
    uint64_t max_bypass_ticks;  /* admission runs in arrival order, but a request that does not
                                   fit yet may be passed by later ones for this many ticks; then
                                   nothing is admitted past it until it fits (0 = default 16) */
} nh_server_params;

typedef struct {
    uint64_t rng_seed;             /* WyHash seed */
    double speciation_probability;
    double turnover_rate;
    double delta_t;
    uint64_t step_slice;
    int32_t dedup_mode;
    double dedup_factor;
    uint64_t dedup_capacity;
    int32_t event_slice_mode;
    double event_slice_factor;
    uint64_t event_slice_capacity;
    int32_t has_pause_before;
    double pause_before;
    const nh_lineage *lineages;    /* copied at submission */
    size_t n_lineages;
} nh_request;

typedef struct {
    uint64_t request_id;
    int32_t status;                /* NH_OK or the request's error */
    int32_t paused;
    double time;
    uint64_t steps;
    uint64_t n_events;
    double submit_s, admit_s, done_s; /* host monotonic clock (s) */
} nh_request_result;

int nh_server_create(nh_ctx *ctx, const nh_tables *tables, const nh_server_params *params,
                     nh_server **out);
/* Device bytes a request of n_lineages lineages and cache_buckets dedup-cache
 * buckets takes from the pool: 130 B per lineage (two lineage-state copies
 * and one event slot), 32 B per bucket and 4608 B of alignment, rounded up to
 * a multiple of 256 B.  Requests whose footprints sum to at most the pool
 * size fit in it together. */
int nh_server_request_footprint(uint64_t n_lineages, uint64_t cache_buckets, uint64_t *bytes);
/* Queue a simulation; returns immediately with its id.  The request is
 * checked like nh_job_create + nh_job_add_lineages (invalid arguments:
 * NH_E_INVALID; a lineage that nh_classify_lineages does not class OK:
 * NH_E_OUT_OF_DEME; a footprint above the pool or more lineages than
 * max_lineages: NH_E_CAPACITY), so no kernel ever runs on a bad request. */
int nh_server_submit(nh_server *srv, const nh_request *req, uint64_t *request_id);
/* Advance the server by one kernel tick: admit queued requests that fit,
 * run the shared persistent kernel over every admitted request, retire
 * finished ones.  *n_active / *n_queued report the load afterwards.  A
 * request whose own step fails (upload, compaction, event sort) is finished
 * with that error in its result's status and the others go on; the call then
 * returns the first such error.  If the shared kernel tick itself fails,
 * every active request is finished with the error. */
int nh_server_tick(nh_server *srv, uint64_t *n_active, uint64_t *n_queued);
/* Pop one finished request (status NH_E_STATE if none is ready). */
int nh_server_poll(nh_server *srv, nh_request_result *result);
/* Sorted speciation events of a finished, polled request. */
int nh_server_request_events(nh_server *srv, uint64_t request_id, nh_speciation_event *buf,
                             size_t cap, size_t *n_out);
int nh_server_release(nh_server *srv, uint64_t request_id);
int nh_server_destroy(nh_server *srv);

/* ---- host helpers (scenario construction for standalone C/C++/Python users;
 *      the Rust glue uses the reference's own scenario code instead) ------- */

/* SeedableRng::seed_from_u64 + WyHash::from_seed (necsim/core/src/cogs/rng.rs:33-63) */
uint64_t nh_seed_from_u64(uint64_t seed);
/* InMemoryPackedSeparableAliasDispersalSampler::new (packed_separable_alias/mod.rs:64-203)
 * or InMemoryPackedAliasDispersalSampler::new (packed_alias/mod.rs:58-118) from a
 * sparse dispersal map: row r's entries (col ascending) are
 * col[row_ptr[r]..row_ptr[r+1]], weight[...]; entries absent from a row are 0.0.
 * Writes alias_ranges[width*height], self_dispersal[width*height] (separable
 * only, may be NULL otherwise) and atoms (capacity atoms_cap); *n_atoms
 * receives the atom count (query with atoms == NULL). */
int nh_build_alias_tables(uint32_t width, uint32_t height, const uint32_t *habitat,
                          const uint64_t *row_ptr, const uint32_t *col, const double *weight,
                          int32_t separable, nh_alias_range *alias_ranges,
                          nh_self_dispersal *self_dispersal, nh_alias_atom *atoms,
                          size_t atoms_cap, size_t *n_atoms);
/* OriginPreSampler::all().percentage(p).partition(rank, size) fed into the
 * scenario's origin sampler (non_spatial.rs:89-112, in_memory.rs:89-119,
 * singleton_demes/circle.rs:123-156).  Query with out == NULL. */
int nh_sample_origins(const nh_scenario *scenario, double percentage, uint32_t rank,
                      uint32_t size, uint32_t circle_centre_x, uint32_t circle_centre_y,
                      uint32_t circle_radius, nh_lineage *out, size_t cap, size_t *n_out);
/* The same origin samplers on the device, over resident tables: the lineages
 * nh_sample_origins gives, in its order, written to the DEVICE buffer d_out
 * (query the count with d_out == NULL; NH_E_CAPACITY if cap is smaller).
 * Replaces the origin walk the reference runs before simulating
 * (rustcoalescence/algorithms/independent/src/initialiser/genesis.rs:39-72, the
 * samplers above and pre_sampler.rs:71-114).  Circle radii up to 2^22. */
int nh_sample_origins_device(nh_ctx *ctx, const nh_tables *tables, double percentage, uint32_t rank,
                             uint32_t size, uint32_t circle_centre_x, uint32_t circle_centre_y,
                             uint32_t circle_radius, nh_lineage *d_out, size_t cap, size_t *n_out);

/* ---- event log ------------------------------------------------------------ */
/* EventLogRecorder::sort_and_write_segment
 * (necsim/impls/std/src/event_log/recorder.rs:174-208) for speciation events:
 * writes one event-log segment file, the bincode 1.3 (default options) image
 * of EventLogHeader (event_log/mod.rs:9-19, with_speciation = true,
 * with_dispersal = false) followed by n 56-byte PackedEvents
 * (necsim/core/src/event.rs:16-29).  The events must be non-empty and in
 * PackedEvent order, as nh_job_drain_events returns them; the reference's
 * replay (event_log/replay) reads such segments. */
int nh_event_log_write_segment(const char *path, const nh_speciation_event *events, size_t n);
/* The same for PackedEvents (speciations and dispersals as they are), with
 * the header's with_speciation / with_dispersal flags given by the caller
 * (the reporter's ReportSpeciation / ReportDispersal, recorder.rs:146-152). */
int nh_event_log_write_packed_segment(const char *path, const nh_packed_event *events, size_t n,
                                      int32_t with_speciation, int32_t with_dispersal);

/* ---- lineage files (pause / resume) -------------------------------------- */
/* LineageFileSaver (necsim/impls/std/src/lineage_file/saver.rs:54-75): writes
 * lineages as the bincode 1.3 `bincode::options()` (varint) image of a
 * Vec<Lineage>; the file must not exist yet.  Paused lineages from
 * nh_job_take_lineages round-trip through it into nh_job_add_lineages. */
int nh_lineage_file_write(const char *path, const nh_lineage *lineages, size_t n);
/* LineageFileLoader (lineage_file/loader.rs:33-45).  With buf == NULL only the
 * count is returned in *n_out. */
int nh_lineage_file_read(const char *path, nh_lineage *buf, size_t cap, size_t *n_out);

/* ---- diagnostics ---------------------------------------------------------
 * Evaluate one device primitive elementwise on the GPU (no counterpart in the
 * reference: used to pin the device MathsCore and RNG against the CPU).
 * op: 0 ln, 1 exp, 2 sin, 3 cos, 4 sqrt, 5 floor, 6 round, 7 divide
 * (in[2i] / in[2i+1]), 8 wymum (u64 pairs), 9 seahash_diffuse (u64),
 * 10 prime (in[3i..3i+2] = seed, location, time; u64), 11 uniform_closed_open
 * after prime (same inputs, f64 out), 12 the Poisson event-time sampler's
 * next event (in[3i..3i+2] = seed, location, time bits for i < n / 2, then
 * bits of lambda and delta_t; out[2i] = event time bits, out[2i + 1] = the
 * WyHash state primed for that event's draws).  in / out are host buffers
 * of 8-byte words; n is the number of outputs. */
int nh_debug_eval(nh_ctx *ctx, int32_t op, const uint64_t *in, uint64_t *out, size_t n);
/* Diagnostics, no counterpart in the reference: the first n ranks of the
 * percentage pre-sampler (OriginPreSampler::all().percentage(p),
 * pre_sampler.rs:71-101) computed on the device into the host buffer out:
 * from the closed form of the golden-ratio sequence in parallel (serial = 0,
 * what the origin samplers use) or by the reference's f64 loop on one thread
 * (serial = 1).  0 < p < 1. */
int nh_debug_presample(nh_ctx *ctx, double percentage, uint64_t n, int32_t serial, uint64_t *out);
/* Diagnostics, no counterpart in the reference: *out = the sequential sum
 * x + x + ... + x of n terms folded from 0.0, as the host evaluates the
 * water-level rate sum Σλ(ν) over n lineages of one rate
 * (parallelisation/independent/monolithic/mod.rs:139-164), bit for bit equal
 * to the fold but computed binade by binade.  Host only. */
int nh_debug_repeated_sum(double x, uint64_t n, double *out);

/* Diagnostics, no counterpart in the reference: times `reps` launches of a
 * plain 16-byte-vector copy kernel (k_copy_calibrate) that reads and writes
 * `bytes` bytes each, on the context's stream, after a fill and a warm-up
 * launch; *ms_per_copy = mean duration.  Its known byte count calibrates HBM
 * byte counters (FETCH_SIZE / WRITE_SIZE) and gives the achievable copy
 * bandwidth.  Allocates 2 x bytes of device memory for the call. */
int nh_copy_calibrate(nh_ctx *ctx, size_t bytes, int32_t reps, double *ms_per_copy);

/* Diagnostics: sorts n events (host buffers) on the device by the
 * PackedEvent order (necsim/core/src/event.rs:256-280), the order in which
 * LiveWaterLevelReporterProxy::advance_water_level reports each water-level
 * round's events (monolithic/reporter/live.rs:88-115), with the eager sort
 * of nh_job_step_round / nh_server_tick (device_sized == 0) or the
 * graph-capturable variant (device_sized != 0).  Allocates device memory. */
int nh_debug_sort_events(nh_ctx *ctx, const nh_speciation_event *in, size_t n, int32_t device_sized,
                         nh_speciation_event *out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The sort of nh_debug_sort_events over DEVICE buffers:
 * d_in holds n speciation events, d_out (distinct from d_in) receives them in
 * PackedEvent order.  Its sort buffers live in one file-scope device array of
 * 25 MiB per GPU, held until the process ends; nothing is allocated or freed,
 * and the call returns after the sort has finished on the context's stream.
 * n above 1,048,576 events fails with NH_E_CAPACITY.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several
 * threads or contexts) race and must be serialised by the caller. */
int nh_sort_events_noalloc(nh_ctx *ctx, const nh_speciation_event *d_in, size_t n, int32_t device_sized,
                           nh_speciation_event *d_out);
/* The same sort over PackedEvents (dispersal reporting's records). */
int nh_debug_sort_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n, int32_t device_sized,
                                nh_packed_event *out);
/* The device EventCounterReporter of device_event_counter jobs over n
 * PackedEvents in PackedEvent order, from a fresh reporter. */
int nh_debug_count_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n, nh_event_counts *out);
/* n PackedEvents in any order, sorted and counted as a round of a
 * device_event_counter job is (the water-level sort's permutation, the counts
 * taken in its windows of equal event times, or through the permutation where
 * a run mixes classes), from a fresh reporter.  Diagnostic: the reporter side
 * of LiveWaterLevelReporterProxy::advance_water_level (live.rs:88-115) feeding
 * EventCounterReporter (event_counter.rs:58-109). */
int nh_debug_sort_count_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n, nh_event_counts *out);

#ifdef __cplusplus
}
#endif

#endif /* NECSIM_HIP_H */
