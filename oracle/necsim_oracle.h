/*
 * necsim_oracle.h -- CPU restatement of necsim-rust's Independent coalescence
 * algorithm (Monolithic parallelism mode, speciation-only reporting).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against, and the CPU baseline that bench.py times.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links, loads or calls anything in this directory.
 *
 * Every function cites the reference file:line that it restates.  Paths are
 * relative to the root of juntyr/necsim-rust.
 *
 * What pins it (tests/test_oracle_pins.py, tests/test_golden.py): published
 * seahash, rand_core seed_from_u64 and FNV-1a vectors; wymum against exact
 * 128-bit products; closed forms of the WyHash composition (operands zeroed
 * through the inverse of seahash_diffuse give prime_with = 16 ^ P5, a
 * location-independent state, seed/location translation, and sample_u64 of
 * a zeroed first operand); exp bit for bit against V8's fdlibm port (node),
 * and fdlibm-order twins of log/sin/cos built from this file's constants and
 * pi/2 reduction bit for bit against V8 on 2e5 inputs per range
 * (tests/golden/v8_maths.json), with log/sin/cos within 1 ulp of their twins
 * and of high-precision values; alias tables against their weights; the
 * Poisson sampler's contract and rate; origin order and partitions; and the
 * whole-path invariants (dedup keeps the species and the unique events,
 * sorted output, pause/resume, slicing independence, partition union).
 *
 * Parity unpinned (no reference-held vector or independent implementation
 * fixes these bits):
 *  - the musl evaluation order of the final assembly of log (and its
 *    reduction threshold) and of the sin/cos kernel polynomials, which crate
 *    libm 0.2.8 follows (reproducible.rs:18-46): V8 evaluates the same
 *    constants in fdlibm's order, so these orders are pinned to within 1 ulp
 *    of the V8-pinned twins, not bit for bit;
 *  - the absolute values of WyHash::prime_with / sample_u64 (wyhash.rs:53-80):
 *    the closed forms fix which constant meets which word and where the
 *    rotations sit, but no vector of the crate's own outputs exists here;
 *  - IntrinsicsMathsCore (glibc log/exp), the CPU Independent algorithm's
 *    default, is not restated at all.
 */
#ifndef NECSIM_ORACLE_H
#define NECSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* necsim/core/src/lineage.rs:97-109 (Lineage, 32 B repr(C)) */
typedef struct {
    uint64_t global_reference;
    double last_event_time;
    uint32_t x, y, index;
    uint32_t _pad;
} orc_lineage;

/* necsim/core/src/event.rs:58-71 (SpeciationEvent, 40 B repr(C)) */
typedef struct {
    uint64_t global_reference;
    double prior_time;
    double event_time;
    uint32_t x, y, index;
    uint32_t _pad;
} orc_speciation_event;

enum {
    ORC_NON_SPATIAL = 0,           /* rustcoalescence/scenarios/src/non_spatial.rs:39-89 */
    ORC_ALMOST_INFINITE_NORMAL = 1, /* .../almost_infinite/normal.rs:43-94 */
    ORC_SPATIALLY_EXPLICIT = 2,    /* .../spatially_explicit/turnover/uniform.rs:49-117 (separable alias) */
    ORC_SPATIALLY_EXPLICIT_PACKED = 3 /* packed_alias/mod.rs:49-119 (not wired in the reference) */
};

typedef struct {
    int kind;
    /* non-spatial: extent width x height at origin (0,0), deme size */
    uint32_t width, height, deme;
    /* almost-infinite: Gaussian dispersal sigma */
    double sigma;
    /* spatially explicit: habitat[width*height] row-major, dispersal rows as CSR
     * (row_ptr[width*height+1], col[], weight[]) in ascending column order.
     * A dense Array2D row iterated in column order and filtered by weight > 0
     * yields exactly the CSR entries of that row. */
    const uint32_t *habitat;
    const uint64_t *row_ptr;
    const uint32_t *col;
    const double *weight;
} orc_scenario_desc;

typedef struct orc_scenario orc_scenario;

enum { ORC_DEDUP_NONE = 0, ORC_DEDUP_RELATIVE = 1, ORC_DEDUP_ABSOLUTE = 2 };
enum { ORC_SLICE_RELATIVE = 1, ORC_SLICE_ABSOLUTE = 2 };

typedef struct {
    uint64_t seed;                  /* Seed(u64) as given in the RON config */
    double speciation_probability;  /* nu */
    double turnover_rate;           /* lambda (UniformTurnoverRate default 0.5) */
    double delta_t;                 /* IndependentArguments.delta_t (default 2.0) */
    uint64_t step_slice;            /* default 10 */
    int dedup_mode;                 /* default Relative(1.0) */
    double dedup_factor;
    uint64_t dedup_capacity;
    int event_slice_mode;           /* default Relative(2.0) */
    double event_slice_factor;
    uint64_t event_slice_capacity;
    int has_pause;
    double pause_before;
    double min_event_time;          /* resume_after; 0.0 for a fresh simulation */
    int report_dispersal;           /* R::ReportDispersal: every event is reported */
} orc_params;

/* necsim/core/src/event.rs:16-29 (PackedEvent, 56 B repr(C)): speciation is
 * prior(-) event(+), target = origin; the Independent algorithm's dispersal
 * (LineageInteraction::Maybe) is prior(+) event(-); coalescence = the
 * lineage's own reference for both (event.rs:103-155). */
typedef struct {
    uint64_t global_reference;
    double prior_time;
    double event_time;
    uint32_t ox, oy, oindex;
    uint32_t tx, ty, tindex;
    uint64_t coalescence;
} orc_packed_event;

typedef struct {
    int paused;        /* Status::Paused */
    double time;       /* max_time */
    uint64_t steps;    /* total_steps */
    size_t n_events;
    size_t n_remaining;
} orc_outcome;

/* ---- MathsCore (ReproducibleMathsCore = crate libm 0.2.8) ---- */
double orc_ln(double x);
double orc_exp(double x);
double orc_sin(double x);
double orc_cos(double x);
/* test-only fdlibm-order twins sharing the constants above, and a vector
 * evaluator of ln/exp/sin/cos and the twins (op 0..6) */
double orc_fdlibm_log(double x);
double orc_fdlibm_sin(double x);
double orc_fdlibm_cos(double x);
int orc_maths_vec(int op, const double *in, double *out, size_t n);
double orc_sqrt(double x);
double orc_floor(double x);
double orc_round(double x);

/* ---- RNG ---- */
uint64_t orc_seed_from_u64(uint64_t state);
uint64_t orc_seahash_diffuse(uint64_t x);
uint64_t orc_wymum(uint64_t a, uint64_t b);
/* returns the state after prime_with(location_index, time_index) */
uint64_t orc_prime(uint64_t seed, uint64_t location_index, uint64_t time_index);
/* sample_u64: advances *state, returns output */
uint64_t orc_sample_u64(uint64_t *state);
double orc_uniform_closed_open(uint64_t *state);
double orc_uniform_open_closed(uint64_t *state);
uint64_t orc_fnv1a64(const uint8_t *bytes, size_t n);
uint64_t orc_speciation_sample_hash(double u, double time, uint32_t x, uint32_t y, uint32_t i);

/* Poisson event time sampler (single call), returns event time and the
 * RNG state after the final re-prime. */
double orc_poisson_next_event_time(uint64_t seed, uint64_t location_id, double time,
                                   double lambda, double delta_t, uint64_t *state_out);

/* Alias method (necsim/impls/no-std/src/alias/packed.rs:59-163).
 * atoms_u[n], atoms_e[n], atoms_k[n] receive the created atoms. */
void orc_alias_create(const uint64_t *events, const double *weights, size_t n,
                      double *atoms_u, uint64_t *atoms_e, uint64_t *atoms_k);
uint64_t orc_alias_sample(const double *atoms_u, const uint64_t *atoms_e, const uint64_t *atoms_k,
                          size_t n, double limit, double u01);

/* ---- scenario ---- */
orc_scenario *orc_scenario_new(const orc_scenario_desc *desc);
void orc_scenario_free(orc_scenario *s);
/* Separable alias tables of a spatially explicit scenario, for cross-checks:
 * per cell: range start/end, self-dispersal probability, non-self event;
 * atoms (u, e, k).  Returns the number of atoms; arrays may be NULL to query. */
size_t orc_scenario_tables(const orc_scenario *s, uint64_t *range_start, uint64_t *range_end,
                           double *self_p, uint64_t *self_event, double *atom_u,
                           uint64_t *atom_e, uint64_t *atom_k);
uint64_t orc_location_id(const orc_scenario *s, uint32_t x, uint32_t y, uint32_t i);
uint32_t orc_deme_at(const orc_scenario *s, uint32_t x, uint32_t y);
/* one dispersal sample from (x,y) with the rng state; writes target */
void orc_sample_dispersal(const orc_scenario *s, uint32_t x, uint32_t y, uint64_t *state,
                          uint32_t *tx, uint32_t *ty);

/* Origin sampling: OriginPreSampler::all().percentage(p).partition(rank, size)
 * fed into the scenario's origin sampler.  circle_* are used by the
 * almost-infinite scenario (Sample::Circle).  Returns the number of lineages
 * written (at most cap); *total receives the full count. */
size_t orc_sample_origins(const orc_scenario *s, double percentage, uint32_t rank, uint32_t size,
                          uint32_t circle_cx, uint32_t circle_cy, uint32_t circle_radius,
                          orc_lineage *out, size_t cap, size_t *total);

/* ---- the algorithm ---- */
/* events_out must hold n entries, remaining_out n entries (may be NULL).
 * Speciation events are returned sorted by the PackedEvent order
 * (necsim/core/src/event.rs:256-280). Returns 0 on success. */
int orc_simulate(const orc_scenario *s, const orc_params *p, const orc_lineage *lineages,
                 size_t n, orc_speciation_event *events_out, orc_lineage *remaining_out,
                 orc_outcome *outcome);

/* The same run with every reported event as a PackedEvent, sorted by the
 * PackedEvent order.  With p->report_dispersal the water level advances by
 * event_slice / sum(lambda) (monolithic/mod.rs:139-148) and every dispersal
 * is reported; otherwise only speciations.  *events_out is malloc'd
 * (release with orc_free). */
int orc_simulate_packed(const orc_scenario *s, const orc_params *p, const orc_lineage *lineages,
                        size_t n, orc_packed_event **events_out, size_t *n_events, orc_lineage *remaining_out,
                        orc_outcome *outcome);
void orc_free(void *ptr);

/* Biodiversity reporter over a sorted stream (necsim/plugins/common/src/biodiversity.rs:33-41) */
uint64_t orc_biodiversity(const orc_speciation_event *events, size_t n);

/* IsolatedIndividuals CPU baseline: T threads, thread r simulates partition r
 * of `partitions` (>= T) of the origin pre-sampler, so T == partitions covers
 * every lineage and T < partitions a bounded sample.  Returns total steps;
 * *seconds = wall time of the simulate phase (origin sampling excluded). */
uint64_t orc_simulate_partitioned(const orc_scenario *s, const orc_params *p, double percentage,
                                  uint32_t circle_cx, uint32_t circle_cy, uint32_t circle_radius,
                                  uint32_t threads, uint32_t partitions, double *seconds,
                                  uint64_t *n_events);

#ifdef __cplusplus
}
#endif

#endif
