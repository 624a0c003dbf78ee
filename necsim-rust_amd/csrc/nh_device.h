// nh_device.h -- device data layout and the per-lineage step of the
// Independent algorithm, shared by the plain and the persistent kernels.
//
// One lineage per lane.  A step is one event of
// Simulation::simulate_incremental_early_stop (necsim/core/src/simulation/mod.rs:53-134):
//   1. next event time weakly after max(last_event_time, min_event_time)
//      (IndependentActiveLineageSampler::pop_active_lineage_and_event_time,
//       active_lineage_sampler/independent/sampler.rs:69-116, with the
//       PoissonEventTimeSampler, event_time_sampler/poisson.rs:30-109);
//   2. the speciation draw and min-speciation-sample tracking
//      (event_sampler/independent.rs:112-172, event_sampler/tracking.rs:39-56);
//   3. otherwise dispersal, the coalescence draw and the indexed target
//      (event_sampler/independent.rs:174-221, coalescence_sampler/independent.rs:38-76).
#pragma once

#include "nh_rng.h"

namespace nh {

enum : int {
    K_NON_SPATIAL = 0,
    K_ALMOST_INFINITE = 1,
    K_SEPARABLE = 2,
    K_PACKED = 3,
};

// One 32-byte record per habitat cell: everything a step needs about the
// lineage's current cell (injective-id prefix and deme for RNG priming and the
// coalescence draw, alias range and self-dispersal for the next dispersal).
struct __attribute__((aligned(32))) CellRec {
    uint64_t prefix;     // InMemoryHabitat u64_injection[cell] (habitat/in_memory.rs:151-160)
    uint32_t habitat;    // deme size
    uint32_t atom_start; // alias range start
    uint32_t atom_len;   // alias range length
    uint32_t non_self;   // SeparableAliasSelfDispersal::non_self_dispersal_event
    double self_p;       // SeparableAliasSelfDispersal::self_dispersal
};
static_assert(sizeof(CellRec) == 32, "cell record is 32 B");

// AliasMethodSamplerAtom<usize> with u32 events (16 B instead of 24 B)
struct __attribute__((aligned(16))) Atom {
    double u;
    uint32_t e;
    uint32_t k;
};
static_assert(sizeof(Atom) == 16, "atom is 16 B");

struct DevScenario {
    int kind;
    uint32_t width, height, ox, oy, deme;
    uint64_t area;   // width * height (non-spatial dispersal index range)
    double area_f;   // (double)area
    double deme_f;   // (double)deme
    double sigma;
    const CellRec *cells;
    const Atom *atoms;
};

constexpr int RATIO_N = 256; // Poisson inversion ratios lambda*dt/k cached for k < RATIO_N
// The inversion's running (acc_k, prod_k) for k < PTBL, computed on the host
// by the same sequence of IEEE operations as the loop: the count is found by
// comparisons against the table and the loop only runs past it.
constexpr int PTBL = 8;
// ratio[RATIO_N], acc[PTBL], prod[PTBL], then floor(acc[k] * 2^53) as u64 bit
// patterns: a U[0,1) draw is m * 2^-53 for an integer m < 2^53, so
// U > acc[k] exactly when m > floor(acc[k] * 2^53)
constexpr int POISSON_TABLE_N = RATIO_N + 3 * PTBL;

// Per-simulation constants
struct DevParams {
    uint64_t seed;          // WyHash seed
    double nu;              // speciation probability
    double lambda_dt;       // lambda * delta_t (poisson.rs:41)
    double p0;              // exp(-lambda_dt) (poisson.rs:42)
    double lambda_dt_sqrt;  // sqrt(lambda_dt) for the normal fallback (poisson.rs:70)
    double delta_t;
    double inv_delta_t;     // 1/delta_t when delta_t is a power of two
    int dt_pow2;
    double min_event_time;  // resume_after
    uint64_t nu_m;          // ceil(nu * 2^53): a U[0,1) draw m * 2^-53 is below nu exactly when m < nu_m
    uint64_t step_slice;
    const double *ratio;    // POISSON_TABLE_N doubles: ratio[k] = lambda_dt / k (k in 1..RATIO_N-1),
                            // then acc[k] and prod[k] for k < PTBL (device, global)
    uint32_t prod_zero_at;  // first k < PTBL with prod[k] == 0, PTBL if none
};

// Host side: the table of DevParams::ratio (the loop of poisson.rs:53-63
// unrolled for k < PTBL); returns prod_zero_at.
inline uint32_t poisson_table(double lambda_dt, double p0, double *t) {
    t[0] = 0.0;
    for (int k = 1; k < RATIO_N; k++) t[k] = lambda_dt / (double)k;
    double *acc = t + RATIO_N, *prod = acc + PTBL, *acc_m = prod + PTBL;
    double pr = p0, ac = p0;
    uint32_t zero_at = PTBL;
    for (int k = 0; k < PTBL; k++) {
        if (k > 0) {
            pr *= t[k];
            ac += pr;
        }
        acc[k] = ac;
        prod[k] = pr;
        if (!(pr > 0.0) && zero_at == PTBL) zero_at = (uint32_t)k;
        // ac * 2^53 is exact; past 2^53 no draw exceeds it
        const double a = ac * 9007199254740992.0;
        acc_m[k] = from_bits(a >= 9007199254740992.0 ? (1ull << 53) : (uint64_t)__builtin_floor(a));
    }
    return zero_at;
}

// ceil(nu * 2^53) for DevParams::nu_m (nu in [0, 1])
inline uint64_t draw_threshold_below(double nu) {
    const double a = nu * 9007199254740992.0;
    return a >= 9007199254740992.0 ? (1ull << 53) : (a > 0.0 ? (uint64_t)__builtin_ceil(a) : 0ull);
}

constexpr uint8_t FLAG_EXACT = 1; // next_t / next_st hold the exact next event
constexpr uint8_t FLAG_DEAD = 2;  // the lineage speciated or was dropped as a duplicate

// Lineage state held in HBM between launches (SoA, coalesced per field).
struct LineageSoA {
    uint64_t *ref;      // global_reference
    double *t_last;     // last_event_time
    uint32_t *x, *y, *idx;
    double *next_t;     // next event time (pessimistic lower bound, or exact if flag bit 0)
    uint64_t *next_st;  // prime key of next_t's event: its draws use prime(seed, id, key) (exact only)
    uint8_t *flags;     // FLAG_EXACT, FLAG_DEAD
};

// Speciation event, the reference's SpeciationEvent layout (40 B)
struct DevEvent {
    uint64_t ref;
    double prior_time;
    double event_time;
    uint32_t x, y, index, pad;
};
static_assert(sizeof(DevEvent) == 40, "event is 40 B");

// PackedEvent (necsim/core/src/event.rs:16-29), 56 B: the reported events when
// dispersal is reported too.  Speciation: prior_time sign bit set, target =
// origin; the Independent algorithm's dispersal (LineageInteraction::Maybe):
// event_time sign bit set; coalescence = the lineage's own reference for
// both (event.rs:103-155).
struct DevPacked {
    uint64_t ref;
    double prior_time;
    double event_time;
    uint32_t ox, oy, oindex;
    uint32_t tx, ty, tindex;
    uint64_t coalescence;
};
static_assert(sizeof(DevPacked) == 56, "packed event is 56 B");

// Lane-resident lineage.  a/b hold (x, y) for the almost-infinite habitat and
// (cell index, unused) for finite extents; `lkey` is the injective location
// id hashed with the seed for priming the RNG (prime_key).
template <int KIND> struct Lane {
    uint64_t ref;
    double t_last;
    uint32_t a, b, i;
    uint64_t lkey;
    CellRec rec;        // finite in-memory kinds: record of the current cell
    double next_t;
    uint64_t next_st; // prime key of the next event
    bool have_next;
    // an unfinished window-by-window search for the next event
    bool searching;
    uint64_t search_ts;
    double search_before;
    // lane_search_step: the open window's remaining event draws.  While the
    // search runs, next_st holds the RNG state before the window's next draw
    // and next_t the earliest valid event time so far (both only mean the
    // next event once have_next is set), which saves registers.
    uint32_t win_left; // draws left (0: no window open)
    uint32_t win_e;    // index of the next draw
    uint32_t win_bi;   // index of the earliest valid event so far, ~0u if none
    // min speciation sample of the current slice
    bool ms_some;
    double ms_u, ms_t;
    uint32_t ms_a, ms_b, ms_i;
};

NH_HD bool is_cell_kind(int kind) { return kind != K_ALMOST_INFINITE; }

// ---------------------------------------------------------------------------
// location helpers
template <int KIND>
__device__ __forceinline__ void location_xy(const DevScenario &S, uint32_t a, uint32_t b, uint32_t &x,
                                            uint32_t &y) {
    if (KIND == K_ALMOST_INFINITE) {
        x = a;
        y = b;
    } else {
        x = S.ox + a % S.width;
        y = S.oy + a / S.width;
    }
}

// Load a lineage's location into lane form.  Returns false if the location is
// outside the habitat extent.
template <int KIND>
__device__ __forceinline__ void lane_set_location(Lane<KIND> &L, const DevScenario &S, uint32_t x,
                                                  uint32_t y, uint32_t i, uint64_t seed) {
    L.i = i;
    if (KIND == K_ALMOST_INFINITE) {
        L.a = x;
        L.b = y;
        L.lkey = prime_key(seed, ((uint64_t)y << 32) | (uint64_t)x); // habitat/almost_infinite.rs:67-71
    } else {
        uint32_t cell = (y - S.oy) * S.width + (x - S.ox);
        L.a = cell;
        L.b = 0;
        if (KIND == K_NON_SPATIAL) {
            // habitat/non_spatial.rs:93-109
            L.lkey = prime_key(seed, (uint64_t)cell * (uint64_t)S.deme + (uint64_t)i);
        } else {
            L.rec = S.cells[cell];
            L.lkey = prime_key(seed, L.rec.prefix + (uint64_t)i); // habitat/in_memory.rs:71-81
        }
    }
}

// ---------------------------------------------------------------------------
// PoissonEventTimeSampler::next_event_time_at_indexed_location_weakly_after
// (event_time_sampler/poisson.rs:30-109), one time window at a time: the
// events of window `ts` (its RNG primed with (location, ts)) are drawn, and the
// earliest one after `time` is the next event (poisson_window = prime,
// poisson_count, poisson_events).  The event's own draws use prime(id, key)
// with the key set by poisson_events.
//
// poisson_count: the number of events of the window whose primed RNG is
// `state`; leaves `state` before the first event's time draw.
// `acc_t` points at the table's acc / prod / acc_m part: ratio + RATIO_N, or a
// copy of that part in LDS (one pointer per call site, never a choice made at
// run time, so that LDS reads compile to LDS instructions, not flat ones).
__device__ __forceinline__ uint32_t poisson_count(const DevParams &P, const double *ratio, uint64_t &state,
                                                  const double *acc_t) {
    uint32_t number;
    if (P.p0 > 0.0) {
        // while u > acc_k && prod_k > 0: k += 1 (acc is non-decreasing, so
        // the loop ends at min(first k with u <= acc_k, first k with prod_k == 0))
        const double *prod_t = acc_t + PTBL, *acc_m = prod_t + PTBL;
        // the draw U = mu * 2^-53 (uniform_closed_open), compared as an integer
        const uint64_t mu = sample_u64(state) >> 11;
        // m = #{k < PTBL : U > acc_k}, by binary search over the sorted table
        static_assert(PTBL == 8, "three search steps");
        uint32_t m = 0;
        m += mu > bits(acc_m[m + 3]) ? 4u : 0u;
        m += mu > bits(acc_m[m + 1]) ? 2u : 0u;
        m += mu > bits(acc_m[m]) ? 1u : 0u;
        m += (m == PTBL - 1 && mu > bits(acc_m[PTBL - 1])) ? 1u : 0u;
        uint32_t k = m < P.prod_zero_at ? m : P.prod_zero_at;
        if (k == PTBL) { // both past the table: continue the loop from its last entry
            const double u = (double)mu * TWO_POW_M53;
            k = PTBL - 1;
            double prod = prod_t[k];
            double acc = acc_t[k];
            while (u > acc && prod > 0.0) {
                k += 1;
                prod *= (k < RATIO_N) ? ratio[k] : P.lambda_dt / (double)k;
                acc += prod;
            }
        }
        number = k;
    } else {
        double a, b;
        sample_2d_normal(state, P.lambda_dt, P.lambda_dt_sqrt, a, b);
        number = as_u32(a); // f64::max(a, 0.0) as u32: NaN and negatives give 0
    }
    return number;
}

__device__ __forceinline__ uint32_t poisson_count(const DevParams &P, const double *ratio, uint64_t &state) {
    return poisson_count(P, ratio, state, ratio + RATIO_N);
}

// The window's `number` event times (ts + U[0,1)) * delta_t, drawn in order
// from `state`; the first strictly smallest one after `time` is the event.
__device__ __forceinline__ bool poisson_events(const DevParams &P, uint64_t state, uint32_t number, uint64_t ts,
                                               double time, double &best, uint64_t &key) {
    bool found = false;
    uint32_t best_index = 0;
    for (uint32_t e = 0; e < number; e++) {
        double t = ((double)ts + uniform_closed_open(state)) * P.delta_t;
        if (t > time && (!found || best > t)) {
            best = t;
            best_index = e;
            found = true;
        }
    }
    // the event's own draws use prime(id, key) (poisson.rs:102-106); the
    // prime itself is left to the execution, where the lanes run together
    if (found) key = ts + INV_PHI * (uint64_t)(best_index + 1);
    return found;
}

// One window at the location keyed as lkey (prime_key): false if it
// holds no event after `time`, else sets best / key.
__device__ __forceinline__ bool poisson_window(const DevParams &P, const double *ratio, uint64_t lkey,
                                               double time, uint64_t ts, double &best, uint64_t &key) {
    uint64_t state = prime_keyed(lkey, ts);
    const uint32_t number = poisson_count(P, ratio, state);
    return poisson_events(P, state, number, ts, time, best, key);
}

__device__ __forceinline__ uint64_t poisson_first_window(const DevParams &P, double time) {
    const double w = floor_(P.dt_pow2 ? time * P.inv_delta_t : time / P.delta_t);
    // below 2^32 (every practical simulation) one 32-bit conversion gives
    // the saturating u64 cast's value
    if (w >= 0.0 && w < 4294967296.0) return (uint64_t)(uint32_t)w;
    return as_u64(w);
}

// (double)ts, with one 32-bit conversion below 2^32
__device__ __forceinline__ double window_f64(uint64_t ts) {
    if (ts < (1ull << 32)) return (double)(uint32_t)ts;
    return (double)ts;
}

// The whole search: windows ts, ts + 1, ... until one holds an event.
__device__ __forceinline__ double poisson_next(const DevParams &P, const double *ratio, uint64_t lkey,
                                               double time, uint64_t &key) {
    double best = 0.0;
    for (uint64_t ts = poisson_first_window(P, time);; ts++)
        if (poisson_window(P, ratio, lkey, time, ts, best, key)) return best;
}

// Make sure L.next_t / L.next_st hold the next event (the `pop` peek).
template <int KIND>
__device__ __forceinline__ void lane_peek(Lane<KIND> &L, const DevParams &P, const double *ratio) {
    if (L.have_next) return;
    // NonNegativeF64::max of last_event_time and min_event_time (sampler.rs:87)
    double before = P.min_event_time > L.t_last ? P.min_event_time : L.t_last;
    uint64_t st;
    double et = poisson_next(P, ratio, L.lkey, before, st);
    L.next_t = max_after(before, et);
    L.next_st = st;
    L.have_next = true;
}

// The same peek advanced a little per call, so that a wave steps all of its
// lanes' searches in lockstep instead of waiting for its slowest lane's
// whole search.  A call opens the next window if none is open (prime and
// count draw) and makes at most STEP_DRAWS of its event draws; a window that
// holds no events is followed by the next one in the same call, up to
// STEP_OPENS windows.  A window holds a Poisson number of events, so drawing
// all of them in one call makes the wave wait for its lane with the most;
// here a lane with many events finishes its window in later calls while the
// other lanes go on.  The draws, their order and the pick (first strictly
// smallest time after the search's lower bound) are those of poisson_window,
// so L.next_t / L.next_st end up equal to lane_peek's.
constexpr uint32_t STEP_DRAWS = 2;
constexpr uint32_t STEP_OPENS = 2;
// `acc_t` as for poisson_count.
template <int KIND>
__device__ __forceinline__ void lane_search_step(Lane<KIND> &L, const DevParams &P, const double *ratio,
                                                 const double *acc_t) {
    if (L.have_next) return;
    // NonNegativeF64::max of last_event_time and min_event_time (sampler.rs:87),
    // recomputed per call instead of held in a register
    const double before = P.min_event_time > L.t_last ? P.min_event_time : L.t_last;
    if (!L.searching) {
        L.search_ts = poisson_first_window(P, before);
        L.searching = true;
        L.win_left = 0;
    }
    // open windows until one holds events (at most STEP_OPENS per call)
#pragma unroll
    for (uint32_t o = 0; o < STEP_OPENS && L.win_left == 0; o++) {
        uint64_t st = prime_keyed(L.lkey, L.search_ts);
        L.win_left = poisson_count(P, ratio, st, acc_t);
        L.next_st = st;
        L.win_e = 0;
        L.win_bi = ~0u;
        if (L.win_left == 0) L.search_ts++;
    }
    if (L.win_left == 0) return;
    const double tsd = window_f64(L.search_ts);
#pragma unroll
    for (uint32_t k = 0; k < STEP_DRAWS; k++) {
        if (L.win_left == 0) break;
        const double t = (tsd + uniform_closed_open(L.next_st)) * P.delta_t;
        if (t > before && (L.win_bi == ~0u || L.next_t > t)) {
            L.next_t = t;
            L.win_bi = L.win_e;
        }
        L.win_e++;
        L.win_left--;
    }
    if (L.win_left != 0) return;
    if (L.win_bi != ~0u) {
        L.next_t = max_after(before, L.next_t);
        L.next_st = L.search_ts + INV_PHI * (uint64_t)(L.win_bi + 1);
        L.have_next = true;
        L.searching = false;
    } else {
        L.search_ts++;
    }
}

template <int KIND>
__device__ __forceinline__ void lane_search_step(Lane<KIND> &L, const DevParams &P, const double *ratio) {
    lane_search_step(L, P, ratio, ratio + RATIO_N);
}

enum : int { STEP_DISPERSED = 0, STEP_SPECIATED = 1 };

// Execute the peeked event (L.have_next must hold and L.next_t < level).
// On speciation the event is written to *ev, all but its reference (the
// lineage's global reference never changes, so callers that do not keep it
// in a register read it from the lineage store).
// TRACK_MS = false skips the slice's min speciation sample (only the dedup
// cache reads it).  FLAT takes the straight-line sin / cos (sincos_flat).
template <int KIND, bool TRACK_MS = true, bool FLAT = false>
__device__ __forceinline__ int lane_execute(Lane<KIND> &L, const DevScenario &S, const DevParams &P,
                                            DevEvent &ev) {
    double net = L.next_t;
    uint64_t st = prime_keyed(L.lkey, L.next_st); // next_st holds the event's prime key
    L.have_next = false;

    // the speciation draw U = m * 2^-53 (uniform_closed_open); U < nu is
    // m < nu_m, and only the dedup cache's min sample needs U itself
    const uint64_t m_spec = sample_u64(st) >> 11;
    if (TRACK_MS) {
        const double u_spec = (double)m_spec * TWO_POW_M53;
        // SpeciationSample::update_min keeps the first minimum (tracking.rs:46-55)
        if (!(L.ms_some && L.ms_u <= u_spec)) {
            L.ms_some = true;
            L.ms_u = u_spec;
            L.ms_t = net;
            L.ms_a = L.a;
            L.ms_b = L.b;
            L.ms_i = L.i;
        }
    }
    if (m_spec < P.nu_m) {
        uint32_t x, y;
        location_xy<KIND>(S, L.a, L.b, x, y);
        ev.prior_time = L.t_last; // ev.ref is the caller's: the lineage's reference
        ev.event_time = net;
        ev.x = x;
        ev.y = y;
        ev.index = L.i;
        ev.pad = 0;
        L.t_last = net;
        return STEP_SPECIATED;
    }

    if (KIND == K_NON_SPATIAL) {
        // NonSpatialDispersalSampler (dispersal_sampler/non_spatial.rs:41-70):
        // sample_index_u64(W*H) (rng.rs:132-144).  The target cell index is the
        // index itself, so no division is needed here.
        double u = uniform_closed_open(st);
        uint64_t v = as_u64(floor_(u * S.area_f));
        uint64_t cell = v < S.area - 1 ? v : S.area - 1;
        double uc = uniform_closed_open(st);
        uint32_t ci = as_u32(floor_(uc * S.deme_f));
        L.a = (uint32_t)cell;
        L.i = ci;
        L.lkey = prime_key(P.seed, cell * (uint64_t)S.deme + (uint64_t)ci);
    } else if (KIND == K_ALMOST_INFINITE) {
        // AlmostInfiniteNormalDispersalSampler + clamp_round_dispersal
        // (almost_infinite_normal.rs:52-69, habitat/almost_infinite.rs:98-117).
        // ((x + dx) % 2^32 + 2^32) % 2^32 is the low word of x + dx.
        double dx, dy;
        sample_2d_normal<FLAT>(st, 0.0, S.sigma, dx, dy);
        // `as i64` of the rounded offsets; below 2^31 in magnitude (every
        // practical sigma) that is the 32-bit conversion, one instruction
        const double rx = round_(dx), ry = round_(dy);
        int64_t ix, iy;
        if (__builtin_fabs(rx) < 2147483648.0 && __builtin_fabs(ry) < 2147483648.0) {
            ix = (int32_t)rx;
            iy = (int32_t)ry;
        } else {
            ix = as_i64(rx);
            iy = as_i64(ry);
        }
        (void)uniform_closed_open(st); // CoalescenceRngSample::new, unused by singleton demes
        L.a = (uint32_t)((uint64_t)L.a + (uint64_t)ix);
        L.b = (uint32_t)((uint64_t)L.b + (uint64_t)iy);
        L.i = 0;
        L.lkey = prime_key(P.seed, ((uint64_t)L.b << 32) | (uint64_t)L.a);
    } else {
        // InMemoryPackedSeparableAliasDispersalSampler (packed_separable_alias/dispersal.rs:16-77)
        // and InMemoryPackedAliasDispersalSampler (packed_alias/dispersal.rs:16-52)
        uint32_t cell = L.a;
        uint32_t target;
        bool alias;
        if (KIND == K_SEPARABLE) {
            if (L.rec.self_p == 1.0) {
                alias = false;
            } else if (L.rec.self_p == 0.0 || L.rec.non_self != cell) {
                alias = true;
            } else {
                alias = uniform_closed_open(st) >= L.rec.self_p;
            }
        } else {
            alias = true;
        }
        if (alias) {
            // AliasMethodSamplerAtom::sample_event_with_cdf_limit, limit 1.0 (alias/packed.rs:135-163)
            double f = uniform_closed_open(st) * 1.0 * (double)L.rec.atom_len;
            uint64_t k = as_u64(floor_(f));
            double yv = f - (double)k;
            Atom at = S.atoms[L.rec.atom_start + (uint32_t)k];
            target = yv < at.u ? at.e : at.k;
        } else {
            target = cell;
        }
        CellRec t = S.cells[target];
        double uc = uniform_closed_open(st);
        uint32_t ci = as_u32(floor_(uc * (double)t.habitat));
        L.a = target;
        L.i = ci;
        L.rec = t;
        L.lkey = prime_key(P.seed, t.prefix + (uint64_t)ci);
    }
    L.t_last = net;
    return STEP_DISPERSED;
}

// ---------------------------------------------------------------------------
// Dedup cache buckets: 4 x u64 (time bits, sample bits, x | y << 32,
// index | check << 32).  An all-zero bucket is empty (event times are > 0).
struct Sample {
    double u, t;
    uint32_t x, y, i;
};

NH_HD uint32_t sample_check(const Sample &s) {
    uint64_t h = wymum(bits(s.t) ^ 0x9e3779b97f4a7c15ull, bits(s.u) ^ (((uint64_t)s.x << 32) | s.y));
    h = wymum(h ^ s.i, 0xeb44accab455d165ull);
    return (uint32_t)(h >> 32) | 1u;
}

__device__ __forceinline__ bool bucket_equals(const uint64_t *b, const Sample &s) {
    return b[0] == bits(s.t) && b[1] == bits(s.u) && b[2] == (((uint64_t)s.y << 32) | s.x) &&
           b[3] == (((uint64_t)sample_check(s) << 32) | s.i);
}

__device__ __forceinline__ void bucket_store(uint64_t *b, const Sample &s) {
    b[0] = bits(s.t);
    b[1] = bits(s.u);
    b[2] = ((uint64_t)s.y << 32) | s.x;
    b[3] = ((uint64_t)sample_check(s) << 32) | s.i;
}

} // namespace nh
