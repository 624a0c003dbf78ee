// nh_kernels.h -- launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <vector>

#include "../../include/necsim_hip.h"
#include "nh_device.h"

struct nh_ctx;

namespace nh {

// Device-side counters of one simulation (or one server).
struct Counters {
    unsigned long long head;          // persistent work-queue head
    unsigned long long n_out;         // lineages written to the next round
    unsigned long long n_events;      // events appended (dispersal reporting: this round only)
    unsigned long long steps;         // completed events
    unsigned long long max_time_bits; // max event time (positive f64 bits)
    unsigned long long n_dead;        // persistent form: lineages that died this round
    unsigned long long ev_overflow;   // dispersal reporting: the round's event buffer was full
    unsigned long long n_ghost;       // dispersal reporting: reserved event slots left unused
    unsigned long long compact_live;  // gated lineage compaction: live lineages it counted
    unsigned long long compact_done;  // ... and whether it compacted them into the other buffer
};

// Persistent dispersal reporting reserves event slots EVENT_CHUNK at a time
// per wave; the unused rest of a wave's last chunk holds ghost records
// (event time +inf, sorted after every event) counted in Counters::n_ghost.
constexpr unsigned long long EVENT_CHUNK = 256;

// Per-round inputs read by the kernels from device memory (so that a captured
// graph can be replayed with new values).
struct RoundDesc {
    double level;
    unsigned long long n_in;
    unsigned long long ev_start; // events appended before this round
    unsigned long long chunk;    // work indices claimed per wave atomic (0: per idle lane)
};

// Largest claim of the persistent kernels' guided work queue: per-lane
// claims while there are few lineages per lane, up to 1024 indices per claim
// when there are many.
inline unsigned long long claim_chunk(unsigned long long n, unsigned long long lanes) {
    unsigned long long per_wave = lanes ? n / (lanes / 64 * 4 + 1) : 0;
    if (per_wave < 256) return 0;
    return 1024;
}

// plain-form element status (one byte per pass element)
enum : uint8_t {
    ST_SKIP = 0,     // not processed: pessimistic next event >= water level
    ST_SPEC = 1,     // speciated
    ST_SLOW = 2,     // stopped by the step slice, next event below the level
    ST_FAST = 3,     // next event at or above the level
    ST_DUP = 4,      // dropped by the dedup cache
    ST_KIND_MASK = 7,
    ST_HAS_SAMPLE = 8, // executed >= 1 step (has a min speciation sample)
};

struct PersistArgs {
    DevScenario S;
    DevParams P;
    LineageSoA lin; // updated in place
    const RoundDesc *round;
    uint64_t *cache;          // 4 u64 per bucket, null if dedup is off
    unsigned long long cache_cap;
    DevEvent *events;
    Counters *ctr;
    unsigned inner_steps;          // iterations per lane between wave-level refills (>= 1)
    unsigned exec_fill;            // eighths of the active lanes that must be ready to execute (1..8)
    DevPacked *packed;             // dispersal reporting: the round's PackedEvents, else null
    unsigned long long packed_cap;
    unsigned coop;                 // lanes that are not searching search ahead for those that are
    unsigned pool;                 // the pooled kernel's lineages per wave (0: one lineage per lane)
    // dispersal reporting, or null: each PackedEvent's time key and slot for
    // the round's sort, and the keys' [min, max] (ghost slots excluded)
    unsigned long long *ev_keys;
    uint32_t *ev_vals;
    unsigned long long *ev_range;
};

struct PlainArgs {
    DevScenario S;
    DevParams P;
    LineageSoA L;
    const uint32_t *pass;     // lineage slots in processing order
    unsigned long long m;     // pass size, or with m_dev the launch bound
    const unsigned long long *m_dev; // device pass size (batched passes), else null
    const RoundDesc *round;
    uint8_t *status;          // [m]
    // min speciation samples of the slice, per pass element
    double *s_u, *s_t;
    uint32_t *s_x, *s_y, *s_i;
    unsigned long long *s_bucket;
    unsigned long long cache_cap; // 0: dedup off
    DevEvent *events;
    Counters *ctr;
    DevPacked *packed;             // dispersal reporting: the round's PackedEvents, else null
    unsigned long long packed_cap;
};

// ---- serving: one persistent kernel over the lineages of many requests ----
constexpr int SERVE_MAX_REQS = 1024;

// per-request device counters
struct ReqCounters {
    unsigned long long n_dead;        // lineages that died this tick
    unsigned long long n_below;       // set if a lineage stopped below the water level (tick limit)
    unsigned long long n_events;      // speciation events appended (all ticks)
    unsigned long long steps;         // completed events this tick
    unsigned long long max_time_bits; // max event time this tick
    unsigned long long pad[3];
};

struct ServeReq {
    uint64_t seed;
    double nu;
    uint64_t nu_m; // ceil(nu * 2^53) (DevParams::nu_m)
    double level;
    LineageSoA lin; // updated in place
    DevEvent *events;
    uint64_t *cache;
    unsigned long long cache_cap;
    unsigned long long n_in;
    unsigned long long work_base; // first global work index of this request
    ReqCounters *ctr;
};

struct ServeArgs {
    DevScenario S;
    DevParams P; // shared: lambda*dt, p0, ratio table, delta_t, step_slice, min_event_time
    const ServeReq *reqs;
    int n_reqs;
    unsigned long long total_work;
    unsigned long long *head;
    unsigned long long tick_steps;
    unsigned long long chunk;
};

void launch_serve(int kind, const ServeArgs &a, int blocks, int threads, hipStream_t stream);
int serve_max_blocks_per_cu(int kind, int threads);

// kernel launchers (all asynchronous on `stream`)
void launch_persistent(int kind, const PersistArgs &a, int blocks, int threads, hipStream_t stream);
int persistent_max_blocks_per_cu(int kind, int threads, bool dedup, bool report_dispersal, bool coop);
// resident one-wave blocks per CU of the pooled kernel with `pool` lineages per wave
int pooled_max_blocks_per_cu(int kind, unsigned pool, bool report_dispersal);
void launch_plain(int kind, const PlainArgs &a, int threads, hipStream_t stream);

// m: pass size, or with m_dev the launch bound of a pass whose size is *m_dev
void launch_dedup_push(const uint8_t *status, const unsigned long long *bucket, uint32_t *head,
                       uint32_t *link, uint64_t m, const unsigned long long *m_dev, hipStream_t stream);
void launch_dedup_resolve(uint8_t *status, const unsigned long long *bucket, const double *s_u,
                          const double *s_t, const uint32_t *s_x, const uint32_t *s_y,
                          const uint32_t *s_i, uint32_t *head, const uint32_t *link,
                          uint64_t *cache, uint64_t m, const unsigned long long *m_dev, hipStream_t stream);

// Ordered two-way compaction of pass elements: slots with status ST_SLOW go to
// out_a (next pass), status ST_FAST to out_b + *b_base (appended).  Counts are
// written to counts[0] (a) and counts[1] (b).
struct CompactBuffers {
    uint32_t *blk_a, *blk_b; // [nblocks + 1]
    unsigned int *counts;    // [2] device
};
uint32_t compact_blocks(uint64_t m);
void launch_compact(const uint8_t *status, const uint32_t *pass, uint64_t m, uint32_t *out_a,
                    uint32_t *out_b, uint64_t b_base, CompactBuffers &cb, hipStream_t stream);

// Batched passes of the plain form: the pass size and the fast-list length
// live on the device, so passes run back to back without a host round trip.
struct PlainLoop {
    unsigned long long m;     // size of the current pass
    unsigned long long f_len; // lineages already in the fast list
};
// one compaction of a pass whose size is loop->m (launch bound m_bound):
// slow lineages to out_a, fast ones appended to out_b at loop->f_len; then
// loop->m and loop->f_len advance
void launch_compact_loop(const uint8_t *status, const uint32_t *pass, uint64_t m_bound, PlainLoop *loop,
                         uint32_t *out_a, uint32_t *out_b, CompactBuffers &cb, hipStream_t stream);

// Event sort: sorts events[0..n) by the PackedEvent order into out[0..n).
constexpr int SORT_MINMAX_WORDS = 32;
struct SortBuffers {
    unsigned long long *keys[2];
    uint32_t *vals[2];
    uint32_t *hist;         // [hist_words]: digit counts, then scan chunk sums
    unsigned long long *minmax; // [SORT_MINMAX_WORDS] device: key ranges, tie count
    uint64_t capacity;
    uint64_t hist_words;    // sort_hist_words(capacity), or sort_hist_words_staged(capacity) for staged passes
};
uint32_t sort_blocks(uint64_t n);
size_t sort_hist_words(uint64_t n);      // words of SortBuffers::hist for n events (8-bit digits)
size_t sort_hist_words_staged(uint64_t n); // ... with room for the staged passes of eager sorts
// keys_ready: the round's kernel wrote the time keys, identity values and
// key range into sb (k_sort_init first; dispersal-reporting persistent rounds)
template <class E>
void sort_events(const E *events, uint64_t n, E *out, SortBuffers &sb, hipStream_t stream, bool keys_ready);
// Same, with the event range [*start, *end) read from device memory when the
// kernels run (fixed grid for up to `cap` events; capturable in a graph).
template <class E>
void sort_events_device_sized(const E *events, const unsigned long long *end, const unsigned long long *start,
                              uint64_t cap, E *out, SortBuffers &sb, hipStream_t stream);

// Ordered compaction of the live (not dead) lineages of in[0, n) into out;
// the live count is written to blk[soa_compact_blocks(n)].
uint32_t soa_compact_blocks(uint64_t n);
void launch_soa_compact(const LineageSoA &in, const LineageSoA &out, uint64_t n, uint32_t *blk,
                        hipStream_t stream);
// The same, decided on the device (capturable in a graph): the slot count is
// *n_dev (at most cap), the live lineages are counted, and only if dead slots
// make up at least an eighth of the buffer are the live ones compacted into
// `out`; ctr->compact_live / compact_done report the count and the decision.
// blk holds soa_compact_blocks(cap) + 1 words.
void launch_soa_compact_gated(const LineageSoA &in, const LineageSoA &out, uint64_t cap,
                              const unsigned long long *n_dev, uint32_t *blk, Counters *ctr,
                              hipStream_t stream);

// EventCounterReporter (necsim/plugins/common/src/event_counter.rs:12-26)
// kept on the device: its eight counters, and the prior time of the last
// counted event (the reporter's last_parent_prior_time) carried from one
// water-level round to the next.
enum {
    EC_RAW_TOTAL, EC_SPECIATION, EC_OUT_DISPERSAL, EC_SELF_DISPERSAL, EC_OUT_COALESCENCE,
    EC_SELF_COALESCENCE, EC_LATE_DISPERSAL_COALESCENCE, EC_LATE_COALESCENCE, EC_COUNTERS
};
struct EventCounts {
    unsigned long long count[EC_COUNTERS];
    unsigned long long carry_has;   // an event was counted before this round
    unsigned long long carry_prior; // its prior time (magnitude bits)
    // the round counted inside the sort's run windows, committed afterwards
    // (k_count_finish)
    unsigned long long round_count[EC_COUNTERS];
    unsigned long long round_carry_prior;
};
// Counts one round's sorted PackedEvents ev[0, n) (or ev[perm[0, n)] in the
// order of a sort permutation): with `end`, n is read from *end when the
// kernels run (clamped to `cap`; capturable in a graph).  `ghosts` (or
// null): the number of trailing ghost records to leave out.
void launch_count_events(const DevPacked *ev, const uint32_t *perm, const unsigned long long *end,
                         uint64_t n_or_cap, const unsigned long long *ghosts, EventCounts *c, hipStream_t stream);
// the round's `slots` records sorted to a permutation and the first n_events
// of that order counted
void sort_and_count_packed(const DevPacked *events, uint64_t slots, uint64_t n_events, bool keys_ready,
                           SortBuffers &sb, EventCounts *c, hipStream_t stream);

// resets a sort's range words (before a kernel that writes keys_ready keys)
void launch_sort_init(SortBuffers &sb, hipStream_t stream);
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t stream);
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t stream);
// reset_events: also the event counter and its overflow flag (dispersal
// reporting keeps one round's events at a time)
void launch_reset_counters(Counters *c, bool reset_events, hipStream_t stream);

} // namespace nh

// Resident scenario tables (opaque nh_tables of the C ABI).
struct nh_tables {
    nh_ctx *ctx = nullptr;
    int kind = 0;
    uint32_t width = 0, height = 0, ox = 0, oy = 0, deme = 1;
    double sigma = 0.0;
    std::vector<uint32_t> habitat; // host copy for lineage validation
    nh::CellRec *d_cells = nullptr;
    nh::Atom *d_atoms = nullptr;
    uint64_t n_atoms = 0;
    uint64_t bytes = 0;
    nh::DevScenario dev{};
};

namespace nh {

struct Ranks;

// The device origin sampler of one scenario (nh_origins.hip): prepare()
// counts the lineages OriginPreSampler::all().percentage(p).partition(rank,
// size) gives (generating the percentage pre-sampler's ranks on the device),
// emit() writes them into a lineage store or a reference-layout array.
struct OriginPlan {
    const nh_tables *tables = nullptr;
    hipStream_t stream = nullptr;
    uint64_t n = 0; // lineages
    uint32_t rank = 0, size = 1, cx = 0, cy = 0, radius = 0;
    Ranks *ranks = nullptr;     // percentage ranks, null for `all`
    uint64_t *d_rows = nullptr; // circle: cells within the radius before each row
    uint64_t n_rows = 0;
    int prepare(const nh_tables *t, double percentage, uint32_t rank, uint32_t size, uint32_t cx, uint32_t cy,
                uint32_t radius, hipStream_t st);
    int emit(LineageSoA *soa, nh_lineage *aos) const;
    OriginPlan() = default;
    OriginPlan(const OriginPlan &) = delete;
    OriginPlan &operator=(const OriginPlan &) = delete;
    ~OriginPlan();
};

// the first n ranks of the percentage pre-sampler (no partition) into a host
// buffer, from the closed form of the golden-ratio chain or (serial) by the
// reference's f64 loop on one thread
int debug_presample(double percentage, uint64_t n, bool serial, uint64_t *out, hipStream_t st);

} // namespace nh

// NH_LINEAGE_* class of a lineage against the tables' scenario (defined in
// nh_runtime.hip)
int nh_lineage_class(const nh_tables *t, const nh_lineage &l);
