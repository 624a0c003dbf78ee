// nh_kernels.hip -- gfx950 kernels of the Independent algorithm.
//
//  * k_persistent<KIND, DEDUP>: one water-level round.  Wave64 lanes each
//    hold one lineage in registers and refill from a guided device work queue
//    (one atomic per wave claim, ballot/mbcnt distribution) whenever their
//    lineage speciates, is dropped as a duplicate, or reaches the water level.
//    Next-event searches advance a bounded amount per iteration and found
//    events execute in batches.  The DEDUP variant inserts into the dedup
//    cache in-kernel at every step-slice boundary.
//  * k_serve<KIND>: the persistent kernel over the lineages of many requests.
//  * k_plain<KIND>: one launch per step slice over the reference's queue
//    order (parallelisation/independent/monolithic/mod.rs:183-251); dedup and
//    reclassification follow in separate order-preserving kernels.
//  * ordered dedup, ordered compaction and an LSD radix sort (wave ballot
//    ranking; a full per-field pass set orders equal event times) for the
//    water-level event sort (monolithic/reporter/live.rs:88-115).
#include "nh_kernels.h"

namespace nh {

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// ---------------------------------------------------------------------------
// lineage load / store
// All fields, flags included, are read with independent loads so that a
// refill waits for one memory round trip; returns the slot's flags.
template <int KIND>
__device__ __forceinline__ uint8_t lane_load(Lane<KIND> &L, const DevScenario &S, const LineageSoA &s,
                                             uint64_t slot, uint64_t seed) {
    uint8_t f = s.flags[slot];
    L.ref = s.ref[slot];
    L.t_last = s.t_last[slot];
    uint32_t x = s.x[slot], y = s.y[slot], i = s.idx[slot];
    L.next_t = s.next_t[slot];
    uint64_t nst = s.next_st[slot];
    L.have_next = (f & FLAG_EXACT) != 0;
    L.next_st = L.have_next ? nst : 0;
    L.searching = false;
    L.ms_some = false;
    if (!(f & FLAG_DEAD)) lane_set_location<KIND>(L, S, x, y, i, seed);
    return f;
}

template <int KIND>
__device__ __forceinline__ void lane_store(const Lane<KIND> &L, const DevScenario &S,
                                           const LineageSoA &s, uint64_t slot) {
    uint32_t x, y;
    location_xy<KIND>(S, L.a, L.b, x, y);
    s.t_last[slot] = L.t_last; // the reference never changes: not written back
    s.x[slot] = x;
    s.y[slot] = y;
    s.idx[slot] = L.i;
    s.next_t[slot] = L.next_t;
    s.next_st[slot] = L.next_st;
    s.flags[slot] = L.have_next ? 1 : 0;
}

template <int KIND>
__device__ __forceinline__ Sample lane_sample(const Lane<KIND> &L, const DevScenario &S) {
    Sample s;
    s.u = L.ms_u;
    s.t = L.ms_t;
    location_xy<KIND>(S, L.ms_a, L.ms_b, s.x, s.y);
    s.i = L.ms_i;
    return s;
}

// DirectMappedCache::insert (necsim/impls/no-std/src/cache.rs:39-54):
// returns true if the bucket already held an equal sample (a duplicate).
__device__ __forceinline__ bool cache_insert_concurrent(uint64_t *cache, unsigned long long cap,
                                                        const Sample &s) {
    uint64_t h = speciation_sample_hash(s.u, s.t, s.x, s.y, s.i);
    uint64_t *b = cache + 4 * (h % cap);
    bool dup = bucket_equals(b, s);
    bucket_store(b, s);
    return dup;
}

__device__ __forceinline__ void load_ratio(double *s_ratio, const double *ratio) {
    for (int k = threadIdx.x; k < POISSON_TABLE_N; k += blockDim.x) s_ratio[k] = ratio[k];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Dispersal reporting: every executed step is reported as a PackedEvent.
// Each step first reserves its slot in the round's event buffer; a lane whose
// slot lies beyond the buffer does not execute, parks its lineage and flags
// the round, so nothing is ever written past the buffer.
//
// wave_reserve: one atomic per wave for all lanes executing together (the
// plain kernel: no unused slots).
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long *ctr) {
    const uint64_t act = __ballot(1);
    const uint32_t leader = __ffsll((unsigned long long)act) - 1;
    unsigned long long base = 0;
    if (lane_id() == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(act));
    base = __shfl(base, leader);
    return base + (unsigned long long)__popcll(act & lanemask_lt());
}

constexpr uint64_t SIGN_BIT = 0x8000000000000000ull;

// wave_take_slot: slots come from a wave-private chunk [ws[0], ws[1]) that is
// refilled EVENT_CHUNK at a time with one atomic (same-address atomics from
// every executing wave would serialise at the L2).  ws lives in LDS and is
// only touched by the wave's lowest active lane.
__device__ __forceinline__ unsigned long long wave_take_slot(volatile unsigned long long *ws,
                                                             unsigned long long *ctr) {
    const uint64_t act = __ballot(1);
    const uint32_t leader = __ffsll((unsigned long long)act) - 1;
    const unsigned long long need = (unsigned long long)__popcll(act);
    const unsigned long long rank = (unsigned long long)__popcll(act & lanemask_lt());
    unsigned long long old_base = 0, rem = 0, new_base = 0;
    if (lane_id() == leader) {
        const unsigned long long nx = ws[0], en = ws[1];
        rem = en - nx < need ? en - nx : need;
        old_base = nx;
        if (rem < need) {
            new_base = atomicAdd(ctr, EVENT_CHUNK);
            ws[0] = new_base + (need - rem);
            ws[1] = new_base + EVENT_CHUNK;
        } else {
            ws[0] = nx + need;
        }
    }
    old_base = __shfl(old_base, leader);
    rem = __shfl(rem, leader);
    new_base = __shfl(new_base, leader);
    return rank < rem ? old_base + rank : new_base + (rank - rem);
}

// at wave exit (all lanes): ghost records into the unused slots below cap,
// with their sort keys (time +inf) when the kernel writes them
__device__ __forceinline__ void wave_fill_ghosts(volatile unsigned long long *ws, DevPacked *out,
                                                 unsigned long long cap, unsigned long long *n_ghost,
                                                 unsigned long long *keys, uint32_t *vals) {
    const unsigned long long nx = ws[0], en0 = ws[1];
    const unsigned long long en = en0 < cap ? en0 : cap;
    if (en <= nx) return;
    DevPacked g{};
    g.event_time = from_bits(0x7ff0000000000000ull);
    for (unsigned long long s = nx + lane_id(); s < en; s += 64) {
        out[s] = g;
        if (keys) {
            keys[s] = 0x7ff0000000000000ull;
            vals[s] = (uint32_t)s;
        }
    }
    if (lane_id() == 0) atomicAdd(n_ghost, en - nx);
}

// a reported event's sort key (the magnitude of its time) and slot
__device__ __forceinline__ void emit_key(const PersistArgs &A, unsigned long long eslot, double t,
                                         unsigned long long &lo, unsigned long long &hi) {
    const unsigned long long k = bits(t) & 0x7fffffffffffffffull;
    A.ev_keys[eslot] = k;
    A.ev_vals[eslot] = (uint32_t)eslot;
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
}

// the wave's key range into A.ev_range (one atomic pair per wave)
__device__ __forceinline__ void wave_flush_range(const PersistArgs &A, unsigned long long lo,
                                                 unsigned long long hi) {
    for (int off = 32; off; off >>= 1) {
        const unsigned long long l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    if (lane_id() == 0 && lo <= hi) {
        atomicMin(&A.ev_range[0], lo);
        atomicMax(&A.ev_range[1], hi);
    }
}

struct Origin {
    uint32_t x, y, i;
    double prior;
};

template <int KIND>
__device__ __forceinline__ Origin lane_origin(const Lane<KIND> &L, const DevScenario &S) {
    Origin o;
    location_xy<KIND>(S, L.a, L.b, o.x, o.y);
    o.i = L.i;
    o.prior = L.t_last;
    return o;
}

// the event the step that left the lineage in L reported: a speciation
// (prior(-), event(+), target = origin) or a dispersal with
// LineageInteraction::Maybe (prior(+), event(-)); coalescence = the
// lineage's own reference (necsim/core/src/event.rs:103-155)
template <int KIND>
__device__ __forceinline__ void emit_packed(DevPacked *out, const Origin &o, const Lane<KIND> &L,
                                            const DevScenario &S, bool speciated) {
    DevPacked e;
    e.ref = L.ref;
    e.coalescence = L.ref;
    e.ox = o.x;
    e.oy = o.y;
    e.oindex = o.i;
    if (speciated) {
        e.prior_time = from_bits(bits(o.prior) | SIGN_BIT);
        e.event_time = L.t_last;
        e.tx = o.x;
        e.ty = o.y;
        e.tindex = o.i;
    } else {
        e.prior_time = o.prior;
        e.event_time = from_bits(bits(L.t_last) | SIGN_BIT);
        location_xy<KIND>(S, L.a, L.b, e.tx, e.ty);
        e.tindex = L.i;
    }
    *out = e;
}

// ---------------------------------------------------------------------------
// Wave-level work claiming: each wave owns a private range [next, end) of
// work indices and refills it with ONE global atomic per claim, so lane
// refill costs a ballot and a popcount, not a contended atomic.  Claims are
// guided: a wave takes about remaining / (2 * waves) indices (at most
// `chunk`, at least its idle lanes), estimating `remaining` from the queue
// head its last claim returned.  Large claims early keep the atomic rate
// low; claims shrink to single lanes as the queue drains, so every wave runs
// out of work at about the same time.  With chunk == 0 a wave always claims
// exactly as many indices as it has idle lanes.

struct WaveQueue {
    unsigned long long next, end;
    unsigned long long seen; // queue head returned by this wave's last claim
    bool drained;   // the global queue is empty
    bool exhausted; // ... and so is this wave's range
};

__device__ __forceinline__ unsigned long long grid_waves() {
    return (unsigned long long)gridDim.x * ((blockDim.x + 63) / 64);
}

// Hands this lane the next work index if it needs one (need = ballot of
// lanes without work).  Returns ~0ull when the lane gets nothing.
__device__ __forceinline__ unsigned long long wave_claim(WaveQueue &q, uint64_t need, bool wants,
                                                         unsigned long long *head,
                                                         unsigned long long n_total,
                                                         unsigned long long chunk) {
    const unsigned lane = lane_id();
    if (q.next >= q.end && !q.drained) {
        uint32_t leader = __ffsll((unsigned long long)need) - 1;
        unsigned long long idle = (unsigned long long)__popcll(need);
        unsigned long long take = idle;
        if (chunk) {
            unsigned long long rem = n_total > q.seen ? n_total - q.seen : 0;
            unsigned long long g = rem / (2 * grid_waves());
            g = g < chunk ? g : chunk;
            take = g > idle ? g : idle;
        }
        // A sparse round spreads its lineages over every wave: the lanes of
        // a wave step in lockstep, so lineages sharing a wave add up their
        // dependent chains, while a wave with one lineage runs at the
        // latency of one chain.
        const unsigned long long per_wave = (n_total + grid_waves() - 1) / grid_waves();
        if (per_wave < take) take = per_wave > 0 ? per_wave : 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(head, take);
        base = __shfl(base, leader);
        q.seen = base + take;
        if (base >= n_total) {
            q.drained = true;
        } else {
            q.next = base;
            q.end = base + take < n_total ? base + take : n_total;
        }
    }
    unsigned long long avail = q.end - q.next;
    unsigned long long mine = ~0ull;
    if (wants) {
        unsigned long long r = __popcll(need & lanemask_lt());
        if (r < avail) mine = q.next + r;
    }
    unsigned long long cnt = __popcll(need);
    q.next += cnt < avail ? cnt : avail;
    if (q.drained && q.next >= q.end) q.exhausted = true;
    return mine;
}

__device__ __forceinline__ void block_flush(unsigned long long steps, double max_t, unsigned long long dead,
                                            Counters *ctr) {
    __shared__ unsigned long long s_steps, s_max, s_dead;
    if (threadIdx.x == 0) {
        s_steps = 0;
        s_max = 0;
        s_dead = 0;
    }
    __syncthreads();
    if (steps) atomicAdd(&s_steps, steps);
    if (max_t > 0.0) atomicMax(&s_max, (unsigned long long)bits(max_t));
    if (dead) atomicAdd(&s_dead, dead);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_steps) atomicAdd(&ctr->steps, s_steps);
        if (s_max) atomicMax(&ctr->max_time_bits, s_max);
        if (s_dead) atomicAdd(&ctr->n_dead, s_dead);
    }
}

// ---------------------------------------------------------------------------
// Persistent round kernel.  Lineages are updated in place: a lineage that
// reaches the water level is written back to its slot, one that speciates or
// is dropped as a duplicate only gets its dead flag.  Each inner iteration
// advances every searching lane's next-event search a little (at most one
// Poisson window opened and two of its event draws, lane_search_step);
// once at least exec_fill eighths of the wave's active lanes hold their next
// event, those lanes handle it (leave at the level, dedup, execute) in the
// same iteration.  A search takes a variable number of windows, so stepping
// the searches in lockstep and batching the executions keeps the lanes of
// the expensive execution path mostly active, where a whole search followed
// by an execution per lane would wait for the wave's longest search.
// DEDUP selects the variant with the in-kernel dedup cache (min speciation
// sample per slice).
// ---------------------------------------------------------------------------
// Cooperative next-event search.  A window's events depend only on the
// location and the window index (the RNG is primed with both), so any lane
// can evaluate any lineage's next window.  In each iteration the lanes that
// are not searching (waiting for their wave to execute, or without a
// lineage) are spread over the searching ones: helper r takes searcher
// r mod s at window offset 1 + r / s, where s is the number of searchers.
// Every lane then evaluates one window, and each searcher takes the first of
// its group's windows that holds an event, which is the window the
// sequential search would have stopped at.  Without helpers this is one
// window per searcher, as before.
struct CoopLds {
    uint8_t tab[256 / 64][64];     // per wave: searcher lanes by rank
    unsigned long long win[256];   // per searcher lane: (offset << 6) | lane of its first window with an event
    double et[256];                // a helper's event time
    uint64_t key[256];             // a helper's event key
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lanes may hold lineages of different simulations (serving): the location
// key a helper takes from its searcher carries the searcher's seed.
template <int KIND>
__device__ __forceinline__ void coop_search_window(Lane<KIND> &L, bool srch, const DevParams &P,
                                                   const double *ratio, CoopLds &c) {
    const uint64_t ms = __ballot(srch);
    if (!ms) return;
    const uint64_t mh = __ballot(!srch);
    const unsigned lane = lane_id(), wave = threadIdx.x / 64, base = wave * 64;
    const unsigned ns = (unsigned)__popcll(ms), nh = (unsigned)__popcll(mh);
    unsigned tgt = lane;
    uint64_t off = 0;
    if (nh) {
        if (srch) {
            c.tab[wave][__popcll(ms & lanemask_lt())] = (uint8_t)lane;
            c.win[threadIdx.x] = ~0ull;
        }
        wave_sync_lds();
        if (!srch) {
            const unsigned r = (unsigned)__popcll(mh & lanemask_lt());
            tgt = c.tab[wave][r % ns];
            off = 1 + r / ns;
        }
    }
    const uint64_t lkey = __shfl(L.lkey, tgt);
    const double before = __shfl(L.search_before, tgt);
    const uint64_t ts = __shfl(L.search_ts, tgt);
    double et = 0.0;
    uint64_t key = 0;
    bool found = poisson_window(P, ratio, lkey, before, ts + off, et, key);
    if (nh) {
        if (!srch && found) {
            c.et[threadIdx.x] = et;
            c.key[threadIdx.x] = key;
            atomicMin(&c.win[base + tgt], (unsigned long long)((off << 6) | lane));
        }
        wave_sync_lds();
    }
    if (!srch) return;
    if (!found && nh) {
        const unsigned long long w = c.win[threadIdx.x];
        if (w != ~0ull) {
            const unsigned src = base + (unsigned)(w & 63u);
            et = c.et[src];
            key = c.key[src];
            found = true;
        }
    }
    if (found) {
        L.next_t = max_after(L.search_before, et);
        L.next_st = key;
        L.have_next = true;
        L.searching = false;
    } else {
        const unsigned k = (unsigned)__popcll(ms & lanemask_lt());
        L.search_ts += 1 + (nh ? nh / ns + (k < nh % ns ? 1u : 0u) : 0u);
    }
}

__device__ __forceinline__ bool wave_ready(bool ready, unsigned fill_eighths) {
    uint32_t act = (uint32_t)__popcll(__ballot(true));
    uint32_t rdy = (uint32_t)__popcll(__ballot(ready));
    return rdy != 0 && rdy * 8u >= act * fill_eighths;
}

// The dedup variant holds the slice's min speciation sample as well: it gets
// registers for 3 waves per SIMD instead of 4, which measured faster than
// spilling.  The DISP variant reports every step as a PackedEvent into the
// round's buffer (speciations included) instead of speciations only.  The
// COOP variant searches cooperatively (coop_search_window): it pays off where
// waves hold few live lineages (coalescing populations, small samples) and
// costs the dense kernel registers, so it is a variant of its own.
template <int KIND, bool DEDUP, bool DISP, bool COOP>
__global__ void __launch_bounds__(256, (DEDUP || DISP) ? 3 : 4) k_persistent(PersistArgs A) {
    __shared__ double s_ratio[POISSON_TABLE_N];
    __shared__ typename std::conditional<COOP, CoopLds, char>::type s_coop;
    __shared__ unsigned long long s_slots[DISP ? 2 * (256 / 64) : 1]; // per wave: event chunk [next, end)
    volatile unsigned long long *ws = &s_slots[DISP ? 2 * (threadIdx.x / 64) : 0];
    if (DISP && lane_id() == 0) {
        ws[0] = 0;
        ws[1] = 0;
    }
    load_ratio(s_ratio, A.P.ratio);

    const double level = A.round->level;
    const unsigned long long n_in = A.round->n_in;

    Lane<KIND> L;
    L.have_next = false;
    L.searching = false;
    L.ms_some = false;
    bool has = false;
    unsigned long long slot = 0;
    WaveQueue q{0, 0, 0, false, false};
    uint64_t slice_steps = 0;
    unsigned long long my_steps = 0, my_dead = 0;
    double my_max = 0.0;
    unsigned long long key_lo = ~0ull, key_hi = 0ull; // DISP: the range of the keys this lane wrote

    for (;;) {
        if (!q.exhausted) {
            uint64_t need = __ballot(!has);
            if (need) {
                unsigned long long w = wave_claim(q, need, !has, &A.ctr->head, n_in, A.round->chunk);
                if (w != ~0ull && !(lane_load<KIND>(L, A.S, A.lin, w, A.P.seed) & FLAG_DEAD)) {
                    slot = w;
                    has = true;
                    slice_steps = 0;
                }
            }
        }
        if (!__any(has)) {
            if (q.exhausted) break;
            continue;
        }

        // up to inner_steps iterations per lane between wave-level refills;
        // with COOP, lanes whose lineage ended stay in the loop to help the
        // searches of the others
        if (has || COOP) {
#pragma unroll 1
            for (unsigned inner = 0; inner < A.inner_steps; inner++) {
                if constexpr (COOP) {
                    if (!__any(has)) break;
                    const bool srch = has && !L.have_next;
                    if (srch && !L.searching) {
                        L.search_before = A.P.min_event_time > L.t_last ? A.P.min_event_time : L.t_last;
                        L.search_ts = poisson_first_window(A.P, L.search_before);
                        L.searching = true;
                    }
                    coop_search_window<KIND>(L, srch, A.P, s_ratio, s_coop);
                    if (!has) continue;
                } else if (!L.have_next) {
                    lane_search_step<KIND>(L, A.P, s_ratio);
                }
                if (!wave_ready(L.have_next, A.exec_fill) || !L.have_next) continue;
                // Leaving this round (water level) and reaching the step slice
                // both end the slice: its min speciation sample is inserted like
                // on reclassification (monolithic/mod.rs:205-219).
                bool leave = L.next_t >= level;
                if (DEDUP) {
                    if (leave || slice_steps >= A.P.step_slice) {
                        bool dup = L.ms_some &&
                                   cache_insert_concurrent(A.cache, A.cache_cap, lane_sample<KIND>(L, A.S));
                        L.ms_some = false;
                        slice_steps = 0;
                        if (dup) {
                            A.lin.flags[slot] = FLAG_DEAD;
                            my_dead++;
                            has = false;
                            if (COOP) continue;
                            break;
                        }
                    }
                }
                if (leave) {
                    lane_store<KIND>(L, A.S, A.lin, slot);
                    has = false;
                    if (COOP) continue;
                    break;
                }
                Origin o;
                unsigned long long eslot = 0;
                if (DISP) {
                    eslot = wave_take_slot(ws, &A.ctr->n_events);
                    if (eslot >= A.packed_cap) {
                        A.ctr->ev_overflow = 1;
                        lane_store<KIND>(L, A.S, A.lin, slot);
                        has = false;
                        if (COOP) continue;
                        break;
                    }
                    o = lane_origin<KIND>(L, A.S);
                }
                DevEvent ev;
                int r = lane_execute<KIND, DEDUP, COOP>(L, A.S, A.P, ev);
                my_steps++;
                slice_steps++;
                my_max = L.t_last > my_max ? L.t_last : my_max;
                if (DISP) {
                    emit_packed<KIND>(A.packed + eslot, o, L, A.S, r == STEP_SPECIATED);
                    if (A.ev_keys) emit_key(A, eslot, L.t_last, key_lo, key_hi);
                }
                if (r == STEP_SPECIATED) {
                    if (!DISP) { // rare
                        ev.ref = A.lin.ref[slot];
                        A.events[atomicAdd(&A.ctr->n_events, 1ull)] = ev;
                    }
                    A.lin.flags[slot] = FLAG_DEAD;
                    my_dead++;
                    has = false;
                    if (COOP) continue;
                    break;
                }
            }
        }
    }
    if (DISP) {
        wave_fill_ghosts(ws, A.packed, A.packed_cap, &A.ctr->n_ghost, A.ev_keys, A.ev_vals);
        if (A.ev_keys) wave_flush_range(A, key_lo, key_hi);
    }
    block_flush(my_steps, my_max, my_dead, A.ctr);
}

// ---------------------------------------------------------------------------
// Pooled persistent round kernel (dense rounds: no dedup cache).  Each wave
// (one per block) keeps POOL lineages in LDS,
// POOL / 64 per lane, and works in two phases per iteration:
//   search:  up to 64 of the lineages still looking for their next event
//            each advance by one lane_search_step, one lineage per lane;
//   execute: once 64 lineages hold their next event (or, when the queue is
//            empty, exec_fill eighths of those left), 64 of them execute it,
//            one per lane.
// Lanes pick lineages by rank in LDS, so both phases run with full waves
// however unevenly the searches progress; in the one-lineage-per-lane
// kernel a lane whose lineage is ready waits for the wave's execution and a
// lane whose lineage still searches idles through it.  Each lineage makes
// exactly the computations of k_persistent, so the event stream is the same.
// Entries are owned by lanes for bookkeeping (entry lane + 64 j), refilled
// from the same guided work queue.  The DISP variant reports every executed
// step as a PackedEvent (dispersal reporting): a dispersal-reporting round
// runs each lineage for about two steps, so lineages enter and leave the
// pool all the time, which the one-lineage-per-lane kernel pays for with
// idle lanes.
constexpr int POOL_MAX = 192;
// PE_LEAVE (dispersal reporting): the next event passes the water level; the
// entry waits for a leave pass, which stores it to the lineage store
enum : uint8_t { PE_EMPTY = 0, PE_SEARCH = 1, PE_READY = 2, PE_LEAVE = 3 };
constexpr uint32_t PE_NOT_STARTED = 0xffffffffu; // win_left of a search that has not opened a window

template <int POOL, bool REF> struct PoolLds {
    uint64_t lkey[POOL];
    uint64_t st[POOL];    // searching: the open window's RNG state; ready: the event's prime key
    uint64_t ts[POOL];    // searching: the window
    double t_last[POOL];
    double next_t[POOL];  // searching: the earliest valid time of the open window so far; ready: the event
    uint32_t a[POOL], b[POOL], i[POOL];
    uint32_t slot[POOL];
    uint32_t win_left[POOL], win_e[POOL], win_bi[POOL];
    uint8_t state[POOL];
    uint8_t assign[64];
    double tab[3 * PTBL]; // the Poisson table's acc / prod / acc_m
    // dispersal reporting: each event carries the lineage's reference (last,
    // so that the other arrays keep offsets that pair LDS accesses)
    uint64_t ref[REF ? POOL : 1];
};

// Lane r (r < the number of selected entries, at most 64) gets the r-th
// selected entry, counting from word `first` round the words of `sel`.
template <int NW>
__device__ __forceinline__ uint32_t pool_pick(const uint64_t (&sel)[NW], int first, uint8_t *assign) {
    const unsigned lane = lane_id();
    uint32_t base = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        const int j = (first + k) % NW;
        if ((sel[j] >> lane) & 1ull) {
            const uint32_t r = base + (uint32_t)__popcll(sel[j] & lanemask_lt());
            if (r < 64) assign[r] = (uint8_t)(lane + 64 * j);
        }
        base += (uint32_t)__popcll(sel[j]);
    }
    wave_sync_lds();
    return base < 64 ? base : 64;
}

template <int KIND, int POOL, bool DISP>
__global__ void __launch_bounds__(64) k_pooled(PersistArgs A) {
    static_assert(POOL % 64 == 0 && POOL <= POOL_MAX, "pool of whole words");
    constexpr int NW = POOL / 64;
    __shared__ PoolLds<POOL, DISP> p;
    __shared__ unsigned long long s_slots[2]; // DISP: the wave's event chunk [next, end)
    volatile unsigned long long *ws = s_slots;
    const unsigned lane = lane_id();
    if (DISP && lane == 0) {
        ws[0] = 0;
        ws[1] = 0;
    }
    if (lane < 3 * PTBL) p.tab[lane] = A.P.ratio[RATIO_N + lane];
#pragma unroll
    for (int j = 0; j < NW; j++) p.state[lane + 64 * j] = PE_EMPTY;
    wave_sync_lds();

    const double level = A.round->level;
    const unsigned long long n_in = A.round->n_in;
    WaveQueue q{0, 0, 0, false, false};
    unsigned long long my_steps = 0, my_dead = 0;
    double my_max = 0.0;
    unsigned long long key_lo = ~0ull, key_hi = 0ull; // DISP: the range of the keys this lane wrote
    uint32_t iter = 0;

    uint32_t n_empty = POOL;
    for (;; iter++) {
        // refill the lane's empty entries from the work queue, one per lane
        // and call (a dead slot leaves its entry empty).  Refills wait until
        // a quarter of the pool is empty (dispersal reporting: half), the
        // pool runs low or 16 iterations have passed: every claim is an
        // atomic on the one queue head, and a refill waits for its loads,
        // which on this hardware also waits for every store issued before
        // them (the event records of dispersal reporting, where a quarter
        // measured 4 % slower; half was 1 % slower on the dense headline).
        const bool refill = !q.exhausted && (n_empty * (DISP ? 2u : 4u) >= POOL || n_empty + 64 > POOL ||
                                             (iter & 15) == 0);
        uint8_t own[NW];
#pragma unroll
        for (int j = 0; j < NW; j++) {
            const int e = lane + 64 * j;
            own[j] = p.state[e];
            if (!refill || q.exhausted) continue;
            const bool want = own[j] == PE_EMPTY;
            const uint64_t need = __ballot(want);
            if (!need) continue;
            const unsigned long long w = wave_claim(q, need, want, &A.ctr->head, n_in, A.round->chunk);
            if (w == ~0ull) continue;
            Lane<KIND> L;
            const uint8_t f = lane_load<KIND>(L, A.S, A.lin, w, A.P.seed);
            if (f & FLAG_DEAD) continue;
            p.lkey[e] = L.lkey;
            if (DISP) p.ref[e] = L.ref;
            p.t_last[e] = L.t_last;
            p.a[e] = L.a;
            p.b[e] = L.b;
            p.i[e] = L.i;
            p.slot[e] = (uint32_t)w;
            p.next_t[e] = L.next_t;
            p.st[e] = L.next_st;
            p.win_left[e] = PE_NOT_STARTED;
            own[j] = L.have_next ? ((DISP && L.next_t >= level) ? PE_LEAVE : PE_READY) : PE_SEARCH;
            p.state[e] = own[j];
        }
        uint64_t srch[NW], rdy[NW], lv[NW];
        uint32_t n_occ = 0, n_rdy = 0, n_lv = 0;
#pragma unroll
        for (int j = 0; j < NW; j++) {
            srch[j] = __ballot(own[j] == PE_SEARCH);
            rdy[j] = __ballot(own[j] == PE_READY);
            lv[j] = DISP ? __ballot(own[j] == PE_LEAVE) : 0ull;
            n_occ += (uint32_t)__popcll(srch[j] | rdy[j] | lv[j]);
            n_rdy += (uint32_t)__popcll(rdy[j]);
            n_lv += (uint32_t)__popcll(lv[j]);
        }
        n_empty = POOL - n_occ;
        if (n_occ == 0) {
            if (q.exhausted) break;
            continue;
        }

        // search phase: one step for up to 64 searching lineages
        if (n_occ > n_rdy + n_lv) {
            const uint32_t n = pool_pick<NW>(srch, (int)(iter % NW), p.assign);
            if (lane < n) {
                const int e = p.assign[lane];
                Lane<KIND> L;
                L.lkey = p.lkey[e];
                L.t_last = p.t_last[e];
                const uint32_t wl = p.win_left[e];
                L.searching = wl != PE_NOT_STARTED;
                L.win_left = L.searching ? wl : 0u;
                L.win_e = p.win_e[e];
                L.win_bi = p.win_bi[e];
                L.search_ts = p.ts[e];
                L.next_st = p.st[e];
                L.next_t = p.next_t[e];
                L.have_next = false;
                lane_search_step<KIND>(L, A.P, A.P.ratio, p.tab);
                p.next_t[e] = L.next_t;
                p.st[e] = L.next_st;
                if (L.have_next) {
                    p.state[e] = (DISP && L.next_t >= level) ? PE_LEAVE : PE_READY;
                } else {
                    p.ts[e] = L.search_ts;
                    p.win_left[e] = L.win_left;
                    p.win_e[e] = L.win_e;
                    p.win_bi[e] = L.win_bi;
                }
            }
            wave_sync_lds();
            n_rdy = 0;
            n_lv = 0;
#pragma unroll
            for (int j = 0; j < NW; j++) {
                const uint8_t st = p.state[lane + 64 * j];
                rdy[j] = __ballot(st == PE_READY);
                n_rdy += (uint32_t)__popcll(rdy[j]);
                if (DISP) {
                    lv[j] = __ballot(st == PE_LEAVE);
                    n_lv += (uint32_t)__popcll(lv[j]);
                }
            }
        }

        // leave phase (dispersal reporting, where lineages reach the water
        // level every few steps): the entries whose next event passes it are
        // stored in a pass of their own, so execution passes do not run with
        // lanes that only store (full waves of 64: measured faster than 16 or
        // 32 per pass)
        if constexpr (DISP) {
            if (n_lv >= 64 || (n_lv > 0 && (n_lv + n_rdy == n_occ || q.exhausted))) {
                const uint32_t n = pool_pick<NW>(lv, (int)(iter % NW), p.assign);
                if (lane < n) {
                    const int e = p.assign[lane];
                    Lane<KIND> L;
                    L.t_last = p.t_last[e];
                    L.a = p.a[e];
                    L.b = p.b[e];
                    L.i = p.i[e];
                    L.next_t = p.next_t[e];
                    L.next_st = p.st[e];
                    L.have_next = true;
                    lane_store<KIND>(L, A.S, A.lin, p.slot[e]);
                    p.state[e] = PE_EMPTY;
                }
                wave_sync_lds();
                n_occ -= n;
                n_lv -= n;
            }
        }

        // execute phase: full waves, or the rest once the queue is empty
        const uint32_t n_act = n_occ - n_lv; // searching or ready
        const bool go = n_rdy >= 64 ||
                        (n_rdy > 0 && (n_rdy == n_act || (q.exhausted && n_rdy * 8u >= n_act * A.exec_fill)));
        if (!go) continue;
        const uint32_t n = pool_pick<NW>(rdy, (int)(iter % NW), p.assign);
        if (lane < n) {
            const int e = p.assign[lane];
            const uint32_t slot = p.slot[e];
            Lane<KIND> L;
            L.lkey = p.lkey[e];
            L.t_last = p.t_last[e];
            L.a = p.a[e];
            L.b = p.b[e];
            L.i = p.i[e];
            L.next_t = p.next_t[e];
            L.next_st = p.st[e];
            L.have_next = true;
            if (is_cell_kind(KIND) && KIND != K_NON_SPATIAL) L.rec = A.S.cells[L.a];
            bool leave = L.next_t >= level; // the water level: back to the lineage store
            unsigned long long eslot = 0;
            Origin o;
            if (DISP && !leave) {
                eslot = wave_take_slot(ws, &A.ctr->n_events);
                if (eslot >= A.packed_cap) { // the round's buffer is full: not executed
                    A.ctr->ev_overflow = 1;
                    leave = true;
                } else {
                    // from LDS: a global load here would also wait for the
                    // previous passes' event stores (7 % of the step kernel)
                    L.ref = p.ref[e];
                    o = lane_origin<KIND>(L, A.S);
                }
            }
            if (leave) {
                lane_store<KIND>(L, A.S, A.lin, slot);
                p.state[e] = PE_EMPTY;
            } else {
                DevEvent ev;
                // straight-line sin / cos: a full wave of Box-Muller angles
                // takes every branch of the case-by-case reduction
                const int r = lane_execute<KIND, false, true>(L, A.S, A.P, ev);
                my_steps++;
                my_max = L.t_last > my_max ? L.t_last : my_max;
                if (DISP) {
                    emit_packed<KIND>(A.packed + eslot, o, L, A.S, r == STEP_SPECIATED);
                    if (A.ev_keys) emit_key(A, eslot, L.t_last, key_lo, key_hi);
                }
                if (r == STEP_SPECIATED) {
                    if (!DISP) {
                        ev.ref = A.lin.ref[slot];
                        A.events[atomicAdd(&A.ctr->n_events, 1ull)] = ev;
                    }
                    A.lin.flags[slot] = FLAG_DEAD;
                    my_dead++;
                    p.state[e] = PE_EMPTY;
                } else {
                    p.lkey[e] = L.lkey;
                    p.t_last[e] = L.t_last;
                    p.a[e] = L.a;
                    p.b[e] = L.b;
                    p.i[e] = L.i;
                    p.win_left[e] = PE_NOT_STARTED;
                    p.state[e] = PE_SEARCH;
                }
            }
        }
        wave_sync_lds();
    }
    if (DISP) {
        wave_fill_ghosts(ws, A.packed, A.packed_cap, &A.ctr->n_ghost, A.ev_keys, A.ev_vals);
        if (A.ev_keys) wave_flush_range(A, key_lo, key_hi);
    }
    block_flush(my_steps, my_max, my_dead, A.ctr);
}

// ---------------------------------------------------------------------------
// Serving kernel: the persistent round kernel over the lineages of many
// requests at once.  Global work index w belongs to the request whose
// [work_base, work_base + n_in) contains it; each lane keeps its request's
// seed, speciation probability and water level in registers and refills from
// the shared queue whichever request the next lineage belongs to.  Per-request
// step / time / death tallies are kept per lane and flushed when the lane
// moves to another request.
template <int KIND>
__global__ void __launch_bounds__(256) k_serve(ServeArgs A) {
    __shared__ double s_ratio[POISSON_TABLE_N];
    __shared__ unsigned long long s_base[SERVE_MAX_REQS];
    __shared__ CoopLds s_coop; // cooperative next-event search: requests are often small
    load_ratio(s_ratio, A.P.ratio);
    for (int r = threadIdx.x; r < A.n_reqs; r += blockDim.x) s_base[r] = A.reqs[r].work_base;
    __syncthreads();

    Lane<KIND> L;
    L.have_next = false;
    L.searching = false;
    L.ms_some = false;
    DevParams P = A.P;
    double level = 0.0;
    int req = -1;
    bool has = false;
    unsigned long long slot = 0;
    WaveQueue q{0, 0, 0, false, false};
    uint64_t slice_steps = 0, tick = 0;
    unsigned long long acc_steps = 0, acc_dead = 0;
    double acc_max = 0.0;
    auto flush = [&](int r) {
        if (r < 0) return;
        ReqCounters *c = A.reqs[r].ctr;
        if (acc_steps) atomicAdd(&c->steps, acc_steps);
        if (acc_max > 0.0) atomicMax(&c->max_time_bits, (unsigned long long)bits(acc_max));
        if (acc_dead) atomicAdd(&c->n_dead, acc_dead);
        acc_steps = 0;
        acc_dead = 0;
        acc_max = 0.0;
    };

    for (;;) {
        if (!q.exhausted) {
            uint64_t need = __ballot(!has);
            if (need) {
                unsigned long long w = wave_claim(q, need, !has, A.head, A.total_work, A.chunk);
                if (w != ~0ull) {
                    int lo = 0, hi = A.n_reqs - 1; // last request with work_base <= w
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (s_base[mid] <= w) lo = mid;
                        else hi = mid - 1;
                    }
                    const ServeReq &R = A.reqs[lo];
                    unsigned long long s = w - s_base[lo];
                    if (!(lane_load<KIND>(L, A.S, R.lin, s, R.seed) & FLAG_DEAD)) {
                        if (lo != req) {
                            flush(req);
                            req = lo;
                            P.seed = R.seed;
                            P.nu = R.nu;
                            P.nu_m = R.nu_m;
                            level = R.level;
                        }
                        slot = s;
                        has = true;
                        slice_steps = 0;
                        tick = 0;
                    }
                }
            }
        }
        if (!__any(has)) {
            if (q.exhausted) break;
            continue;
        }
        {
            // every lane iterates: those without a lineage help the searches
            const ServeReq &R = A.reqs[req < 0 ? 0 : req];
#pragma unroll 1
            for (int inner = 0; inner < 4; inner++) {
                if (!__any(has)) break;
                const bool srch = has && !L.have_next;
                if (srch && !L.searching) {
                    L.search_before = P.min_event_time > L.t_last ? P.min_event_time : L.t_last;
                    L.search_ts = poisson_first_window(P, L.search_before);
                    L.searching = true;
                }
                coop_search_window<KIND>(L, srch, P, s_ratio, s_coop);
                if (!has) continue;
                if (!wave_ready(L.have_next, 5) || !L.have_next) continue;
                // The tick limit ends a lane's work only at a step-slice
                // boundary (or before the slice's first step), where the
                // slice's min sample has just been inserted: the next tick
                // starts a fresh slice, so slices (and the dedup inserts) fall
                // where they do in an uninterrupted round.
                const bool slice_end = slice_steps >= P.step_slice;
                bool leave = L.next_t >= level ||
                             (A.tick_steps && tick >= A.tick_steps && (slice_end || slice_steps == 0));
                if (leave || slice_end) {
                    bool dup = R.cache && L.ms_some &&
                               cache_insert_concurrent(R.cache, R.cache_cap, lane_sample<KIND>(L, A.S));
                    L.ms_some = false;
                    slice_steps = 0;
                    if (dup) {
                        R.lin.flags[slot] = FLAG_DEAD;
                        acc_dead++;
                        has = false;
                        continue;
                    }
                    if (leave) {
                        lane_store<KIND>(L, A.S, R.lin, slot);
                        if (L.next_t < level) R.ctr->n_below = 1; // the round continues next tick
                        has = false;
                        continue;
                    }
                }
                DevEvent ev;
                int r = lane_execute<KIND, true, true>(L, A.S, P, ev);
                acc_steps++;
                slice_steps++;
                tick++;
                acc_max = L.t_last > acc_max ? L.t_last : acc_max;
                if (r == STEP_SPECIATED) {
                    ev.ref = R.lin.ref[slot];
                    R.events[atomicAdd(&R.ctr->n_events, 1ull)] = ev;
                    R.lin.flags[slot] = FLAG_DEAD;
                    acc_dead++;
                    has = false;
                    continue;
                }
            }
        }
    }
    flush(req);
}

// ---------------------------------------------------------------------------
// Ordered compaction of a lineage SoA (drops dead slots, keeps slot order)
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__global__ void __launch_bounds__(SC_THREADS) k_soa_count(const uint8_t *flags, uint64_t n,
                                                           const unsigned long long *n_dev, uint32_t *blk) {
    __shared__ uint32_t c;
    if (n_dev) n = *n_dev < n ? *n_dev : n;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    uint32_t a = 0;
    for (int j = 0; j < SC_ITEMS; j++) {
        uint64_t q = base + (uint64_t)j * SC_THREADS + threadIdx.x;
        if (q < n) a += !(flags[q] & FLAG_DEAD);
    }
    if (a) atomicAdd(&c, a);
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = c;
}

__global__ void __launch_bounds__(SC_THREADS) k_soa_scatter(LineageSoA in, LineageSoA out, uint64_t n,
                                                            const unsigned long long *n_dev, const Counters *gate,
                                                            const uint32_t *off) {
    __shared__ uint32_t wcnt[SC_THREADS / 64];
    __shared__ uint32_t run;
    if (gate && !gate->compact_done) return;
    if (n_dev) n = *n_dev < n ? *n_dev : n;
    if ((uint64_t)blockIdx.x * SC_TILE >= n) return;
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
    if (threadIdx.x == 0) run = off[blockIdx.x];
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    for (int j = 0; j < SC_ITEMS; j++) {
        uint64_t q = base + (uint64_t)j * SC_THREADS + threadIdx.x;
        bool alive = q < n && !(in.flags[q] & FLAG_DEAD);
        uint64_t m = __ballot(alive);
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        uint32_t pos = run + __popcll(m & lanemask_lt());
        for (unsigned w = 0; w < wave; w++) pos += wcnt[w];
        if (alive) {
            out.ref[pos] = in.ref[q];
            out.t_last[pos] = in.t_last[q];
            out.x[pos] = in.x[q];
            out.y[pos] = in.y[q];
            out.idx[pos] = in.idx[q];
            out.next_t[pos] = in.next_t[q];
            out.next_st[pos] = in.next_st[q];
            out.flags[pos] = in.flags[q];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (unsigned w = 0; w < SC_THREADS / 64; w++) run += wcnt[w];
        __syncthreads();
    }
}

// the compaction decision of the gated compaction: dead slots are at least an
// eighth of the n slots (blk[nb] holds the live count after the scan)
__global__ void k_soa_gate(const uint32_t *blk, uint32_t nb, uint64_t cap, const unsigned long long *n_dev,
                           Counters *ctr) {
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    const uint64_t live = blk[nb];
    ctr->compact_live = live;
    ctr->compact_done = (n > live && (n - live) * 8 >= n) ? 1ull : 0ull;
}

// ---------------------------------------------------------------------------
// Plain slice kernel: element q of the pass runs one step slice
template <int KIND, bool DISP>
__global__ void __launch_bounds__(256) k_plain(PlainArgs A) {
    __shared__ double s_ratio[POISSON_TABLE_N];
    load_ratio(s_ratio, A.P.ratio);
    const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long m = A.m_dev ? *A.m_dev : A.m;
    const double level = A.round->level;
    unsigned long long my_steps = 0;
    double my_max = 0.0;
    bool emit = false;
    DevEvent ev;
    if (q < m) {
        uint32_t slot = A.pass[q];
        // the pop loop's check on the stored (pessimistic) next event time
        // (monolithic/mod.rs:186-196)
        double stored_next = A.L.next_t[slot];
        if (!(stored_next < level)) {
            A.status[q] = ST_SKIP;
        } else {
            Lane<KIND> L;
            lane_load<KIND>(L, A.S, A.L, slot, A.P.seed);
            uint8_t st;
            for (;;) {
                lane_peek<KIND>(L, A.P, s_ratio);
                if (L.next_t >= level) {
                    st = ST_FAST;
                    break;
                }
                if (my_steps >= A.P.step_slice) {
                    st = ST_SLOW;
                    break;
                }
                Origin o;
                unsigned long long eslot = 0;
                if (DISP) {
                    eslot = wave_reserve(&A.ctr->n_events);
                    if (eslot >= A.packed_cap) {
                        A.ctr->ev_overflow = 1;
                        st = ST_SLOW;
                        break;
                    }
                    o = lane_origin<KIND>(L, A.S);
                }
                int r = lane_execute<KIND>(L, A.S, A.P, ev);
                my_steps++;
                if (DISP) emit_packed<KIND>(A.packed + eslot, o, L, A.S, r == STEP_SPECIATED);
                if (r == STEP_SPECIATED) {
                    ev.ref = L.ref;
                    st = ST_SPEC;
                    emit = !DISP;
                    break;
                }
            }
            if (my_steps) my_max = L.t_last;
            if (st != ST_SPEC) {
                lane_store<KIND>(L, A.S, A.L, slot);
                if (L.ms_some) {
                    st |= ST_HAS_SAMPLE;
                    if (A.s_u) { // the slice's sample for the dedup cache (device or host)
                        Sample s = lane_sample<KIND>(L, A.S);
                        A.s_u[q] = s.u;
                        A.s_t[q] = s.t;
                        A.s_x[q] = s.x;
                        A.s_y[q] = s.y;
                        A.s_i[q] = s.i;
                        if (A.cache_cap)
                            A.s_bucket[q] = speciation_sample_hash(s.u, s.t, s.x, s.y, s.i) % A.cache_cap;
                    }
                }
            }
            A.status[q] = st;
        }
    }
    const unsigned lane = lane_id();
    uint64_t em = __ballot(emit);
    if (em) {
        uint32_t leader = __ffsll((unsigned long long)em) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&A.ctr->n_events, (unsigned long long)__popcll(em));
        base = __shfl(base, leader);
        if (emit) A.events[base + __popcll(em & lanemask_lt())] = ev;
    }
    __shared__ unsigned long long s_steps;
    __shared__ unsigned long long s_max;
    if (threadIdx.x == 0) {
        s_steps = 0;
        s_max = 0;
    }
    __syncthreads();
    if (my_steps) atomicAdd(&s_steps, my_steps);
    if (my_max > 0.0) atomicMax(&s_max, (unsigned long long)bits(my_max));
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_steps) atomicAdd(&A.ctr->steps, s_steps);
        if (s_max) atomicMax(&A.ctr->max_time_bits, s_max);
    }
}

// ---------------------------------------------------------------------------
// Ordered dedup of one pass (exact DirectMappedCache semantics in pass order)
constexpr uint32_t NO_LINK = 0xffffffffu;
constexpr int DEDUP_LIST_MAX = 64;

__global__ void k_dedup_push(const uint8_t *status, const unsigned long long *bucket, uint32_t *head,
                             uint32_t *link, uint64_t m, const unsigned long long *m_dev) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (m_dev ? *m_dev : m)) return;
    if (!(status[q] & ST_HAS_SAMPLE)) return;
    uint32_t old = atomicExch(&head[bucket[q]], (uint32_t)q);
    link[q] = old;
}

// One insert into a cache bucket held in registers: marks the element a
// duplicate if the bucket holds an equal sample (cache.rs:39-54).
__device__ __forceinline__ void bucket_replay(uint64_t &w0, uint64_t &w1, uint64_t &w2, uint64_t &w3, uint32_t e,
                                              uint8_t *status, const double *s_u, const double *s_t,
                                              const uint32_t *s_x, const uint32_t *s_y, const uint32_t *s_i) {
    Sample s{s_u[e], s_t[e], s_x[e], s_y[e], s_i[e]};
    uint64_t n0 = bits(s.t), n1 = bits(s.u), n2 = ((uint64_t)s.y << 32) | s.x;
    uint64_t n3 = ((uint64_t)sample_check(s) << 32) | s.i;
    if (w0 == n0 && w1 == n1 && w2 == n2 && w3 == n3) status[e] = (uint8_t)((status[e] & ~ST_KIND_MASK) | ST_DUP);
    w0 = n0; w1 = n1; w2 = n2; w3 = n3;
}

// The last pusher of a bucket owns it: it orders the bucket's inserts by
// pass position and replays them sequentially against the cache bucket.
// Lists longer than DEDUP_LIST_MAX (tiny caches) are replayed by repeated
// minimum searches over the list, O(length^2) but still in pass order.
__global__ void k_dedup_resolve(uint8_t *status, const unsigned long long *bucket, const double *s_u,
                                const double *s_t, const uint32_t *s_x, const uint32_t *s_y,
                                const uint32_t *s_i, uint32_t *head, const uint32_t *link,
                                uint64_t *cache, uint64_t m, const unsigned long long *m_dev) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (m_dev ? *m_dev : m)) return;
    if (!(status[q] & ST_HAS_SAMPLE)) return;
    unsigned long long b = bucket[q];
    if (head[b] != (uint32_t)q) return;
    uint32_t list[DEDUP_LIST_MAX];
    int n = 0;
    bool longer = false;
    for (uint32_t e = (uint32_t)q; e != NO_LINK; e = link[e]) {
        if (n == DEDUP_LIST_MAX) {
            longer = true;
            break;
        }
        list[n++] = e;
    }
    if (longer) {
        uint64_t *cb = cache + 4 * b;
        uint64_t w0 = cb[0], w1 = cb[1], w2 = cb[2], w3 = cb[3];
        uint32_t prev = 0;
        for (bool first = true;; first = false) {
            uint32_t next = NO_LINK;
            for (uint32_t e = (uint32_t)q; e != NO_LINK; e = link[e])
                if ((first || e > prev) && e < next) next = e;
            if (next == NO_LINK) break;
            bucket_replay(w0, w1, w2, w3, next, status, s_u, s_t, s_x, s_y, s_i);
            prev = next;
        }
        cb[0] = w0; cb[1] = w1; cb[2] = w2; cb[3] = w3;
        head[b] = NO_LINK;
        return;
    }
    for (int a = 1; a < n; a++) { // insertion sort by pass position
        uint32_t v = list[a];
        int c = a - 1;
        while (c >= 0 && list[c] > v) {
            list[c + 1] = list[c];
            c--;
        }
        list[c + 1] = v;
    }
    uint64_t *cb = cache + 4 * b;
    uint64_t w0 = cb[0], w1 = cb[1], w2 = cb[2], w3 = cb[3];
    for (int a = 0; a < n; a++) bucket_replay(w0, w1, w2, w3, list[a], status, s_u, s_t, s_x, s_y, s_i);
    cb[0] = w0; cb[1] = w1; cb[2] = w2; cb[3] = w3;
    head[b] = NO_LINK;
}

// ---------------------------------------------------------------------------
// Ordered two-way compaction
constexpr int CT_THREADS = 256;
constexpr int CT_ITEMS = 8;
constexpr int CT_TILE = CT_THREADS * CT_ITEMS;

__global__ void __launch_bounds__(CT_THREADS) k_compact_count(const uint8_t *status, uint64_t m,
                                                              const unsigned long long *m_dev,
                                                              uint32_t *blk_a, uint32_t *blk_b) {
    __shared__ uint32_t ca, cbn;
    if (m_dev) m = *m_dev;
    if (threadIdx.x == 0) { ca = 0; cbn = 0; }
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * CT_TILE;
    uint32_t a = 0, b = 0;
    for (int j = 0; j < CT_ITEMS; j++) {
        uint64_t q = base + (uint64_t)j * CT_THREADS + threadIdx.x;
        if (q < m) {
            uint8_t k = status[q] & ST_KIND_MASK;
            a += k == ST_SLOW;
            b += k == ST_FAST;
        }
    }
    if (a) atomicAdd(&ca, a);
    if (b) atomicAdd(&cbn, b);
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_a[blockIdx.x] = ca;
        blk_b[blockIdx.x] = cbn;
    }
}

// exclusive scan of n values in one block; data[n] receives the total
__global__ void __launch_bounds__(1024) k_scan_exclusive(uint32_t *data, uint64_t n) {
    __shared__ uint32_t partial[1024];
    const int T = blockDim.x;
    uint64_t per = (n + T - 1) / T;
    uint64_t lo = per * threadIdx.x, hi = lo + per < n ? lo + per : n;
    uint32_t sum = 0;
    for (uint64_t j = lo; j < hi; j++) sum += data[j];
    partial[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? partial[threadIdx.x - off] : 0;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? partial[threadIdx.x - 1] : 0;
    for (uint64_t j = lo; j < hi; j++) {
        uint32_t v = data[j];
        data[j] = run;
        run += v;
    }
    if (threadIdx.x == T - 1) data[n] = partial[T - 1];
}

__global__ void __launch_bounds__(CT_THREADS) k_compact_scatter(const uint8_t *status, const uint32_t *pass,
                                                                uint64_t m, const PlainLoop *loop,
                                                                const uint32_t *off_a, const uint32_t *off_b,
                                                                uint32_t *out_a, uint32_t *out_b, uint64_t b_base) {
    __shared__ uint32_t wa[CT_THREADS / 64], wb[CT_THREADS / 64];
    __shared__ uint32_t run_a, run_b;
    if (loop) {
        m = loop->m;
        b_base = loop->f_len;
    }
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
    if (threadIdx.x == 0) {
        run_a = off_a[blockIdx.x];
        run_b = off_b[blockIdx.x];
    }
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * CT_TILE;
    for (int j = 0; j < CT_ITEMS; j++) {
        uint64_t q = base + (uint64_t)j * CT_THREADS + threadIdx.x;
        uint8_t k = q < m ? (status[q] & ST_KIND_MASK) : 0xff;
        bool ia = k == ST_SLOW, ib = k == ST_FAST;
        uint64_t ma = __ballot(ia), mb = __ballot(ib);
        if (lane == 0) {
            wa[wave] = __popcll(ma);
            wb[wave] = __popcll(mb);
        }
        __syncthreads();
        uint32_t pa = run_a, pb = run_b;
        for (unsigned w = 0; w < wave; w++) {
            pa += wa[w];
            pb += wb[w];
        }
        if (ia) out_a[pa + __popcll(ma & lanemask_lt())] = pass[q];
        if (ib) out_b[b_base + pb + __popcll(mb & lanemask_lt())] = pass[q];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (unsigned w = 0; w < CT_THREADS / 64; w++) {
                run_a += wa[w];
                run_b += wb[w];
            }
        }
        __syncthreads();
    }
}

__global__ void k_compact_totals(const uint32_t *off_a, const uint32_t *off_b, uint32_t nb,
                                 unsigned int *counts) {
    counts[0] = off_a[nb];
    counts[1] = off_b[nb];
}

// the next pass: its size, and the fast lineages appended by this one
__global__ void k_plain_loop_advance(const uint32_t *off_a, const uint32_t *off_b, uint32_t nb, PlainLoop *loop) {
    loop->m = off_a[nb];
    loop->f_len += off_b[nb];
}

// ---------------------------------------------------------------------------
// LSD radix sort of (u64 key, u32 value) by 8-bit digits, stable
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

// Sort sizes may come from device memory (graph replays): n = *end - *start
// and the events start at events + *start.  A non-null `gate` makes every
// kernel of a pass a no-op while *gate == 0 (the tie passes of a captured
// sort, which only run when the round has equal event times).
struct SortSize {
    uint64_t n; // the event count, or with `end` the capacity of the range
    const unsigned long long *end;
    const unsigned long long *start;
    const unsigned long long *gate;
    __device__ __forceinline__ uint64_t get() const {
        if (!end) return n;
        uint64_t c = *end - (start ? *start : 0ull);
        return c < n ? c : n; // an overflowed round reports its overflow; its sort stays in bounds
    }
    __device__ __forceinline__ uint64_t offset() const { return start ? *start : 0ull; }
    // radix tiles actually used (the launch grid may be sized for capacity)
    __device__ __forceinline__ uint32_t tiles() const { return (uint32_t)((get() + RS_TILE - 1) / RS_TILE); }
    __device__ __forceinline__ bool skip() const { return gate && *gate == 0ull; }
};

// The PackedEvent order (necsim/core/src/event.rs:256-280) compares
// event_time, then the origin (x, y, index), the target, prior_time and the
// lineage reference, times by magnitude.  As LSD radix keys, from the least
// to the most significant field; positive f64 sort like their bit patterns.
// DevEvent (speciation only) has target = origin, so its key skips the
// target.  Each record's total digit count is even, so the gated tie passes
// of a captured sort leave the result in the same buffer whether they run.
constexpr unsigned long long ABS_BITS = 0x7fffffffffffffffull;
template <class E> struct EventOrder;
template <> struct EventOrder<DevEvent> {
    static constexpr int N = 6; // ref, prior, index, y, x, time
    static constexpr int TIME = 5;
    __host__ __device__ static constexpr int digits(int f) { return f <= 1 || f == 5 ? 8 : 4; }
    __device__ static unsigned long long field(const DevEvent &e, int f) {
        switch (f) {
        case 0: return e.ref;
        case 1: return bits(e.prior_time) & ABS_BITS;
        case 2: return e.index;
        case 3: return e.y;
        case 4: return e.x;
        default: return bits(e.event_time) & ABS_BITS;
        }
    }
};
template <> struct EventOrder<DevPacked> {
    static constexpr int N = 9; // ref, prior, target index/y/x, origin index/y/x, time
    static constexpr int TIME = 8;
    __host__ __device__ static constexpr int digits(int f) { return f <= 1 || f == 8 ? 8 : 4; }
    __device__ static unsigned long long field(const DevPacked &e, int f) {
        switch (f) {
        case 0: return e.ref;
        case 1: return bits(e.prior_time) & ABS_BITS;
        case 2: return e.tindex;
        case 3: return e.ty;
        case 4: return e.tx;
        case 5: return e.oindex;
        case 6: return e.oy;
        case 7: return e.ox;
        default: return bits(e.event_time) & ABS_BITS;
        }
    }
};
static_assert((8 + 8 + 4 + 4 + 4 + 8) % 2 == 0 && (8 + 8 + 6 * 4 + 8) % 2 == 0,
              "the gated tie passes must not change the ping-pong buffer");
constexpr int SORT_FIELDS_MAX = 9;

// Word layout of SortBuffers::minmax
constexpr int MM_TIME = 0;   // [min, max] of the event-time keys
constexpr int MM_TIES = 2;   // number of adjacent equal event times after the time sort
constexpr int MM_FIELDS = 4; // [min, max] per field
constexpr int MM_TIE_MIXED = MM_FIELDS + 2 * SORT_FIELDS_MAX; // a run of equal times has several origins / targets
constexpr int MM_TIE_LONG = MM_TIE_MIXED + 1;                  // a run is longer than TIE_RUN_MAX
constexpr int MM_WORDS = MM_TIE_LONG + 1;
static_assert(MM_WORDS <= SORT_MINMAX_WORDS, "sort range words");

// Per-element sort kernels process KS_ITEMS elements per thread and combine
// their results per block, so that each block issues one atomic per result
// (a same-address atomic from every wave serialises at the L2).
constexpr int KS_THREADS = 256;
constexpr int KS_ITEMS = 16;
constexpr int KS_TILE = KS_THREADS * KS_ITEMS;
// time key of a reserved but unused event slot: +inf sorts after every event
constexpr unsigned long long GHOST_TIME_BITS = 0x7ff0000000000000ull;

// [min, max] over a block (lo > hi when the block saw no key): one atomic pair
__device__ __forceinline__ void block_minmax(unsigned long long lo, unsigned long long hi, unsigned long long *mm) {
    __shared__ unsigned long long s_lo[KS_THREADS / 64], s_hi[KS_THREADS / 64];
    for (int off = 32; off; off >>= 1) {
        unsigned long long l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    const unsigned wave = threadIdx.x / 64;
    if (lane_id() == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / 64; w++) {
            lo = s_lo[w] < lo ? s_lo[w] : lo;
            hi = s_hi[w] > hi ? s_hi[w] : hi;
        }
        if (lo <= hi) {
            atomicMin(&mm[0], lo);
            atomicMax(&mm[1], hi);
        }
    }
    __syncthreads();
}

// sum over a block, added to *out with one atomic
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long *out) {
    __shared__ unsigned long long s_v[KS_THREADS / 64];
    for (int off = 32; off; off >>= 1) v += __shfl_xor(v, off);
    if (lane_id() == 0) s_v[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / 64; w++) v += s_v[w];
        if (v) atomicAdd(out, v);
    }
    __syncthreads();
}

__global__ void k_sort_init(unsigned long long *mm) {
    int t = threadIdx.x;
    // minima start at ~0, maxima, counts and flags at 0
    const bool field_min = t >= MM_FIELDS && t < MM_FIELDS + 2 * SORT_FIELDS_MAX && ((t - MM_FIELDS) & 1) == 0;
    if (t < MM_WORDS) mm[t] = (t == MM_TIME || field_min) ? ~0ull : 0ull;
}

// time keys, identity values and the time key range
template <class E>
__global__ void __launch_bounds__(KS_THREADS) k_event_keys(const E *ev0, SortSize sz, unsigned long long *keys,
                                                           uint32_t *vals, unsigned long long *mm) {
    const uint64_t n = sz.get();
    const uint64_t base = (uint64_t)blockIdx.x * KS_TILE;
    if (base >= n) return;
    const E *ev = ev0 + sz.offset();
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int j = 0; j < KS_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * KS_THREADS + threadIdx.x;
        if (p < n) {
            const unsigned long long k = EventOrder<E>::field(ev[p], EventOrder<E>::TIME);
            keys[p] = k;
            vals[p] = (uint32_t)p;
            if (k != GHOST_TIME_BITS) { // unused slots sort last, outside the range
                lo = k < lo ? k : lo;
                hi = k > hi ? k : hi;
            }
        }
    }
    block_minmax(lo, hi, mm + MM_TIME);
}

// number of adjacent equal keys in the sorted time keys (unused slots excluded)
__global__ void __launch_bounds__(KS_THREADS) k_tie_count(const unsigned long long *keys, SortSize sz,
                                                          unsigned long long *count) {
    const uint64_t n = sz.get();
    const uint64_t base = (uint64_t)blockIdx.x * KS_TILE;
    if (base >= n) return;
    unsigned long long ties = 0;
    for (int j = 0; j < KS_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * KS_THREADS + threadIdx.x;
        if (p + 1 < n) {
            const unsigned long long k = keys[p];
            ties += (k == keys[p + 1] && k != GHOST_TIME_BITS) ? 1 : 0;
        }
    }
    block_add(ties, count);
}

template <class E>
__global__ void __launch_bounds__(KS_THREADS) k_field_minmax(const E *ev0, SortSize sz, unsigned long long *mm) {
    using O = EventOrder<E>;
    const uint64_t n = sz.get();
    const uint64_t base = (uint64_t)blockIdx.x * KS_TILE;
    if (base >= n) return;
    const E *ev = ev0 + sz.offset();
    unsigned long long lo[O::N], hi[O::N];
    for (int f = 0; f < O::N; f++) {
        lo[f] = ~0ull;
        hi[f] = 0ull;
    }
    for (int j = 0; j < KS_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * KS_THREADS + threadIdx.x;
        if (p < n) {
            const E e = ev[p];
            for (int f = 0; f < O::N; f++) {
                const unsigned long long k = O::field(e, f);
                lo[f] = k < lo[f] ? k : lo[f];
                hi[f] = k > hi[f] ? k : hi[f];
            }
        }
    }
    for (int f = 0; f < O::N; f++) block_minmax(lo[f], hi[f], mm + MM_FIELDS + 2 * f);
}

// ---------------------------------------------------------------------------
// Runs of equal event times after the time sort.  Events of one run come
// from lineages that share a location and a Poisson event (coalesced
// duplicates), so they share origin and target, and the PackedEvent order
// within the run reduces to (prior time, reference).  Runs of up to
// TIE_RUN_MAX events are sorted in LDS (k_tie_runs): the workgroup of window
// w sorts the runs that start in positions [w * TIE_WINDOW, (w + 1) *
// TIE_WINDOW), which with their tails span fewer than TIE_WINDOW +
// TIE_RUN_MAX positions, reading each event's record once and checking that
// assumption as it goes.  Longer runs are ordered by radix passes over the
// reference, prior and time fields only;
// runs with several origins or targets by passes over every field.
constexpr uint64_t TIE_RUN_MAX = 1024;
// A window's span is its positions plus the tail of the run that crosses its
// end; 896 positions keep spans with tails up to 128 in a 1024-element
// network (1024 positions made nearly every span need 2048).
constexpr uint64_t TIE_WINDOW = 896;
constexpr int TIE_SORT_N = 2048; // >= TIE_WINDOW + TIE_RUN_MAX - 1, a power of two
constexpr int TIE_THREADS = 512;
// a window element is (run << 11 | span position, prior, reference), 20 B:
// the span's run index orders like its time (the span is time-sorted)
constexpr int TIE_POS_BITS = 11;
constexpr uint32_t TIE_POS_MASK = (1u << TIE_POS_BITS) - 1;
static_assert(TIE_SORT_N <= (1 << TIE_POS_BITS), "span positions and run indices in 11 bits");

// An event's class for the device EventCounterReporter: bit 0 prior time
// negative, bit 1 event time negative, bit 2 origin == target.
__device__ __forceinline__ uint8_t event_class(const DevPacked &e) {
    const bool self = e.ox == e.tx && e.oy == e.ty && e.oindex == e.tindex;
    return (uint8_t)(((bits(e.prior_time) >> 63) ? 1 : 0) | ((bits(e.event_time) >> 63) ? 2 : 0) | (self ? 4 : 0));
}
__device__ __forceinline__ uint8_t event_class(const DevEvent &) { return 0; }

// The counter an event that is not a duplicate adds to (event_counter.rs:
// 58-109 and the LineageInteraction of the signs, event.rs:186-229)
__device__ __forceinline__ int first_event_counter(bool speciation, bool prior_pos, bool event_pos, bool self) {
    if (speciation) return EC_SPECIATION;
    if (prior_pos && !event_pos) return EC_LATE_DISPERSAL_COALESCENCE; // Maybe
    const bool coal = prior_pos && event_pos;
    return self ? (coal ? EC_SELF_COALESCENCE : EC_SELF_DISPERSAL) : (coal ? EC_OUT_COALESCENCE : EC_OUT_DISPERSAL);
}
__device__ __forceinline__ int first_event_counter(uint8_t cls) {
    const bool prior_neg = cls & 1, event_neg = cls & 2;
    return first_event_counter(prior_neg && !event_neg, !prior_neg, !event_neg, (cls & 4) != 0);
}

// Flags runs whose members differ in origin or target (flags[0]) and runs
// longer than TIE_RUN_MAX (flags[1], checked at each run's head): the
// check of a round whose runs k_tie_runs could not take
template <class E>
__global__ void __launch_bounds__(KS_THREADS) k_tie_check(const E *ev, const unsigned long long *keys,
                                                          const uint32_t *vals, uint64_t n,
                                                          unsigned long long *flags) {
    using O = EventOrder<E>;
    constexpr int NF = O::N - 3; // the origin / target fields 2 .. N - 2 (32-bit values)
    // each event's origin / target fields, for the thread that compares the
    // next position with it (a ring of three iterations: a slot is rewritten
    // only after every thread has passed the barrier that follows its reads)
    __shared__ uint32_t s_f[3][KS_THREADS][NF];
    const uint64_t base = (uint64_t)blockIdx.x * KS_TILE;
    if (base >= n) return;
    unsigned long long mixed = 0, longer = 0;
    for (int j = 0; j < KS_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * KS_THREADS + threadIdx.x;
        const bool have = p < n;
        uint32_t f[NF];
        if (have) {
            const E e = ev[vals[p]];
#pragma unroll
            for (int k = 0; k < NF; k++) f[k] = s_f[j % 3][threadIdx.x][k] = (uint32_t)O::field(e, 2 + k);
        }
        __syncthreads();
        if (!have) continue;
        const unsigned long long k = keys[p];
        const bool ghost = k == GHOST_TIME_BITS; // unused event slots: runs of one
        const bool head = ghost || p == 0 || keys[p - 1] != k;
        if (!head) {
            // position p - 1: the previous thread, the last thread of the
            // previous iteration, or (the block's first position) memory
            uint32_t q[NF];
            if (threadIdx.x > 0 || j > 0) {
                const int qt = threadIdx.x > 0 ? (int)threadIdx.x - 1 : KS_THREADS - 1;
                const int qj = threadIdx.x > 0 ? j : j - 1;
#pragma unroll
                for (int kk = 0; kk < NF; kk++) q[kk] = s_f[qj % 3][qt][kk];
            } else {
                const E qe = ev[vals[p - 1]];
#pragma unroll
                for (int kk = 0; kk < NF; kk++) q[kk] = (uint32_t)O::field(qe, 2 + kk);
            }
#pragma unroll
            for (int kk = 0; kk < NF; kk++) mixed |= f[kk] != q[kk] ? 1ull : 0ull;
        } else if (!ghost && p + TIE_RUN_MAX < n && keys[p + TIE_RUN_MAX] == k) {
            longer = 1;
        }
    }
    block_add(mixed, flags);
    block_add(longer, flags + 1);
}

// first run heads in [lo, hi) and [hi, tail) (hi and tail if none), found
// by the whole block in one pass over the keys
__device__ __forceinline__ void block_first_heads(const unsigned long long *keys, uint64_t lo, uint64_t hi,
                                                  uint64_t tail, uint64_t &s, uint64_t &e) {
    __shared__ unsigned long long s_first[2];
    if (threadIdx.x == 0) {
        s_first[0] = hi;
        s_first[1] = tail;
    }
    __syncthreads();
    for (uint64_t p = lo + threadIdx.x; p < tail; p += blockDim.x)
        if (p == 0 || keys[p] != keys[p - 1] || keys[p] == GHOST_TIME_BITS)
            atomicMin(&s_first[p < hi ? 0 : 1], (unsigned long long)p);
    __syncthreads();
    s = s_first[0];
    e = s_first[1];
    __syncthreads();
}

// first run head in [lo, hi) (hi if none), found by the whole block
__device__ __forceinline__ uint64_t block_first_head(const unsigned long long *keys, uint64_t lo, uint64_t hi) {
    __shared__ unsigned long long s_first;
    if (threadIdx.x == 0) s_first = hi;
    __syncthreads();
    for (uint64_t p = lo + threadIdx.x; p < hi; p += blockDim.x)
        if (p == 0 || keys[p] != keys[p - 1] || keys[p] == GHOST_TIME_BITS)
            atomicMin(&s_first, (unsigned long long)p);
    __syncthreads();
    const uint64_t r = s_first;
    __syncthreads();
    return r;
}

// Sorts the runs of window w (positions [w * TIE_WINDOW, (w + 1) *
// TIE_WINDOW) and the tail of the run crossing its end) by (prior,
// reference) and writes their permutation.  Each position's record is read
// once, through the time sort's permutation: prior, reference, counter class
// and the origin / target fields, which must equal those of the previous
// position of the run.  A window whose runs break an assumption (several
// origins or targets: flags[0]; a run longer than TIE_RUN_MAX: flags[1])
// writes nothing and the round is sorted by field passes instead.  A span
// without equal times keeps its order; unused event slots (ghost keys) sort
// after every event and are runs of one.
// With `c` (a device_event_counter round, REF = false) the window counts its
// events instead of writing the permutation: a run's first event adds to its
// class's counter, every later one is a duplicate, and the run adds one late
// coalescence per distinct prior time after the first (event_counter.rs:
// 58-109 over the run in PackedEvent order, where equal prior times are
// adjacent).  Events [0, n_real) are real.  A run whose events differ in
// class is flagged like one with several origins (flags[0]), and the round
// is sorted by field passes and counted through that permutation.
template <class E, bool REF = true>
__global__ void __launch_bounds__(TIE_THREADS) k_tie_runs(const E *ev, const unsigned long long *keys,
                                                          const uint32_t *vals, uint64_t n, uint32_t *out_vals,
                                                          unsigned long long n_real, EventCounts *c,
                                                          unsigned long long *flags) {
    using O = EventOrder<E>;
    constexpr int NF = O::N - 3; // the origin / target fields 2 .. N - 2 (32-bit values)
    constexpr int NWV = TIE_THREADS / 64;
    __shared__ unsigned long long s_p[TIE_SORT_N], s_r[REF ? TIE_SORT_N : 1];
    __shared__ uint32_t s_m[TIE_SORT_N];     // run << 11 | span position
    __shared__ uint8_t s_cls[TIE_SORT_N];    // counter class, by span position
    __shared__ uint32_t s_hpos[REF ? 1 : TIE_SORT_N]; // !REF: span position of each run's first event
    __shared__ unsigned s_nodd;
    __shared__ unsigned long long s_carry;
    __shared__ uint32_t s_edge[NWV + 1][NF]; // origin / target of each wave's last lane; [NWV]: the chunk's last
    __shared__ uint32_t s_wsum[NWV];
    __shared__ unsigned long long s_cnt[EC_COUNTERS];
    __shared__ unsigned s_bad; // bit 0: several origins / targets in a run; bit 1: a long run
    const uint64_t lo = (uint64_t)blockIdx.x * TIE_WINDOW;
    if (lo >= n) return;
    const uint64_t hi = lo + TIE_WINDOW < n ? lo + TIE_WINDOW : n;
    const uint64_t tail_hi = hi + TIE_RUN_MAX < n ? hi + TIE_RUN_MAX : n;
    uint64_t s, e;
    block_first_heads(keys, lo, hi, tail_hi, s, e);
    if (s >= hi) return; // every position of the window continues an earlier run
    if (e == tail_hi && tail_hi < n && keys[tail_hi] == keys[tail_hi - 1] && keys[tail_hi] != GHOST_TIME_BITS) {
        // the run crossing the window's end goes on past its tail
        if (threadIdx.x == 0) atomicAdd(&flags[1], 1ull);
        return;
    }
    const int len = (int)(e - s);
    int nsort = 64; // the network spans the next power of two >= len (padded with ~0 keys)
    while (nsort < len) nsort <<= 1;
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    if (threadIdx.x == 0) s_bad = 0;
    if (c && threadIdx.x < EC_COUNTERS) s_cnt[threadIdx.x] = 0;
    // the thread's positions i = threadIdx.x + TIE_THREADS * k: every index,
    // then every record, in flight at once (positions past the span read
    // its last one and are ignored)
    constexpr int CH = TIE_SORT_N / TIE_THREADS;
    uint32_t vi[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = (int)threadIdx.x + TIE_THREADS * k;
        vi[k] = vals[s + (uint64_t)(i < len ? i : len - 1)];
    }
    unsigned long long pr[CH], rf[CH];
    uint32_t fl[CH][NF];
    uint8_t cl[CH];
    bool hd[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = (int)threadIdx.x + TIE_THREADS * k;
        const uint64_t p = s + (uint64_t)(i < len ? i : len - 1);
        const E rec = ev[vi[k]];
        pr[k] = O::field(rec, 1);
        rf[k] = REF ? O::field(rec, 0) : 0ull;
        cl[k] = event_class(rec);
#pragma unroll
        for (int t = 0; t < NF; t++) fl[k][t] = (uint32_t)O::field(rec, 2 + t);
        const unsigned long long key = keys[p], before = keys[p > 0 ? p - 1 : 0];
        hd[k] = i == 0 || key == GHOST_TIME_BITS || before != key;
    }
    int tie = 0;
    uint32_t run_base = 0; // runs that started in earlier chunks
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int c0 = TIE_THREADS * k;
        if (c0 >= len) break;
        const int i = c0 + (int)threadIdx.x;
        const bool have = i < len;
        uint32_t f[NF];
#pragma unroll
        for (int t = 0; t < NF; t++) f[t] = have ? fl[k][t] : 0u;
        const bool head = !have || hd[k];
        // the previous position's origin / target: the lane below, the last
        // lane of the wave below, or the previous chunk's last position
        uint32_t below[NF];
#pragma unroll
        for (int t = 0; t < NF; t++) below[t] = (uint32_t)__shfl_up((int)f[t], 1);
        if (lane == 63) {
#pragma unroll
            for (int t = 0; t < NF; t++) s_edge[wave][t] = f[t];
        }
        const uint64_t hb = __ballot(have && head);
        if (lane == 0) s_wsum[wave] = (uint32_t)__popcll(hb);
        __syncthreads();
        if (have && !head) { // the previous position is in this run: same origin and target
            const uint32_t *edge = wave > 0 ? s_edge[wave - 1] : s_edge[NWV];
            bool mixed = false;
#pragma unroll
            for (int t = 0; t < NF; t++) mixed |= f[t] != (lane > 0 ? below[t] : edge[t]);
            if (mixed) atomicOr(&s_bad, 1u);
            tie = 1;
        }
        // the position's run: heads of the span up to it, minus one
        const uint32_t incl = (uint32_t)__popcll(hb & (lanemask_lt() | (1ull << lane)));
        uint32_t before = run_base, chunk = 0;
        for (unsigned w = 0; w < NWV; w++) {
            before += w < wave ? s_wsum[w] : 0u;
            chunk += s_wsum[w];
        }
        if (have) {
            s_p[i] = pr[k];
            if (REF) s_r[i] = rf[k];
            s_m[i] = ((before + incl - 1) << TIE_POS_BITS) | (uint32_t)i;
            s_cls[i] = cl[k];
            if (!REF && head) s_hpos[before + incl - 1] = (uint32_t)i;
        }
        run_base += chunk;
        __syncthreads(); // s_edge and s_wsum are rewritten by the next chunk
        if (threadIdx.x == TIE_THREADS - 1) {
#pragma unroll
            for (int t = 0; t < NF; t++) s_edge[NWV][t] = f[t];
        }
    }
    if constexpr (!REF) {
        // A counted round: the counters need per run only its length, its
        // class and the number of distinct prior times (the later events of
        // a run are duplicates, each a late coalescence where its prior time
        // differs from the one before it in sorted order: one per distinct
        // prior time after the first).  Events whose prior time differs from
        // their run's first event are few; they alone are sorted, by (run,
        // prior time), and their distinct pairs counted.  A run whose events
        // differ in class goes, with the round, to the field passes.
        __syncthreads();
        if (s_bad) {
            if (threadIdx.x == 0) atomicAdd(&flags[(s_bad & 1u) ? 0 : 1], 1ull);
            return;
        }
        const int real = n_real > s ? (int)(n_real - s < (uint64_t)len ? n_real - s : (uint64_t)len) : 0;
        const int last_run = (s + (uint64_t)real == n_real && real > 0) ? (int)(s_m[real - 1] >> TIE_POS_BITS) : -1;
        if (threadIdx.x == 0) {
            s_nodd = 0;
            s_carry = 0;
        }
        __syncthreads();
        unsigned long long cnt[EC_COUNTERS];
#pragma unroll
        for (int k = 0; k < EC_COUNTERS; k++) cnt[k] = 0;
        bool mixed = false;
        unsigned long long op[CH];
        uint32_t orun[CH];
        bool odd[CH];
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int i = (int)threadIdx.x + TIE_THREADS * k;
            odd[k] = false;
            if (i >= real) continue;
            const uint32_t run = s_m[i] >> TIE_POS_BITS, h = s_hpos[run];
            const unsigned long long pi = s_p[i];
            cnt[EC_RAW_TOTAL]++;
            if (h == (uint32_t)i) {
                const int w = first_event_counter(s_cls[i]);
#pragma unroll
                for (int q = EC_SPECIATION; q < EC_COUNTERS; q++) cnt[q] += q == w ? 1 : 0;
            } else {
                mixed |= s_cls[i] != s_cls[h];
                odd[k] = pi != s_p[h];
            }
            op[k] = pi;
            orun[k] = run;
            if ((int)run == last_run) atomicMax(&s_carry, pi);
        }
        if (__syncthreads_or(mixed ? 1 : 0)) {
            if (threadIdx.x == 0) atomicAdd(&flags[0], 1ull);
            return;
        }
        // every read of the span is done: the odd events go to the front
#pragma unroll
        for (int k = 0; k < CH; k++) {
            if (!odd[k]) continue;
            const unsigned q = atomicAdd(&s_nodd, 1u);
            s_p[q] = op[k];
            s_m[q] = (orun[k] << TIE_POS_BITS) | q;
        }
        __syncthreads();
        const int nodd = (int)s_nodd;
        if (nodd > 1) {
            int nso = 2;
            while (nso < nodd) nso <<= 1;
            for (int i = nodd + (int)threadIdx.x; i < nso; i += TIE_THREADS) {
                s_p[i] = ~0ull;
                s_m[i] = ~0u;
            }
            __syncthreads();
            for (int k = 2; k <= nso; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int q = threadIdx.x; q < nso / 2; q += TIE_THREADS) {
                        const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                        const int l = i + j;
                        const bool ascending = (i & k) == 0;
                        const uint32_t mi = s_m[i], ml = s_m[l];
                        const unsigned long long pi = s_p[i], pl = s_p[l];
                        const uint32_t ui = mi >> TIE_POS_BITS, ul = ml >> TIE_POS_BITS;
                        const bool l_less = ul < ui || (ul == ui && (pl < pi || (pl == pi && ml < mi)));
                        if (l_less == ascending) {
                            s_p[i] = pl; s_p[l] = pi;
                            s_m[i] = ml; s_m[l] = mi;
                        }
                    }
                    __syncthreads();
                }
            }
        }
        for (int q = threadIdx.x; q < nodd; q += TIE_THREADS)
            if (q == 0 || (s_m[q] >> TIE_POS_BITS) != (s_m[q - 1] >> TIE_POS_BITS) || s_p[q] != s_p[q - 1])
                cnt[EC_LATE_COALESCENCE]++;
#pragma unroll
        for (int k = 0; k < EC_COUNTERS; k++) {
            unsigned long long v = cnt[k];
            for (int off = 32; off; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0 && v) atomicAdd(&s_cnt[k], v);
        }
        __syncthreads();
        if (threadIdx.x == 0 && last_run >= 0) c->round_carry_prior = s_carry;
        if (threadIdx.x < EC_COUNTERS && s_cnt[threadIdx.x]) atomicAdd(&c->round_count[threadIdx.x], s_cnt[threadIdx.x]);
        return;
    }
    for (int i = len + (int)threadIdx.x; i < nsort; i += TIE_THREADS) {
        s_p[i] = ~0ull;
        if (REF) s_r[i] = ~0ull;
        s_m[i] = ~0u;
    }
    tie = __syncthreads_or(tie);
    if (s_bad) {
        if (threadIdx.x == 0) atomicAdd(&flags[(s_bad & 1u) ? 0 : 1], 1ull);
        return;
    }
    if (tie) {
        // bitonic network: in stage (k, j), pair q compares i and i + j
        for (int k = 2; k <= nsort; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int q = threadIdx.x; q < nsort / 2; q += TIE_THREADS) {
                    const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                    const int l = i + j;
                    const bool ascending = (i & k) == 0;
                    const uint32_t mi = s_m[i], ml = s_m[l];
                    const unsigned long long pi = s_p[i], pl = s_p[l];
                    const unsigned long long ri = REF ? s_r[i] : 0ull, rl = REF ? s_r[l] : 0ull;
                    const uint32_t ui = mi >> TIE_POS_BITS, ul = ml >> TIE_POS_BITS;
                    const bool l_less =
                        ul < ui || (ul == ui && (pl < pi || (pl == pi && (rl < ri || (rl == ri && ml < mi)))));
                    if (l_less == ascending) {
                        s_p[i] = pl; s_p[l] = pi;
                        if (REF) {
                            s_r[i] = rl;
                            s_r[l] = ri;
                        }
                        s_m[i] = ml; s_m[l] = mi;
                    }
                }
                __syncthreads();
            }
        }
    }
    for (int i = threadIdx.x; i < len; i += TIE_THREADS) out_vals[s + i] = vals[s + (s_m[i] & TIE_POS_MASK)];
}

template <class E>
__global__ void k_field_keys(const E *ev0, const uint32_t *vals, unsigned long long *keys, int f, SortSize sz) {
    if (sz.skip()) return;
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const E *ev = ev0 + sz.offset();
    if (p < sz.get()) keys[p] = EventOrder<E>::field(ev[vals[p]], f);
}

// Digits are those of min(key, top) - base: a key range [base, top) sorted
// in fewer passes, keys at or above top (unused slots) after it; base 0 and
// top ~0 sort the keys themselves.
__device__ __forceinline__ uint32_t radix_digit(unsigned long long key, unsigned long long base,
                                                unsigned long long top, int shift) {
    return (uint32_t)(((key < top ? key : top) - base) >> shift) & 255u;
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const unsigned long long *keys, SortSize sz,
                                                           int shift, unsigned long long key_base,
                                                           unsigned long long top, uint32_t *hist) {
    __shared__ uint32_t h[256];
    if (sz.skip()) return;
    const uint64_t n = sz.get();
    const uint32_t nb = sz.tiles();
    if (blockIdx.x >= nb) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; j++) {
        uint64_t p = base + (uint64_t)j * RS_THREADS + threadIdx.x;
        if (p < n) atomicAdd(&h[radix_digit(keys[p], key_base, top, shift)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const unsigned long long *kin,
                                                              const uint32_t *vin,
                                                              unsigned long long *kout, uint32_t *vout,
                                                              SortSize sz, int shift, unsigned long long key_base,
                                                              unsigned long long top, const uint32_t *offsets) {
    constexpr int W = RS_THREADS / 64;
    if (sz.skip()) return;
    const uint64_t n = sz.get();
    const uint32_t nb = sz.tiles();
    if (blockIdx.x >= nb) return;
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[W][256];
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x / 64;
    running[threadIdx.x] = offsets[(uint64_t)threadIdx.x * nb + blockIdx.x];
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; j++) {
        for (int w = 0; w < W; w++) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        uint64_t p = base + (uint64_t)j * RS_THREADS + threadIdx.x;
        bool valid = p < n;
        unsigned long long key = valid ? kin[p] : 0;
        uint32_t val = valid ? vin[p] : 0;
        uint32_t digit = radix_digit(key, key_base, top, shift);
        uint64_t peers = __ballot(valid);
        for (int bit = 0; bit < 8; bit++) {
            bool set = (digit >> bit) & 1u;
            uint64_t mset = __ballot(valid && set);
            peers &= set ? mset : ~mset;
        }
        uint32_t rank = __popcll(peers & lanemask_lt());
        if (valid && rank == 0) wcnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[digit] + rank;
            for (unsigned w = 0; w < wave; w++) pos += wcnt[w][digit];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < W; w++) add += wcnt[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Staged LSD passes for eager sorts (the time keys of a round).  A tile of
// SR_TILE keys is ranked by digit in LDS (stable: wave-major, then round,
// then lane, which is the tile's own order) and written out in digit order,
// so each digit's run of the tile (16 keys on average) goes to consecutive
// addresses.  Measured on 25.1M-event rounds (radix time per round, the
// digit width and tile chosen from): 8-bit digits in tiles of 4096 keys
// (two workgroups per CU) 1.48 ms; 9-bit 1.52; 10-bit 1.59; 10-bit in tiles
// of 8192 (one workgroup per CU) 2.01; 11-bit 2.31; the per-pass 8-bit sort
// without staging 2.60.
constexpr int SR_BITS = 8;
constexpr int SR_BINS = 1 << SR_BITS;
constexpr int SR_THREADS = 512;
constexpr int SR_WAVES = SR_THREADS / 64;
constexpr int SR_ITEMS = 8;
constexpr int SR_TILE = SR_THREADS * SR_ITEMS;
constexpr int SR_DPT = SR_BINS >= SR_THREADS ? SR_BINS / SR_THREADS : 1; // digits per thread in the tile scan
static_assert(64 * SR_ITEMS < (1 << 16) && SR_BINS <= (1 << 16), "ranks and digits packed in 16 bits each");
static_assert(SR_BINS % SR_THREADS == 0 || SR_THREADS % SR_BINS == 0, "digits spread evenly over threads");

__device__ __forceinline__ uint32_t sradix_digit(unsigned long long key, unsigned long long base,
                                                 unsigned long long top, int shift) {
    return (uint32_t)(((key < top ? key : top) - base) >> shift) & (uint32_t)(SR_BINS - 1);
}

// lanes of the wave whose digit equals this lane's (valid lanes only)
__device__ __forceinline__ uint64_t sradix_peers(uint32_t digit, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < SR_BITS; bit++) {
        const bool set = (digit >> bit) & 1u;
        const uint64_t mset = __ballot(valid && set);
        peers &= set ? mset : ~mset;
    }
    return peers;
}

__device__ __forceinline__ uint32_t sradix_tiles(uint64_t n) { return (uint32_t)((n + SR_TILE - 1) / SR_TILE); }

// Workgroups go to the 8 XCDs round-robin, each XCD with its own L2.  Block
// b takes tile (b % 8) * per + b / 8, so one XCD works through consecutive
// tiles and the digit-major histogram words of neighbouring tiles (adjacent
// in memory) meet in one L2.  The grid has 8 * per blocks; those past the
// last tile exit.
constexpr uint32_t XCDS = 8;
__host__ __device__ __forceinline__ uint32_t xcd_grid(uint32_t tiles) { return XCDS * ((tiles + XCDS - 1) / XCDS); }
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t tiles) {
    return (b % XCDS) * ((tiles + XCDS - 1) / XCDS) + b / XCDS;
}

__global__ void __launch_bounds__(SR_THREADS) k_sradix_hist(const unsigned long long *keys, SortSize sz, int shift,
                                                            unsigned long long key_base, unsigned long long top,
                                                            uint32_t *hist) {
    __shared__ uint32_t h[SR_BINS];
    if (sz.skip()) return;
    const uint64_t n = sz.get();
    const uint32_t nb = sradix_tiles(n);
    const uint32_t tile = xcd_tile(blockIdx.x, nb);
    if (tile >= nb) return;
    for (int d = threadIdx.x; d < SR_BINS; d += SR_THREADS) h[d] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)tile * SR_TILE;
    // every load in flight before the first is used: positions past the end
    // read the last key (unconditional loads are not waited for one by one)
    unsigned long long key[SR_ITEMS];
#pragma unroll
    for (int j = 0; j < SR_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * SR_THREADS + threadIdx.x;
        key[j] = keys[p < n ? p : n - 1];
    }
#pragma unroll
    for (int j = 0; j < SR_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * SR_THREADS + threadIdx.x;
        const bool valid = p < n;
        const uint32_t digit = valid ? sradix_digit(key[j], key_base, top, shift) : 0u;
        // one LDS atomic per distinct digit of the wave (sorted keys repeat digits)
        const uint64_t peers = sradix_peers(digit, valid);
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[digit], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < SR_BINS; d += SR_THREADS) hist[(uint64_t)d * nb + tile] = h[d];
}

// `offsets`: the exclusive scan of the digit-major histogram of k_sradix_hist
__global__ void __launch_bounds__(SR_THREADS) k_sradix_scatter(const unsigned long long *kin, const uint32_t *vin,
                                                               unsigned long long *kout, uint32_t *vout,
                                                               SortSize sz, int shift, unsigned long long key_base,
                                                               unsigned long long top, const uint32_t *offsets) {
    __shared__ unsigned long long s_key[SR_TILE];
    __shared__ uint32_t s_val[SR_TILE];
    __shared__ uint16_t s_wc[SR_WAVES][SR_BINS]; // per wave: digit counts, then the counts of the waves before
    __shared__ uint32_t s_start[SR_BINS];        // the digit's first position in the tile
    __shared__ uint32_t s_dst[SR_BINS];          // global position of the tile's position 0 in the digit's run
    __shared__ uint32_t s_wsum[SR_WAVES];
    if (sz.skip()) return;
    const uint64_t n = sz.get();
    const uint32_t nb = sradix_tiles(n);
    const uint32_t tile = xcd_tile(blockIdx.x, nb);
    if (tile >= nb) return;
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    // the global offsets of the thread's digits, read while the tile is ranked
    const bool has_digits = (int)threadIdx.x * SR_DPT < SR_BINS;
    uint32_t goff[SR_DPT];
#pragma unroll
    for (int k = 0; k < SR_DPT; k++)
        goff[k] = has_digits ? offsets[(uint64_t)(SR_DPT * threadIdx.x + k) * nb + tile] : 0u;
    for (int i = threadIdx.x; i < SR_WAVES * SR_BINS; i += SR_THREADS) (&s_wc[0][0])[i] = 0;
    __syncthreads();
    const uint64_t tbase = (uint64_t)tile * SR_TILE;
    const uint64_t wbase = tbase + (uint64_t)wave * (64 * SR_ITEMS);
    unsigned long long key[SR_ITEMS];
    uint32_t val[SR_ITEMS], dr[SR_ITEMS];
#pragma unroll
    for (int j = 0; j < SR_ITEMS; j++) {
        const uint64_t p = wbase + (uint64_t)j * 64 + lane;
        key[j] = p < n ? kin[p] : 0ull;
        val[j] = p < n ? vin[p] : 0u;
    }
    // rank of each key among the wave's keys of its digit, in wave order
#pragma unroll
    for (int j = 0; j < SR_ITEMS; j++) {
        const bool valid = wbase + (uint64_t)j * 64 + lane < n;
        const uint32_t digit = valid ? sradix_digit(key[j], key_base, top, shift) : 0u;
        const uint64_t peers = sradix_peers(digit, valid);
        uint32_t r = 0;
        if (valid) {
            const uint32_t c = s_wc[wave][digit];
            r = c + (uint32_t)__popcll(peers & lanemask_lt());
            if ((peers & lanemask_lt()) == 0) s_wc[wave][digit] = (uint16_t)(c + (uint32_t)__popcll(peers));
        }
        dr[j] = digit | (r << 16);
    }
    __syncthreads();
    // digits DPT t .. DPT t + DPT - 1: counts of the waves before each wave,
    // tile totals, and their exclusive scan over the tile's digits
    uint32_t tot[SR_DPT], sum = 0;
#pragma unroll
    for (int k = 0; k < SR_DPT; k++) {
        const int d = SR_DPT * threadIdx.x + k;
        uint32_t run = 0;
        if (has_digits) {
            for (int w = 0; w < SR_WAVES; w++) {
                const uint32_t c = s_wc[w][d];
                s_wc[w][d] = (uint16_t)run;
                run += c;
            }
        }
        tot[k] = run;
        sum += run;
    }
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= (unsigned)off) incl += y;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (unsigned w = 0; w < wave; w++) run += s_wsum[w];
#pragma unroll
    for (int k = 0; k < SR_DPT; k++) {
        const int d = SR_DPT * threadIdx.x + k;
        if (has_digits) {
            s_start[d] = run;
            s_dst[d] = goff[k] - run;
        }
        run += tot[k];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SR_ITEMS; j++) {
        if (wbase + (uint64_t)j * 64 + lane < n) {
            const uint32_t d = dr[j] & 0xffffu;
            const uint32_t pos = s_start[d] + s_wc[wave][d] + (dr[j] >> 16);
            s_key[pos] = key[j];
            s_val[pos] = val[j];
        }
    }
    __syncthreads();
    const uint32_t cnt = (uint32_t)(n - tbase < (uint64_t)SR_TILE ? n - tbase : (uint64_t)SR_TILE);
    for (uint32_t i = threadIdx.x; i < cnt; i += SR_THREADS) {
        const unsigned long long k = s_key[i];
        const uint32_t dst = s_dst[sradix_digit(k, key_base, top, shift)] + i;
        kout[dst] = k;
        vout[dst] = s_val[i];
    }
}

template <class E> __global__ void k_gather_events(const E *ev0, const uint32_t *vals, E *out, SortSize sz) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const E *ev = ev0 + sz.offset();
    if (p < sz.get()) out[p] = ev[vals[p]];
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_reset_counters(Counters *c, bool reset_events) {
    c->head = 0;
    c->n_out = 0;
    c->n_dead = 0;
    c->compact_done = 0;
    if (reset_events) {
        c->n_events = 0;
        c->ev_overflow = 0;
        c->n_ghost = 0;
    }
}

// ---------------------------------------------------------------------------
// launchers
static inline unsigned grid_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

template <bool DEDUP, bool DISP, bool COOP>
static void launch_persistent_t(int kind, const PersistArgs &a, int blocks, int threads, hipStream_t stream) {
    switch (kind) {
    case K_NON_SPATIAL: hipLaunchKernelGGL((k_persistent<K_NON_SPATIAL, DEDUP, DISP, COOP>), dim3(blocks), dim3(threads), 0, stream, a); break;
    case K_ALMOST_INFINITE: hipLaunchKernelGGL((k_persistent<K_ALMOST_INFINITE, DEDUP, DISP, COOP>), dim3(blocks), dim3(threads), 0, stream, a); break;
    case K_SEPARABLE: hipLaunchKernelGGL((k_persistent<K_SEPARABLE, DEDUP, DISP, COOP>), dim3(blocks), dim3(threads), 0, stream, a); break;
    default: hipLaunchKernelGGL((k_persistent<K_PACKED, DEDUP, DISP, COOP>), dim3(blocks), dim3(threads), 0, stream, a); break;
    }
}

template <int POOL, bool DISP>
static void launch_pooled_t(int kind, const PersistArgs &a, int blocks, hipStream_t stream) {
    switch (kind) {
    case K_NON_SPATIAL: hipLaunchKernelGGL((k_pooled<K_NON_SPATIAL, POOL, DISP>), dim3(blocks), dim3(64), 0, stream, a); break;
    case K_ALMOST_INFINITE: hipLaunchKernelGGL((k_pooled<K_ALMOST_INFINITE, POOL, DISP>), dim3(blocks), dim3(64), 0, stream, a); break;
    case K_SEPARABLE: hipLaunchKernelGGL((k_pooled<K_SEPARABLE, POOL, DISP>), dim3(blocks), dim3(64), 0, stream, a); break;
    default: hipLaunchKernelGGL((k_pooled<K_PACKED, POOL, DISP>), dim3(blocks), dim3(64), 0, stream, a); break;
    }
}

template <int POOL, bool DISP> static int pooled_max_blocks_per_cu_t(int kind) {
    int n = 0;
    switch (kind) {
    case K_NON_SPATIAL: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pooled<K_NON_SPATIAL, POOL, DISP>, 64, 0); break;
    case K_ALMOST_INFINITE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pooled<K_ALMOST_INFINITE, POOL, DISP>, 64, 0); break;
    case K_SEPARABLE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pooled<K_SEPARABLE, POOL, DISP>, 64, 0); break;
    default: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pooled<K_PACKED, POOL, DISP>, 64, 0); break;
    }
    return n > 0 ? n : 1;
}

int pooled_max_blocks_per_cu(int kind, unsigned pool, bool report_dispersal) {
    if (report_dispersal)
        return pool == 192 ? pooled_max_blocks_per_cu_t<192, true>(kind) : pooled_max_blocks_per_cu_t<128, true>(kind);
    return pool == 192 ? pooled_max_blocks_per_cu_t<192, false>(kind) : pooled_max_blocks_per_cu_t<128, false>(kind);
}

// dispersal reporting has no cooperative variant (its rounds are dense); the
// pooled kernel runs dense rounds without dedup cache (one-wave blocks)
void launch_persistent(int kind, const PersistArgs &a, int blocks, int threads, hipStream_t stream) {
    if (a.pool && !a.cache && !a.coop) {
        if (a.packed) {
            if (a.pool == 192) launch_pooled_t<192, true>(kind, a, blocks, stream);
            else launch_pooled_t<128, true>(kind, a, blocks, stream);
        } else {
            if (a.pool == 192) launch_pooled_t<192, false>(kind, a, blocks, stream);
            else launch_pooled_t<128, false>(kind, a, blocks, stream);
        }
        return;
    }
    if (a.packed) {
        if (a.cache) launch_persistent_t<true, true, false>(kind, a, blocks, threads, stream);
        else launch_persistent_t<false, true, false>(kind, a, blocks, threads, stream);
    } else if (a.coop) {
        if (a.cache) launch_persistent_t<true, false, true>(kind, a, blocks, threads, stream);
        else launch_persistent_t<false, false, true>(kind, a, blocks, threads, stream);
    } else {
        if (a.cache) launch_persistent_t<true, false, false>(kind, a, blocks, threads, stream);
        else launch_persistent_t<false, false, false>(kind, a, blocks, threads, stream);
    }
}

template <bool DEDUP, bool DISP, bool COOP> static int persistent_max_blocks_per_cu_t(int kind, int threads) {
    int n = 0;
    switch (kind) {
    case K_NON_SPATIAL: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<K_NON_SPATIAL, DEDUP, DISP, COOP>, threads, 0); break;
    case K_ALMOST_INFINITE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<K_ALMOST_INFINITE, DEDUP, DISP, COOP>, threads, 0); break;
    case K_SEPARABLE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<K_SEPARABLE, DEDUP, DISP, COOP>, threads, 0); break;
    default: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<K_PACKED, DEDUP, DISP, COOP>, threads, 0); break;
    }
    return n > 0 ? n : 1;
}

// resident blocks per CU of the persistent kernel variant with / without
// dedup, dispersal reporting and the cooperative search
int persistent_max_blocks_per_cu(int kind, int threads, bool dedup, bool report_dispersal, bool coop) {
    if (report_dispersal)
        return dedup ? persistent_max_blocks_per_cu_t<true, true, false>(kind, threads)
                     : persistent_max_blocks_per_cu_t<false, true, false>(kind, threads);
    if (coop)
        return dedup ? persistent_max_blocks_per_cu_t<true, false, true>(kind, threads)
                     : persistent_max_blocks_per_cu_t<false, false, true>(kind, threads);
    return dedup ? persistent_max_blocks_per_cu_t<true, false, false>(kind, threads)
                 : persistent_max_blocks_per_cu_t<false, false, false>(kind, threads);
}

void launch_serve(int kind, const ServeArgs &a, int blocks, int threads, hipStream_t stream) {
    switch (kind) {
    case K_NON_SPATIAL: hipLaunchKernelGGL(k_serve<K_NON_SPATIAL>, dim3(blocks), dim3(threads), 0, stream, a); break;
    case K_ALMOST_INFINITE: hipLaunchKernelGGL(k_serve<K_ALMOST_INFINITE>, dim3(blocks), dim3(threads), 0, stream, a); break;
    case K_SEPARABLE: hipLaunchKernelGGL(k_serve<K_SEPARABLE>, dim3(blocks), dim3(threads), 0, stream, a); break;
    default: hipLaunchKernelGGL(k_serve<K_PACKED>, dim3(blocks), dim3(threads), 0, stream, a); break;
    }
}

int serve_max_blocks_per_cu(int kind, int threads) {
    int n = 0;
    switch (kind) {
    case K_NON_SPATIAL: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_serve<K_NON_SPATIAL>, threads, 0); break;
    case K_ALMOST_INFINITE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_serve<K_ALMOST_INFINITE>, threads, 0); break;
    case K_SEPARABLE: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_serve<K_SEPARABLE>, threads, 0); break;
    default: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_serve<K_PACKED>, threads, 0); break;
    }
    return n > 0 ? n : 1;
}

template <bool DISP>
static void launch_plain_t(int kind, const PlainArgs &a, unsigned g, int threads, hipStream_t stream) {
    switch (kind) {
    case K_NON_SPATIAL: hipLaunchKernelGGL((k_plain<K_NON_SPATIAL, DISP>), dim3(g), dim3(threads), 0, stream, a); break;
    case K_ALMOST_INFINITE: hipLaunchKernelGGL((k_plain<K_ALMOST_INFINITE, DISP>), dim3(g), dim3(threads), 0, stream, a); break;
    case K_SEPARABLE: hipLaunchKernelGGL((k_plain<K_SEPARABLE, DISP>), dim3(g), dim3(threads), 0, stream, a); break;
    default: hipLaunchKernelGGL((k_plain<K_PACKED, DISP>), dim3(g), dim3(threads), 0, stream, a); break;
    }
}

void launch_plain(int kind, const PlainArgs &a, int threads, hipStream_t stream) {
    unsigned g = grid_for(a.m, threads);
    if (!g) return;
    if (a.packed) launch_plain_t<true>(kind, a, g, threads, stream);
    else launch_plain_t<false>(kind, a, g, threads, stream);
}

void launch_dedup_push(const uint8_t *status, const unsigned long long *bucket, uint32_t *head,
                       uint32_t *link, uint64_t m, const unsigned long long *m_dev, hipStream_t stream) {
    if (!m) return;
    hipLaunchKernelGGL(k_dedup_push, dim3(grid_for(m, 256)), dim3(256), 0, stream, status, bucket, head, link, m,
                       m_dev);
}

void launch_dedup_resolve(uint8_t *status, const unsigned long long *bucket, const double *s_u,
                          const double *s_t, const uint32_t *s_x, const uint32_t *s_y,
                          const uint32_t *s_i, uint32_t *head, const uint32_t *link,
                          uint64_t *cache, uint64_t m, const unsigned long long *m_dev, hipStream_t stream) {
    if (!m) return;
    hipLaunchKernelGGL(k_dedup_resolve, dim3(grid_for(m, 256)), dim3(256), 0, stream, status, bucket, s_u,
                       s_t, s_x, s_y, s_i, head, link, cache, m, m_dev);
}

uint32_t compact_blocks(uint64_t m) { return (uint32_t)((m + CT_TILE - 1) / CT_TILE); }

void launch_compact(const uint8_t *status, const uint32_t *pass, uint64_t m, uint32_t *out_a,
                    uint32_t *out_b, uint64_t b_base, CompactBuffers &cb, hipStream_t stream) {
    uint32_t nb = compact_blocks(m);
    if (nb == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, (uint32_t *)cb.counts, 0u, (uint64_t)2);
        return;
    }
    hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(CT_THREADS), 0, stream, status, m,
                       (const unsigned long long *)nullptr, cb.blk_a, cb.blk_b);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, cb.blk_a, (uint64_t)nb);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, cb.blk_b, (uint64_t)nb);
    hipLaunchKernelGGL(k_compact_scatter, dim3(nb), dim3(CT_THREADS), 0, stream, status, pass, m,
                       (const PlainLoop *)nullptr, cb.blk_a, cb.blk_b, out_a, out_b, b_base);
    hipLaunchKernelGGL(k_compact_totals, dim3(1), dim3(1), 0, stream, cb.blk_a, cb.blk_b, nb, cb.counts);
}

// Blocks of the launch bound past the device pass size count zero, so the
// scans over the bound's blocks give the pass's totals.
void launch_compact_loop(const uint8_t *status, const uint32_t *pass, uint64_t m_bound, PlainLoop *loop,
                         uint32_t *out_a, uint32_t *out_b, CompactBuffers &cb, hipStream_t stream) {
    uint32_t nb = compact_blocks(m_bound);
    if (nb == 0) return;
    hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(CT_THREADS), 0, stream, status, m_bound, &loop->m, cb.blk_a,
                       cb.blk_b);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, cb.blk_a, (uint64_t)nb);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, cb.blk_b, (uint64_t)nb);
    hipLaunchKernelGGL(k_compact_scatter, dim3(nb), dim3(CT_THREADS), 0, stream, status, pass, m_bound,
                       (const PlainLoop *)loop, cb.blk_a, cb.blk_b, out_a, out_b, (uint64_t)0);
    hipLaunchKernelGGL(k_plain_loop_advance, dim3(1), dim3(1), 0, stream, cb.blk_a, cb.blk_b, nb, loop);
}

uint32_t sort_blocks(uint64_t n) { return (uint32_t)((n + RS_TILE - 1) / RS_TILE); }

// exclusive scan of the 256 x tiles digit histogram of a sort
// Exclusive scan of the digit-major histogram (256 x tiles counts) in three
// phases: per-chunk sums, a scan of the chunk sums, and each chunk's scan
// offset by its sum prefix.  The chunk sums live after the histogram.
constexpr int HS_THREADS = 256;
constexpr int HS_ITEMS = 16;
constexpr int HS_CHUNK = HS_THREADS * HS_ITEMS;

// digit-major histogram words of a sort: BINS digits x tiles of TILE keys
template <int BINS, int TILE> __device__ __forceinline__ uint64_t hist_words(const SortSize &sz) {
    return (uint64_t)BINS * ((sz.get() + TILE - 1) / TILE);
}

template <int BINS = 256, int TILE = RS_TILE>
__global__ void __launch_bounds__(HS_THREADS) k_hist_chunk_sums(const uint32_t *h, SortSize sz, uint32_t *sums) {
    if (sz.skip()) return;
    const uint64_t n = hist_words<BINS, TILE>(sz);
    const uint64_t base = (uint64_t)blockIdx.x * HS_CHUNK;
    if (base >= n) return;
    unsigned long long s = 0;
    for (int j = 0; j < HS_ITEMS; j++) {
        const uint64_t p = base + (uint64_t)j * HS_THREADS + threadIdx.x;
        if (p < n) s += h[p];
    }
    __shared__ unsigned long long total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    block_add(s, &total);
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = (uint32_t)total;
}

// one block: exclusive scan of the chunk sums in place
template <int BINS = 256, int TILE = RS_TILE>
__global__ void __launch_bounds__(1024) k_scan_chunk_sums(uint32_t *sums, SortSize sz) {
    __shared__ uint32_t partial[1024];
    if (sz.skip()) return;
    const uint64_t n = (hist_words<BINS, TILE>(sz) + HS_CHUNK - 1) / HS_CHUNK;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t lo = per * threadIdx.x, hi = lo + per < n ? lo + per : n;
    uint32_t sum = 0;
    for (uint64_t j = lo; j < hi; j++) sum += sums[j];
    partial[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? partial[threadIdx.x - off] : 0;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? partial[threadIdx.x - 1] : 0;
    for (uint64_t j = lo; j < hi; j++) {
        uint32_t v = sums[j];
        sums[j] = run;
        run += v;
    }
}

// each thread scans HS_ITEMS consecutive counts; the block scans the thread
// totals (wave shuffles, then the wave totals)
template <int BINS = 256, int TILE = RS_TILE>
__global__ void __launch_bounds__(HS_THREADS) k_hist_scan_apply(uint32_t *h, SortSize sz, const uint32_t *sums) {
    __shared__ uint32_t s_wave[HS_THREADS / 64];
    if (sz.skip()) return;
    const uint64_t n = hist_words<BINS, TILE>(sz);
    const uint64_t base = (uint64_t)blockIdx.x * HS_CHUNK;
    if (base >= n) return;
    const uint64_t first = base + (uint64_t)threadIdx.x * HS_ITEMS;
    uint32_t v[HS_ITEMS];
    uint32_t total = 0;
    for (int j = 0; j < HS_ITEMS; j++) {
        v[j] = first + j < n ? h[first + j] : 0u;
        total += v[j];
    }
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    uint32_t incl = total;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off);
        if (lane >= (unsigned)off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t run = sums[blockIdx.x] + incl - total;
    for (unsigned w = 0; w < wave; w++) run += s_wave[w];
    for (int j = 0; j < HS_ITEMS; j++) {
        if (first + j < n) h[first + j] = run;
        run += v[j];
    }
}

size_t sort_hist_words(uint64_t n) {
    uint64_t h = (uint64_t)256 * sort_blocks(n);
    return (size_t)(h + 1 + (h + HS_CHUNK - 1) / HS_CHUNK + 1);
}

static size_t staged_hist_words(uint64_t n) {
    uint64_t h = (uint64_t)SR_BINS * ((n + SR_TILE - 1) / SR_TILE);
    return (size_t)(h + 1 + (h + HS_CHUNK - 1) / HS_CHUNK + 1);
}

size_t sort_hist_words_staged(uint64_t n) {
    const size_t a = sort_hist_words(n), b = staged_hist_words(n);
    return a > b ? a : b;
}

static int radix_passes(SortSize sz, uint64_t cap, SortBuffers &sb, hipStream_t stream, int first, int last,
                        int cur, unsigned long long base = 0, unsigned long long top = ~0ull) {
    uint32_t nb = sort_blocks(cap);
    uint32_t nc = (uint32_t)(((uint64_t)256 * nb + HS_CHUNK - 1) / HS_CHUNK);
    uint32_t *chunk_sums = sb.hist + (size_t)256 * nb + 1;
    for (int d = first; d <= last; d++) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(RS_THREADS), 0, stream, sb.keys[cur], sz, 8 * d, base, top,
                           sb.hist);
        hipLaunchKernelGGL((k_hist_chunk_sums<256, RS_TILE>), dim3(nc), dim3(HS_THREADS), 0, stream, sb.hist, sz, chunk_sums);
        hipLaunchKernelGGL((k_scan_chunk_sums<256, RS_TILE>), dim3(1), dim3(1024), 0, stream, chunk_sums, sz);
        hipLaunchKernelGGL((k_hist_scan_apply<256, RS_TILE>), dim3(nc), dim3(HS_THREADS), 0, stream, sb.hist, sz, chunk_sums);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(RS_THREADS), 0, stream, sb.keys[cur], sb.vals[cur],
                           sb.keys[cur ^ 1], sb.vals[cur ^ 1], sz, 8 * d, base, top, sb.hist);
        cur ^= 1;
    }
    return cur;
}

// `passes` SR_BITS-bit digit passes from bit 0 (eager sorts: n known on the host)
static int radix_passes_staged(uint64_t n, SortBuffers &sb, hipStream_t stream, int passes, int cur,
                             unsigned long long base, unsigned long long top) {
    SortSize sz{n, nullptr, nullptr, nullptr};
    const uint32_t nb = (uint32_t)((n + SR_TILE - 1) / SR_TILE), grid = xcd_grid(nb);
    const uint64_t words = (uint64_t)SR_BINS * nb;
    const uint32_t nc = (uint32_t)((words + HS_CHUNK - 1) / HS_CHUNK);
    uint32_t *chunk_sums = sb.hist + words + 1;
    for (int d = 0; d < passes; d++) {
        hipLaunchKernelGGL(k_sradix_hist, dim3(grid), dim3(SR_THREADS), 0, stream, sb.keys[cur], sz, SR_BITS * d, base,
                           top, sb.hist);
        hipLaunchKernelGGL((k_hist_chunk_sums<SR_BINS, SR_TILE>), dim3(nc), dim3(HS_THREADS), 0, stream, sb.hist, sz,
                           chunk_sums);
        hipLaunchKernelGGL((k_scan_chunk_sums<SR_BINS, SR_TILE>), dim3(1), dim3(1024), 0, stream, chunk_sums, sz);
        hipLaunchKernelGGL((k_hist_scan_apply<SR_BINS, SR_TILE>), dim3(nc), dim3(HS_THREADS), 0, stream, sb.hist, sz,
                           chunk_sums);
        hipLaunchKernelGGL(k_sradix_scatter, dim3(grid), dim3(SR_THREADS), 0, stream, sb.keys[cur], sb.vals[cur],
                           sb.keys[cur ^ 1], sb.vals[cur ^ 1], sz, SR_BITS * d, base, top, sb.hist);
        cur ^= 1;
    }
    return cur;
}

static int highest_digit(unsigned long long diff) {
    for (int d = 7; d >= 0; d--)
        if ((diff >> (8 * d)) & 255ull) return d;
    return -1;
}

__global__ void k_count_round_clear(EventCounts *c);

// Sort of one round's events by the PackedEvent order: an LSD radix sort of
// the event times; if two events share a time (duplicate lineages when dedup
// is off), the runs of equal times sorted in LDS windows (k_tie_runs), or,
// for runs with several origins or longer than TIE_RUN_MAX, a full LSD radix
// sort over every field of the order.  Digits above the highest one in which
// a key range differs are equal for every key and are skipped.  Returns the
// sorted permutation (one of sb.vals).  With `keys_ready` the round's kernel
// has already written the time keys, the identity values and the key range
// (sb.keys[0], sb.vals[0], sb.minmax, initialised by k_sort_init).
// With `counts` (PackedEvents of a device_event_counter round, n_real of the
// n records real), a round whose runs of equal times go through the LDS
// windows is counted there, into counts->round_*, and *counted is set; the
// caller then commits it (k_count_finish).
template <class E>
static const uint32_t *sort_permutation(const E *events, uint64_t n, SortBuffers &sb, hipStream_t stream,
                                        bool keys_ready, EventCounts *counts = nullptr, uint64_t n_real = 0,
                                        bool *counted = nullptr) {
    if (counted) *counted = false;
    using O = EventOrder<E>;
    SortSize sz{n, nullptr, nullptr, nullptr};
    unsigned g = grid_for(n, 256), gk = grid_for(n, KS_TILE);
    if (!keys_ready) {
        hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(64), 0, stream, sb.minmax);
        hipLaunchKernelGGL(k_event_keys<E>, dim3(gk), dim3(KS_THREADS), 0, stream, events, sz, sb.keys[0],
                           sb.vals[0], sb.minmax);
    }
    unsigned long long mm[MM_WORDS];
    (void)hipMemcpyAsync(mm, sb.minmax, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, stream);
    (void)hipStreamSynchronize(stream);
    // the times' range [lo, hi] (unused slots excluded): digits of key - lo,
    // unused slots as hi + 1
    const unsigned long long lo = mm[MM_TIME], hi = mm[MM_TIME + 1];
    if (lo > hi) return sb.vals[0]; // only unused slots, in slot order
    const unsigned long long base = lo, top = hi + 1;
    int cur;
    if (sb.hist_words >= staged_hist_words(n) && n < (1ull << 32)) {
        // with room for their histogram, staged passes: ceil(bits / SR_BITS) of them
        const int bits = 64 - __builtin_clzll(top - base);
        cur = radix_passes_staged(n, sb, stream, (bits + SR_BITS - 1) / SR_BITS, 0, base, top);
    } else {
        cur = radix_passes(sz, n, sb, stream, 0, highest_digit(top - base), 0, base, top);
    }
    if (n < 2) return sb.vals[cur];
    // the runs of equal times sorted in LDS windows (a span without equal
    // times keeps its order)
    static_assert(MM_TIE_LONG == MM_TIE_MIXED + 1, "tie flags are adjacent");
    // a counted round's permutation is read by the counters only: its runs
    // are ordered by prior time alone (k_tie_runs<E, false>)
    if (counts)
        hipLaunchKernelGGL((k_tie_runs<E, false>), dim3(grid_for(n, TIE_WINDOW)), dim3(TIE_THREADS), 0, stream,
                           events, sb.keys[cur], sb.vals[cur], n, sb.vals[cur ^ 1], (unsigned long long)n_real, counts,
                           sb.minmax + MM_TIE_MIXED);
    else
        hipLaunchKernelGGL((k_tie_runs<E, true>), dim3(grid_for(n, TIE_WINDOW)), dim3(TIE_THREADS), 0, stream,
                           events, sb.keys[cur], sb.vals[cur], n, sb.vals[cur ^ 1], (unsigned long long)n_real, counts,
                           sb.minmax + MM_TIE_MIXED);
    (void)hipMemcpyAsync(&mm[MM_TIE_MIXED], sb.minmax + MM_TIE_MIXED, 2 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, stream);
    (void)hipStreamSynchronize(stream);
    if (!mm[MM_TIE_MIXED] && !mm[MM_TIE_LONG]) {
        if (counted) *counted = counts != nullptr;
        return sb.vals[cur ^ 1];
    }
    // field passes: the window counts are dropped, and windows that stopped
    // at a long run did not check the rest for several origins / targets
    if (counts) hipLaunchKernelGGL(k_count_round_clear, dim3(1), dim3(1), 0, stream, counts);
    if (!mm[MM_TIE_MIXED])
        hipLaunchKernelGGL(k_tie_check<E>, dim3(gk), dim3(KS_THREADS), 0, stream, events, sb.keys[cur], sb.vals[cur],
                           n, sb.minmax + MM_TIE_MIXED);
    hipLaunchKernelGGL(k_field_minmax<E>, dim3(gk), dim3(KS_THREADS), 0, stream, events, sz, sb.minmax);
    (void)hipMemcpyAsync(mm, sb.minmax, sizeof(mm), hipMemcpyDeviceToHost, stream);
    (void)hipStreamSynchronize(stream);
    // runs sharing origin and target need the reference, prior and time passes only
    const bool uniform_runs = !mm[MM_TIE_MIXED];
    for (int f = 0; f < O::N; f++) {
        if (uniform_runs && f >= 2 && f < O::N - 1) continue;
        const int lf = highest_digit(mm[MM_FIELDS + 2 * f] ^ mm[MM_FIELDS + 2 * f + 1]);
        if (lf < 0) continue;
        hipLaunchKernelGGL(k_field_keys<E>, dim3(g), dim3(256), 0, stream, events, sb.vals[cur], sb.keys[cur], f, sz);
        cur = radix_passes(sz, n, sb, stream, 0, lf, cur);
    }
    return sb.vals[cur];
}

template <class E>
void sort_events(const E *events, uint64_t n, E *out, SortBuffers &sb, hipStream_t stream, bool keys_ready) {
    if (n == 0) return;
    const uint32_t *perm = sort_permutation(events, n, sb, stream, keys_ready);
    SortSize sz{n, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_gather_events<E>, dim3(grid_for(n, 256)), dim3(256), 0, stream, events, perm, out, sz);
}

// Capturable variant: the event range is read from device memory, grids are
// sized for `cap` events, every digit pass is launched, and the tie passes
// are gated on the device-side tie count (an even number of passes, so the
// result lies in the same buffer whether they run or not).
template <class E>
void sort_events_device_sized(const E *events, const unsigned long long *end, const unsigned long long *start,
                              uint64_t cap, E *out, SortBuffers &sb, hipStream_t stream) {
    using O = EventOrder<E>;
    if (cap == 0) return;
    SortSize sz{cap, end, start, nullptr};
    unsigned g = grid_for(cap, 256), gk = grid_for(cap, KS_TILE);
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(64), 0, stream, sb.minmax);
    hipLaunchKernelGGL(k_event_keys<E>, dim3(gk), dim3(KS_THREADS), 0, stream, events, sz, sb.keys[0], sb.vals[0],
                       sb.minmax);
    int cur = radix_passes(sz, cap, sb, stream, 0, 7, 0);
    hipLaunchKernelGGL(k_tie_count, dim3(gk), dim3(KS_THREADS), 0, stream, sb.keys[cur], sz, sb.minmax + MM_TIES);
    SortSize gz = sz;
    gz.gate = sb.minmax + MM_TIES;
    for (int f = 0; f < O::N; f++) {
        hipLaunchKernelGGL(k_field_keys<E>, dim3(g), dim3(256), 0, stream, events, sb.vals[cur], sb.keys[cur], f,
                           gz);
        cur = radix_passes(gz, cap, sb, stream, 0, O::digits(f) - 1, cur);
    }
    hipLaunchKernelGGL(k_gather_events<E>, dim3(g), dim3(256), 0, stream, events, sb.vals[cur], out, sz);
}

template void sort_events<DevEvent>(const DevEvent *, uint64_t, DevEvent *, SortBuffers &, hipStream_t, bool);
template void sort_events<DevPacked>(const DevPacked *, uint64_t, DevPacked *, SortBuffers &, hipStream_t, bool);
template void sort_events_device_sized<DevEvent>(const DevEvent *, const unsigned long long *,
                                                 const unsigned long long *, uint64_t, DevEvent *, SortBuffers &,
                                                 hipStream_t);
template void sort_events_device_sized<DevPacked>(const DevPacked *, const unsigned long long *,
                                                  const unsigned long long *, uint64_t, DevPacked *, SortBuffers &,
                                                  hipStream_t);

// ---------------------------------------------------------------------------
// EventCounterReporter::report_{speciation,dispersal} (necsim/plugins/common/
// src/event_counter.rs:58-109) over one round's events in PackedEvent order.
// The reporter compares each event with the last one of its own kind
// (SpeciationEvent: equal origin and event time; DispersalEvent: equal
// origin, event time and target, event.rs:301-320) and with the prior time of
// the event just before it.  In sorted order, events with equal origin and
// time are contiguous, so the last event of the same kind is equal exactly
// when one lies in the contiguous run of equal keys that ends at this event:
// each event scans back through that run (one step unless coalesced copies
// share the key).  Equal times never straddle two rounds (all events below
// the water level are in one round), so only the last prior time is carried.
__device__ __forceinline__ bool packed_is_speciation(const DevPacked &e) {
    return (bits(e.prior_time) >> 63) != 0 && (bits(e.event_time) >> 63) == 0; // p(-), e(+)
}

__device__ __forceinline__ bool same_event_key(const DevPacked &a, const DevPacked &b, bool speciation) {
    return (bits(a.event_time) & ABS_BITS) == (bits(b.event_time) & ABS_BITS) && a.ox == b.ox && a.oy == b.oy &&
           a.oindex == b.oindex && (speciation || (a.tx == b.tx && a.ty == b.ty && a.tindex == b.tindex));
}

// the fields of a PackedEvent the counter reads (all but the reference and
// the coalescence)
struct CountRec {
    uint64_t prior, time; // bit patterns, signs included
    uint32_t ox, oy, oi, tx, ty, ti;
};

__device__ __forceinline__ CountRec count_rec(const DevPacked &e) {
    return CountRec{bits(e.prior_time), bits(e.event_time), e.ox, e.oy, e.oindex, e.tx, e.ty, e.tindex};
}
__device__ __forceinline__ bool rec_is_speciation(const CountRec &e) {
    return (e.prior >> 63) != 0 && (e.time >> 63) == 0; // p(-), e(+)
}
__device__ __forceinline__ bool rec_same_key(const CountRec &a, const CountRec &b, bool speciation) {
    return (a.time & ABS_BITS) == (b.time & ABS_BITS) && a.ox == b.ox && a.oy == b.oy && a.oi == b.oi &&
           (speciation || (a.tx == b.tx && a.ty == b.ty && a.ti == b.ti));
}

// Each thread reads its own event once (through the permutation, a random
// read) and shares it in LDS: the previous events in sorted order come from
// the neighbouring threads, and from memory only across the block's start.
__global__ void __launch_bounds__(256) k_count_events(const DevPacked *ev, const uint32_t *perm, SortSize sz,
                                                     const unsigned long long *ghosts, EventCounts *c,
                                                     const unsigned long long *gate) {
    if (gate && !*gate) return; // counted in the sort's windows
    // the k-th event in sorted order: ev[k], or ev[perm[k]] from the sort's permutation
    auto at = [&](uint64_t k) -> CountRec { return count_rec(ev[perm ? perm[k] : k]); };
    __shared__ unsigned long long s_cnt[EC_COUNTERS];
    __shared__ CountRec s_rec[256];
    if (threadIdx.x < EC_COUNTERS) s_cnt[threadIdx.x] = 0;
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t i = first + threadIdx.x;
    const uint64_t n = sz.get() - (ghosts ? *ghosts : 0ull);
    CountRec e{};
    if (i < n) s_rec[threadIdx.x] = e = at(i);
    __syncthreads();
    auto rec = [&](uint64_t k) -> CountRec { return k >= first ? s_rec[k - first] : at(k); };
    int which = -1; // the counter this event adds to, besides the raw total
    if (i < n) {
        const bool sp = rec_is_speciation(e);
        const bool has_prev = i > 0 || c->carry_has != 0;
        const uint64_t prev_prior = i > 0 ? (rec(i - 1).prior & ABS_BITS) : c->carry_prior;
        bool dup = false;
        for (uint64_t j = i; j-- > 0;) {
            const CountRec p = rec(j);
            if (!rec_same_key(p, e, sp)) break;
            if (rec_is_speciation(p) == sp) {
                dup = true;
                break;
            }
        }
        if (dup) {
            if (!has_prev || prev_prior != (e.prior & ABS_BITS)) which = EC_LATE_COALESCENCE;
        } else if (sp) {
            which = EC_SPECIATION;
        } else {
            // LineageInteraction from the signs (event.rs:186-229)
            const bool prior_pos = (e.prior >> 63) == 0, event_pos = (e.time >> 63) == 0;
            if (prior_pos && !event_pos) {
                which = EC_LATE_DISPERSAL_COALESCENCE; // Maybe
            } else {
                const bool self = e.ox == e.tx && e.oy == e.ty && e.oi == e.ti;
                const bool coal = prior_pos && event_pos;
                which = self ? (coal ? EC_SELF_COALESCENCE : EC_SELF_DISPERSAL)
                             : (coal ? EC_OUT_COALESCENCE : EC_OUT_DISPERSAL);
            }
        }
    }
    const uint64_t valid = __ballot(i < n);
    if (lane_id() == 0 && valid) atomicAdd(&s_cnt[EC_RAW_TOTAL], (unsigned long long)__popcll(valid));
    for (int k = EC_SPECIATION; k < EC_COUNTERS; k++) {
        const uint64_t m = __ballot(which == k);
        if (lane_id() == 0 && m) atomicAdd(&s_cnt[k], (unsigned long long)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < EC_COUNTERS && s_cnt[threadIdx.x]) atomicAdd(&c->count[threadIdx.x], s_cnt[threadIdx.x]);
}

// after the round's count: the last prior time for the next round
__global__ void k_count_carry(const DevPacked *ev, const uint32_t *perm, SortSize sz,
                              const unsigned long long *ghosts, EventCounts *c, const unsigned long long *gate) {
    if (gate && !*gate) return;
    const uint64_t n = sz.get() - (ghosts ? *ghosts : 0ull);
    if (n) {
        c->carry_has = 1;
        c->carry_prior = bits(ev[perm ? perm[n - 1] : n - 1].prior_time) & ABS_BITS;
    }
}

void launch_count_events(const DevPacked *ev, const uint32_t *perm, const unsigned long long *end,
                         uint64_t n_or_cap, const unsigned long long *ghosts, EventCounts *c, hipStream_t stream) {
    if (n_or_cap == 0) return;
    SortSize sz{n_or_cap, end, nullptr, nullptr};
    hipLaunchKernelGGL(k_count_events, dim3(grid_for(n_or_cap, 256)), dim3(256), 0, stream, ev, perm, sz, ghosts,
                       c, (const unsigned long long *)nullptr);
    hipLaunchKernelGGL(k_count_carry, dim3(1), dim3(1), 0, stream, ev, perm, sz, ghosts, c,
                       (const unsigned long long *)nullptr);
}

// commits a round counted in the sort's windows
__global__ void k_count_finish(EventCounts *c) {
    for (int k = 0; k < EC_COUNTERS; k++) c->count[k] += c->round_count[k];
    c->carry_has = 1;
    c->carry_prior = c->round_carry_prior;
    for (int k = 0; k < EC_COUNTERS; k++) c->round_count[k] = 0;
}
// a round whose window counts are dropped (its runs went to field passes)
__global__ void k_count_round_clear(EventCounts *c) {
    for (int k = 0; k < EC_COUNTERS; k++) c->round_count[k] = 0;
}

// A round of a device_event_counter job: the events' sorted permutation, then
// the counters read the events through it (no sorted copy is written)
void sort_and_count_packed(const DevPacked *events, uint64_t slots, uint64_t n_events, bool keys_ready,
                           SortBuffers &sb, EventCounts *c, hipStream_t stream) {
    if (slots == 0 || n_events == 0) return;
    bool counted = false;
    const uint32_t *perm = sort_permutation(events, slots, sb, stream, keys_ready, c, n_events, &counted);
    if (!counted) {
        launch_count_events(events, perm, nullptr, n_events, nullptr, c, stream);
        return;
    }
    // counted in the windows
    hipLaunchKernelGGL(k_count_finish, dim3(1), dim3(1), 0, stream, c);
}

uint32_t soa_compact_blocks(uint64_t n) { return (uint32_t)((n + SC_TILE - 1) / SC_TILE); }

void launch_soa_compact(const LineageSoA &in, const LineageSoA &out, uint64_t n, uint32_t *blk,
                        hipStream_t stream) {
    uint32_t nb = soa_compact_blocks(n);
    if (!nb) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, blk, 0u, (uint64_t)1);
        return;
    }
    hipLaunchKernelGGL(k_soa_count, dim3(nb), dim3(SC_THREADS), 0, stream, in.flags, n,
                       (const unsigned long long *)nullptr, blk);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, blk, (uint64_t)nb);
    hipLaunchKernelGGL(k_soa_scatter, dim3(nb), dim3(SC_THREADS), 0, stream, in, out, n,
                       (const unsigned long long *)nullptr, (const Counters *)nullptr, blk);
}

void launch_soa_compact_gated(const LineageSoA &in, const LineageSoA &out, uint64_t cap,
                              const unsigned long long *n_dev, uint32_t *blk, Counters *ctr,
                              hipStream_t stream) {
    uint32_t nb = soa_compact_blocks(cap);
    if (!nb) return;
    // blocks past *n_dev count zero, so the scan total is the live count
    hipLaunchKernelGGL(k_soa_count, dim3(nb), dim3(SC_THREADS), 0, stream, in.flags, cap, n_dev, blk);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, stream, blk, (uint64_t)nb);
    hipLaunchKernelGGL(k_soa_gate, dim3(1), dim3(1), 0, stream, blk, nb, cap, n_dev, ctr);
    hipLaunchKernelGGL(k_soa_scatter, dim3(nb), dim3(SC_THREADS), 0, stream, in, out, cap, n_dev,
                       (const Counters *)ctr, blk);
}

void launch_sort_init(SortBuffers &sb, hipStream_t stream) {
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(64), 0, stream, sb.minmax);
}

void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, v, n);
}
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, v, n);
}
void launch_reset_counters(Counters *c, bool reset_events, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, stream, c, reset_events);
}

} // namespace nh
