// nh_server.hip -- continuous batching of Independent simulations on resident
// scenario tables.
//
// Requests arrive with their own seed, speciation probability, dedup setting
// and pause time (and lineages).  A request is admitted as soon as its device
// footprint fits the HBM budget and the lane-state capacity; every kernel tick
// runs ONE persistent kernel over the lineages of all admitted requests, each
// towards its own water level (parallelisation/independent/monolithic/mod.rs:135-177
// per request).  Requests that arrive while others run are admitted at the
// next tick, their work overlapping with the others' on the device.  A request
// retires when it has no lineages left or reached its pause time; its events
// are sorted per water-level round like the live reporter (reporter/live.rs:88-115).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/necsim_hip.h"
#include "nh_internal.h"
#include "nh_kernels.h"

using namespace nh;


// the table's device view (defined in nh_runtime.hip)
const DevScenario *nh_tables_device_view(const nh_tables *t, int *kind);

namespace {

double mono_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Request memory comes from ONE device allocation of hbm_budget_bytes made
// when the server is created: each admitted request takes one contiguous
// block (first fit over a free list ordered by offset; freed blocks merge
// with their neighbours), so admission and retirement never call hipMalloc
// or hipFree (hipFree synchronises the device, which would stall the tick
// loop).
constexpr uint64_t ARENA_ALIGN = 256;
constexpr int REQ_ARRAYS = 2 * 8 + 2; // two SoA copies, the event slots, the dedup cache
constexpr uint64_t BYTES_PER_LINEAGE = 2 * 45 + 40; // two SoA copies + one event slot
constexpr uint64_t BYTES_PER_BUCKET = 32;
constexpr uint64_t REQ_SLACK = REQ_ARRAYS * ARENA_ALIGN; // alignment of the sub-arrays

uint64_t align_up(uint64_t v) { return (v + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

struct Arena {
    char *base = nullptr;
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_; // offset -> length
    int create(uint64_t bytes) {
        size = align_up(bytes);
        if (size == 0) return NH_OK;
        NH_TRY(dalloc(&base, size));
        free_[0] = size;
        return NH_OK;
    }
    void destroy() { dfree(base); }
    bool alloc(uint64_t len, uint64_t *off) {
        len = align_up(len ? len : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            *off = it->first;
            uint64_t rest = it->second - len, at = it->first + len;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t len) {
        len = align_up(len ? len : 1);
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + len == nx->first) { // merge with the next block
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) { // ... and the previous one
                pv->second += len;
                return;
            }
        }
        free_[off] = len;
    }
};

// Carves a request's arrays out of its block (every sub-array 256-aligned).
struct Carver {
    char *p;
    template <class T> T *take(uint64_t n) {
        T *r = reinterpret_cast<T *>(p);
        p += align_up(n * sizeof(T));
        return r;
    }
};

struct Req {
    uint64_t id = 0;
    nh_request p{};
    std::vector<nh_lineage> lineages; // until admission
    uint64_t n_initial = 0;
    uint64_t event_slice = 1;
    uint64_t cache_cap = 0;
    uint64_t footprint = 0;
    uint64_t blocked_ticks = 0; // ticks it waited for memory or lanes
    // device state (inside the arena block [arena_off, arena_off + footprint))
    bool admitted = false;
    uint64_t arena_off = 0;
    LineageSoA soa[2]{};
    int cur = 0;
    uint64_t n_in = 0;  // slots of the current buffer
    uint64_t alive = 0; // live lineages
    DevEvent *d_events = nullptr;
    uint64_t *d_cache = nullptr;
    double level = 0.0;
    bool round_open = false;
    uint64_t round_event_start = 0;
    uint64_t n_events = 0, steps = 0;
    double max_time = 0.0;
    // results
    std::vector<nh_speciation_event> events;
    nh_request_result result{};
};

// the device footprint of a request: its arena block, a multiple of the
// block alignment (so that requests whose footprints sum to the pool size
// fit it together)
uint64_t request_footprint(uint64_t n, uint64_t cache_cap) {
    return align_up(n * BYTES_PER_LINEAGE + cache_cap * BYTES_PER_BUCKET + REQ_SLACK);
}

void carve(Req *r, char *block) {
    Carver c{block};
    uint64_t n = r->n_initial;
    for (LineageSoA &s : r->soa) {
        s.ref = c.take<uint64_t>(n);
        s.t_last = c.take<double>(n);
        s.next_t = c.take<double>(n);
        s.next_st = c.take<uint64_t>(n);
        s.x = c.take<uint32_t>(n);
        s.y = c.take<uint32_t>(n);
        s.idx = c.take<uint32_t>(n);
        s.flags = c.take<uint8_t>(n);
    }
    r->d_events = c.take<DevEvent>(n);
    r->d_cache = r->cache_cap ? c.take<uint64_t>(4 * r->cache_cap) : nullptr;
}

} // namespace

struct nh_server {
    nh_ctx *ctx = nullptr;
    const nh_tables *tables = nullptr;
    DevScenario S{};
    int kind = 0;
    nh_server_params prm{};
    uint64_t max_bypass_ticks = 16;
    DevParams P{};
    double *d_ratio = nullptr;
    double lambda_dt = 0.0, delta_t = 0.0, turnover = 0.0;
    uint64_t step_slice = 0;
    bool shared_set = false;
    int threads = 256, blocks = 0;
    uint64_t next_id = 1;
    std::deque<Req *> queued;
    std::vector<Req *> active;
    std::deque<Req *> finished; // not yet polled
    std::map<uint64_t, Req *> polled;
    uint64_t used_bytes = 0, used_lineages = 0;
    Arena arena;
    // shared device buffers
    ServeReq *d_reqs = nullptr;
    ReqCounters *d_ctr = nullptr; // SERVE_MAX_REQS
    unsigned long long *d_head = nullptr;
    uint32_t *d_blk = nullptr; // SoA compaction block counts
    DevEvent *d_sorted = nullptr;
    SortBuffers sb{};
    std::vector<ServeReq> h_reqs;
    std::vector<ReqCounters> h_ctr;
    ~nh_server() {
        for (Req *r : queued) delete r;
        for (Req *r : active) delete r;
        for (Req *r : finished) delete r;
        for (auto &kv : polled) delete kv.second;
        arena.destroy();
        dfree(d_ratio); dfree(d_reqs); dfree(d_ctr); dfree(d_head); dfree(d_blk); dfree(d_sorted);
        dfree(sb.keys[0]); dfree(sb.keys[1]); dfree(sb.vals[0]); dfree(sb.vals[1]);
        dfree(sb.hist); dfree(sb.minmax);
    }
};

extern "C" int nh_server_create(nh_ctx *ctx, const nh_tables *tables, const nh_server_params *prm,
                                nh_server **out) {
    if (!ctx || !tables || !prm || !out) return set_error(NH_E_INVALID, "nh_server_create: null argument");
    *out = nullptr;
    if (prm->max_lineages == 0 || prm->max_lineages > 0xffffffffull)
        return set_error(NH_E_INVALID, "max_lineages must lie in 1..2^32-1");
    if (prm->block_size && (prm->block_size % 64 || prm->block_size > 256))
        return set_error(NH_E_INVALID, "block size must be 64, 128, 192 or 256");
    if (prm->hbm_budget_bytes == 0) return set_error(NH_E_INVALID, "hbm_budget_bytes must be positive");
    NH_HIP(hipSetDevice(ctx->device));
    nh_server *s = new nh_server();
    s->ctx = ctx;
    s->tables = tables;
    s->S = *nh_tables_device_view(tables, &s->kind);
    s->prm = *prm;
    if (prm->max_bypass_ticks) s->max_bypass_ticks = prm->max_bypass_ticks;
    s->threads = prm->block_size ? (int)prm->block_size : 256;
    // one block per CU unless asked otherwise: the serving stream is latency
    // bound (the last lineages of each request run serially, searched
    // cooperatively), and more resident waves per SIMD slow every wave down;
    // on the 256x256 map stream this halves p99 against full occupancy
    int per_cu = serve_max_blocks_per_cu(s->kind, s->threads);
    per_cu = std::min<int>(per_cu, prm->blocks_per_cu ? (int)prm->blocks_per_cu : 1);
    s->blocks = per_cu * (int)ctx->cus;
    uint64_t cap = prm->max_lineages;
    auto fail = [&](int rc) { delete s; return rc; };
    int rc;
    if ((rc = s->arena.create(prm->hbm_budget_bytes))) return fail(rc);
    if ((rc = dalloc(&s->d_ratio, POISSON_TABLE_N))) return fail(rc);
    if ((rc = dalloc(&s->d_reqs, SERVE_MAX_REQS))) return fail(rc);
    if ((rc = dalloc(&s->d_ctr, SERVE_MAX_REQS))) return fail(rc);
    if ((rc = dalloc(&s->d_head, 1))) return fail(rc);
    if ((rc = dalloc(&s->d_blk, soa_compact_blocks(cap) + 1))) return fail(rc);
    if ((rc = dalloc(&s->d_sorted, cap))) return fail(rc);
    s->sb.capacity = cap;
    s->sb.hist_words = sort_hist_words_staged(cap);
    if ((rc = dalloc(&s->sb.keys[0], cap)) || (rc = dalloc(&s->sb.keys[1], cap)) ||
        (rc = dalloc(&s->sb.vals[0], cap)) || (rc = dalloc(&s->sb.vals[1], cap)) ||
        (rc = dalloc(&s->sb.hist, s->sb.hist_words)) || (rc = dalloc(&s->sb.minmax, SORT_MINMAX_WORDS)))
        return fail(rc);
    s->h_reqs.resize(SERVE_MAX_REQS);
    s->h_ctr.resize(SERVE_MAX_REQS);
    *out = s;
    return NH_OK;
}

extern "C" int nh_server_request_footprint(uint64_t n_lineages, uint64_t cache_buckets, uint64_t *bytes) {
    if (!bytes) return set_error(NH_E_INVALID, "nh_server_request_footprint: null argument");
    *bytes = request_footprint(n_lineages, cache_buckets);
    return NH_OK;
}

extern "C" int nh_server_submit(nh_server *s, const nh_request *rq, uint64_t *id) {
    if (!s || !rq || !id || (!rq->lineages && rq->n_lineages)) return set_error(NH_E_INVALID, "nh_server_submit: null argument");
    // the argument checks of nh_job_create
    if (!(rq->speciation_probability > 0.0 && rq->speciation_probability <= 1.0))
        return set_error(NH_E_INVALID, "speciation probability must lie in (0, 1]");
    if (!(rq->turnover_rate > 0.0) || std::isinf(rq->turnover_rate) || !(rq->delta_t > 0.0) ||
        std::isinf(rq->delta_t) || rq->step_slice == 0)
        return set_error(NH_E_INVALID, "turnover rate, delta_t and step_slice must be positive");
    if (rq->dedup_mode < NH_DEDUP_NONE || rq->dedup_mode > NH_DEDUP_ABSOLUTE) return set_error(NH_E_INVALID, "dedup mode");
    if (rq->dedup_mode == NH_DEDUP_RELATIVE && !(rq->dedup_factor > 0.0)) return set_error(NH_E_INVALID, "dedup factor");
    if (rq->event_slice_mode != NH_SLICE_RELATIVE && rq->event_slice_mode != NH_SLICE_ABSOLUTE)
        return set_error(NH_E_INVALID, "event slice mode");
    if (rq->event_slice_mode == NH_SLICE_RELATIVE && !(rq->event_slice_factor > 0.0))
        return set_error(NH_E_INVALID, "event slice factor");
    if (rq->event_slice_mode == NH_SLICE_ABSOLUTE && rq->event_slice_capacity == 0)
        return set_error(NH_E_INVALID, "event slice capacity");
    if (rq->has_pause_before && std::isnan(rq->pause_before)) return set_error(NH_E_INVALID, "pause_before is NaN");
    // lambda, delta_t and step_slice are shared by every request of a server
    if (s->shared_set &&
        (rq->turnover_rate != s->turnover || rq->delta_t != s->delta_t || rq->step_slice != s->step_slice))
        return set_error(NH_E_INVALID, "turnover rate, delta_t and step_slice must match the server's first request");
    if (rq->n_lineages > s->prm.max_lineages)
        return set_error(NH_E_CAPACITY, "request of %zu lineages exceeds the server capacity", rq->n_lineages);
    // the job path's lineage filter (nh_job_add_lineages): no kernel ever
    // sees a lineage outside the scenario's habitat or its deme
    for (size_t k = 0; k < rq->n_lineages; k++) {
        const nh_lineage &l = rq->lineages[k];
        if (nh_lineage_class(s->tables, l) != NH_LINEAGE_OK)
            return set_error(NH_E_OUT_OF_DEME, "lineage %zu at (%u, %u, %u) is outside the habitat or its deme", k,
                             l.x, l.y, l.index);
        if (!(l.last_event_time >= 0.0) || std::isinf(l.last_event_time))
            return set_error(NH_E_INVALID, "lineage %zu: last_event_time must be finite and non-negative", k);
    }
    uint64_t n = rq->n_lineages, event_slice, cache_cap = 0;
    if (rq->event_slice_mode == NH_SLICE_ABSOLUTE) event_slice = std::max<uint64_t>(rq->event_slice_capacity, 1);
    else event_slice = std::max<uint64_t>(as_u64((double)n * rq->event_slice_factor), 1);
    if (rq->dedup_mode == NH_DEDUP_RELATIVE) cache_cap = as_u64((double)n * rq->dedup_factor);
    else if (rq->dedup_mode == NH_DEDUP_ABSOLUTE) cache_cap = rq->dedup_capacity;
    if (cache_cap > (UINT64_MAX - REQ_SLACK - ARENA_ALIGN - n * BYTES_PER_LINEAGE) / BYTES_PER_BUCKET ||
        request_footprint(n, cache_cap) > s->arena.size)
        return set_error(NH_E_CAPACITY, "request footprint exceeds the HBM budget");
    if (!s->shared_set) {
        s->turnover = rq->turnover_rate;
        s->delta_t = rq->delta_t;
        s->step_slice = rq->step_slice;
        DevParams &P = s->P;
        P.lambda_dt = rq->turnover_rate * rq->delta_t;
        P.p0 = nh::exp(-P.lambda_dt);
        P.lambda_dt_sqrt = sqrt_(P.lambda_dt);
        P.delta_t = rq->delta_t;
        uint64_t b = bits(rq->delta_t);
        P.dt_pow2 = ((b >> 52) & 0x7ff) != 0 && (b & 0xfffffffffffffull) == 0;
        P.inv_delta_t = P.dt_pow2 ? 1.0 / rq->delta_t : 0.0;
        P.min_event_time = 0.0;
        P.step_slice = rq->step_slice;
        std::vector<double> ratio(POISSON_TABLE_N, 0.0);
        P.prod_zero_at = poisson_table(P.lambda_dt, P.p0, ratio.data());
        NH_HIP(hipMemcpy(s->d_ratio, ratio.data(), POISSON_TABLE_N * sizeof(double), hipMemcpyHostToDevice));
        P.ratio = s->d_ratio;
        s->shared_set = true;
    }
    Req *r = new Req();
    r->id = s->next_id++;
    r->p = *rq;
    r->p.lineages = nullptr;
    r->lineages.assign(rq->lineages, rq->lineages + rq->n_lineages);
    r->n_initial = n;
    r->event_slice = event_slice;
    r->cache_cap = cache_cap;
    r->footprint = request_footprint(n, cache_cap);
    if (n) {
        double mx = r->lineages[0].last_event_time, mn = mx;
        for (auto &l : r->lineages) {
            mx = std::max(mx, l.last_event_time);
            mn = std::min(mn, l.last_event_time);
        }
        r->max_time = mx;
        r->level = mn;
    }
    r->result.request_id = r->id;
    r->result.submit_s = mono_s();
    s->queued.push_back(r);
    *id = r->id;
    return NH_OK;
}

// gives a request's arena block back
static void release_block(nh_server *s, Req *r) {
    if (!r->admitted) return;
    s->arena.release(r->arena_off, r->footprint);
    s->used_bytes -= r->footprint;
    s->used_lineages -= r->n_initial;
    r->admitted = false;
    r->soa[0] = r->soa[1] = LineageSoA{};
    r->d_events = nullptr;
    r->d_cache = nullptr;
}

// ends a request (status NH_OK, or the error that ended it) and queues its
// result for nh_server_poll
static void finish(nh_server *s, Req *r, int status) {
    r->result.status = status;
    r->result.paused = r->alive > 0;
    r->result.time = r->max_time;
    r->result.steps = r->steps;
    r->result.n_events = r->events.size();
    r->result.done_s = mono_s();
    release_block(s, r);
    s->finished.push_back(r);
}

// uploads an admitted request's lineages into its arena block
static int admit(nh_server *s, Req *r, uint64_t off) {
    hipStream_t st = s->ctx->stream;
    uint64_t n = r->n_initial;
    r->admitted = true;
    r->arena_off = off;
    s->used_bytes += r->footprint;
    s->used_lineages += n;
    carve(r, s->arena.base + off);
    r->result.admit_s = mono_s();
    if (r->cache_cap) launch_fill_u64(r->d_cache, 0, 4 * r->cache_cap, st);
    if (n) {
        std::vector<uint64_t> ref(n), zero(n, 0);
        std::vector<double> t(n);
        std::vector<uint32_t> x(n), y(n), idx(n);
        std::vector<uint8_t> fl(n, 0);
        for (uint64_t k = 0; k < n; k++) {
            const nh_lineage &l = r->lineages[k];
            ref[k] = l.global_reference; t[k] = l.last_event_time; x[k] = l.x; y[k] = l.y; idx[k] = l.index;
        }
        LineageSoA &d = r->soa[0];
        NH_HIP(hipMemcpyAsync(d.ref, ref.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.t_last, t.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.next_t, t.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.next_st, zero.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.x, x.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.y, y.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.idx, idx.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.flags, fl.data(), n, hipMemcpyHostToDevice, st));
        NH_HIP(hipStreamSynchronize(st));
    } else {
        NH_HIP(hipGetLastError());
    }
    r->lineages.clear();
    r->lineages.shrink_to_fit();
    r->n_in = n;
    r->alive = n;
    r->cur = 0;
    return NH_OK;
}

static bool request_finished(const Req *r) {
    return r->alive == 0 || (r->p.has_pause_before && !(r->level < r->p.pause_before));
}

// open the next water-level round of a request (monolithic/mod.rs:139-174)
static void open_round(nh_server *s, Req *r) {
    double total_rate = repeated_sum(r->p.turnover_rate * r->p.speciation_probability, r->alive);
    r->level = r->level + (double)r->event_slice / total_rate;
    if (r->p.has_pause_before && r->p.pause_before < r->level) r->level = r->p.pause_before;
    r->round_open = true;
    r->round_event_start = r->n_events;
}

// Admission: queued requests in arrival order, each as soon as its block and
// its lanes fit.  A request that does not fit yet is passed by later ones
// until it has waited max_bypass_ticks ticks; from then on nothing is
// admitted past it, so the memory it needs drains.
static void admit_queued(nh_server *s) {
    size_t k = 0;
    bool blocked_head = false;
    while (k < s->queued.size() && s->active.size() < (size_t)SERVE_MAX_REQS) {
        Req *r = s->queued[k];
        uint64_t off = 0;
        if (s->used_lineages + r->n_initial > s->prm.max_lineages || !s->arena.alloc(r->footprint, &off)) {
            if (!blocked_head) {
                blocked_head = true;
                if (++r->blocked_ticks > s->max_bypass_ticks) break; // waited long enough: let it drain
            }
            k++;
            continue;
        }
        s->queued.erase(s->queued.begin() + k);
        int rc = admit(s, r, off);
        if (rc != NH_OK) {
            finish(s, r, rc);
            continue;
        }
        if (request_finished(r)) finish(s, r, NH_OK);
        else s->active.push_back(r);
    }
}

// the end of a request's water-level round: drop dead slots and sort the
// round's events like the live water-level reporter
static int close_round(nh_server *s, Req *r) {
    hipStream_t st = s->ctx->stream;
    r->round_open = false;
    if (r->n_in > r->alive) {
        launch_soa_compact(r->soa[r->cur], r->soa[r->cur ^ 1], r->n_in, s->d_blk, st);
        NH_HIP(hipGetLastError());
        uint32_t live = 0;
        NH_HIP(hipMemcpyAsync(&live, s->d_blk + soa_compact_blocks(r->n_in), 4, hipMemcpyDeviceToHost, st));
        NH_HIP(hipStreamSynchronize(st));
        if (live != r->alive)
            return set_error(NH_E_STATE, "request %llu: %u live lineages, expected %llu", (unsigned long long)r->id,
                             live, (unsigned long long)r->alive);
        r->cur ^= 1;
        r->n_in = live;
    }
    uint64_t ne = r->n_events - r->round_event_start;
    if (ne) {
        sort_events(r->d_events + r->round_event_start, ne, s->d_sorted, s->sb, st, false);
        NH_HIP(hipGetLastError());
        size_t old = r->events.size();
        r->events.resize(old + ne);
        NH_HIP(hipMemcpyAsync(r->events.data() + old, s->d_sorted, ne * sizeof(DevEvent), hipMemcpyDeviceToHost, st));
        NH_HIP(hipStreamSynchronize(st));
    }
    return NH_OK;
}

extern "C" int nh_server_tick(nh_server *s, uint64_t *n_active, uint64_t *n_queued) {
    if (!s) return set_error(NH_E_INVALID, "null server");
    NH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    admit_queued(s);
    // every active request without an open round starts its next one
    for (Req *r : s->active)
        if (!r->round_open) open_round(s, r);
    uint64_t total = 0;
    int nr = (int)s->active.size();
    for (int k = 0; k < nr; k++) {
        Req *r = s->active[k];
        ServeReq &d = s->h_reqs[k];
        d.seed = r->p.rng_seed;
        d.nu = r->p.speciation_probability;
        d.nu_m = draw_threshold_below(d.nu);
        d.level = r->level;
        d.lin = r->soa[r->cur];
        d.events = r->d_events;
        d.cache = r->d_cache;
        d.cache_cap = r->cache_cap;
        d.n_in = r->n_in;
        d.work_base = total;
        d.ctr = s->d_ctr + k;
        total += r->n_in;
        s->h_ctr[k] = ReqCounters{0, 0, r->n_events, 0, 0, {0, 0, 0}};
        if (r->alive == 0) d.n_in = 0;
    }
    int rc = NH_OK;
    if (nr > 0 && total > 0) {
        auto shared_tick = [&]() -> int {
            NH_HIP(hipMemcpyAsync(s->d_reqs, s->h_reqs.data(), nr * sizeof(ServeReq), hipMemcpyHostToDevice, st));
            NH_HIP(hipMemcpyAsync(s->d_ctr, s->h_ctr.data(), nr * sizeof(ReqCounters), hipMemcpyHostToDevice, st));
            NH_HIP(hipMemsetAsync(s->d_head, 0, 8, st));
            ServeArgs a{};
            a.S = s->S;
            a.P = s->P;
            a.reqs = s->d_reqs;
            a.n_reqs = nr;
            a.total_work = total;
            a.head = s->d_head;
            a.tick_steps = s->prm.tick_steps;
            a.chunk = claim_chunk(total, (unsigned long long)s->blocks * s->threads);
            launch_serve(s->kind, a, s->blocks, s->threads, st);
            NH_HIP(hipGetLastError());
            NH_HIP(hipMemcpyAsync(s->h_ctr.data(), s->d_ctr, nr * sizeof(ReqCounters), hipMemcpyDeviceToHost, st));
            NH_HIP(hipStreamSynchronize(st));
            return NH_OK;
        };
        rc = shared_tick();
        if (rc != NH_OK) {
            // the shared tick failed: no request's state can be trusted, so
            // every active request ends with the error
            for (Req *r : s->active) finish(s, r, rc);
            s->active.clear();
            if (n_active) *n_active = 0;
            if (n_queued) *n_queued = s->queued.size();
            return rc;
        }
    }
    std::vector<Req *> still;
    for (int k = 0; k < nr; k++) {
        Req *r = s->active[k];
        const ReqCounters &c = s->h_ctr[k];
        r->alive -= c.n_dead;
        r->n_events = c.n_events;
        r->steps += c.steps;
        if (c.max_time_bits && from_bits(c.max_time_bits) > r->max_time) r->max_time = from_bits(c.max_time_bits);
        if (c.n_below && r->alive > 0) { // the round continues next tick
            still.push_back(r);
            continue;
        }
        int rr = close_round(s, r);
        if (rr != NH_OK) {
            finish(s, r, rr); // this request fails; the others go on
            if (rc == NH_OK) rc = rr;
            continue;
        }
        if (request_finished(r)) finish(s, r, NH_OK);
        else still.push_back(r);
    }
    s->active.swap(still);
    if (n_active) *n_active = s->active.size();
    if (n_queued) *n_queued = s->queued.size();
    return rc;
}

extern "C" int nh_server_poll(nh_server *s, nh_request_result *res) {
    if (!s || !res) return set_error(NH_E_INVALID, "null argument");
    if (s->finished.empty()) return set_error(NH_E_STATE, "no finished request");
    Req *r = s->finished.front();
    s->finished.pop_front();
    *res = r->result;
    s->polled[r->id] = r;
    return NH_OK;
}

extern "C" int nh_server_request_events(nh_server *s, uint64_t id, nh_speciation_event *buf, size_t cap,
                                        size_t *n_out) {
    if (!s || !n_out) return set_error(NH_E_INVALID, "null argument");
    auto it = s->polled.find(id);
    if (it == s->polled.end()) return set_error(NH_E_STATE, "request %llu is not a polled result", (unsigned long long)id);
    const std::vector<nh_speciation_event> &ev = it->second->events;
    *n_out = ev.size();
    if (!buf) return NH_OK;
    if (cap < ev.size()) return set_error(NH_E_CAPACITY, "event buffer too small");
    std::memcpy(buf, ev.data(), ev.size() * sizeof(nh_speciation_event));
    return NH_OK;
}

extern "C" int nh_server_release(nh_server *s, uint64_t id) {
    if (!s) return set_error(NH_E_INVALID, "null server");
    auto it = s->polled.find(id);
    if (it == s->polled.end()) return set_error(NH_E_STATE, "request %llu is not a polled result", (unsigned long long)id);
    delete it->second;
    s->polled.erase(it);
    return NH_OK;
}

extern "C" int nh_server_destroy(nh_server *s) {
    if (!s) return NH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return NH_OK;
}
