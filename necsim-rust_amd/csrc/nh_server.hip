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

struct Soa {
    LineageSoA s{};
    int alloc(uint64_t n) {
        NH_TRY(dalloc(&s.ref, n)); NH_TRY(dalloc(&s.t_last, n)); NH_TRY(dalloc(&s.x, n));
        NH_TRY(dalloc(&s.y, n)); NH_TRY(dalloc(&s.idx, n)); NH_TRY(dalloc(&s.next_t, n));
        NH_TRY(dalloc(&s.next_st, n)); NH_TRY(dalloc(&s.flags, n));
        return NH_OK;
    }
    void release() {
        dfree(s.ref); dfree(s.t_last); dfree(s.x); dfree(s.y); dfree(s.idx);
        dfree(s.next_t); dfree(s.next_st); dfree(s.flags);
    }
};

constexpr uint64_t BYTES_PER_LINEAGE = 2 * 45 + 40; // two SoA copies + event slot
constexpr uint64_t BYTES_PER_BUCKET = 32;

struct Req {
    uint64_t id = 0;
    nh_request p{};
    std::vector<nh_lineage> lineages; // until admission
    uint64_t n_initial = 0;
    uint64_t event_slice = 1;
    uint64_t cache_cap = 0;
    uint64_t footprint = 0;
    // device state
    Soa soa[2];
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
    bool done = false;
    void release() {
        soa[0].release();
        soa[1].release();
        dfree(d_events);
        dfree(d_cache);
    }
};

} // namespace

struct nh_server {
    nh_ctx *ctx = nullptr;
    const nh_tables *tables = nullptr;
    DevScenario S{};
    int kind = 0;
    nh_server_params prm{};
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
        for (Req *r : active) { r->release(); delete r; }
        for (Req *r : finished) { r->release(); delete r; }
        for (auto &kv : polled) { kv.second->release(); delete kv.second; }
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
    NH_HIP(hipSetDevice(ctx->device));
    nh_server *s = new nh_server();
    s->ctx = ctx;
    s->tables = tables;
    s->S = *nh_tables_device_view(tables, &s->kind);
    s->prm = *prm;
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
    if ((rc = dalloc(&s->d_ratio, POISSON_TABLE_N))) return fail(rc);
    if ((rc = dalloc(&s->d_reqs, SERVE_MAX_REQS))) return fail(rc);
    if ((rc = dalloc(&s->d_ctr, SERVE_MAX_REQS))) return fail(rc);
    if ((rc = dalloc(&s->d_head, 1))) return fail(rc);
    if ((rc = dalloc(&s->d_blk, soa_compact_blocks(cap) + 1))) return fail(rc);
    if ((rc = dalloc(&s->d_sorted, cap))) return fail(rc);
    s->sb.capacity = cap;
    if ((rc = dalloc(&s->sb.keys[0], cap)) || (rc = dalloc(&s->sb.keys[1], cap)) ||
        (rc = dalloc(&s->sb.vals[0], cap)) || (rc = dalloc(&s->sb.vals[1], cap)) ||
        (rc = dalloc(&s->sb.hist, sort_hist_words(cap))) || (rc = dalloc(&s->sb.minmax, SORT_MINMAX_WORDS)))
        return fail(rc);
    s->h_reqs.resize(SERVE_MAX_REQS);
    s->h_ctr.resize(SERVE_MAX_REQS);
    *out = s;
    return NH_OK;
}

extern "C" int nh_server_submit(nh_server *s, const nh_request *rq, uint64_t *id) {
    if (!s || !rq || !id || (!rq->lineages && rq->n_lineages)) return set_error(NH_E_INVALID, "nh_server_submit: null argument");
    if (!(rq->speciation_probability > 0.0 && rq->speciation_probability <= 1.0))
        return set_error(NH_E_INVALID, "speciation probability must lie in (0, 1]");
    if (!(rq->turnover_rate > 0.0) || !(rq->delta_t > 0.0) || rq->step_slice == 0)
        return set_error(NH_E_INVALID, "turnover rate, delta_t and step_slice must be positive");
    // lambda, delta_t and step_slice are shared by every request of a server
    if (!s->shared_set) {
        s->turnover = rq->turnover_rate;
        s->delta_t = rq->delta_t;
        s->step_slice = rq->step_slice;
        s->shared_set = true;
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
    } else if (rq->turnover_rate != s->turnover || rq->delta_t != s->delta_t || rq->step_slice != s->step_slice) {
        return set_error(NH_E_INVALID, "turnover rate, delta_t and step_slice must match the server's first request");
    }
    if (rq->n_lineages > s->prm.max_lineages)
        return set_error(NH_E_CAPACITY, "request of %zu lineages exceeds the server capacity", rq->n_lineages);
    Req *r = new Req();
    r->id = s->next_id++;
    r->p = *rq;
    r->p.lineages = nullptr;
    r->lineages.assign(rq->lineages, rq->lineages + rq->n_lineages);
    r->n_initial = rq->n_lineages;
    uint64_t n = r->n_initial;
    if (rq->event_slice_mode == NH_SLICE_ABSOLUTE) r->event_slice = std::max<uint64_t>(rq->event_slice_capacity, 1);
    else r->event_slice = std::max<uint64_t>(as_u64((double)n * rq->event_slice_factor), 1);
    if (rq->dedup_mode == NH_DEDUP_RELATIVE) r->cache_cap = as_u64((double)n * rq->dedup_factor);
    else if (rq->dedup_mode == NH_DEDUP_ABSOLUTE) r->cache_cap = rq->dedup_capacity;
    r->footprint = n * BYTES_PER_LINEAGE + r->cache_cap * BYTES_PER_BUCKET;
    if (r->footprint > s->prm.hbm_budget_bytes) {
        delete r;
        return set_error(NH_E_CAPACITY, "request footprint exceeds the HBM budget");
    }
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

static int admit(nh_server *s, Req *r) {
    hipStream_t st = s->ctx->stream;
    uint64_t n = r->n_initial;
    NH_TRY(r->soa[0].alloc(n));
    NH_TRY(r->soa[1].alloc(n));
    NH_TRY(dalloc(&r->d_events, n));
    if (r->cache_cap) {
        NH_TRY(dalloc(&r->d_cache, 4 * r->cache_cap));
        launch_fill_u64(r->d_cache, 0, 4 * r->cache_cap, st);
    }
    if (n) {
        std::vector<uint64_t> ref(n), zero(n, 0);
        std::vector<double> t(n);
        std::vector<uint32_t> x(n), y(n), idx(n);
        std::vector<uint8_t> fl(n, 0);
        for (uint64_t k = 0; k < n; k++) {
            const nh_lineage &l = r->lineages[k];
            ref[k] = l.global_reference; t[k] = l.last_event_time; x[k] = l.x; y[k] = l.y; idx[k] = l.index;
        }
        LineageSoA &d = r->soa[0].s;
        NH_HIP(hipMemcpyAsync(d.ref, ref.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.t_last, t.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.next_t, t.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.next_st, zero.data(), n * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.x, x.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.y, y.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.idx, idx.data(), n * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(d.flags, fl.data(), n, hipMemcpyHostToDevice, st));
        NH_HIP(hipStreamSynchronize(st));
    }
    r->lineages.clear();
    r->lineages.shrink_to_fit();
    r->n_in = n;
    r->alive = n;
    r->cur = 0;
    r->result.admit_s = mono_s();
    s->used_bytes += r->footprint;
    s->used_lineages += n;
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

static int retire(nh_server *s, Req *r) {
    r->done = true;
    r->result.status = NH_OK;
    r->result.paused = r->alive > 0;
    r->result.time = r->max_time;
    r->result.steps = r->steps;
    r->result.n_events = r->events.size();
    r->result.done_s = mono_s();
    s->used_bytes -= r->footprint;
    s->used_lineages -= r->n_initial;
    r->release();
    s->finished.push_back(r);
    return NH_OK;
}

extern "C" int nh_server_tick(nh_server *s, uint64_t *n_active, uint64_t *n_queued) {
    if (!s) return set_error(NH_E_INVALID, "null server");
    NH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    // admission in arrival order while memory and lane state allow
    while (!s->queued.empty() && s->active.size() < (size_t)SERVE_MAX_REQS) {
        Req *r = s->queued.front();
        if (s->used_bytes + r->footprint > s->prm.hbm_budget_bytes ||
            s->used_lineages + r->n_initial > s->prm.max_lineages)
            break;
        s->queued.pop_front();
        NH_TRY(admit(s, r));
        if (request_finished(r)) {
            NH_TRY(retire(s, r));
            continue;
        }
        s->active.push_back(r);
    }
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
        d.lin = r->soa[r->cur].s;
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
    if (nr > 0 && total > 0) {
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
        // the round is complete: drop dead slots, sort the round's events
        // (live water-level reporter)
        r->round_open = false;
        if (r->n_in > r->alive) {
            launch_soa_compact(r->soa[r->cur].s, r->soa[r->cur ^ 1].s, r->n_in, s->d_blk, st);
            NH_HIP(hipGetLastError());
            uint32_t live = 0;
            NH_HIP(hipMemcpyAsync(&live, s->d_blk + soa_compact_blocks(r->n_in), 4, hipMemcpyDeviceToHost, st));
            NH_HIP(hipStreamSynchronize(st));
            if (live != r->alive) return set_error(NH_E_STATE, "request %llu: %u live lineages, expected %llu",
                                                  (unsigned long long)r->id, live, (unsigned long long)r->alive);
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
        if (request_finished(r)) NH_TRY(retire(s, r));
        else still.push_back(r);
    }
    s->active.swap(still);
    if (n_active) *n_active = s->active.size();
    if (n_queued) *n_queued = s->queued.size();
    return NH_OK;
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
