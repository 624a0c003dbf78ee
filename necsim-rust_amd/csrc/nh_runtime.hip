// nh_runtime.hip -- C ABI of the HIP Independent algorithm: contexts,
// resident scenario tables and simulation jobs (the water-level loop of
// necsim/impls/no-std/src/parallelisation/independent/monolithic/mod.rs:43-270
// driven on the host, every per-lineage step on the device).
#include <algorithm>
#include <deque>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/necsim_hip.h"
#include "nh_internal.h"
#include "nh_kernels.h"

using namespace nh;

namespace {

int kind_of(int32_t scenario_kind) {
    switch (scenario_kind) {
    case NH_SCENARIO_NON_SPATIAL: return K_NON_SPATIAL;
    case NH_SCENARIO_ALMOST_INFINITE_NORMAL: return K_ALMOST_INFINITE;
    case NH_SCENARIO_SPATIALLY_EXPLICIT: return K_SEPARABLE;
    case NH_SCENARIO_SPATIALLY_EXPLICIT_PACKED: return K_PACKED;
    default: return -1;
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool is_pow2_double(double x) {
    if (!(x > 0.0) || std::isinf(x)) return false;
    uint64_t b = bits(x);
    uint64_t exp = (b >> 52) & 0x7ff;
    return exp != 0 && exp != 0x7ff && (b & 0xfffffffffffffull) == 0;
}

struct LineageBuffers {
    LineageSoA s{};
    uint64_t cap = 0;
    int alloc(uint64_t n) {
        cap = n;
        NH_TRY(dalloc(&s.ref, n));
        NH_TRY(dalloc(&s.t_last, n));
        NH_TRY(dalloc(&s.x, n));
        NH_TRY(dalloc(&s.y, n));
        NH_TRY(dalloc(&s.idx, n));
        NH_TRY(dalloc(&s.next_t, n));
        NH_TRY(dalloc(&s.next_st, n));
        NH_TRY(dalloc(&s.flags, n));
        return NH_OK;
    }
    void release() {
        dfree(s.ref); dfree(s.t_last); dfree(s.x); dfree(s.y); dfree(s.idx);
        dfree(s.next_t); dfree(s.next_st); dfree(s.flags);
    }
    static uint64_t bytes_per_lineage() { return 8 + 8 + 4 + 4 + 4 + 8 + 8 + 1; }
};

// upload host lineages (reference layout) into SoA slots [0, n)
int upload_lineages(const std::vector<nh_lineage> &v, LineageBuffers &b, hipStream_t st) {
    size_t n = v.size();
    if (!n) return NH_OK;
    std::vector<uint64_t> ref(n), nst(n, 0);
    std::vector<double> t(n);
    std::vector<uint32_t> x(n), y(n), idx(n);
    std::vector<uint8_t> fl(n, 0);
    for (size_t j = 0; j < n; j++) {
        ref[j] = v[j].global_reference;
        t[j] = v[j].last_event_time;
        x[j] = v[j].x;
        y[j] = v[j].y;
        idx[j] = v[j].index;
    }
    // the pessimistic next event time of a fresh lineage is its
    // last_event_time (monolithic/mod.rs:94-103)
    NH_HIP(hipMemcpyAsync(b.s.ref, ref.data(), n * 8, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.t_last, t.data(), n * 8, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.next_t, t.data(), n * 8, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.next_st, nst.data(), n * 8, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.x, x.data(), n * 4, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.y, y.data(), n * 4, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.idx, idx.data(), n * 4, hipMemcpyHostToDevice, st));
    NH_HIP(hipMemcpyAsync(b.s.flags, fl.data(), n, hipMemcpyHostToDevice, st));
    NH_HIP(hipStreamSynchronize(st));
    return NH_OK;
}

// download SoA slots (listed by `slots`, or [0, n) if null) into host lineages
int download_lineages(const LineageBuffers &b, const uint32_t *d_slots, uint64_t n,
                      std::vector<nh_lineage> &out, hipStream_t st) {
    out.resize(n);
    if (!n) return NH_OK;
    std::vector<uint64_t> ref(b.cap);
    std::vector<double> t(b.cap);
    std::vector<uint32_t> x(b.cap), y(b.cap), idx(b.cap), slots(n);
    uint64_t m = d_slots ? b.cap : n;
    NH_HIP(hipMemcpyAsync(ref.data(), b.s.ref, m * 8, hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(t.data(), b.s.t_last, m * 8, hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(x.data(), b.s.x, m * 4, hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(y.data(), b.s.y, m * 4, hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(idx.data(), b.s.idx, m * 4, hipMemcpyDeviceToHost, st));
    if (d_slots) NH_HIP(hipMemcpyAsync(slots.data(), d_slots, n * 4, hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    for (uint64_t j = 0; j < n; j++) {
        uint64_t s = d_slots ? slots[j] : j;
        out[j] = nh_lineage{ref[s], t[s], x[s], y[s], idx[s], 0};
    }
    return NH_OK;
}

} // namespace

// ===========================================================================
// Scratch of the allocation-free event sort (nh_sort_events_noalloc): one
// file-scope device array, present on every device the library runs on and
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host it is only a stand-in symbol): nh_init resolves its
// address on each device once, with that device current, and every use goes
// through that address.  The sort's kernels initialise every word they read,
// so nothing carries over from the previous call.
constexpr uint64_t NOALLOC_MAX_EVENTS = 1ull << 20;
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 25ull << 20; // 25 MiB
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

uint64_t align256(uint64_t v) { return (v + 255) / 256 * 256; }

// carves the sort buffers for n events (and the two range words of the
// device-sized variant) out of the scratch array at `base`; returns the bytes used
uint64_t carve_sort_scratch(uintptr_t base, uint64_t n, SortBuffers &sb, unsigned long long **range) {
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uintptr_t p = base + off;
        off += align256(bytes);
        return p;
    };
    sb.keys[0] = reinterpret_cast<unsigned long long *>(take(8 * n));
    sb.keys[1] = reinterpret_cast<unsigned long long *>(take(8 * n));
    sb.vals[0] = reinterpret_cast<uint32_t *>(take(4 * n));
    sb.vals[1] = reinterpret_cast<uint32_t *>(take(4 * n));
    sb.hist = reinterpret_cast<uint32_t *>(take(4 * (uint64_t)sort_hist_words(n)));
    sb.hist_words = sort_hist_words(n);
    sb.minmax = reinterpret_cast<unsigned long long *>(take(8 * (uint64_t)SORT_MINMAX_WORDS));
    *range = reinterpret_cast<unsigned long long *>(take(16));
    sb.capacity = n;
    return off;
}

// the scratch array's address on `device` (current), resolved once
int resolve_scratch(int device, void **out) {
    if (device >= MAX_DEVICES) return set_error(NH_E_INVALID, "device %d above the scratch table", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NH_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return NH_OK;
}
} // namespace

extern "C" int nh_sort_events_noalloc(nh_ctx *ctx, const nh_speciation_event *d_in, size_t n,
                                      int32_t device_sized, nh_speciation_event *d_out) {
    static_assert(sizeof(nh_speciation_event) == sizeof(DevEvent), "event record");
    if (!ctx || (n && (!d_in || !d_out))) return set_error(NH_E_INVALID, "nh_sort_events_noalloc: null argument");
    if (n == 0) return NH_OK;
    if (n > NOALLOC_MAX_EVENTS)
        return set_error(NH_E_CAPACITY, "nh_sort_events_noalloc: %zu events, at most %llu", n,
                         (unsigned long long)NOALLOC_MAX_EVENTS);
    if (d_in == d_out) return set_error(NH_E_INVALID, "nh_sort_events_noalloc: input and output must differ");
    if (!ctx->scratch) return set_error(NH_E_STATE, "nh_sort_events_noalloc: scratch not resolved");
    NH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    SortBuffers sb{};
    unsigned long long *range = nullptr;
    carve_sort_scratch(reinterpret_cast<uintptr_t>(ctx->scratch), n, sb, &range);
    const DevEvent *in = reinterpret_cast<const DevEvent *>(d_in);
    DevEvent *out = reinterpret_cast<DevEvent *>(d_out);
    if (device_sized) {
        // the range words are written by fill kernels, never by a memset
        launch_fill_u64(reinterpret_cast<uint64_t *>(range), 0, 1, st);
        launch_fill_u64(reinterpret_cast<uint64_t *>(range) + 1, n, 1, st);
        sort_events_device_sized(in, range + 1, range, n, out, sb, st);
    } else {
        sort_events(in, n, out, sb, st, false);
    }
    NH_HIP(hipGetLastError());
    NH_HIP(hipStreamSynchronize(st));
    return NH_OK;
}

// ===========================================================================
// context
extern "C" int nh_abi_version(void) { return NH_ABI_VERSION; }

extern "C" int nh_last_error(char *buf, size_t cap) {
    if (!buf || !cap) return NH_E_INVALID;
    std::snprintf(buf, cap, "%s", nh::g_last_error);
    return NH_OK;
}

extern "C" int nh_init(int device, nh_ctx **out) {
    if (!out) return set_error(NH_E_INVALID, "nh_init: out is null");
    *out = nullptr;
    int count = 0;
    NH_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return set_error(NH_E_INVALID, "device %d of %d", device, count);
    NH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    NH_HIP(hipGetDeviceProperties(&prop, device));
    nh_ctx *c = new nh_ctx();
    c->device = device;
    c->cus = (uint32_t)prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_error(NH_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    {
        // the scratch of nh_sort_events_noalloc must hold the sort buffers of
        // its largest input
        SortBuffers sb{};
        unsigned long long *range = nullptr;
        if (carve_sort_scratch(0, NOALLOC_MAX_EVENTS, sb, &range) > NOALLOC_SCRATCH_BYTES) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return set_error(NH_E_STATE, "scratch array smaller than the sort buffers");
        }
    }
    int rs = resolve_scratch(device, &c->scratch);
    if (rs != NH_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rs;
    }
    *out = c;
    return NH_OK;
}

extern "C" int nh_ctx_destroy(nh_ctx *ctx) {
    if (!ctx) return NH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NH_OK;
}

extern "C" int nh_device_info(nh_ctx *ctx, uint32_t *cus, char *name, size_t cap) {
    if (!ctx) return set_error(NH_E_INVALID, "null context");
    if (cus) *cus = ctx->cus;
    if (name && cap) std::snprintf(name, cap, "%s", ctx->name.c_str());
    return NH_OK;
}

// ===========================================================================
// tables
extern "C" int nh_tables_create(nh_ctx *ctx, const nh_scenario *s, nh_tables **out) {
    if (!ctx || !s || !out) return set_error(NH_E_INVALID, "nh_tables_create: null argument");
    *out = nullptr;
    int kind = kind_of(s->kind);
    if (kind < 0) return set_error(NH_E_INVALID, "unknown scenario kind %d", s->kind);
    NH_HIP(hipSetDevice(ctx->device));
    nh_tables *t = new nh_tables();
    t->ctx = ctx;
    t->kind = kind;
    t->width = s->width;
    t->height = s->height;
    t->ox = s->origin_x;
    t->oy = s->origin_y;
    t->deme = s->deme;
    t->sigma = s->sigma;
    DevScenario &d = t->dev;
    d.kind = kind;
    d.width = s->width;
    d.height = s->height;
    d.ox = s->origin_x;
    d.oy = s->origin_y;
    d.deme = s->deme;
    d.sigma = s->sigma;
    d.area = (uint64_t)s->width * s->height;
    d.area_f = (double)d.area;
    d.deme_f = (double)s->deme;
    auto fail = [&](int code) {
        nh_tables_destroy(t);
        return code;
    };
    if (kind == K_NON_SPATIAL) {
        if (!s->width || !s->height || !s->deme) return fail(set_error(NH_E_INVALID, "empty non-spatial habitat"));
        if (d.area > 0xffffffffull) return fail(set_error(NH_E_INVALID, "non-spatial extent above 2^32 cells"));
    } else if (kind == K_ALMOST_INFINITE) {
        if (!(s->sigma >= 0.0)) return fail(set_error(NH_E_INVALID, "negative sigma"));
    } else {
        uint64_t cells = d.area;
        if (!cells || cells > 0xffffffffull || !s->habitat || !s->alias_ranges)
            return fail(set_error(NH_E_INVALID, "in-memory scenario needs habitat and alias ranges"));
        if (kind == K_SEPARABLE && !s->self_dispersal)
            return fail(set_error(NH_E_INVALID, "separable scenario needs self-dispersal"));
        if (s->n_atoms > 0xffffffffull) return fail(set_error(NH_E_INVALID, "more than 2^32 atoms"));
        t->habitat.assign(s->habitat, s->habitat + cells);
        std::vector<CellRec> rec(cells);
        uint64_t prefix = 0;
        for (uint64_t c = 0; c < cells; c++) {
            const nh_alias_range &r = s->alias_ranges[c];
            if (r.end < r.start || r.end > s->n_atoms)
                return fail(set_error(NH_E_INVALID, "alias range of cell %llu out of bounds", (unsigned long long)c));
            if (s->habitat[c] > 0 && r.end == r.start &&
                !(kind == K_SEPARABLE && s->self_dispersal[c].self_dispersal == 1.0))
                return fail(set_error(NH_E_CONTRACT, "habitable cell %llu without dispersal", (unsigned long long)c));
            rec[c].prefix = prefix;
            rec[c].habitat = s->habitat[c];
            rec[c].atom_start = (uint32_t)r.start;
            rec[c].atom_len = (uint32_t)(r.end - r.start);
            if (kind == K_SEPARABLE) {
                rec[c].self_p = s->self_dispersal[c].self_dispersal;
                uint64_t nse = s->self_dispersal[c].non_self_dispersal_event;
                if (nse >= cells) return fail(set_error(NH_E_INVALID, "non-self dispersal event out of bounds"));
                rec[c].non_self = (uint32_t)nse;
            } else {
                rec[c].self_p = 0.0;
                rec[c].non_self = (uint32_t)c;
            }
            prefix += s->habitat[c];
        }
        if (prefix == 0) return fail(set_error(NH_E_INVALID, "habitat map is empty"));
        std::vector<Atom> at(s->n_atoms);
        for (size_t j = 0; j < s->n_atoms; j++) {
            if (s->atoms[j].e >= cells || s->atoms[j].k >= cells)
                return fail(set_error(NH_E_INVALID, "alias event out of bounds"));
            at[j] = Atom{s->atoms[j].u, (uint32_t)s->atoms[j].e, (uint32_t)s->atoms[j].k};
        }
        int r = dalloc(&t->d_cells, cells);
        if (r) return fail(r);
        r = dalloc(&t->d_atoms, at.size());
        if (r) return fail(r);
        hipError_t e = hipMemcpy(t->d_cells, rec.data(), cells * sizeof(CellRec), hipMemcpyHostToDevice);
        if (e == hipSuccess && !at.empty())
            e = hipMemcpy(t->d_atoms, at.data(), at.size() * sizeof(Atom), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(set_error(NH_E_HIP, "table upload: %s", hipGetErrorString(e)));
        t->n_atoms = at.size();
        t->bytes = cells * sizeof(CellRec) + at.size() * sizeof(Atom);
        d.cells = t->d_cells;
        d.atoms = t->d_atoms;
    }
    *out = t;
    return NH_OK;
}

extern "C" int nh_tables_destroy(nh_tables *t) {
    if (!t) return NH_OK;
    dfree(t->d_cells);
    dfree(t->d_atoms);
    delete t;
    return NH_OK;
}

const DevScenario *nh_tables_device_view(const nh_tables *t, int *kind) {
    *kind = t->kind;
    return &t->dev;
}

extern "C" int nh_tables_footprint(const nh_tables *t, uint64_t *bytes) {
    if (!t || !bytes) return set_error(NH_E_INVALID, "null argument");
    *bytes = t->bytes;
    return NH_OK;
}

// IndependentActiveLineageSampler's resume filters
// (active_lineage_sampler/independent/mod.rs:98-117): a lineage whose
// location is not habitable (outside the extent, or a cell of habitat 0,
// necsim/core/src/cogs/habitat.rs:25-27) is OutOfHabitat, one whose index is
// not below the location's habitat is OutOfDeme.
int nh_lineage_class(const nh_tables *t, const nh_lineage &l) {
    if (t->kind == K_ALMOST_INFINITE) return l.index == 0 ? NH_LINEAGE_OK : NH_LINEAGE_OUT_OF_DEME;
    uint32_t cx = l.x - t->ox, cy = l.y - t->oy;
    if (cx >= t->width || cy >= t->height) return NH_LINEAGE_OUT_OF_HABITAT;
    uint32_t deme = t->kind == K_NON_SPATIAL ? t->deme : t->habitat[(uint64_t)cy * t->width + cx];
    if (deme == 0) return NH_LINEAGE_OUT_OF_HABITAT;
    return l.index < deme ? NH_LINEAGE_OK : NH_LINEAGE_OUT_OF_DEME;
}

extern "C" int nh_classify_lineages(const nh_tables *t, const nh_lineage *l, size_t n, int32_t *kinds,
                                    size_t *n_exceptional) {
    if (!t || (n && (!l || !kinds))) return set_error(NH_E_INVALID, "nh_classify_lineages: null argument");
    size_t bad = 0;
    for (size_t k = 0; k < n; k++) {
        kinds[k] = nh_lineage_class(t, l[k]);
        bad += kinds[k] != NH_LINEAGE_OK;
    }
    if (n_exceptional) *n_exceptional = bad;
    return NH_OK;
}

// ===========================================================================
// jobs
static_assert(sizeof(nh_packed_event) == sizeof(DevPacked) &&
                  offsetof(nh_packed_event, coalescence) == offsetof(DevPacked, coalescence) &&
                  offsetof(nh_packed_event, target_x) == offsetof(DevPacked, tx),
              "nh_packed_event is the device PackedEvent record");

struct nh_job {
    nh_ctx *ctx = nullptr;
    const nh_tables *tables = nullptr;
    nh_job_params p{};
    DevParams P{};
    double *d_ratio = nullptr;
    std::vector<nh_lineage> pending;
    // nh_job_add_origins: lineages sampled on the device, waiting for the start
    LineageBuffers origins{};
    uint64_t n_origins = 0;
    bool has_origins = false;
    bool started = false, finished = false;
    uint64_t n_initial = 0;
    uint64_t event_slice = 1;
    uint64_t cache_cap = 0;
    double level_time = 0.0;
    double max_time = 0.0;
    uint64_t rounds = 0, launches = 0;
    uint64_t steps = 0;

    LineageBuffers lin[2];
    int cur = 0;
    uint64_t cur_n = 0; // slots of the current buffer (plain: queue length)
    uint64_t alive = 0; // live lineages (persistent: cur_n minus dead slots)
    uint32_t *d_blk = nullptr; // SoA compaction block counts
    Counters *d_ctr = nullptr;
    RoundDesc *d_round = nullptr;
    DevEvent *d_events = nullptr, *d_sorted = nullptr;
    uint64_t n_events = 0;
    std::vector<nh_speciation_event> ready;
    size_t ready_pos = 0;
    // dispersal reporting: one round's PackedEvents at a time
    bool disp = false;
    bool failed = false;
    DevPacked *d_packed = nullptr, *d_packed_sorted = nullptr;
    uint64_t packed_cap = 0;
    std::vector<nh_packed_event> ready_packed;
    size_t packed_pos = 0;
    // the last round's sorted events not drained yet, still in d_packed_sorted
    // [dev_pos, dev_n): drained straight into the caller's buffer, moved to
    // ready_packed before the next round overwrites them
    uint64_t dev_pos = 0, dev_n = 0;
    EventCounts *d_counts = nullptr; // device_event_counter jobs
    SortBuffers sb{};
    uint64_t *d_cache = nullptr;

    // plain form
    uint32_t *d_S = nullptr, *d_F = nullptr, *d_pass[2] = {nullptr, nullptr};
    uint8_t *d_status = nullptr;
    double *s_u = nullptr, *s_t = nullptr;
    uint32_t *s_x = nullptr, *s_y = nullptr, *s_i = nullptr;
    unsigned long long *s_bucket = nullptr;
    uint32_t *d_head = nullptr, *d_link = nullptr;
    CompactBuffers cb{};
    PlainLoop *d_loop = nullptr; // batched passes

    // task-list form: the slow / fast queues and the dedup cache on the host
    struct Queued {
        nh_lineage l;
        double next; // pessimistic next event time (monolithic/mod.rs:94-103)
    };
    std::deque<Queued> slow_q, fast_q;
    struct HostSample {
        bool some;
        double u, t;
        uint32_t x, y, i;
    };
    std::vector<HostSample> host_cache;
    uint64_t task_cap = 0;

    // persistent form
    int threads = 256;
    int blocks = 0;
    bool coop = false; // cooperative next-event search
    bool pooled = false; // the pooled kernel (nh_job_params.pool)

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // persistent round, input buffer 0 / 1
    double kernel_ms = 0.0;
    uint64_t timed = 0;
    RoundDesc round_host{};

    ~nh_job() {
        for (auto &l : lin) l.release();
        origins.release();
        dfree(d_ratio); dfree(d_ctr); dfree(d_round); dfree(d_events); dfree(d_sorted);
        dfree(d_packed); dfree(d_packed_sorted); dfree(d_counts);
        dfree(sb.keys[0]); dfree(sb.keys[1]); dfree(sb.vals[0]); dfree(sb.vals[1]);
        dfree(sb.hist); dfree(sb.minmax); dfree(d_cache);
        dfree(d_S); dfree(d_F); dfree(d_pass[0]); dfree(d_pass[1]); dfree(d_status);
        dfree(s_u); dfree(s_t); dfree(s_x); dfree(s_y); dfree(s_i); dfree(s_bucket);
        dfree(d_head); dfree(d_link); dfree(cb.blk_a); dfree(cb.blk_b); dfree(cb.counts); dfree(d_blk);
        dfree(d_loop);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g);
    }
};

extern "C" int nh_job_create(nh_ctx *ctx, const nh_tables *tables, const nh_job_params *p, nh_job **out) {
    if (!ctx || !tables || !p || !out) return set_error(NH_E_INVALID, "nh_job_create: null argument");
    *out = nullptr;
    if (!(p->speciation_probability > 0.0 && p->speciation_probability <= 1.0))
        return set_error(NH_E_INVALID, "speciation probability must lie in (0, 1]");
    if (!(p->turnover_rate > 0.0) || std::isinf(p->turnover_rate))
        return set_error(NH_E_INVALID, "turnover rate must be positive");
    if (!(p->delta_t > 0.0) || std::isinf(p->delta_t)) return set_error(NH_E_INVALID, "delta_t must be positive");
    if (p->step_slice == 0) return set_error(NH_E_INVALID, "step_slice must be positive");
    if (p->exec_fill > 8) return set_error(NH_E_INVALID, "exec_fill is a number of eighths (0..8)");
    if (p->coop_search > 2) return set_error(NH_E_INVALID, "coop_search is 0 (automatic), 1 (off) or 2 (on)");
    if (p->device_event_counter && !p->report_dispersal)
        return set_error(NH_E_INVALID, "device_event_counter counts the events of a report_dispersal job");
    if (p->form != NH_FORM_PLAIN && p->form != NH_FORM_PERSISTENT && p->form != NH_FORM_TASK_LIST)
        return set_error(NH_E_INVALID, "unknown form");
    if (p->form == NH_FORM_TASK_LIST && p->report_dispersal)
        return set_error(NH_E_INVALID, "the task-list form reports speciation events only");
    if (p->dedup_mode < NH_DEDUP_NONE || p->dedup_mode > NH_DEDUP_ABSOLUTE) return set_error(NH_E_INVALID, "dedup mode");
    if (p->dedup_mode == NH_DEDUP_RELATIVE && !(p->dedup_factor > 0.0)) return set_error(NH_E_INVALID, "dedup factor");
    if (p->event_slice_mode == NH_SLICE_RELATIVE && !(p->event_slice_factor > 0.0))
        return set_error(NH_E_INVALID, "event slice factor");
    if (p->event_slice_mode == NH_SLICE_ABSOLUTE && p->event_slice_capacity == 0)
        return set_error(NH_E_INVALID, "event slice capacity");
    if (p->event_slice_mode != NH_SLICE_RELATIVE && p->event_slice_mode != NH_SLICE_ABSOLUTE)
        return set_error(NH_E_INVALID, "event slice mode");
    if (!(p->resume_after >= 0.0)) return set_error(NH_E_INVALID, "resume_after must be non-negative");
    if (p->block_size && (p->block_size % 64 || p->block_size > 256))
        return set_error(NH_E_INVALID, "block size must be 64, 128, 192 or 256");
    if (p->pool != 0 && p->pool != 128 && p->pool != 192)
        return set_error(NH_E_INVALID, "pool is 0 (off), 128 or 192 lineages per wave");
    NH_HIP(hipSetDevice(ctx->device));
    nh_job *j = new nh_job();
    j->ctx = ctx;
    j->tables = tables;
    j->p = *p;
    j->disp = p->report_dispersal != 0;
    *out = j;
    return NH_OK;
}

extern "C" int nh_job_add_lineages(nh_job *j, const nh_lineage *l, size_t n) {
    if (!j || (!l && n)) return set_error(NH_E_INVALID, "nh_job_add_lineages: null argument");
    if (j->started) return set_error(NH_E_STATE, "lineages must be added before the first round");
    if (j->has_origins) return set_error(NH_E_STATE, "the job's lineages were sampled on the device");
    for (size_t k = 0; k < n; k++) {
        if (nh_lineage_class(j->tables, l[k]) != NH_LINEAGE_OK)
            return set_error(NH_E_OUT_OF_DEME, "lineage %zu at (%u, %u, %u) is outside the habitat or its deme", k,
                             l[k].x, l[k].y, l[k].index);
        if (!(l[k].last_event_time >= 0.0)) return set_error(NH_E_INVALID, "negative last_event_time");
    }
    j->pending.insert(j->pending.end(), l, l + n);
    return NH_OK;
}

extern "C" int nh_job_add_origins(nh_job *j, double percentage, uint32_t rank, uint32_t size, uint32_t cx,
                                  uint32_t cy, uint32_t radius) {
    if (!j) return set_error(NH_E_INVALID, "nh_job_add_origins: null job");
    if (j->started) return set_error(NH_E_STATE, "lineages must be added before the first round");
    if (j->has_origins || !j->pending.empty())
        return set_error(NH_E_STATE, "a job takes host lineages or one device origin sample, not both");
    if (!(percentage == percentage)) return set_error(NH_E_INVALID, "percentage is NaN");
    NH_HIP(hipSetDevice(j->ctx->device));
    hipStream_t st = j->ctx->stream;
    OriginPlan plan;
    NH_TRY(plan.prepare(j->tables, percentage, rank, size, cx, cy, radius, st));
    if (j->p.form == NH_FORM_TASK_LIST) {
        // the task-list form keeps its queues on the host
        nh_lineage *d = nullptr;
        NH_TRY(dalloc(&d, plan.n));
        int rc = plan.emit(nullptr, d);
        if (rc == NH_OK) {
            j->pending.resize(plan.n);
            hipError_t e = hipMemcpy(j->pending.data(), d, plan.n * sizeof(nh_lineage), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = set_error(NH_E_HIP, "origin download: %s", hipGetErrorString(e));
        }
        dfree(d);
        return rc;
    }
    LineageBuffers b{};
    int rc = b.alloc(plan.n);
    if (rc == NH_OK) rc = plan.emit(&b.s, nullptr);
    if (rc != NH_OK) {
        b.release();
        return rc;
    }
    j->origins = b;
    j->n_origins = plan.n;
    j->has_origins = true;
    return NH_OK;
}

extern "C" int nh_sample_origins_device(nh_ctx *ctx, const nh_tables *tables, double percentage, uint32_t rank,
                                        uint32_t size, uint32_t cx, uint32_t cy, uint32_t radius, nh_lineage *d_out,
                                        size_t cap, size_t *n_out) {
    if (!ctx || !tables || !n_out) return set_error(NH_E_INVALID, "nh_sample_origins_device: null argument");
    if (!(percentage == percentage)) return set_error(NH_E_INVALID, "percentage is NaN");
    NH_HIP(hipSetDevice(ctx->device));
    OriginPlan plan;
    NH_TRY(plan.prepare(tables, percentage, rank, size, cx, cy, radius, ctx->stream));
    *n_out = plan.n;
    if (!d_out) return NH_OK;
    if (cap < plan.n) return set_error(NH_E_CAPACITY, "origin buffer holds %zu of %llu lineages", cap,
                                       (unsigned long long)plan.n);
    return plan.emit(nullptr, d_out);
}

static int job_start(nh_job *j) {
    hipStream_t st = j->ctx->stream;
    const nh_job_params &p = j->p;
    uint64_t n = j->has_origins ? j->n_origins : j->pending.size();
    j->n_initial = n;
    j->started = true;

    // DevParams: the Poisson constants are evaluated once, on the host, with
    // the same MathsCore as the device (poisson.rs:39-42)
    DevParams &P = j->P;
    P.seed = p.rng_seed;
    P.nu = p.speciation_probability;
    P.nu_m = draw_threshold_below(P.nu);
    P.lambda_dt = p.turnover_rate * p.delta_t;
    P.p0 = nh::exp(-P.lambda_dt);
    P.lambda_dt_sqrt = sqrt_(P.lambda_dt);
    P.delta_t = p.delta_t;
    P.dt_pow2 = is_pow2_double(p.delta_t);
    P.inv_delta_t = P.dt_pow2 ? 1.0 / p.delta_t : 0.0; // exact for powers of two
    P.min_event_time = p.resume_after;
    P.step_slice = p.step_slice;
    std::vector<double> ratio(POISSON_TABLE_N, 0.0);
    P.prod_zero_at = poisson_table(P.lambda_dt, P.p0, ratio.data());
    NH_TRY(dalloc(&j->d_ratio, POISSON_TABLE_N));
    NH_HIP(hipMemcpy(j->d_ratio, ratio.data(), POISSON_TABLE_N * sizeof(double), hipMemcpyHostToDevice));
    P.ratio = j->d_ratio;

    // EventSlice::capacity and DedupCache::construct (parallelisation/independent/mod.rs:34-82)
    if (p.event_slice_mode == NH_SLICE_ABSOLUTE) j->event_slice = p.event_slice_capacity;
    else j->event_slice = std::max<uint64_t>(as_u64((double)n * p.event_slice_factor), 1);
    if (p.dedup_mode == NH_DEDUP_RELATIVE) j->cache_cap = as_u64((double)n * p.dedup_factor);
    else if (p.dedup_mode == NH_DEDUP_ABSOLUTE) j->cache_cap = p.dedup_capacity;
    else j->cache_cap = 0;

    // max_time / level_time start at the max / min last_event_time (monolithic/mod.rs:120-133)
    // (device-sampled origins all start at 0, the initial max / min)
    if (!j->pending.empty()) {
        double mx = j->pending[0].last_event_time, mn = mx;
        for (auto &l : j->pending) {
            mx = l.last_event_time > mx ? l.last_event_time : mx;
            mn = l.last_event_time < mn ? l.last_event_time : mn;
        }
        j->max_time = mx;
        j->level_time = mn;
    }

    NH_TRY(dalloc(&j->d_ctr, 1));
    NH_TRY(dalloc(&j->d_round, 1));
    NH_HIP(hipMemsetAsync(j->d_ctr, 0, sizeof(Counters), st));
    if (j->cache_cap) {
        NH_TRY(dalloc(&j->d_cache, 4 * j->cache_cap));
        launch_fill_u64(j->d_cache, 0, 4 * j->cache_cap, st);
    }
    NH_HIP(hipEventCreate(&j->ev0));
    NH_HIP(hipEventCreate(&j->ev1));

    if (p.form == NH_FORM_TASK_LIST) {
        // the lineages stay on the host, in the slow queue with their
        // last_event_time as the pessimistic next event time
        j->task_cap = p.task_list_size ? p.task_list_size : 4096;
        NH_TRY(j->lin[0].alloc(j->task_cap));
        for (auto &l : j->pending) j->slow_q.push_back({l, l.last_event_time});
    } else if (j->has_origins) {
        j->lin[0] = j->origins; // already in the lineage store layout
        j->origins = LineageBuffers{};
    } else {
        NH_TRY(j->lin[0].alloc(n));
        NH_TRY(upload_lineages(j->pending, j->lin[0], st));
    }
    j->cur = 0;
    j->cur_n = n;
    j->alive = n;
    j->threads = p.block_size ? (int)p.block_size : (p.form == NH_FORM_TASK_LIST ? 64 : 256);
    if (p.form == NH_FORM_TASK_LIST) {
        const uint64_t T = j->task_cap;
        NH_TRY(dalloc(&j->d_status, T));
        NH_TRY(dalloc(&j->d_S, T));
        std::vector<uint32_t> iota(T);
        for (uint64_t k = 0; k < T; k++) iota[k] = (uint32_t)k;
        NH_HIP(hipMemcpy(j->d_S, iota.data(), T * 4, hipMemcpyHostToDevice));
        if (j->cache_cap) {
            NH_TRY(dalloc(&j->s_u, T));
            NH_TRY(dalloc(&j->s_t, T));
            NH_TRY(dalloc(&j->s_x, T));
            NH_TRY(dalloc(&j->s_y, T));
            NH_TRY(dalloc(&j->s_i, T));
            j->host_cache.assign(j->cache_cap, nh_job::HostSample{false, 0.0, 0.0, 0, 0, 0});
        }
    } else if (p.form == NH_FORM_PERSISTENT) {
        NH_TRY(j->lin[1].alloc(n));
        NH_TRY(dalloc(&j->d_blk, soa_compact_blocks(n) + 1));
        int per_cu = persistent_max_blocks_per_cu(j->tables->kind, j->threads, j->cache_cap != 0, j->disp, false);
        if (p.blocks_per_cu) per_cu = std::min<int>(per_cu, (int)p.blocks_per_cu);
        // the cooperative search pays off where waves will hold few live
        // lineages: coalescing populations (dedup) and small samples
        const uint64_t lanes = (uint64_t)per_cu * j->ctx->cus * (uint64_t)j->threads;
        if (p.coop_search == 2) j->coop = true;
        else if (p.coop_search == 0) j->coop = !j->disp && (j->cache_cap != 0 || n < lanes);
        if (j->coop) {
            per_cu = persistent_max_blocks_per_cu(j->tables->kind, j->threads, j->cache_cap != 0, false, true);
            if (p.blocks_per_cu) per_cu = std::min<int>(per_cu, (int)p.blocks_per_cu);
        }
        // the pooled kernel: dense rounds without dedup cache, one wave per block
        j->pooled = p.pool != 0 && j->cache_cap == 0 && p.coop_search != 2;
        if (j->pooled) {
            j->coop = false;
            j->threads = 64;
            per_cu = pooled_max_blocks_per_cu(j->tables->kind, p.pool, j->disp);
            if (p.blocks_per_cu) per_cu = std::min<int>(per_cu, (int)p.blocks_per_cu);
        }
        j->blocks = per_cu * (int)j->ctx->cus;
    } else {
        NH_TRY(dalloc(&j->d_S, n));
        NH_TRY(dalloc(&j->d_F, n));
        NH_TRY(dalloc(&j->d_pass[0], n));
        NH_TRY(dalloc(&j->d_pass[1], n));
        NH_TRY(dalloc(&j->d_status, n));
        std::vector<uint32_t> iota(n);
        for (uint64_t k = 0; k < n; k++) iota[k] = (uint32_t)k;
        if (n) NH_HIP(hipMemcpy(j->d_S, iota.data(), n * 4, hipMemcpyHostToDevice));
        uint32_t nb = compact_blocks(n);
        NH_TRY(dalloc(&j->cb.blk_a, nb + 1));
        NH_TRY(dalloc(&j->cb.blk_b, nb + 1));
        NH_TRY(dalloc(&j->cb.counts, 2));
        NH_TRY(dalloc(&j->d_loop, 1));
        if (j->cache_cap) {
            NH_TRY(dalloc(&j->s_u, n));
            NH_TRY(dalloc(&j->s_t, n));
            NH_TRY(dalloc(&j->s_x, n));
            NH_TRY(dalloc(&j->s_y, n));
            NH_TRY(dalloc(&j->s_i, n));
            NH_TRY(dalloc(&j->s_bucket, n));
            NH_TRY(dalloc(&j->d_link, n));
            NH_TRY(dalloc(&j->d_head, j->cache_cap));
            launch_fill_u32(j->d_head, 0xffffffffu, j->cache_cap, st);
        }
    }
    // speciation-only reporting appends to one buffer for the whole run (each
    // lineage speciates at most once); dispersal reporting keeps one round's
    // events, about event_slice of them on average (the level advances by
    // event_slice / sum(lambda))
    uint64_t sort_cap = n;
    if (j->disp) {
        // persistent waves reserve slots EVENT_CHUNK at a time: each may leave
        // up to a chunk unused per round
        uint64_t waves = p.form == NH_FORM_PERSISTENT ? (uint64_t)j->blocks * (uint64_t)(j->threads / 64) : 0;
        j->packed_cap = p.event_buffer ? p.event_buffer : 2 * j->event_slice + n + 65536 + waves * EVENT_CHUNK;
        NH_TRY(dalloc(&j->d_packed, j->packed_cap));
        NH_TRY(dalloc(&j->d_packed_sorted, j->packed_cap));
        sort_cap = j->packed_cap;
        if (p.device_event_counter) {
            NH_TRY(dalloc(&j->d_counts, 1));
            NH_HIP(hipMemsetAsync(j->d_counts, 0, sizeof(EventCounts), st));
        }
    } else {
        NH_TRY(dalloc(&j->d_events, n));
        NH_TRY(dalloc(&j->d_sorted, n));
    }
    j->sb.capacity = sort_cap;
    NH_TRY(dalloc(&j->sb.keys[0], sort_cap));
    NH_TRY(dalloc(&j->sb.keys[1], sort_cap));
    NH_TRY(dalloc(&j->sb.vals[0], sort_cap));
    NH_TRY(dalloc(&j->sb.vals[1], sort_cap));
    j->sb.hist_words = sort_hist_words_staged(sort_cap);
    NH_TRY(dalloc(&j->sb.hist, j->sb.hist_words));
    NH_TRY(dalloc(&j->sb.minmax, SORT_MINMAX_WORDS));
    j->pending.clear();
    j->pending.shrink_to_fit();
    NH_HIP(hipStreamSynchronize(st));
    return NH_OK;
}

static int read_counters(nh_job *j, Counters &c) {
    NH_HIP(hipMemcpyAsync(&c, j->d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, j->ctx->stream));
    NH_HIP(hipStreamSynchronize(j->ctx->stream));
    return NH_OK;
}

static void account_kernel_time(nh_job *j) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, j->ev0, j->ev1) == hipSuccess) {
        j->kernel_ms += ms;
        j->timed++;
    }
}

// one water-level round of the plain form: one launch per step slice over the
// reference's FIFO queue order (monolithic/mod.rs:183-251).  The first pass
// is ordered on the host; with plain_batch > 1 the later passes run in
// batches whose pass size and fast-list length stay on the device (one host
// synchronisation per batch instead of two per pass).
static int plain_round(nh_job *j) {
    hipStream_t st = j->ctx->stream;
    uint64_t m = j->cur_n;
    const uint32_t *plist = j->d_S;
    uint64_t f_len = 0;
    int pb = 0;
    const uint32_t batch = j->p.plain_batch ? j->p.plain_batch : 16;
    auto pass_args = [&](uint64_t m_or_bound, const unsigned long long *m_dev) {
        PlainArgs a{};
        a.S = j->tables->dev;
        a.P = j->P;
        a.L = j->lin[0].s;
        a.pass = plist;
        a.m = m_or_bound;
        a.m_dev = m_dev;
        a.round = j->d_round;
        a.status = j->d_status;
        a.s_u = j->s_u; a.s_t = j->s_t; a.s_x = j->s_x; a.s_y = j->s_y; a.s_i = j->s_i;
        a.s_bucket = j->s_bucket;
        a.cache_cap = j->cache_cap;
        a.events = j->d_events;
        a.ctr = j->d_ctr;
        a.packed = j->d_packed;
        a.packed_cap = j->packed_cap;
        return a;
    };
    auto dedup = [&](uint64_t m_or_bound, const unsigned long long *m_dev) {
        if (!j->cache_cap) return;
        launch_dedup_push(j->d_status, j->s_bucket, j->d_head, j->d_link, m_or_bound, m_dev, st);
        launch_dedup_resolve(j->d_status, j->s_bucket, j->s_u, j->s_t, j->s_x, j->s_y, j->s_i, j->d_head,
                             j->d_link, j->d_cache, m_or_bound, m_dev, st);
    };
    auto overflowed = [&](bool *out) -> int {
        // a dispersal-reporting lineage that found the event buffer full made no progress
        *out = false;
        if (!j->disp) return NH_OK;
        Counters c;
        NH_TRY(read_counters(j, c));
        *out = c.ev_overflow != 0;
        return NH_OK;
    };
    for (uint64_t pass = 1; m > 0; pass++) {
        uint32_t *next = j->d_pass[pb];
        if (pass > 1 && batch > 1) {
            // batched passes: launch bounds from the last known pass size
            const uint64_t bound = m;
            PlainLoop loop{m, f_len};
            NH_HIP(hipMemcpyAsync(j->d_loop, &loop, sizeof(loop), hipMemcpyHostToDevice, st));
            NH_HIP(hipEventRecord(j->ev0, st));
            for (uint32_t b = 0; b < batch; b++) {
                launch_plain(j->tables->kind, pass_args(bound, &j->d_loop->m), j->threads, st);
                dedup(bound, &j->d_loop->m);
                launch_compact_loop(j->d_status, plist, bound, j->d_loop, j->d_pass[pb], j->d_F, j->cb, st);
                NH_HIP(hipGetLastError());
                plist = j->d_pass[pb];
                pb ^= 1;
                j->launches++;
            }
            NH_HIP(hipEventRecord(j->ev1, st));
            NH_HIP(hipMemcpyAsync(&loop, j->d_loop, sizeof(loop), hipMemcpyDeviceToHost, st));
            NH_HIP(hipStreamSynchronize(st));
            account_kernel_time(j);
            m = loop.m;
            f_len = loop.f_len;
            pass += batch - 1;
            bool ov;
            NH_TRY(overflowed(&ov));
            if (ov) return NH_OK;
            continue;
        }
        NH_HIP(hipEventRecord(j->ev0, st));
        launch_plain(j->tables->kind, pass_args(m, nullptr), j->threads, st);
        NH_HIP(hipGetLastError());
        NH_HIP(hipEventRecord(j->ev1, st));
        j->launches++;
        bool ov;
        NH_TRY(overflowed(&ov));
        if (ov) return NH_OK;
        dedup(m, nullptr);
        NH_HIP(hipGetLastError());
        uint64_t next_m;
        if (pass == 1) {
            // The first pass pops every lineage of the round: skipped ones go to
            // the fast queue at once, a processed one only when the next
            // processed lineage is popped (monolithic/mod.rs:186-219).
            std::vector<uint8_t> status(m);
            std::vector<uint32_t> slots(m), nextv, fv;
            NH_HIP(hipMemcpyAsync(status.data(), j->d_status, m, hipMemcpyDeviceToHost, st));
            NH_HIP(hipMemcpyAsync(slots.data(), plist, m * 4, hipMemcpyDeviceToHost, st));
            NH_HIP(hipStreamSynchronize(st));
            nextv.reserve(m);
            fv.reserve(m);
            int64_t pending = -1;
            for (uint64_t q = 0; q < m; q++) {
                uint8_t k = status[q] & ST_KIND_MASK;
                if (k == ST_SKIP) {
                    fv.push_back(slots[q]);
                    continue;
                }
                if (pending >= 0 && (status[pending] & ST_KIND_MASK) == ST_FAST) fv.push_back(slots[pending]);
                pending = (int64_t)q;
                if (k == ST_SLOW) nextv.push_back(slots[q]);
            }
            if (pending >= 0 && (status[pending] & ST_KIND_MASK) == ST_FAST) fv.push_back(slots[pending]);
            if (!nextv.empty())
                NH_HIP(hipMemcpyAsync(next, nextv.data(), nextv.size() * 4, hipMemcpyHostToDevice, st));
            if (!fv.empty()) NH_HIP(hipMemcpyAsync(j->d_F, fv.data(), fv.size() * 4, hipMemcpyHostToDevice, st));
            NH_HIP(hipStreamSynchronize(st));
            next_m = nextv.size();
            f_len = fv.size();
        } else {
            launch_compact(j->d_status, plist, m, next, j->d_F, f_len, j->cb, st);
            NH_HIP(hipGetLastError());
            unsigned int counts[2];
            NH_HIP(hipMemcpyAsync(counts, j->cb.counts, sizeof(counts), hipMemcpyDeviceToHost, st));
            NH_HIP(hipStreamSynchronize(st));
            next_m = counts[0];
            f_len += counts[1];
        }
        account_kernel_time(j);
        plist = next;
        pb ^= 1;
        m = next_m;
    }
    std::swap(j->d_S, j->d_F);
    j->cur_n = f_len;
    j->alive = f_len;
    return NH_OK;
}

// One water-level round of the task-list form, the reference CUDA algorithm's
// host loop (cuda/src/parallelisation/monolithic.rs:222-318): fill a task
// list from the front of the slow queue (lineages already at the level go to
// the fast queue), upload it, run one step slice per task, download the
// tasks' state, next event times and slice samples, then insert the samples
// into the dedup cache and requeue the surviving tasks, in task order.
static int task_list_round(nh_job *j) {
    hipStream_t st = j->ctx->stream;
    const double level = j->level_time;
    const uint64_t T = j->task_cap;
    std::vector<nh_lineage> tasks;
    std::vector<uint64_t> ref(T), nst(T, 0);
    std::vector<double> t(T), nt(T), su, stt;
    std::vector<uint32_t> x(T), y(T), idx(T), sx, sy, si;
    std::vector<uint8_t> fl(T, 0), status(T);
    if (j->cache_cap) {
        su.resize(T); stt.resize(T); sx.resize(T); sy.resize(T); si.resize(T);
    }
    while (!j->slow_q.empty()) {
        tasks.clear();
        while (tasks.size() < T && !j->slow_q.empty()) {
            nh_job::Queued q = j->slow_q.front();
            j->slow_q.pop_front();
            if (q.next < level) {
                tasks.push_back(q.l);
                nt[tasks.size() - 1] = q.next;
            } else {
                j->fast_q.push_back(q);
            }
        }
        const uint64_t m = tasks.size();
        if (!m) break;
        for (uint64_t k = 0; k < m; k++) {
            ref[k] = tasks[k].global_reference;
            t[k] = tasks[k].last_event_time;
            x[k] = tasks[k].x;
            y[k] = tasks[k].y;
            idx[k] = tasks[k].index;
        }
        const LineageSoA &s = j->lin[0].s;
        NH_HIP(hipMemcpyAsync(s.ref, ref.data(), m * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.t_last, t.data(), m * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.next_t, nt.data(), m * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.next_st, nst.data(), m * 8, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.x, x.data(), m * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.y, y.data(), m * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.idx, idx.data(), m * 4, hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(s.flags, fl.data(), m, hipMemcpyHostToDevice, st));
        PlainArgs a{};
        a.S = j->tables->dev;
        a.P = j->P;
        a.L = s;
        a.pass = j->d_S;
        a.m = m;
        a.round = j->d_round;
        a.status = j->d_status;
        a.s_u = j->s_u; a.s_t = j->s_t; a.s_x = j->s_x; a.s_y = j->s_y; a.s_i = j->s_i;
        a.cache_cap = 0; // the cache is the host's
        a.events = j->d_events;
        a.ctr = j->d_ctr;
        NH_HIP(hipEventRecord(j->ev0, st));
        launch_plain(j->tables->kind, a, j->threads, st);
        NH_HIP(hipGetLastError());
        NH_HIP(hipEventRecord(j->ev1, st));
        j->launches++;
        NH_HIP(hipMemcpyAsync(status.data(), j->d_status, m, hipMemcpyDeviceToHost, st));
        NH_HIP(hipMemcpyAsync(t.data(), s.t_last, m * 8, hipMemcpyDeviceToHost, st));
        NH_HIP(hipMemcpyAsync(nt.data(), s.next_t, m * 8, hipMemcpyDeviceToHost, st));
        NH_HIP(hipMemcpyAsync(x.data(), s.x, m * 4, hipMemcpyDeviceToHost, st));
        NH_HIP(hipMemcpyAsync(y.data(), s.y, m * 4, hipMemcpyDeviceToHost, st));
        NH_HIP(hipMemcpyAsync(idx.data(), s.idx, m * 4, hipMemcpyDeviceToHost, st));
        if (j->cache_cap) {
            NH_HIP(hipMemcpyAsync(su.data(), j->s_u, m * 8, hipMemcpyDeviceToHost, st));
            NH_HIP(hipMemcpyAsync(stt.data(), j->s_t, m * 8, hipMemcpyDeviceToHost, st));
            NH_HIP(hipMemcpyAsync(sx.data(), j->s_x, m * 4, hipMemcpyDeviceToHost, st));
            NH_HIP(hipMemcpyAsync(sy.data(), j->s_y, m * 4, hipMemcpyDeviceToHost, st));
            NH_HIP(hipMemcpyAsync(si.data(), j->s_i, m * 4, hipMemcpyDeviceToHost, st));
        }
        NH_HIP(hipStreamSynchronize(st));
        account_kernel_time(j);
        for (uint64_t k = 0; k < m; k++) {
            // DirectMappedCache::insert (cache.rs:39-54): false if the bucket
            // already holds an equal sample
            bool dup = false;
            if (j->cache_cap && (status[k] & ST_HAS_SAMPLE)) {
                nh_job::HostSample &b =
                    j->host_cache[speciation_sample_hash(su[k], stt[k], sx[k], sy[k], si[k]) % j->cache_cap];
                dup = b.some && b.u == su[k] && b.t == stt[k] && b.x == sx[k] && b.y == sy[k] && b.i == si[k];
                if (!dup) b = nh_job::HostSample{true, su[k], stt[k], sx[k], sy[k], si[k]};
            }
            const uint8_t kind = status[k] & ST_KIND_MASK;
            if (kind == ST_SPEC || dup) continue;
            nh_job::Queued q{nh_lineage{ref[k], t[k], x[k], y[k], idx[k], 0}, nt[k]};
            if (q.next < level) j->slow_q.push_back(q);
            else j->fast_q.push_back(q);
        }
    }
    std::swap(j->slow_q, j->fast_q);
    j->cur_n = j->slow_q.size();
    j->alive = j->cur_n;
    return NH_OK;
}

// ordered compaction of the persistent form's live lineages into the other buffer
static int compact_lineages(nh_job *j) {
    hipStream_t st = j->ctx->stream;
    launch_soa_compact(j->lin[j->cur].s, j->lin[j->cur ^ 1].s, j->cur_n, j->d_blk, st);
    NH_HIP(hipGetLastError());
    uint32_t live = 0;
    NH_HIP(hipMemcpyAsync(&live, j->d_blk + soa_compact_blocks(j->cur_n), 4, hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    if (live != j->alive) return set_error(NH_E_STATE, "compaction found %u live lineages, expected %llu", live, (unsigned long long)j->alive);
    j->cur ^= 1;
    j->cur_n = live;
    return NH_OK;
}

static PersistArgs persist_args(nh_job *j, int cur) {
    PersistArgs a{};
    a.S = j->tables->dev;
    a.P = j->P;
    a.lin = j->lin[cur].s;
    a.round = j->d_round;
    a.cache = j->d_cache;
    a.cache_cap = j->cache_cap;
    a.events = j->d_events;
    a.ctr = j->d_ctr;
    a.packed = j->d_packed;
    a.packed_cap = j->packed_cap;
    a.inner_steps = j->p.inner_steps ? j->p.inner_steps : 16;
    a.coop = j->coop ? 1u : 0u;
    // default fill: five eighths; the alias kinds execute once an eighth of
    // a wave is ready (their execution is cheap next to the search: the 256 x
    // 256 map window, 1.6 lineages per lane, 14.3 -> 13.4 ms)
    const bool alias_kind = j->tables->dev.kind == K_SEPARABLE || j->tables->dev.kind == K_PACKED;
    a.exec_fill = j->p.exec_fill ? j->p.exec_fill : (alias_kind ? 1u : 5u);
    a.pool = j->pooled ? j->p.pool : 0u;
    if (j->disp) { // the kernel writes the round sort's keys, values and key range
        a.ev_keys = j->sb.keys[0];
        a.ev_vals = j->sb.vals[0];
        a.ev_range = j->sb.minmax;
    }
    return a;
}

// One round of the persistent form.  With use_graph, the round's step ->
// event-key -> radix-sort -> gather cycle is captured once per input buffer
// and replayed; the round's level, size and event offset are read from
// device memory (RoundDesc, Counters).  Sets *sorted when the round's events
// are already sorted into d_sorted.  Dispersal reporting captures the step
// kernel only: its rounds are large (about event_slice events), so the eager
// sort, which reads the key ranges back and skips constant digits, costs one
// synchronisation per round and saves many digit passes.
static int persistent_round(nh_job *j, bool *sorted) {
    hipStream_t st = j->ctx->stream;
    *sorted = false;
    if (j->p.use_graph) {
        int g = j->cur;
        if (!j->graph[g]) {
            hipGraph_t graph = nullptr;
            NH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_reset_counters(j->d_ctr, j->disp, st);
            if (j->disp) launch_sort_init(j->sb, st);
            launch_persistent(j->tables->kind, persist_args(j, g), j->blocks, j->threads, st);
            if (!j->disp)
                sort_events_device_sized(j->d_events, &j->d_ctr->n_events, &j->d_round->ev_start, j->n_initial,
                                         j->d_sorted, j->sb, st);
            // step -> sort -> compaction: the whole cycle of a round in one graph
            launch_soa_compact_gated(j->lin[g].s, j->lin[g ^ 1].s, j->lin[g].cap, &j->d_round->n_in, j->d_blk,
                                     j->d_ctr, st);
            NH_HIP(hipStreamEndCapture(st, &graph));
            hipError_t e = hipGraphInstantiate(&j->graph[g], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return set_error(NH_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        NH_HIP(hipEventRecord(j->ev0, st));
        NH_HIP(hipGraphLaunch(j->graph[g], st));
        NH_HIP(hipEventRecord(j->ev1, st));
        *sorted = !j->disp;
    } else {
        launch_reset_counters(j->d_ctr, j->disp, st);
        if (j->disp) launch_sort_init(j->sb, st);
        NH_HIP(hipEventRecord(j->ev0, st));
        launch_persistent(j->tables->kind, persist_args(j, j->cur), j->blocks, j->threads, st);
        NH_HIP(hipGetLastError());
        NH_HIP(hipEventRecord(j->ev1, st));
        launch_soa_compact_gated(j->lin[j->cur].s, j->lin[j->cur ^ 1].s, j->lin[j->cur].cap, &j->d_round->n_in,
                                 j->d_blk, j->d_ctr, st);
        NH_HIP(hipGetLastError());
    }
    j->launches++;
    Counters c;
    NH_TRY(read_counters(j, c));
    account_kernel_time(j);
    j->alive -= c.n_dead;
    // the device dropped the dead slots once they made up an eighth of the
    // buffer (launch_soa_compact_gated): the live lineages are in the other one
    if (c.compact_done) {
        if (c.compact_live != j->alive)
            return set_error(NH_E_STATE, "compaction found %llu live lineages, expected %llu", c.compact_live,
                             (unsigned long long)j->alive);
        j->cur ^= 1;
        j->cur_n = c.compact_live;
    }
    return NH_OK;
}

extern "C" int nh_job_start(nh_job *j) {
    if (!j) return set_error(NH_E_INVALID, "null job");
    if (j->started) return NH_OK;
    NH_HIP(hipSetDevice(j->ctx->device));
    return job_start(j);
}

// the undrained rest of the last round's sorted events, copied to the host
static int stash_device_events(nh_job *j) {
    uint64_t k = j->dev_n - j->dev_pos;
    size_t old = j->ready_packed.size();
    j->ready_packed.resize(old + k);
    NH_HIP(hipMemcpy(j->ready_packed.data() + old, j->d_packed_sorted + j->dev_pos, k * sizeof(DevPacked),
                     hipMemcpyDeviceToHost));
    j->dev_pos = j->dev_n = 0;
    return NH_OK;
}

extern "C" int nh_job_step_round(nh_job *j, int32_t *finished) {
    if (!j) return set_error(NH_E_INVALID, "null job");
    NH_HIP(hipSetDevice(j->ctx->device));
    if (!j->started) NH_TRY(job_start(j));
    const nh_job_params &p = j->p;
    auto done = [&]() { return j->alive == 0 || (p.has_pause_before && !(j->level_time < p.pause_before)); };
    if (j->failed) return set_error(NH_E_STATE, "the job failed in an earlier round");
    if (!j->finished && done()) j->finished = true;
    if (j->finished) {
        if (finished) *finished = 1;
        return NH_OK;
    }
    hipStream_t st = j->ctx->stream;
    if (j->dev_pos < j->dev_n) NH_TRY(stash_device_events(j));
    // total rate = sum of the reported event rate over the slow lineages,
    // folded in order: lambda with dispersal reporting (monolithic/mod.rs:139-148),
    // lambda * nu for speciation only (monolithic/mod.rs:150-164)
    double rate = j->disp ? p.turnover_rate : p.turnover_rate * p.speciation_probability;
    double total_rate = repeated_sum(rate, j->alive);
    j->level_time = j->level_time + (double)j->event_slice / total_rate;
    if (p.has_pause_before && p.pause_before < j->level_time) j->level_time = p.pause_before;
    uint64_t ev_start = j->disp ? 0 : j->n_events;
    // the pooled kernel holds p.pool lineages per wave: claims are sized by
    // pool entries, not lanes
    const unsigned long long holders =
        (unsigned long long)j->blocks * (j->pooled ? (unsigned long long)p.pool : (unsigned long long)j->threads);
    j->round_host = RoundDesc{j->level_time, (unsigned long long)j->cur_n, (unsigned long long)ev_start,
                              claim_chunk(j->cur_n, holders)};
    NH_HIP(hipMemcpyAsync(j->d_round, &j->round_host, sizeof(RoundDesc), hipMemcpyHostToDevice, st));
    if (j->disp && p.form == NH_FORM_PLAIN) {
        NH_HIP(hipMemsetAsync(&j->d_ctr->n_events, 0, sizeof(unsigned long long), st));
        NH_HIP(hipMemsetAsync(&j->d_ctr->ev_overflow, 0, sizeof(unsigned long long), st));
        NH_HIP(hipMemsetAsync(&j->d_ctr->n_ghost, 0, sizeof(unsigned long long), st));
    }

    bool sorted = false;
    if (p.form == NH_FORM_PLAIN) NH_TRY(plain_round(j));
    else if (p.form == NH_FORM_TASK_LIST) NH_TRY(task_list_round(j));
    else NH_TRY(persistent_round(j, &sorted));

    Counters c;
    NH_TRY(read_counters(j, c));
    if (c.ev_overflow) {
        j->failed = true;
        return set_error(NH_E_CAPACITY,
                         "a round reported more than %llu events (the dispersal event buffer); "
                         "use a smaller event slice",
                         (unsigned long long)j->packed_cap);
    }
    // dispersal reporting: slots [0, min(reserved, cap)) hold the round's
    // events and the ghost records of unused reserved slots
    uint64_t slots = j->disp ? std::min<uint64_t>(c.n_events, j->packed_cap) : 0;
    uint64_t ne = j->disp ? slots - c.n_ghost : c.n_events - ev_start;
    j->n_events = j->disp ? j->n_events + ne : c.n_events;
    j->steps = c.steps;
    double mt = from_bits(c.max_time_bits);
    if (c.max_time_bits && mt > j->max_time) j->max_time = mt;
    // every round ends with an empty simulate call whose time is
    // max(0, min_event_time) (monolithic/mod.rs:198-242)
    if (p.resume_after > j->max_time) j->max_time = p.resume_after;
    j->rounds++;

    // the live water-level reporter sorts each round's events (live.rs:88-115)
    // (ghost records sort after every event); the persistent form's kernels
    // wrote the sort's time keys as they reported the events
    const bool keys_ready = j->p.form == NH_FORM_PERSISTENT;
    if (ne && j->disp && j->d_counts) {
        // the counters read the round through the sort's permutation: no
        // sorted copy (nothing drains a counted round)
        if (!sorted) sort_and_count_packed(j->d_packed, slots, ne, keys_ready, j->sb, j->d_counts, st);
        NH_HIP(hipGetLastError());
    } else if (ne && j->disp) {
        if (!sorted) sort_events(j->d_packed, slots, j->d_packed_sorted, j->sb, st, keys_ready);
        NH_HIP(hipGetLastError());
        NH_HIP(hipStreamSynchronize(st));
        j->dev_pos = 0;
        j->dev_n = ne;
    } else if (ne) {
        if (!sorted) sort_events(j->d_events + ev_start, ne, j->d_sorted, j->sb, st, false);
        NH_HIP(hipGetLastError());
        size_t old = j->ready.size();
        j->ready.resize(old + ne);
        NH_HIP(hipMemcpyAsync(j->ready.data() + old, j->d_sorted, ne * sizeof(DevEvent), hipMemcpyDeviceToHost, st));
        NH_HIP(hipStreamSynchronize(st));
    }
    if (done()) j->finished = true;
    if (finished) *finished = j->finished ? 1 : 0;
    return NH_OK;
}

extern "C" int nh_job_simulate(nh_job *j) {
    int32_t fin = 0;
    while (!fin) NH_TRY(nh_job_step_round(j, &fin));
    return NH_OK;
}

// hands out the front of a ready list, in order; the list is emptied once read
template <class T> static void drain_ready(std::vector<T> &ready, size_t &pos, T *buf, size_t cap, size_t *n_out) {
    size_t avail = ready.size() - pos;
    if (!buf) {
        *n_out = avail;
        return;
    }
    size_t n = std::min(avail, cap);
    if (n) std::memcpy(buf, ready.data() + pos, n * sizeof(T));
    pos += n;
    if (pos == ready.size()) {
        ready.clear();
        pos = 0;
    }
    *n_out = n;
}

extern "C" int nh_job_drain_packed_events(nh_job *j, nh_packed_event *buf, size_t cap, size_t *n_out) {
    if (!j || !n_out) return set_error(NH_E_INVALID, "null argument");
    if (!j->disp) return set_error(NH_E_STATE, "the job reports speciation events only (nh_job_drain_events)");
    const uint64_t on_device = j->dev_n - j->dev_pos;
    if (!buf) {
        *n_out = j->ready_packed.size() - j->packed_pos + on_device;
        return NH_OK;
    }
    size_t host = 0;
    drain_ready(j->ready_packed, j->packed_pos, buf, cap, &host); // older events first
    size_t k = std::min<uint64_t>(cap - host, on_device);
    if (k) {
        NH_HIP(hipSetDevice(j->ctx->device));
        NH_HIP(hipMemcpy(buf + host, j->d_packed_sorted + j->dev_pos, k * sizeof(DevPacked), hipMemcpyDeviceToHost));
        j->dev_pos += k;
    }
    *n_out = host + k;
    return NH_OK;
}

extern "C" int nh_job_event_counts(nh_job *j, nh_event_counts *out) {
    static_assert(sizeof(nh_event_counts) == sizeof(unsigned long long) * EC_COUNTERS, "event counters");
    if (!j || !out) return set_error(NH_E_INVALID, "null argument");
    if (!j->disp || !j->p.device_event_counter)
        return set_error(NH_E_STATE, "the job does not count its events on the device");
    if (!j->d_counts) { // not started: nothing counted yet
        std::memset(out, 0, sizeof(*out));
        return NH_OK;
    }
    NH_HIP(hipSetDevice(j->ctx->device));
    EventCounts c;
    NH_HIP(hipMemcpyAsync(&c, j->d_counts, sizeof(c), hipMemcpyDeviceToHost, j->ctx->stream));
    NH_HIP(hipStreamSynchronize(j->ctx->stream));
    std::memcpy(out, c.count, sizeof(*out));
    return NH_OK;
}

extern "C" int nh_job_drain_events(nh_job *j, nh_speciation_event *buf, size_t cap, size_t *n_out) {
    if (!j || !n_out) return set_error(NH_E_INVALID, "null argument");
    if (j->disp) return set_error(NH_E_STATE, "the job reports dispersal: drain PackedEvents (nh_job_drain_packed_events)");
    drain_ready(j->ready, j->ready_pos, buf, cap, n_out);
    return NH_OK;
}

extern "C" int nh_job_outcome(const nh_job *j, nh_outcome *o) {
    if (!j || !o) return set_error(NH_E_INVALID, "null argument");
    o->paused = j->started ? (j->alive > 0) : (j->pending.size() + j->n_origins > 0);
    o->time = j->max_time;
    o->steps = j->steps;
    o->n_events = j->n_events;
    o->n_remaining = j->started ? j->alive : j->pending.size() + j->n_origins;
    o->rounds = j->rounds;
    o->launches = j->launches;
    return NH_OK;
}

extern "C" int nh_job_take_lineages(nh_job *j, nh_lineage *buf, size_t cap, size_t *n_out) {
    if (!j || !n_out) return set_error(NH_E_INVALID, "null argument");
    if (!j->started && j->has_origins) {
        *n_out = j->n_origins;
        if (!buf) return NH_OK;
        if (cap < j->n_origins) return set_error(NH_E_CAPACITY, "lineage buffer holds %zu of %llu", cap,
                                                 (unsigned long long)j->n_origins);
        NH_HIP(hipSetDevice(j->ctx->device));
        std::vector<nh_lineage> v;
        NH_TRY(download_lineages(j->origins, nullptr, j->n_origins, v, j->ctx->stream));
        std::memcpy(buf, v.data(), v.size() * sizeof(nh_lineage));
        return NH_OK;
    }
    if (!j->started) {
        *n_out = j->pending.size();
        if (buf) std::memcpy(buf, j->pending.data(), std::min(cap, j->pending.size()) * sizeof(nh_lineage));
        return NH_OK;
    }
    *n_out = j->alive;
    if (!buf) return NH_OK;
    if (cap < j->alive) return set_error(NH_E_CAPACITY, "lineage buffer holds %zu of %llu", cap, (unsigned long long)j->alive);
    NH_HIP(hipSetDevice(j->ctx->device));
    if (j->p.form == NH_FORM_PERSISTENT && j->cur_n != j->alive) NH_TRY(compact_lineages(j));
    std::vector<nh_lineage> v;
    if (j->p.form == NH_FORM_TASK_LIST)
        for (auto &q : j->slow_q) v.push_back(q.l);
    else if (j->p.form == NH_FORM_PLAIN) NH_TRY(download_lineages(j->lin[0], j->d_S, j->cur_n, v, j->ctx->stream));
    else NH_TRY(download_lineages(j->lin[j->cur], nullptr, j->cur_n, v, j->ctx->stream));
    std::memcpy(buf, v.data(), v.size() * sizeof(nh_lineage));
    return NH_OK;
}

extern "C" int nh_job_destroy(nh_job *j) {
    if (!j) return NH_OK;
    (void)hipSetDevice(j->ctx->device);
    (void)hipStreamSynchronize(j->ctx->stream);
    delete j;
    return NH_OK;
}

extern "C" int nh_job_kernel_time(const nh_job *j, double *total_ms, uint64_t *launches) {
    if (!j) return set_error(NH_E_INVALID, "null job");
    if (total_ms) *total_ms = j->kernel_ms;
    if (launches) *launches = j->timed;
    return NH_OK;
}

// ===========================================================================
// autotuner
extern "C" int nh_autotune(nh_ctx *ctx, const nh_tables *tables, const nh_job_params *base,
                           const nh_lineage *sample, size_t n, double window, nh_job_params *tuned,
                           double *best_ms) {
    if (!ctx || !tables || !base || !tuned || (!sample && n) || !(window > 0.0))
        return set_error(NH_E_INVALID, "nh_autotune: bad argument");
    std::vector<uint64_t> slices;
    if (base->step_slice) slices.push_back(base->step_slice);
    else slices = {5, 10, 20, 50, 150};
    double best = 1e300;
    nh_job_params winner = *base;
    winner.form = NH_FORM_PERSISTENT;
    for (uint32_t threads : {128u, 256u}) {
        int max_per_cu = persistent_max_blocks_per_cu(tables->kind, (int)threads,
                                                      base->dedup_mode != NH_DEDUP_NONE,
                                                      base->report_dispersal != 0, false);
        std::vector<uint32_t> residency = {(uint32_t)max_per_cu};
        if (max_per_cu >= 2) residency.push_back((uint32_t)(max_per_cu / 2));
        for (uint32_t per_cu : residency) {
            for (uint32_t inner : {4u, 16u})
            for (uint32_t fill : {1u, 5u, 8u})
            for (uint64_t slice : slices) {
                nh_job_params p = *base;
                p.form = NH_FORM_PERSISTENT;
                p.pool = 0;
                p.block_size = threads;
                p.blocks_per_cu = per_cu;
                p.step_slice = slice;
                p.inner_steps = inner;
                p.exec_fill = fill;
                p.has_pause_before = 1;
                p.pause_before = window;
                p.use_graph = 0;
                nh_job *j = nullptr;
                NH_TRY(nh_job_create(ctx, tables, &p, &j));
                int rc = nh_job_add_lineages(j, sample, n);
                if (rc == NH_OK) rc = nh_job_start(j);
                double t0 = now_s();
                if (rc == NH_OK) rc = nh_job_simulate(j);
                double ms = (now_s() - t0) * 1e3;
                nh_job_destroy(j);
                if (rc != NH_OK) return rc;
                if (ms < best) {
                    best = ms;
                    winner = p;
                }
            }
        }
    }
    // the pooled kernel (dense rounds without dedup cache): one-wave blocks
    // at full residency, each pool size
    if (base->dedup_mode == NH_DEDUP_NONE && base->coop_search != 2) {
        for (uint32_t pool : {128u, 192u})
        for (uint64_t slice : slices) {
            nh_job_params p = *base;
            p.form = NH_FORM_PERSISTENT;
            p.pool = pool;
            p.block_size = 64;
            p.blocks_per_cu = (uint32_t)pooled_max_blocks_per_cu(tables->kind, pool, base->report_dispersal != 0);
            p.step_slice = slice;
            p.inner_steps = 16;
            p.exec_fill = 5;
            p.has_pause_before = 1;
            p.pause_before = window;
            p.use_graph = 0;
            nh_job *j = nullptr;
            NH_TRY(nh_job_create(ctx, tables, &p, &j));
            int rc = nh_job_add_lineages(j, sample, n);
            if (rc == NH_OK) rc = nh_job_start(j);
            double t0 = now_s();
            if (rc == NH_OK) rc = nh_job_simulate(j);
            double ms = (now_s() - t0) * 1e3;
            nh_job_destroy(j);
            if (rc != NH_OK) return rc;
            if (ms < best) {
                best = ms;
                winner = p;
            }
        }
    }
    winner.has_pause_before = base->has_pause_before;
    winner.pause_before = base->pause_before;
    winner.use_graph = base->use_graph;
    *tuned = winner;
    if (best_ms) *best_ms = best;
    return NH_OK;
}
