// nh_host.cpp -- host-side scenario construction of the library's C ABI:
// alias tables from a sparse dispersal map and origin sampling.  These are
// the inputs of the hot path; the Rust glue obtains them from the reference's
// own scenario code instead, standalone C/C++/Python users from here.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/necsim_hip.h"
#include "nh_math.h"
#include "nh_rng.h"
#include "nh_internal.h"

using namespace nh;

extern "C" uint64_t nh_seed_from_u64(uint64_t seed) { return seed_from_u64(seed); }

namespace {

// AliasMethodSamplerAtom::create (necsim/impls/no-std/src/alias/packed.rs:59-120).
// Vose's method with LIFO stacks of overfull / underfull indices; leftovers
// are rounded to full buckets.
void alias_create(const std::vector<uint64_t> &events, const std::vector<double> &weights,
                  std::vector<nh_alias_atom> &atoms) {
    size_t n = events.size();
    atoms.resize(n);
    double total = 0.0; // NonNegativeF64 Sum: fold from 0.0 in order
    for (double w : weights) total = total + w;
    double nf = (double)n;
    for (size_t i = 0; i < n; i++) atoms[i] = nh_alias_atom{weights[i] * nf / total, events[i], events[i]};
    std::vector<size_t> over, under;
    over.reserve(n);
    under.reserve(n);
    for (size_t i = 0; i < n; i++)
        if (atoms[i].u > 1.0) over.push_back(i);
    for (size_t i = 0; i < n; i++)
        if (atoms[i].u < 1.0) under.push_back(i);
    while (!over.empty() && !under.empty()) {
        size_t o = over.back();
        over.pop_back();
        size_t u = under.back();
        under.pop_back();
        atoms[o].u = atoms[o].u + atoms[u].u - 1.0;
        atoms[u].k = atoms[o].e;
        if (atoms[o].u < 1.0) under.push_back(o);
        else if (atoms[o].u > 1.0) over.push_back(o);
    }
    for (size_t i : over) atoms[i].u = 1.0;
    for (size_t i : under) atoms[i].u = 1.0;
}

} // namespace

// InMemoryPackedSeparableAliasDispersalSampler::new (packed_separable_alias/mod.rs:64-203)
// and InMemoryPackedAliasDispersalSampler::new (packed_alias/mod.rs:58-118),
// including the in-memory dispersal contract (in_memory/contract.rs:13-69).
extern "C" int nh_build_alias_tables(uint32_t width, uint32_t height, const uint32_t *habitat,
                                     const uint64_t *row_ptr, const uint32_t *col,
                                     const double *weight, int32_t separable,
                                     nh_alias_range *ranges, nh_self_dispersal *self_dispersal,
                                     nh_alias_atom *atoms_out, size_t atoms_cap, size_t *n_atoms) {
    if (!habitat || !row_ptr || !ranges || !n_atoms || width == 0 || height == 0)
        return set_error(NH_E_INVALID, "nh_build_alias_tables: null argument or empty extent");
    const size_t cells = (size_t)width * height;
    std::vector<uint64_t> ev;
    std::vector<double> w;
    std::vector<nh_alias_atom> atoms;
    size_t total_atoms = 0;
    for (size_t row = 0; row < cells; row++) {
        ev.clear();
        w.clear();
        double self_w = 0.0;
        for (uint64_t j = row_ptr[row]; j < row_ptr[row + 1]; j++) {
            uint32_t c = col[j];
            double p = weight[j];
            if (c >= cells) return set_error(NH_E_INVALID, "dispersal column outside the habitat");
            if (j > row_ptr[row] && c <= col[j - 1])
                return set_error(NH_E_INVALID, "dispersal columns must ascend within a row");
            if (p < 0.0 || (p > 0.0 && (habitat[row] == 0 || habitat[c] == 0)))
                return set_error(NH_E_CONTRACT, "dispersal from/to non-habitat");
            double wt = p * (double)habitat[c];
            if (wt > 0.0) {
                ev.push_back(c);
                w.push_back(wt);
                if (c == row) self_w = wt;
            }
        }
        if (habitat[row] > 0 && ev.empty())
            return set_error(NH_E_CONTRACT, "no dispersal from a habitable cell");
        uint64_t from = total_atoms;
        if (ev.empty()) {
            ranges[row] = nh_alias_range{from, from};
            if (separable && self_dispersal) self_dispersal[row] = nh_self_dispersal{0.0, row};
            continue;
        }
        uint64_t non_self = row;
        double self_prob = 0.0;
        if (!separable) {
            alias_create(ev, w, atoms);
        } else {
            double total = 0.0;
            for (double x : w) total = total + x;
            self_prob = self_w / total;
            double self_u = self_w * (double)ev.size() / total;
            if (self_u < 1.0) {
                alias_create(ev, w, atoms);
                // exactly one atom holds self-dispersal: flip it into the k
                // slot and move it last
                for (size_t j = 0; j < atoms.size(); j++) {
                    if (atoms[j].e == row || atoms[j].k == row) {
                        if (atoms[j].e == row) {
                            std::swap(atoms[j].e, atoms[j].k);
                            atoms[j].u = 1.0 - atoms[j].u;
                        }
                        non_self = atoms[j].e;
                        std::swap(atoms[j], atoms.back());
                        break;
                    }
                }
            } else {
                std::vector<uint64_t> ev2;
                std::vector<double> w2;
                for (size_t j = 0; j < ev.size(); j++)
                    if (ev[j] != row) {
                        ev2.push_back(ev[j]);
                        w2.push_back(w[j]);
                    }
                if (ev2.empty()) atoms.clear();
                else alias_create(ev2, w2, atoms);
            }
            if (self_dispersal) self_dispersal[row] = nh_self_dispersal{self_prob, non_self};
        }
        if (atoms_out) {
            if (total_atoms + atoms.size() > atoms_cap)
                return set_error(NH_E_CAPACITY, "atoms buffer too small");
            std::memcpy(atoms_out + total_atoms, atoms.data(), atoms.size() * sizeof(nh_alias_atom));
        }
        total_atoms += atoms.size();
        ranges[row] = nh_alias_range{from, total_atoms};
    }
    *n_atoms = total_atoms;
    return NH_OK;
}

namespace {

// OriginPreSampler (origin_sampler/pre_sampler.rs:50-114):
// all() -> percentage(p) -> partition(rank, size)
struct PreSampler {
    uint64_t next_base = 0;
    int mode; // 0 none, 1 all, 2 quasi-random geometric skips
    double inv_rate, q = 0.5;
    uint32_t rank, size;
    bool started = false;
    PreSampler(double p, uint32_t r, uint32_t s) : rank(r), size(s ? s : 1) {
        mode = p <= 0.0 ? 0 : (p >= 1.0 ? 1 : 2);
        inv_rate = 1.0 / ln(1.0 - p);
    }
    bool inner(uint64_t &out) {
        if (mode == 0) return false;
        if (mode == 2) {
            q += 6.180339887498949e-1; // INV_PHI
            q -= floor_(q);
            next_base += as_u64(floor_(ln(q) * inv_rate)); // self.nth(skip)
        }
        out = next_base++;
        return true;
    }
    bool next(uint64_t &out) {
        uint64_t v;
        if (!started) {
            started = true;
            for (uint32_t r = 0; r < rank; r++)
                if (!inner(v)) return false;
            return inner(out);
        }
        for (uint32_t r = 1; r < size; r++)
            if (!inner(v)) return false;
        return inner(out);
    }
};

} // namespace

extern "C" int nh_sample_origins(const nh_scenario *s, double percentage, uint32_t rank, uint32_t size,
                                 uint32_t cx, uint32_t cy, uint32_t radius, nh_lineage *out, size_t cap,
                                 size_t *n_out) {
    if (!s || !n_out) return set_error(NH_E_INVALID, "nh_sample_origins: null argument");
    PreSampler pre(percentage, rank, size);
    size_t n = 0;
    uint64_t next_index;
    auto emit = [&](uint32_t x, uint32_t y, uint32_t i, uint64_t ref) {
        if (out && n < cap) out[n] = nh_lineage{ref, 0.0, x, y, i, 0};
        n++;
    };
    if (s->kind == NH_SCENARIO_ALMOST_INFINITE_NORMAL) {
        // SingletonDemesCircleOriginSampler (singleton_demes/circle.rs:52-156):
        // walk the (2r+1)^2 extent row-major, keep wrap-min distances <= r
        uint32_t ox = cx - radius, oy = cy - radius;
        uint64_t d = (uint64_t)radius * 2 + 1, r2 = (uint64_t)radius * radius;
        uint64_t it = 0, end = d * d, last = 0;
        while (pre.next(next_index)) {
            uint64_t diff = next_index - last;
            last = next_index + 1;
            bool got = false;
            while (it < end) {
                uint32_t x = ox + (uint32_t)(it % d), y = oy + (uint32_t)(it / d);
                it++;
                uint64_t dx = std::min<uint32_t>(x - cx, cx - x), dy = std::min<uint32_t>(y - cy, cy - y);
                if (dx * dx + dy * dy <= r2) {
                    if (diff == 0) {
                        emit(x, y, 0, ((uint64_t)y << 32) | x);
                        got = true;
                        break;
                    }
                    diff--;
                }
            }
            if (!got) break;
        }
    } else {
        // NonSpatialOriginSampler (origin_sampler/non_spatial.rs:89-112) and
        // InMemoryOriginSampler (origin_sampler/in_memory.rs:89-119)
        if (s->width == 0 || s->height == 0) return set_error(NH_E_INVALID, "empty extent");
        bool in_memory = s->kind != NH_SCENARIO_NON_SPATIAL;
        if (in_memory && !s->habitat) return set_error(NH_E_INVALID, "habitat missing");
        uint64_t W = s->width, cells = W * s->height, loc = 0, last = 0, prefix = 0;
        uint32_t nli = 0;
        while (pre.next(next_index)) {
            uint64_t diff = next_index - last;
            last = next_index;
            bool ok = true;
            for (;;) {
                if (loc >= cells) {
                    ok = false;
                    break;
                }
                uint32_t h = in_memory ? s->habitat[loc] : s->deme;
                if ((uint64_t)nli + diff >= h) {
                    diff -= (uint64_t)(h - nli);
                    nli = 0;
                    prefix += h;
                    loc++;
                } else break;
            }
            if (!ok) break;
            nli += (uint32_t)diff;
            uint32_t x = s->origin_x + (uint32_t)(loc % W), y = s->origin_y + (uint32_t)(loc / W);
            uint64_t ref = in_memory ? prefix + nli : loc * (uint64_t)s->deme + nli;
            emit(x, y, nli, ref);
        }
    }
    *n_out = n;
    if (out && n > cap) return set_error(NH_E_CAPACITY, "origin buffer too small");
    return NH_OK;
}

namespace {

void put_u64(std::vector<uint8_t> &b, uint64_t v) {
    for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
void put_u32(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}

// PackedEvent order of two speciation events (necsim/core/src/event.rs:256-280):
// event_time, origin, target (= origin), prior_time, reference
bool speciation_less(const nh_speciation_event &a, const nh_speciation_event &b) {
    const uint64_t m = 0x7fffffffffffffffull;
    uint64_t ta = bits(a.event_time) & m, tb = bits(b.event_time) & m;
    if (ta != tb) return ta < tb;
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    if (a.index != b.index) return a.index < b.index;
    uint64_t pa = bits(a.prior_time) & m, pb = bits(b.prior_time) & m;
    if (pa != pb) return pa < pb;
    return a.global_reference < b.global_reference;
}

// PackedEvent::cmp (necsim/core/src/event.rs:256-280): event_time, origin,
// target, prior_time, reference; times by magnitude
bool packed_less(const nh_packed_event &a, const nh_packed_event &b) {
    const uint64_t m = 0x7fffffffffffffffull;
    uint64_t ta = bits(a.event_time) & m, tb = bits(b.event_time) & m;
    if (ta != tb) return ta < tb;
    if (a.origin_x != b.origin_x) return a.origin_x < b.origin_x;
    if (a.origin_y != b.origin_y) return a.origin_y < b.origin_y;
    if (a.origin_index != b.origin_index) return a.origin_index < b.origin_index;
    if (a.target_x != b.target_x) return a.target_x < b.target_x;
    if (a.target_y != b.target_y) return a.target_y < b.target_y;
    if (a.target_index != b.target_index) return a.target_index < b.target_index;
    uint64_t pa = bits(a.prior_time) & m, pb = bits(b.prior_time) & m;
    if (pa != pb) return pa < pb;
    return a.global_reference < b.global_reference;
}

int write_file(const char *path, const std::vector<uint8_t> &b, const char *what) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(NH_E_INVALID, "%s: cannot open %s", what, path);
    size_t w = std::fwrite(b.data(), 1, b.size(), f);
    int c = std::fclose(f);
    if (w != b.size() || c != 0) return set_error(NH_E_INVALID, "%s: write to %s failed", what, path);
    return NH_OK;
}

} // namespace

// EventLogRecorder::sort_and_write_segment (necsim/impls/std/src/event_log/recorder.rs:174-208)
// for speciation events.  bincode 1.3 with its default options (little-endian,
// fixed-width integers, usize as u64, bool as one byte, newtypes as their
// inner value) of EventLogHeader {min_time, max_time, length,
// with_speciation, with_dispersal} (event_log/mod.rs:9-19) followed by each
// PackedEvent (necsim/core/src/event.rs:16-29): reference u64, prior_time
// f64 with the sign bit set (a speciation, event.rs:103-115), event_time f64,
// origin and target (x, y, index u32 each; target = origin), coalescence u64
// (= the reference).
extern "C" int nh_event_log_write_segment(const char *path, const nh_speciation_event *events, size_t n) {
    static_assert(sizeof(double) == 8, "f64");
    if (!path || !events || n == 0)
        return set_error(NH_E_INVALID, "nh_event_log_write_segment: a segment holds at least one event");
    for (size_t i = 1; i < n; i++)
        if (speciation_less(events[i], events[i - 1]))
            return set_error(NH_E_INVALID, "nh_event_log_write_segment: events are not in PackedEvent order");
    std::vector<uint8_t> b;
    b.reserve(26 + 56 * n);
    put_u64(b, bits(events[0].event_time) & 0x7fffffffffffffffull);
    put_u64(b, bits(events[n - 1].event_time) & 0x7fffffffffffffffull);
    put_u64(b, (uint64_t)n);
    b.push_back(1); // with_speciation
    b.push_back(0); // with_dispersal
    for (size_t i = 0; i < n; i++) {
        const nh_speciation_event &e = events[i];
        put_u64(b, e.global_reference);
        put_u64(b, bits(e.prior_time) | 0x8000000000000000ull);
        put_u64(b, bits(e.event_time));
        for (int rep = 0; rep < 2; rep++) {
            put_u32(b, e.x);
            put_u32(b, e.y);
            put_u32(b, e.index);
        }
        put_u64(b, e.global_reference);
    }
    return write_file(path, b, "nh_event_log_write_segment");
}

// The same for PackedEvents as they are (speciations and dispersals), with
// the header's filter flags given by the caller (recorder.rs:146-152: the
// reporter's ReportSpeciation / ReportDispersal); the header times are the
// event times' magnitudes (PackedEvent::event_time, event.rs:33-36).
extern "C" int nh_event_log_write_packed_segment(const char *path, const nh_packed_event *events, size_t n,
                                                 int32_t with_speciation, int32_t with_dispersal) {
    if (!path || !events || n == 0)
        return set_error(NH_E_INVALID, "nh_event_log_write_packed_segment: a segment holds at least one event");
    for (size_t i = 1; i < n; i++)
        if (packed_less(events[i], events[i - 1]))
            return set_error(NH_E_INVALID, "nh_event_log_write_packed_segment: events are not in PackedEvent order");
    std::vector<uint8_t> b;
    b.reserve(26 + 56 * n);
    put_u64(b, bits(events[0].event_time) & 0x7fffffffffffffffull);
    put_u64(b, bits(events[n - 1].event_time) & 0x7fffffffffffffffull);
    put_u64(b, (uint64_t)n);
    b.push_back(with_speciation ? 1 : 0);
    b.push_back(with_dispersal ? 1 : 0);
    for (size_t i = 0; i < n; i++) {
        const nh_packed_event &e = events[i];
        put_u64(b, e.global_reference);
        put_u64(b, bits(e.prior_time));
        put_u64(b, bits(e.event_time));
        put_u32(b, e.origin_x);
        put_u32(b, e.origin_y);
        put_u32(b, e.origin_index);
        put_u32(b, e.target_x);
        put_u32(b, e.target_y);
        put_u32(b, e.target_index);
        put_u64(b, e.coalescence);
    }
    return write_file(path, b, "nh_event_log_write_packed_segment");
}

// ---------------------------------------------------------------------------
// Lineage files (necsim/impls/std/src/lineage_file/{saver,loader}.rs): the
// bincode 1.3 `bincode::options()` image of a Vec<Lineage>, i.e. varint
// integers (v < 251: one byte; else 251 / 252 / 253 followed by the value as
// little-endian u16 / u32 / u64), f64 as 8 little-endian bytes, a sequence as
// its varint length followed by its elements, and Lineage as its fields in
// order: global_reference u64, last_event_time f64, indexed_location
// {x u32, y u32, index u32} (necsim/core/src/lineage.rs:88-108,
// necsim/core/src/landscape/location.rs:70-78).
namespace {

void put_varint(std::vector<uint8_t> &b, uint64_t v) {
    if (v < 251) {
        b.push_back((uint8_t)v);
    } else if (v <= 0xffffull) {
        b.push_back(251);
        for (int i = 0; i < 2; i++) b.push_back((uint8_t)(v >> (8 * i)));
    } else if (v <= 0xffffffffull) {
        b.push_back(252);
        for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
    } else {
        b.push_back(253);
        for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
    }
}

struct Reader {
    const std::vector<uint8_t> &b;
    size_t pos = 0;
    bool ok = true;
    uint64_t fixed(int bytes) {
        if (pos + (size_t)bytes > b.size()) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 0; i < bytes; i++) v |= (uint64_t)b[pos + i] << (8 * i);
        pos += bytes;
        return v;
    }
    uint64_t varint(uint64_t max) {
        uint64_t tag = fixed(1), v;
        if (tag < 251) v = tag;
        else if (tag == 251) v = fixed(2);
        else if (tag == 252) v = fixed(4);
        else if (tag == 253) v = fixed(8);
        else {
            ok = false;
            return 0;
        }
        if (v > max) ok = false;
        return v;
    }
};

} // namespace

// LineageFileSaver::try_new + write (saver.rs:54-75): the file must not exist yet.
extern "C" int nh_lineage_file_write(const char *path, const nh_lineage *lineages, size_t n) {
    if (!path || (n && !lineages)) return set_error(NH_E_INVALID, "nh_lineage_file_write: null argument");
    std::vector<uint8_t> b;
    b.reserve(9 + 32 * n);
    put_varint(b, (uint64_t)n);
    for (size_t i = 0; i < n; i++) {
        const nh_lineage &l = lineages[i];
        put_varint(b, l.global_reference);
        put_u64(b, bits(l.last_event_time));
        put_varint(b, l.x);
        put_varint(b, l.y);
        put_varint(b, l.index);
    }
    FILE *f = std::fopen(path, "wbx"); // create_new
    if (!f) return set_error(NH_E_INVALID, "nh_lineage_file_write: cannot create %s (it must not exist)", path);
    size_t w = std::fwrite(b.data(), 1, b.size(), f);
    int c = std::fclose(f);
    if (w != b.size() || c != 0) return set_error(NH_E_INVALID, "nh_lineage_file_write: write to %s failed", path);
    return NH_OK;
}

// LineageFileLoader::try_new (loader.rs:33-45).  With buf == NULL only the
// lineage count is returned.
extern "C" int nh_lineage_file_read(const char *path, nh_lineage *buf, size_t cap, size_t *n_out) {
    if (!path || !n_out) return set_error(NH_E_INVALID, "nh_lineage_file_read: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_error(NH_E_INVALID, "nh_lineage_file_read: cannot open %s", path);
    std::vector<uint8_t> b;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + got);
    std::fclose(f);
    Reader r{b};
    uint64_t n = r.varint(~0ull);
    if (!r.ok) return set_error(NH_E_INVALID, "nh_lineage_file_read: %s is not a lineage file", path);
    *n_out = (size_t)n;
    if (!buf) return NH_OK;
    if (cap < n) return set_error(NH_E_CAPACITY, "nh_lineage_file_read: buffer holds %zu of %llu lineages", cap,
                                  (unsigned long long)n);
    for (uint64_t i = 0; i < n && r.ok; i++) {
        nh_lineage l{};
        l.global_reference = r.varint(~0ull);
        l.last_event_time = from_bits(r.fixed(8));
        l.x = (uint32_t)r.varint(0xffffffffull);
        l.y = (uint32_t)r.varint(0xffffffffull);
        l.index = (uint32_t)r.varint(0xffffffffull);
        buf[i] = l;
    }
    if (!r.ok || r.pos != b.size())
        return set_error(NH_E_INVALID, "nh_lineage_file_read: %s is truncated or has trailing bytes", path);
    return NH_OK;
}
