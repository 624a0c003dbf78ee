// nh_origins.hip -- the origin samplers on the device.
//
// OriginPreSampler::all().percentage(p).partition(rank, size)
// (necsim/impls/no-std/src/cogs/origin_sampler/pre_sampler.rs:49-114) feeding
// the scenario's origin sampler: NonSpatialOriginSampler (non_spatial.rs:89-112),
// InMemoryOriginSampler (in_memory.rs:89-119) or SingletonDemesCircleOriginSampler
// (singleton_demes/circle.rs:52-156).  nh_sample_origins (nh_host.cpp) walks
// the same samplers on the host; these kernels give the same lineages in the
// same order.
//
// Every origin sampler walks an enumeration of the habitat's individuals and
// emits those whose rank the pre-sampler names, so a lineage is a function of
// its rank alone:
//  - non-spatial: rank e is cell e / deme, index e % deme, reference e;
//  - in-memory: the cell whose injective-id range [prefix, prefix + habitat)
//    holds e (binary search over the cell records), index e - prefix,
//    reference e;
//  - circle: the row of the (2r+1)^2 box whose range of cells within the
//    radius holds e (binary search over per-row counts 2 floor(sqrt(r^2 -
//    dy^2)) + 1), then the column; index 0, reference y << 32 | x.
// The pre-sampler's ranks are 0, 1, 2, ... for `all`, and for a percentage p
// the running sum of the skips floor(ln(q_j) / ln(1 - p)) + 1 minus one,
// over the golden-ratio sequence q_0 = 0.5, q_{j+1} = (q_j + INV_PHI) -
// floor(q_j + INV_PHI) in f64 (pre_sampler.rs:77-98).  The partition keeps
// every size-th rank from the rank-th on.
//
// The golden-ratio sequence is a chain of f64 additions, but an exact one:
// with Q = q 2^53, q_0 = 0.5 is Q_0 = 2^52 and INV_PHI is P = INV_PHI 2^53 =
// 0x13c6ef372fe950, both multiples of 16, so every sum q + INV_PHI < 2 is a
// multiple of 2^-49 and needs no rounding in f64 (spacing 2^-53 below 1,
// 2^-52 in [1, 2)), and subtracting the floor is exact.  Hence
//   Q_j = (2^52 + j P) mod 2^53,
// and each rank's skip is a function of j alone: the ranks are one running
// sum, computed in parallel (a device scan).
#include <algorithm>
#include <cmath>
#include <vector>

#include "nh_internal.h"
#include "nh_kernels.h"

namespace nh {
namespace {

constexpr uint64_t Q_ONE = 1ull << 53, Q_MASK = Q_ONE - 1;
constexpr uint64_t Q_START = Q_ONE >> 1;  // q_0 = 0.5
constexpr int SCAN_T = 256;       // threads per block of the u64 scan
constexpr int SCAN_ITEMS = 4;     // values per thread
constexpr uint64_t SCAN_BLOCK = (uint64_t)SCAN_T * SCAN_ITEMS;
constexpr uint32_t MAX_CIRCLE_RADIUS = 1u << 22;
constexpr double PHI = 6.180339887498949e-1; // INV_PHI of pre_sampler.rs
constexpr uint64_t PHI_Q = 0x13c6ef372fe950ull; // PHI 2^53
static_assert(PHI_Q % 16 == 0 && Q_START % 16 == 0, "the golden-ratio chain is exact");

// q_j 2^53 (the low 53 bits of a wrapping product are those of the exact one)
NH_HD uint64_t golden_q(uint64_t j) { return (Q_START + j * PHI_Q) & Q_MASK; }

// pre_sampler.rs:92-95: skip = floor(ln(q) / ln(1 - p)) as usize, then
// nth(skip) takes skip + 1 values of the inner iterator.  Sample j0 + k uses
// q_{j0 + k + 1}.
__global__ void k_presample_steps(uint64_t j0, uint64_t n, double inv_rate, uint64_t *out) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double q = (double)golden_q(j0 + k + 1) * TWO_POW_M53;
    out[k] = as_u64(floor_(ln(q) * inv_rate)) + 1;
}

// the same steps by the reference's own loop (q += INV_PHI; q -= floor(q)) on
// one thread: the cross-check of the closed form on the device
__global__ void k_presample_steps_serial(uint64_t n, double inv_rate, uint64_t *out) {
    double q = 0.5;
    for (uint64_t k = 0; k < n; k++) {
        q += PHI;
        q -= floor_(q);
        out[k] = as_u64(floor_(ln(q) * inv_rate)) + 1;
    }
}

// inclusive scan of u64 (wrapping): per-block scans, block totals scanned
// recursively, then added back
__global__ void __launch_bounds__(SCAN_T) k_scan64_local(uint64_t *v, uint64_t n, uint64_t *block_sums) {
    __shared__ uint64_t part[SCAN_T];
    const uint64_t base = blockIdx.x * SCAN_BLOCK + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t x[SCAN_ITEMS], run = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        x[i] = base + i < n ? v[base + i] : 0;
        run += x[i];
        x[i] = run;
    }
    part[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        uint64_t a = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += a;
        __syncthreads();
    }
    const uint64_t before = threadIdx.x ? part[threadIdx.x - 1] : 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) v[base + i] = x[i] + before;
    if (threadIdx.x == SCAN_T - 1) block_sums[blockIdx.x] = part[SCAN_T - 1];
}

__global__ void k_scan64_add(uint64_t *v, uint64_t n, const uint64_t *block_prefix) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= n || k < SCAN_BLOCK) return;
    v[k] += block_prefix[k / SCAN_BLOCK - 1];
}

// ranks = base + inclusive sum - 1 (next_base before the batch is `base`)
__global__ void k_offset(uint64_t *v, uint64_t n, uint64_t add) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k < n) v[k] += add;
}

// *out = the number of leading values below key (the ranks ascend)
__global__ void k_count_below(const uint64_t *v, uint64_t n, uint64_t key, uint64_t *out) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (v[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    *out = lo;
}

struct OriginArgs {
    uint64_t n;             // lineages to emit
    uint64_t rank, size;    // the partition
    const uint64_t *ranks;  // percentage: the pre-sampler's ranks, else null (rank = position)
    uint32_t width, ox, oy, deme;
    const CellRec *cells;   // in-memory
    uint64_t n_cells;
    uint32_t cx, cy, radius;
    const uint64_t *row_start; // circle: cells within the radius before each row, [2r + 2]
    uint64_t n_rows;
    LineageSoA soa;         // either the job's lineage store
    nh_lineage *aos;        // or the reference layout
};

enum : int { O_NON_SPATIAL = 0, O_IN_MEMORY = 1, O_CIRCLE = 2 };

// index of the last element of a[0, n) (ascending, a[0] <= e) that is <= e
template <class Get> __device__ uint64_t last_at_most(uint64_t n, uint64_t e, Get get) {
    uint64_t lo = 0, hi = n; // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (get(mid) <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int MODE> __global__ void k_origins(OriginArgs A) {
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    const uint64_t k = A.rank + j * A.size;
    const uint64_t e = A.ranks ? A.ranks[k] : k;
    uint32_t x, y, i;
    uint64_t ref;
    if (MODE == O_CIRCLE) {
        const uint64_t row = last_at_most(A.n_rows, e, [&](uint64_t r) { return A.row_start[r]; });
        const uint64_t h = (A.row_start[row + 1] - A.row_start[row] - 1) / 2;
        x = A.cx - (uint32_t)h + (uint32_t)(e - A.row_start[row]);
        y = A.cy - A.radius + (uint32_t)row;
        i = 0;
        ref = ((uint64_t)y << 32) | x;
    } else {
        uint64_t cell;
        if (MODE == O_IN_MEMORY) {
            cell = last_at_most(A.n_cells, e, [&](uint64_t c) { return A.cells[c].prefix; });
            i = (uint32_t)(e - A.cells[cell].prefix);
        } else {
            cell = e / A.deme;
            i = (uint32_t)(e % A.deme);
        }
        x = A.ox + (uint32_t)(cell % A.width);
        y = A.oy + (uint32_t)(cell / A.width);
        ref = e;
    }
    if (A.aos) {
        A.aos[j] = nh_lineage{ref, 0.0, x, y, i, 0};
    } else {
        // a fresh lineage: last_event_time 0, its pessimistic next event time
        // too (monolithic/mod.rs:94-103), no cached event
        A.soa.ref[j] = ref;
        A.soa.t_last[j] = 0.0;
        A.soa.next_t[j] = 0.0;
        A.soa.next_st[j] = 0;
        A.soa.x[j] = x;
        A.soa.y[j] = y;
        A.soa.idx[j] = i;
        A.soa.flags[j] = 0;
    }
}

uint32_t grid_of(uint64_t n, uint32_t threads) { return (uint32_t)std::max<uint64_t>((n + threads - 1) / threads, 1); }

int scan_u64(uint64_t *v, uint64_t n, hipStream_t st) {
    if (n == 0) return NH_OK;
    const uint64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    uint64_t *sums = nullptr;
    NH_TRY(dalloc(&sums, nb));
    hipLaunchKernelGGL(k_scan64_local, dim3((uint32_t)nb), dim3(SCAN_T), 0, st, v, n, sums);
    int rc = NH_OK;
    if (nb > 1) {
        rc = scan_u64(sums, nb, st);
        if (rc == NH_OK)
            hipLaunchKernelGGL(k_scan64_add, dim3(grid_of(n, 256)), dim3(256), 0, st, v, n, (const uint64_t *)sums);
    }
    hipError_t e = hipStreamSynchronize(st);
    dfree(sums);
    if (rc != NH_OK) return rc;
    if (e != hipSuccess) return set_error(NH_E_HIP, "u64 scan: %s", hipGetErrorString(e));
    return NH_OK;
}

} // namespace

// The pre-sampler's percentage ranks: at least every rank below `below`
// (plus the one at or above it that ends the walk), extended batch by batch.
struct Ranks {
    uint64_t *d = nullptr;
    uint64_t n = 0, cap = 0;
    ~Ranks() { dfree(d); }
};

namespace {

int grow(Ranks &r, uint64_t need, hipStream_t st) {
    if (need <= r.cap) return NH_OK;
    uint64_t cap = std::max<uint64_t>(need, r.cap + r.cap / 2);
    uint64_t *d = nullptr;
    NH_TRY(dalloc(&d, cap));
    if (r.n) NH_HIP(hipMemcpyAsync(d, r.d, r.n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
    NH_HIP(hipStreamSynchronize(st));
    dfree(r.d);
    r.d = d;
    r.cap = cap;
    return NH_OK;
}

// percentage ranks until one reaches `below` (or exactly `count` ranks if
// count > 0); *n_below receives the number of ranks below `below`.  serial:
// the steps by the reference's f64 loop on one thread.
int percentage_ranks(double p, uint64_t below, uint64_t count, bool serial, Ranks &r, uint64_t *n_below,
                     hipStream_t st) {
    const double inv_rate = 1.0 / ln(1.0 - p);
    uint64_t next_base = 0;
    const uint64_t est = count ? count : (uint64_t)((double)below * p * 1.05) + 4096;
    if (serial && !count) return set_error(NH_E_INVALID, "the serial pre-sampler takes a rank count");
    for (;;) {
        const uint64_t batch = count ? count - r.n : (r.n == 0 ? est : std::max<uint64_t>(est / 4, 4096));
        NH_TRY(grow(r, r.n + batch, st));
        uint64_t *out = r.d + r.n;
        if (serial) hipLaunchKernelGGL(k_presample_steps_serial, dim3(1), dim3(1), 0, st, batch, inv_rate, out);
        else hipLaunchKernelGGL(k_presample_steps, dim3(grid_of(batch, 256)), dim3(256), 0, st, r.n, batch, inv_rate, out);
        NH_TRY(scan_u64(out, batch, st));
        hipLaunchKernelGGL(k_offset, dim3(grid_of(batch, 256)), dim3(256), 0, st, out, batch, next_base - 1);
        uint64_t last = 0;
        NH_HIP(hipMemcpyAsync(&last, out + batch - 1, 8, hipMemcpyDeviceToHost, st));
        NH_HIP(hipStreamSynchronize(st));
        next_base = last + 1;
        r.n += batch;
        if (count ? r.n >= count : last >= below) break;
    }
    if (n_below) {
        uint64_t *d_cnt = nullptr;
        NH_TRY(dalloc(&d_cnt, 1));
        hipLaunchKernelGGL(k_count_below, dim3(1), dim3(1), 0, st, (const uint64_t *)r.d, r.n, below, d_cnt);
        hipError_t e = hipMemcpyAsync(n_below, d_cnt, 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        dfree(d_cnt);
        if (e != hipSuccess) return set_error(NH_E_HIP, "percentage pre-sampler: %s", hipGetErrorString(e));
    }
    return NH_OK;
}

uint64_t isqrt_u64(uint64_t v) {
    uint64_t h = (uint64_t)std::sqrt((double)v);
    while (h > 0 && h * h > v) h--;
    while ((h + 1) * (h + 1) <= v) h++;
    return h;
}

} // namespace

int OriginPlan::prepare(const nh_tables *t, double percentage, uint32_t rank, uint32_t size, uint32_t cx,
                        uint32_t cy, uint32_t radius, hipStream_t st) {
    stream = st;
    tables = t;
    this->rank = rank;
    this->size = size ? size : 1;
    this->cx = cx;
    this->cy = cy;
    this->radius = radius;
    // the enumeration's size N
    uint64_t N = 0;
    if (t->kind == K_ALMOST_INFINITE) {
        if (radius > MAX_CIRCLE_RADIUS)
            return set_error(NH_E_INVALID, "device origin sampling takes circle radii up to %u (got %u)",
                             MAX_CIRCLE_RADIUS, radius);
        const uint64_t d = 2 * (uint64_t)radius + 1, r2 = (uint64_t)radius * radius;
        std::vector<uint64_t> rows(d + 1);
        for (uint64_t row = 0; row < d; row++) {
            const uint64_t dy = row > radius ? row - radius : radius - row;
            rows[row] = N;
            N += 2 * isqrt_u64(r2 - dy * dy) + 1;
        }
        rows[d] = N;
        n_rows = d;
        NH_TRY(dalloc(&d_rows, d + 1));
        NH_HIP(hipMemcpyAsync(d_rows, rows.data(), (d + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        NH_HIP(hipStreamSynchronize(st));
    } else if (t->kind == K_NON_SPATIAL) {
        N = (uint64_t)t->width * t->height * t->deme;
    } else {
        for (uint32_t h : t->habitat) N += h;
    }
    // the pre-sampler's ranks below N (percentage.rs takes p <= 0 as none, p >= 1 as all)
    uint64_t M = 0;
    if (percentage >= 1.0) {
        M = N;
    } else if (percentage > 0.0 && N > 0) {
        ranks = new Ranks();
        NH_TRY(percentage_ranks(percentage, N, 0, false, *ranks, &M, st));
    }
    n = M > this->rank ? (M - this->rank - 1) / this->size + 1 : 0;
    return NH_OK;
}

int OriginPlan::emit(LineageSoA *soa, nh_lineage *aos) const {
    if (n == 0) return NH_OK;
    OriginArgs A{};
    A.n = n;
    A.rank = rank;
    A.size = size;
    A.ranks = ranks ? ranks->d : nullptr;
    A.width = tables->width;
    A.ox = tables->ox;
    A.oy = tables->oy;
    A.deme = tables->deme ? tables->deme : 1;
    A.cells = tables->d_cells;
    A.n_cells = (uint64_t)tables->width * tables->height;
    A.cx = cx;
    A.cy = cy;
    A.radius = radius;
    A.row_start = d_rows;
    A.n_rows = n_rows;
    if (soa) A.soa = *soa;
    A.aos = aos;
    const dim3 g(grid_of(n, 256)), b(256);
    if (tables->kind == K_ALMOST_INFINITE) hipLaunchKernelGGL(k_origins<O_CIRCLE>, g, b, 0, stream, A);
    else if (tables->kind == K_NON_SPATIAL) hipLaunchKernelGGL(k_origins<O_NON_SPATIAL>, g, b, 0, stream, A);
    else hipLaunchKernelGGL(k_origins<O_IN_MEMORY>, g, b, 0, stream, A);
    NH_HIP(hipGetLastError());
    NH_HIP(hipStreamSynchronize(stream));
    return NH_OK;
}

OriginPlan::~OriginPlan() {
    delete ranks;
    dfree(d_rows);
}

int debug_presample(double percentage, uint64_t n, bool serial, uint64_t *out, hipStream_t st) {
    Ranks r;
    NH_TRY(percentage_ranks(percentage, 0, n, serial, r, nullptr, st));
    NH_HIP(hipMemcpyAsync(out, r.d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    return NH_OK;
}

} // namespace nh
