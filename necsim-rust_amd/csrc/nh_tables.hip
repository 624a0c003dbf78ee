// nh_tables.hip -- resident tables of the in-memory map scenarios built on
// the GPU from the habitat map and a sparse (row-major, ascending-column)
// dispersal map: the same CellRec / Atom bytes as nh_build_alias_tables
// followed by nh_tables_create, without the host's sequential row loop.
//
// InMemoryPackedSeparableAliasDispersalSampler::new
// (necsim/impls/no-std/src/cogs/dispersal_sampler/in_memory/packed_separable_alias/mod.rs:64-203),
// InMemoryPackedAliasDispersalSampler::new (.../in_memory/packed_alias/mod.rs:58-118),
// AliasMethodSamplerAtom::create (necsim/impls/no-std/src/alias/packed.rs:59-120) and
// the in-memory dispersal contract (.../in_memory/contract.rs:13-69).  One
// thread owns one origin row: the Vose pairing is sequential within a row (its
// LIFO stack order fixes the result) and independent across rows.
#include "../../include/necsim_hip.h"
#include "nh_internal.h"
#include "nh_kernels.h"

using namespace nh;

namespace {

// first failure of a row, in the host builder's order of checks
enum : unsigned long long { MAP_OK = 0, MAP_BAD_COLUMN = 1, MAP_CONTRACT = 2, MAP_NO_DISPERSAL = 3 };

// Pass 1: per row, its first contract failure (lowest row wins, as the host
// loop stops at the first failing row) and its atom count.
__global__ void __launch_bounds__(256) k_map_count(const uint32_t *habitat, const uint64_t *row_ptr,
                                                   const uint32_t *col, const double *weight, uint64_t cells,
                                                   int separable, uint64_t *count,
                                                   unsigned long long *first_error) {
    const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= cells) return;
    count[row] = 0;
    uint64_t n = 0;
    double total = 0.0, self_w = 0.0;
    for (uint64_t j = row_ptr[row]; j < row_ptr[row + 1]; j++) {
        const uint32_t c = col[j];
        const double p = weight[j];
        if (c >= cells || (j > row_ptr[row] && c <= col[j - 1])) {
            atomicMin(first_error, (row << 2) | MAP_BAD_COLUMN);
            return;
        }
        if (p < 0.0 || (p > 0.0 && (habitat[row] == 0 || habitat[c] == 0))) {
            atomicMin(first_error, (row << 2) | MAP_CONTRACT);
            return;
        }
        const double wt = p * (double)habitat[c];
        if (wt > 0.0) {
            n++;
            total = total + wt;
            if (c == row) self_w = wt;
        }
    }
    if (habitat[row] > 0 && n == 0) {
        atomicMin(first_error, (row << 2) | MAP_NO_DISPERSAL);
        return;
    }
    uint64_t m = n;
    // separable rows whose self-dispersal fills a whole bucket drop it from the atoms
    if (separable && n > 0 && !(self_w * (double)n / total < 1.0)) m = n - 1;
    count[row] = m;
}

__global__ void k_widen_u32(const uint32_t *in, uint64_t *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// in-place exclusive scan of data[0, n) by one block; data[n] = total
__global__ void __launch_bounds__(1024) k_scan_u64(uint64_t *data, uint64_t n) {
    __shared__ uint64_t partial[1024];
    const int T = blockDim.x;
    const uint64_t per = (n + T - 1) / T;
    const uint64_t lo = per * threadIdx.x, hi = lo + per < n ? lo + per : n;
    uint64_t sum = 0;
    for (uint64_t j = lo; j < hi; j++) sum += data[j];
    partial[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        uint64_t v = threadIdx.x >= (unsigned)off ? partial[threadIdx.x - off] : 0;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? partial[threadIdx.x - 1] : 0;
    for (uint64_t j = lo; j < hi; j++) {
        uint64_t v = data[j];
        data[j] = run;
        run += v;
    }
    if (threadIdx.x == T - 1) data[n] = partial[T - 1];
}

// AliasMethodSamplerAtom::create over the row's included entries (all
// positive ones, or all but self-dispersal): u_i = w_i * n / sum(w) with the
// sum folded in column order, then pairs popped from the ends of the
// overfull / underfull index stacks; leftovers get u = 1.
__device__ void alias_create_row(Atom *A, uint32_t *over, uint32_t *under, const uint32_t *habitat,
                                 const uint32_t *col, const double *weight, uint64_t j0, uint64_t j1,
                                 uint64_t row, bool skip_self, uint32_t m) {
    double total = 0.0;
    for (uint64_t j = j0; j < j1; j++) {
        const double wt = weight[j] * (double)habitat[col[j]];
        if (wt > 0.0 && !(skip_self && col[j] == row)) total = total + wt;
    }
    const double nf = (double)m;
    uint32_t i = 0;
    for (uint64_t j = j0; j < j1; j++) {
        const uint32_t c = col[j];
        const double wt = weight[j] * (double)habitat[c];
        if (wt > 0.0 && !(skip_self && c == row)) A[i++] = Atom{wt * nf / total, c, c};
    }
    uint32_t no = 0, nu = 0;
    for (uint32_t k = 0; k < m; k++)
        if (A[k].u > 1.0) over[no++] = k;
    for (uint32_t k = 0; k < m; k++)
        if (A[k].u < 1.0) under[nu++] = k;
    while (no > 0 && nu > 0) {
        const uint32_t o = over[--no];
        const uint32_t u = under[--nu];
        A[o].u = A[o].u + A[u].u - 1.0;
        A[u].k = A[o].e;
        if (A[o].u < 1.0) under[nu++] = o;
        else if (A[o].u > 1.0) over[no++] = o;
    }
    for (uint32_t k = 0; k < no; k++) A[over[k]].u = 1.0;
    for (uint32_t k = 0; k < nu; k++) A[under[k]].u = 1.0;
}

// Pass 3: one thread per row writes the row's atoms and its cell record.
__global__ void __launch_bounds__(256) k_map_build(const uint32_t *habitat, const uint64_t *row_ptr,
                                                   const uint32_t *col, const double *weight, uint64_t cells,
                                                   int separable, const uint64_t *atom_start,
                                                   const uint64_t *prefix, Atom *atoms, uint32_t *stack_a,
                                                   uint32_t *stack_b, CellRec *rec_out) {
    const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= cells) return;
    const uint64_t start = atom_start[row];
    const uint32_t m = (uint32_t)(atom_start[row + 1] - start);
    const uint64_t j0 = row_ptr[row], j1 = row_ptr[row + 1];
    CellRec rec;
    rec.prefix = prefix[row];
    rec.habitat = habitat[row];
    rec.atom_start = (uint32_t)start;
    rec.atom_len = m;
    rec.non_self = (uint32_t)row;
    rec.self_p = 0.0;
    Atom *A = atoms + start;
    uint64_t n = 0;
    double total = 0.0, self_w = 0.0;
    for (uint64_t j = j0; j < j1; j++) {
        const double wt = weight[j] * (double)habitat[col[j]];
        if (wt > 0.0) {
            n++;
            total = total + wt;
            if (col[j] == row) self_w = wt;
        }
    }
    if (n > 0) {
        if (!separable) {
            alias_create_row(A, stack_a + start, stack_b + start, habitat, col, weight, j0, j1, row, false, m);
        } else {
            rec.self_p = self_w / total;
            const double self_u = self_w * (double)n / total;
            if (self_u < 1.0) {
                alias_create_row(A, stack_a + start, stack_b + start, habitat, col, weight, j0, j1, row, false,
                                 m);
                // exactly one atom holds self-dispersal: flip it into the k
                // slot and move it last (packed_separable_alias/mod.rs:147-168)
                for (uint32_t k = 0; k < m; k++) {
                    if (A[k].e == row || A[k].k == row) {
                        Atom a = A[k];
                        if (a.e == row) {
                            const uint32_t t = a.e;
                            a.e = a.k;
                            a.k = t;
                            a.u = 1.0 - a.u;
                        }
                        rec.non_self = a.e;
                        A[k] = A[m - 1];
                        A[m - 1] = a;
                        break;
                    }
                }
            } else if (m > 0) {
                alias_create_row(A, stack_a + start, stack_b + start, habitat, col, weight, j0, j1, row, true, m);
            }
        }
    }
    rec_out[row] = rec;
}

} // namespace

extern "C" int nh_tables_create_from_map(nh_ctx *ctx, const nh_map_scenario *s, nh_tables **out) {
    if (!ctx || !s || !out) return set_error(NH_E_INVALID, "nh_tables_create_from_map: null argument");
    *out = nullptr;
    int separable;
    if (s->kind == NH_SCENARIO_SPATIALLY_EXPLICIT) separable = 1;
    else if (s->kind == NH_SCENARIO_SPATIALLY_EXPLICIT_PACKED) separable = 0;
    else return set_error(NH_E_INVALID, "nh_tables_create_from_map: not an in-memory map scenario");
    const uint64_t cells = (uint64_t)s->width * s->height;
    if (!cells || cells > 0xffffffffull || !s->habitat || !s->row_ptr)
        return set_error(NH_E_INVALID, "nh_tables_create_from_map: empty extent or missing map");
    if (s->row_ptr[0] != 0) return set_error(NH_E_INVALID, "dispersal rows must start at 0");
    for (uint64_t r = 0; r < cells; r++)
        if (s->row_ptr[r + 1] < s->row_ptr[r]) return set_error(NH_E_INVALID, "dispersal rows out of order");
    const uint64_t nnz = s->row_ptr[cells];
    if (nnz && (!s->col || !s->weight)) return set_error(NH_E_INVALID, "nh_tables_create_from_map: missing entries");
    NH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;

    uint32_t *d_hab = nullptr, *d_col = nullptr, *d_sa = nullptr, *d_sb = nullptr;
    uint64_t *d_rp = nullptr, *d_count = nullptr, *d_prefix = nullptr;
    double *d_w = nullptr;
    unsigned long long *d_err = nullptr;
    nh_tables *t = new nh_tables();
    auto cleanup = [&]() {
        dfree(d_hab);
        dfree(d_col);
        dfree(d_sa);
        dfree(d_sb);
        dfree(d_rp);
        dfree(d_count);
        dfree(d_prefix);
        dfree(d_w);
        dfree(d_err);
    };
    auto fail = [&](int code) {
        cleanup();
        nh_tables_destroy(t);
        return code;
    };
    int rc;
    if ((rc = dalloc(&d_hab, cells)) || (rc = dalloc(&d_rp, cells + 1)) || (rc = dalloc(&d_col, nnz)) ||
        (rc = dalloc(&d_w, nnz)) || (rc = dalloc(&d_count, cells + 1)) || (rc = dalloc(&d_prefix, cells + 1)) ||
        (rc = dalloc(&d_err, 1)))
        return fail(rc);
    hipError_t e = hipMemcpyAsync(d_hab, s->habitat, cells * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_rp, s->row_ptr, (cells + 1) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(d_col, s->col, nnz * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(d_w, s->weight, nnz * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return fail(set_error(NH_E_HIP, "map upload: %s", hipGetErrorString(e)));
    launch_fill_u64(reinterpret_cast<uint64_t *>(d_err), ~0ull, 1, st);
    const unsigned grid = (unsigned)((cells + 255) / 256);
    hipLaunchKernelGGL(k_map_count, dim3(grid), dim3(256), 0, st, d_hab, d_rp, d_col, d_w, cells, separable,
                       d_count, d_err);
    hipLaunchKernelGGL(k_widen_u32, dim3(grid), dim3(256), 0, st, d_hab, d_prefix, cells);
    hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(1024), 0, st, d_count, cells);
    hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(1024), 0, st, d_prefix, cells);
    unsigned long long err = 0;
    uint64_t n_atoms = 0, hab_total = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&err, d_err, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&n_atoms, d_count + cells, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&hab_total, d_prefix + cells, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(set_error(NH_E_HIP, "map count: %s", hipGetErrorString(e)));
    if (err != ~0ull) {
        const unsigned long long row = err >> 2;
        switch (err & 3) {
        case MAP_BAD_COLUMN:
            return fail(set_error(NH_E_INVALID, "dispersal column outside the habitat or not ascending (row %llu)", row));
        case MAP_CONTRACT: return fail(set_error(NH_E_CONTRACT, "dispersal from/to non-habitat (row %llu)", row));
        default: return fail(set_error(NH_E_CONTRACT, "no dispersal from a habitable cell (row %llu)", row));
        }
    }
    if (n_atoms > 0xffffffffull) return fail(set_error(NH_E_INVALID, "more than 2^32 atoms"));
    if (hab_total == 0) return fail(set_error(NH_E_INVALID, "habitat map is empty"));
    if ((rc = dalloc(&t->d_cells, cells)) || (rc = dalloc(&t->d_atoms, n_atoms)) || (rc = dalloc(&d_sa, n_atoms)) ||
        (rc = dalloc(&d_sb, n_atoms)))
        return fail(rc);
    hipLaunchKernelGGL(k_map_build, dim3(grid), dim3(256), 0, st, d_hab, d_rp, d_col, d_w, cells, separable,
                       d_count, d_prefix, t->d_atoms, d_sa, d_sb, t->d_cells);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(set_error(NH_E_HIP, "map build: %s", hipGetErrorString(e)));
    cleanup();

    t->ctx = ctx;
    t->kind = separable ? K_SEPARABLE : K_PACKED;
    t->width = s->width;
    t->height = s->height;
    t->ox = s->origin_x;
    t->oy = s->origin_y;
    t->habitat.assign(s->habitat, s->habitat + cells);
    t->n_atoms = n_atoms;
    t->bytes = cells * sizeof(CellRec) + n_atoms * sizeof(Atom);
    DevScenario &d = t->dev;
    d.kind = t->kind;
    d.width = s->width;
    d.height = s->height;
    d.ox = s->origin_x;
    d.oy = s->origin_y;
    d.deme = 1;
    d.sigma = 0.0;
    d.area = cells;
    d.area_f = (double)cells;
    d.deme_f = 1.0;
    d.cells = t->d_cells;
    d.atoms = t->d_atoms;
    *out = t;
    return NH_OK;
}

extern "C" int nh_tables_export(const nh_tables *t, void *cells, size_t cells_cap, void *atoms, size_t atoms_cap,
                                size_t *n_cells, size_t *n_atoms) {
    if (!t || !n_cells || !n_atoms) return set_error(NH_E_INVALID, "nh_tables_export: null argument");
    const bool in_memory = t->kind == K_SEPARABLE || t->kind == K_PACKED;
    *n_cells = in_memory ? (size_t)t->width * t->height : 0;
    *n_atoms = in_memory ? (size_t)t->n_atoms : 0;
    if (!in_memory || (!cells && !atoms)) return NH_OK;
    if ((cells && cells_cap < *n_cells) || (atoms && atoms_cap < *n_atoms))
        return set_error(NH_E_CAPACITY, "nh_tables_export: buffers too small");
    NH_HIP(hipSetDevice(t->ctx->device));
    if (cells) NH_HIP(hipMemcpy(cells, t->d_cells, *n_cells * sizeof(CellRec), hipMemcpyDeviceToHost));
    if (atoms && *n_atoms) NH_HIP(hipMemcpy(atoms, t->d_atoms, *n_atoms * sizeof(Atom), hipMemcpyDeviceToHost));
    return NH_OK;
}
