// Labels on the cell grid: core flags, lock-free union-find over cells,
// canonical cluster ids, border attachment, results in input order, meta.
//
// Reference semantics (data_process.cpp:93-266, exact mode at its fixed
// point): cores of one connected component share the id = rank among cores
// (ascending point id) of the component's lowest-id core; a non-core point
// within eps of cores takes the smallest of their ids; else NO_CLUSTER.
//
// Two cores in one cell are always within eps (cell side sqrt(e2/d)
// (1-2^-18)), so a cell with at least one core is one node of the union-find;
// an edge between two core cells needs one core pair that passes the float32
// predicate.
#include <climits>

#include "ndbx_internal.h"

namespace ndbx {

__device__ __forceinline__ int32_t ld_relaxed(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root = smallest index of the set (larger roots are hung under smaller ones).
__device__ int32_t uf_find(int32_t *parent, int32_t x) {
    for (;;) {
        const int32_t p = ld_relaxed(parent + x);
        if (p == x) return x;
        const int32_t gp = ld_relaxed(parent + p);
        if (gp != p) atomicCAS(parent + x, p, gp);  // path halving: gp is an ancestor of x
        x = gp;
    }
}

__device__ void uf_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + b, b, a) == b) return;
    }
}

// Per-cell core presence and lowest core id; core flags as int32 in input
// order for the core-rank scan.
__global__ __launch_bounds__(256) void k_core(const uint8_t *__restrict__ core_s, const uint32_t *__restrict__ perm,
                                              const int32_t *__restrict__ cell_of, int64_t n,
                                              int32_t *__restrict__ cell_core_min, int32_t *__restrict__ cell_has_core,
                                              int32_t *__restrict__ core_i32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool is_core = core_s[i];
    const uint32_t p = perm[i];
    core_i32[p] = is_core;
    if (is_core) {
        const int32_t c = cell_of[i];
        atomicMin(cell_core_min + c, (int32_t)p);
        cell_has_core[c] = 1;
    }
}

__global__ __launch_bounds__(256) void k_iota(int32_t *__restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// One wave per tile of a core cell A: for every neighbour cell B > A with a
// core and a different root, test A's cores of this tile against B's cores
// until one pair passes, then unite A and B.
template <int D>
__global__ __launch_bounds__(256) void k_union(const int2 *__restrict__ tiles, int64_t n_tiles,
                                               const int32_t *__restrict__ cell_start,
                                               const uint64_t *__restrict__ cell_key,
                                               const InstGrid *__restrict__ ig, int lin_bits,
                                               const int32_t *__restrict__ nranges, const int2 *__restrict__ ranges,
                                               int rows_per_cell, const float *__restrict__ soa, int64_t npad,
                                               const uint8_t *__restrict__ core_s,
                                               const int32_t *__restrict__ cell_of,
                                               const int32_t *__restrict__ cell_has_core,
                                               int32_t *__restrict__ parent) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (t >= n_tiles) return;
    const int lane = threadIdx.x & 63;
    const int2 tile = tiles[t];
    const int A = tile.x;
    if (!cell_has_core[A]) return;
    const int q = tile.y + lane;
    const bool active = q < cell_start[A + 1] && core_s[q];
    if (__ballot(active) == 0) return;
    const float e2 = ig[cell_key[A] >> lin_bits].e2;
    float qx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) qx[k] = soa[k * npad + (active ? q : tile.y)];
    const int nr = nranges[A];
    const int2 *rr = ranges + (int64_t)A * rows_per_cell;
    for (int r = 0; r < nr; ++r) {
        const int2 rg = rr[r];
        int j = rg.x;
        while (j < rg.y) {
            const int B = cell_of[j];
            const int bend = cell_start[B + 1];
            if (B > A && cell_has_core[B] && uf_find(parent, A) != uf_find(parent, B)) {
                bool hit = false;
                for (int jj = j; jj < bend; ++jj) {
                    if (!core_s[jj]) continue;
                    float len = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const float df = qx[k] - soa[k * npad + jj];
                        len = __fadd_rn(len, __fmul_rn(df, df));
                    }
                    if (__ballot(active && len <= e2) != 0) {
                        hit = true;
                        break;
                    }
                }
                if (hit && lane == 0) uf_unite(parent, A, B);
            }
            j = bend;
        }
    }
}

// Every core cell's root and the lowest core id of each root's set.
__global__ __launch_bounds__(256) void k_flatten(const int32_t *__restrict__ cell_has_core,
                                                 const int32_t *__restrict__ cell_core_min, int64_t nc,
                                                 int32_t *__restrict__ parent, int32_t *__restrict__ root_min) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc || !cell_has_core[c]) return;
    const int32_t r = uf_find(parent, (int32_t)c);
    atomicMin(root_min + r, cell_core_min[c]);
}

__global__ __launch_bounds__(256) void k_core_base(const int32_t *__restrict__ rank, const int32_t *__restrict__ total,
                                                   int64_t n, InstGrid *__restrict__ ig, int n_inst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const int64_t b = ig[i].point_begin;
    ig[i].core_base = b < n ? rank[b] : *total;
}

// Cluster id of each core cell: rank (within its instance) of the lowest-id
// core of its set.
__global__ __launch_bounds__(256) void k_cell_label(const int32_t *__restrict__ cell_has_core,
                                                    const uint64_t *__restrict__ cell_key,
                                                    const InstGrid *__restrict__ ig, int lin_bits, int64_t nc,
                                                    const int32_t *__restrict__ parent,
                                                    const int32_t *__restrict__ root_min,
                                                    const int32_t *__restrict__ rank, int32_t *__restrict__ cell_label) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (!cell_has_core[c]) {
        cell_label[c] = NDBX_NO_CLUSTER;
        return;
    }
    int32_t r = (int32_t)c;
    for (int32_t p = parent[r]; p != r; p = parent[r]) r = p;
    cell_label[c] = rank[root_min[r]] - ig[cell_key[c] >> lin_bits].core_base;
}

// Labels per point: cores take their cell's id; a non-core point takes the
// smallest id among core cells holding a core within eps (its own cell's
// cores are within eps by construction).
template <int D>
__global__ __launch_bounds__(256) void k_border(const int2 *__restrict__ tiles, int64_t n_tiles,
                                                const int32_t *__restrict__ cell_start,
                                                const uint64_t *__restrict__ cell_key,
                                                const InstGrid *__restrict__ ig, int lin_bits,
                                                const int32_t *__restrict__ nranges, const int2 *__restrict__ ranges,
                                                int rows_per_cell, const float *__restrict__ soa, int64_t npad,
                                                const uint8_t *__restrict__ core_s,
                                                const int32_t *__restrict__ cell_of,
                                                const int32_t *__restrict__ cell_label,
                                                const uint32_t *__restrict__ perm, int32_t *__restrict__ labels) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (t >= n_tiles) return;
    const int lane = threadIdx.x & 63;
    const int2 tile = tiles[t];
    const int A = tile.x;
    const int q = tile.y + lane;
    const bool valid = q < cell_start[A + 1];
    const bool is_core = valid && core_s[q];
    const int32_t LA = cell_label[A];
    int32_t best = LA;  // NO_CLUSTER when A has no core
    bool need = valid && !is_core;
    if (__ballot(need) != 0) {
        const float e2 = ig[cell_key[A] >> lin_bits].e2;
        float qx[D];
#pragma unroll
        for (int k = 0; k < D; ++k) qx[k] = soa[k * npad + (valid ? q : tile.y)];
        const int nr = nranges[A];
        const int2 *rr = ranges + (int64_t)A * rows_per_cell;
        for (int r = 0; r < nr; ++r) {
            const int2 rg = rr[r];
            int j = rg.x;
            while (j < rg.y) {
                const int B = cell_of[j];
                const int bend = cell_start[B + 1];
                const int32_t LB = cell_label[B];
                if (B != A && LB != NDBX_NO_CLUSTER) {
                    bool want = need && LB < best;
                    for (int jj = j; jj < bend && __ballot(want) != 0; ++jj) {
                        if (!core_s[jj]) continue;
                        float len = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) {
                            const float df = qx[k] - soa[k * npad + jj];
                            len = __fadd_rn(len, __fmul_rn(df, df));
                        }
                        if (want && len <= e2) {
                            best = LB;
                            want = false;
                        }
                    }
                }
                j = bend;
            }
        }
    }
    if (valid) labels[perm[q]] = best;
}

// Per-instance (cores, noise, clusters) from input-order flags and labels;
// clusters = core cells that are roots.
__global__ __launch_bounds__(256) void k_meta_points(const uint8_t *__restrict__ core, const int32_t *__restrict__ labels,
                                                     const int64_t *__restrict__ off, int n_inst, int64_t n,
                                                     int32_t *__restrict__ meta) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = i < n;
    const int c = v && core[i];
    const int z = v && labels[i] == NDBX_NO_CLUSTER;
    if (n_inst == 1) {
        int a = c, b = z;
        for (int o = 32; o >= 1; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            if (a) atomicAdd(meta + 0, a);
            if (b) atomicAdd(meta + 1, b);
        }
        return;
    }
    if (!v) return;
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    if (c) atomicAdd(meta + 3 * lo + 0, 1);
    if (z) atomicAdd(meta + 3 * lo + 1, 1);
}

__global__ __launch_bounds__(256) void k_meta_cells(const int32_t *__restrict__ cell_has_core,
                                                    const uint64_t *__restrict__ cell_key, int lin_bits,
                                                    const int32_t *__restrict__ parent, int64_t nc,
                                                    int32_t *__restrict__ meta) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (cell_has_core[c] && parent[c] == (int32_t)c) atomicAdd(meta + 3 * (cell_key[c] >> lin_bits) + 2, 1);
}

template <int D>
static void grid_label_d(Ctx &c) {
    const int64_t n = c.n, nc = c.n_cells, nt = c.n_tiles;
    const hipStream_t s = c.stream;
    const int nb = (int)ceil_div(n, 256), ncb = (int)ceil_div(nc, 256);
    InstGrid *ig = c.inst_d.as<InstGrid>();
    int rows = 1;
    for (int k = 1; k < D; ++k) rows *= 2 * c.reach_max + 1;
    const int64_t npad = ceil_div(n, 64) * 64;

    c.cell_core_min.ensure(nc * 4);
    c.cell_has_core.ensure(nc * 4);
    c.rank.ensure((n + 1) * 4);
    c.label_s.ensure(n * 4);  // core flags as int32, input order
    fill_i32(c.cell_core_min.as<int32_t>(), INT_MAX, nc, s, c.prof);
    fill_i32(c.cell_has_core.as<int32_t>(), 0, nc, s, c.prof);
    {
        Launch L(c.prof, s, "core");
        k_core<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.vals.as<uint32_t>(), c.cell_of.as<int32_t>(), n,
                                  c.cell_core_min.as<int32_t>(), c.cell_has_core.as<int32_t>(),
                                  c.label_s.as<int32_t>());
    }
    int32_t *d_total = c.misc.as<int32_t>();
    scan_exclusive_i32(c.label_s.as<int32_t>(), c.rank.as<int32_t>(), n, d_total, c.scratch, s, c.prof);
    {
        Launch L(c.prof, s, "core_base");
        k_core_base<<<(unsigned)ceil_div(c.n_inst, 256), 256, 0, s>>>(c.rank.as<int32_t>(), d_total, n, ig, c.n_inst);
    }
    c.parent.ensure(nc * 4);
    c.comp_min.ensure(nc * 4);
    c.cell_label.ensure(nc * 4);
    {
        Launch L(c.prof, s, "uf_init");
        k_iota<<<ncb, 256, 0, s>>>(c.parent.as<int32_t>(), nc);
    }
    fill_i32(c.comp_min.as<int32_t>(), INT_MAX, nc, s, c.prof);
    const unsigned tb = (unsigned)ceil_div(nt, 4);
    {
        Launch L(c.prof, s, "union");
        k_union<D><<<tb, 256, 0, s>>>(c.tiles.as<int2>(), nt, c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(),
                                      ig, c.lin_bits, c.nranges.as<int32_t>(), c.ranges.as<int2>(), rows,
                                      c.soa.as<float>(), npad, c.core_s.as<uint8_t>(), c.cell_of.as<int32_t>(),
                                      c.cell_has_core.as<int32_t>(), c.parent.as<int32_t>());
    }
    {
        Launch L(c.prof, s, "flatten");
        k_flatten<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_core_min.as<int32_t>(), nc,
                                      c.parent.as<int32_t>(), c.comp_min.as<int32_t>());
    }
    {
        Launch L(c.prof, s, "cell_label");
        k_cell_label<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_key.as<uint64_t>(), ig, c.lin_bits, nc,
                                         c.parent.as<int32_t>(), c.comp_min.as<int32_t>(), c.rank.as<int32_t>(),
                                         c.cell_label.as<int32_t>());
    }
    {
        Launch L(c.prof, s, "border");
        k_border<D><<<tb, 256, 0, s>>>(c.tiles.as<int2>(), nt, c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(),
                                       ig, c.lin_bits, c.nranges.as<int32_t>(), c.ranges.as<int2>(), rows,
                                       c.soa.as<float>(), npad, c.core_s.as<uint8_t>(), c.cell_of.as<int32_t>(),
                                       c.cell_label.as<int32_t>(), c.vals.as<uint32_t>(), c.labels.as<int32_t>());
    }
    fill_i32(c.meta.as<int32_t>(), 0, 3 * c.n_inst, s, c.prof);
    {
        Launch L(c.prof, s, "meta");
        k_meta_points<<<nb, 256, 0, s>>>(c.core.as<uint8_t>(), c.labels.as<int32_t>(), c.offsets_d.as<int64_t>(),
                                         c.n_inst, n, c.meta.as<int32_t>());
        k_meta_cells<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_key.as<uint64_t>(), c.lin_bits,
                                         c.parent.as<int32_t>(), nc, c.meta.as<int32_t>());
    }
}

void grid_label(Ctx &c) {
    switch (c.d) {
        case 1: grid_label_d<1>(c); break;
        case 2: grid_label_d<2>(c); break;
        case 3: grid_label_d<3>(c); break;
        default: throw Error(NDBX_EUNSUPPORTED, "grid path needs d <= 3");
    }
}

}  // namespace ndbx
