// Chunked runs: classify and label slab by slab, so that the grid pipeline's
// buffers are sized for one slab rather than the whole input (the reference
// bounds its working copies the same way, to chunks of 100M points for the
// counts and for the label sweeps, data_process.cpp:285-302, :335-362).
//
// The points are sorted by x once; slab k owns a range of that order and
// works on its owned points plus every point within the predicate's reach in
// x (ndbx_slabs.h), gathered into a slab buffer in sorted order.  Per-point
// state stays resident for the whole input: counts, core flags, labels, core
// ranks and a union-find over core ranks (17 bytes per point).
//   A  each slab: grid + counts; the owned points' counts are exact.
//   B  each slab: grid with every point's true core flag, the cell
//      union-find of the label pass; each slab cluster joins the global
//      union-find over core ranks (an edge between two cores lies in the slab
//      of either end's owner, so the union of the slab clusters is the
//      clustering).
//   C  the global sets flattened (root = smallest core rank = the cluster
//      id); each slab: the label pass's border search against those ids
//      (cell labels from the ids, no union-find), owned points' labels kept.
// Results equal the resident run's bit for bit.
#include <climits>
#include <cmath>
#include <memory>

#include "ndbx_internal.h"
#include "ndbx_slabs.h"

namespace ndbx {

// float -> unsigned key in the same order
__device__ __forceinline__ uint32_t x_order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_ch_keys(const float *__restrict__ x, int64_t n, int d,
                                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = x_order_key(x[i * d]);
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_ch_sorted_x(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                     int64_t n, int d, float *__restrict__ xs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xs[i] = x[(int64_t)perm[i] * d];
}

// The slab's points (sorted positions [lo, lo + m)) into a buffer of its own,
// with their input indices.
__global__ __launch_bounds__(256) void k_ch_gather(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                   int64_t lo, int64_t m, int d, float *__restrict__ out,
                                                   int32_t *__restrict__ gid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t g = perm[lo + i];
    gid[i] = (int32_t)g;
    for (int k = 0; k < d; ++k) out[i * d + k] = x[g * d + k];
}

// Owned slab points [o0, o1) write their per-point value back in input order.
__global__ __launch_bounds__(256) void k_ch_scatter(const int32_t *__restrict__ v, const int32_t *__restrict__ gid,
                                                    int64_t o0, int64_t o1, int32_t *__restrict__ out) {
    const int64_t i = o0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < o1) out[gid[i]] = v[i];
}

// Core flags in input order (bytes, and int32 for the rank scan).
__global__ __launch_bounds__(256) void k_ch_core(const int32_t *__restrict__ nn, int64_t n, int32_t m,
                                                 uint8_t *__restrict__ core, int32_t *__restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool c = nn[i] >= m;  // data_process.cpp:282
    core[i] = c;
    flag[i] = c;
}

// The slab context's counts and core flags (its input order and its sorted
// order) replaced by the whole input's: halo points keep their true core
// flags, whatever the slab sees of their neighbours.
__global__ __launch_bounds__(256) void k_ch_inject(const int32_t *__restrict__ nn, const int32_t *__restrict__ gid,
                                                   const uint32_t *__restrict__ perm_sub, int64_t ns, int32_t m,
                                                   int32_t *__restrict__ nn_sub, uint8_t *__restrict__ core_sub,
                                                   int32_t *__restrict__ nn_s, uint8_t *__restrict__ core_s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int32_t a = nn[gid[i]];
    nn_sub[i] = a;
    core_sub[i] = a >= m;
    const int32_t b = nn[gid[perm_sub[i]]];
    nn_s[i] = b;
    core_s[i] = b >= m;
}

// Slab index of each slab core rank.
__global__ __launch_bounds__(256) void k_ch_rank_inv(const uint8_t *__restrict__ core_sub,
                                                     const int32_t *__restrict__ rank, int64_t ns,
                                                     int32_t *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns && core_sub[i]) inv[rank[i]] = (int32_t)i;
}

__device__ __forceinline__ int32_t ch_ld(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int32_t ch_find(int32_t *parent, int32_t x) {
    for (;;) {
        const int32_t p = ch_ld(parent + x);
        if (p == x) return x;
        const int32_t gp = ch_ld(parent + p);
        if (gp != p) atomicCAS(parent + x, p, gp);
        x = gp;
    }
}

// Lock-free union: the larger root hangs under the smaller, so a set's root
// is its smallest core rank.
__device__ void ch_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = ch_find(parent, a);
        b = ch_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + b, b, a) == b) return;
    }
}

// Each slab core joins the global set of its slab cluster's representative
// (the slab core whose slab rank is the cluster's id).
__global__ __launch_bounds__(256) void k_ch_union(const uint8_t *__restrict__ core_sub,
                                                  const int32_t *__restrict__ lab_sub, const int32_t *__restrict__ inv,
                                                  const int32_t *__restrict__ gid, const int32_t *__restrict__ grank,
                                                  int64_t ns, int32_t *__restrict__ parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns || !core_sub[i]) return;
    const int32_t rep = inv[lab_sub[i]];
    const int32_t a = grank[gid[i]], b = grank[gid[rep]];
    if (a != b) ch_unite(parent, a, b);
}

__global__ __launch_bounds__(256) void k_ch_iota(int32_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_ch_flatten(int32_t *__restrict__ parent, int64_t nk) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) parent[k] = ch_find(parent, (int32_t)k);
}

// Cluster id of each slab core (the global root of its core rank);
// NO_CLUSTER for the others.
__global__ __launch_bounds__(256) void k_ch_ext(const uint8_t *__restrict__ core_sub, const int32_t *__restrict__ gid,
                                                const int32_t *__restrict__ grank, const int32_t *__restrict__ parent,
                                                int64_t ns, int32_t *__restrict__ ext) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) ext[i] = core_sub[i] ? parent[grank[gid[i]]] : NDBX_NO_CLUSTER;
}

// (cores, noise, clusters) of the whole input (data_process.cpp:50-91).
__global__ __launch_bounds__(256) void k_ch_meta(const uint8_t *__restrict__ core, const int32_t *__restrict__ labels,
                                                 int64_t n, const int32_t *__restrict__ parent, int64_t nk,
                                                 int32_t *__restrict__ meta) {
    int32_t a = 0, b = 0, r = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a += core[i];
        b += labels[i] == NDBX_NO_CLUSTER;
        if (i < nk) r += parent[i] == (int32_t)i;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        r += __shfl_xor(r, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(meta + 0, a);
        atomicAdd(meta + 1, b);
        atomicAdd(meta + 2, r);
    }
}

// ------------------------------------------------------------------ host --
static void merge_prof(Profiler &dst, Profiler &src) {
    for (size_t i = 0; i < src.names.size(); ++i) {
        const int slot = dst.slot_of(src.names[i].c_str());
        dst.ms[slot] += src.ms[i];
        dst.calls[slot] += src.calls[i];
    }
    src.reset();
}

static Ctx &slab_ctx(Ctx &c) {
    if (!c.sub) {
        c.sub = new Ctx();
        Ctx &s = *c.sub;
        s.device = c.device;
        s.stream = c.stream;
        s.own_stream = false;
        s.d = c.d;
        s.n_inst = 1;
        s.flags = NDBX_DEVICE_INPUT;
        s.group_cells = c.group_cells;
        s.h_eps = c.h_eps;
        s.h_m = c.h_m;
    }
    Ctx &s = *c.sub;
    s.prof.on = c.prof.on;
    s.prof.only = c.prof.only;
    return s;
}

// Gather slab k's points and build its grid (with or without counts).
static Ctx &slab_grid(Ctx &c, const Slab &sl, bool count) {
    Ctx &s = slab_ctx(c);
    const int64_t m = sl.sub_hi - sl.sub_lo;
    c.ch_coords.ensure((size_t)m * c.d * 4);
    c.ch_gid.ensure((size_t)m * 4);
    {
        Launch L(c.prof, c.stream, "chunk_gather");
        k_ch_gather<<<(unsigned)ceil_div(m, 256), 256, 0, c.stream>>>(c.coords, c.ch_perm.as<uint32_t>(), sl.sub_lo, m,
                                                                      c.d, c.ch_coords.as<float>(), c.ch_gid.as<int32_t>());
    }
    s.n = m;
    s.h_offsets = {0, m};
    s.coords = c.ch_coords.as<float>();
    s.nn.ensure((size_t)m * 4);
    s.labels.ensure((size_t)m * 4);
    s.core.ensure((size_t)m);
    s.meta.ensure(3 * 4);
    s.no_count = !count;
    s.ext_label = nullptr;
    grid_plan(s);
    if (s.path != PATH_GRID) throw Error(NDBX_EUNSUPPORTED, "chunked run: a slab left the cell grid");
    grid_classify(s);
    return s;
}

// The slab's core flags from the whole input's counts.
static void slab_inject(Ctx &c, Ctx &s) {
    Launch L(c.prof, c.stream, "chunk_inject");
    k_ch_inject<<<(unsigned)ceil_div(s.n, 256), 256, 0, c.stream>>>(
        c.nn.as<int32_t>(), c.ch_gid.as<int32_t>(), s.vals.as<uint32_t>(), s.n, c.h_m[0], s.nn.as<int32_t>(),
        s.core.as<uint8_t>(), s.nn_s.as<int32_t>(), s.core_s.as<uint8_t>());
}

void chunked_classify(Ctx &c) {
    const int64_t n = c.n;
    const hipStream_t st = c.stream;
    grid_plan(c);  // bounds and path over the whole input
    if (c.path != PATH_GRID || c.n_inst != 1)
        throw Error(NDBX_EUNSUPPORTED, "chunked runs take the cell grid: one instance, d <= 3, finite coordinates");
    // points sorted by x; the sorted x values plan the slabs on the host
    c.keys.ensure((size_t)n * 8);
    c.keys2.ensure((size_t)n * 8);
    c.vals.ensure((size_t)n * 4);
    c.vals2.ensure((size_t)n * 4);
    {
        Launch L(c.prof, st, "chunk_sort");
        k_ch_keys<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(c.coords, n, c.d, c.keys.as<uint64_t>(),
                                                              c.vals.as<uint32_t>());
    }
    bool alt = false;
    radix_sort_u64(c.keys.as<uint64_t>(), c.vals.as<uint32_t>(), c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>(), n, 32,
                   c.scratch, st, c.prof, &alt);
    const uint32_t *perm = alt ? c.vals2.as<uint32_t>() : c.vals.as<uint32_t>();
    c.ch_perm.ensure((size_t)n * 4);
    NDBX_HIP(hipMemcpyAsync(c.ch_perm.p, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    float *xs_d = reinterpret_cast<float *>(c.keys.p);  // keys are no longer needed
    {
        Launch L(c.prof, st, "chunk_sort");
        k_ch_sorted_x<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(c.coords, c.ch_perm.as<uint32_t>(), n, c.d, xs_d);
    }
    std::vector<float> xs((size_t)n);
    NDBX_HIP(hipMemcpyAsync(xs.data(), xs_d, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    c.sync();
    // the sort buffers (24 bytes per point) are not kept
    c.keys.release();
    c.keys2.release();
    c.vals.release();
    c.vals2.release();
    // reach: accepted pairs are within sqrt(e2 (1 + 2^-18)) in true distance
    const float e = c.h_eps[0];
    const float e2 = e * e;
    const double reach = std::sqrt((double)e2 * (1.0 + std::ldexp(1.0, -16)));
    std::vector<Slab> slabs;
    int64_t bad = -1;
    if (!plan_slabs(xs.data(), n, reach, c.chunk_points, slabs, &bad))
        throw Error(NDBX_EUNSUPPORTED, "chunked run: the points within eps of x = " + std::to_string(xs[bad]) +
                                           " exceed the slab budget of " + std::to_string(c.chunk_points) +
                                           " points");
    c.n_slabs = (int64_t)slabs.size();
    c.h_slabs = slabs;
    // A: counts, slab by slab
    int64_t cells = 0, tiles = 0, tests = 0;
    for (const Slab &sl : slabs) {
        Ctx &s = slab_grid(c, sl, true);
        cells += s.n_cells;
        tiles += s.n_tiles;
        tests += s.pair_tests;
        const int64_t o0 = sl.own_lo - sl.sub_lo, o1 = sl.own_hi - sl.sub_lo;
        {
            Launch L(c.prof, st, "chunk_scatter");
            k_ch_scatter<<<(unsigned)ceil_div(std::max<int64_t>(o1 - o0, 1), 256), 256, 0, st>>>(
                s.nn.as<int32_t>(), c.ch_gid.as<int32_t>(), o0, o1, c.nn.as<int32_t>());
        }
        s.sync();
        merge_prof(c.prof, s.prof);
    }
    c.n_cells = cells;
    c.n_tiles = tiles;
    c.pair_tests = tests;
    // core flags and core ranks of the whole input
    c.ch_grank.ensure((size_t)(n + 1) * 4);
    c.label_s.ensure((size_t)n * 4);
    {
        Launch L(c.prof, st, "chunk_core");
        k_ch_core<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(c.nn.as<int32_t>(), n, c.h_m[0], c.core.as<uint8_t>(),
                                                              c.label_s.as<int32_t>());
    }
    scan_exclusive_i32(c.label_s.as<int32_t>(), c.ch_grank.as<int32_t>(), n, c.ch_grank.as<int32_t>() + n, c.scratch, st,
                       c.prof);
    c.sync();
}

void chunked_label(Ctx &c) {
    const int64_t n = c.n;
    const hipStream_t st = c.stream;
    int32_t n_cores = 0;
    NDBX_HIP(hipMemcpyAsync(&n_cores, c.ch_grank.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    c.sync();
    c.ch_parent.ensure((size_t)std::max<int32_t>(n_cores, 1) * 4);
    {
        Launch L(c.prof, st, "chunk_union");
        k_ch_iota<<<(unsigned)ceil_div(std::max<int32_t>(n_cores, 1), 256), 256, 0, st>>>(c.ch_parent.as<int32_t>(),
                                                                                         n_cores);
    }
    // B: each slab's clusters join the global sets
    for (const Slab &sl : c.h_slabs) {
        Ctx &s = slab_grid(c, sl, false);
        slab_inject(c, s);
        grid_label(s);
        c.ch_inv.ensure((size_t)s.n * 4);
        {
            Launch L(c.prof, st, "chunk_union");
            k_ch_rank_inv<<<(unsigned)ceil_div(s.n, 256), 256, 0, st>>>(s.core.as<uint8_t>(), s.rank.as<int32_t>(), s.n,
                                                                        c.ch_inv.as<int32_t>());
            k_ch_union<<<(unsigned)ceil_div(s.n, 256), 256, 0, st>>>(s.core.as<uint8_t>(), s.labels.as<int32_t>(),
                                                                     c.ch_inv.as<int32_t>(), c.ch_gid.as<int32_t>(),
                                                                     c.ch_grank.as<int32_t>(), s.n,
                                                                     c.ch_parent.as<int32_t>());
        }
        s.sync();
        merge_prof(c.prof, s.prof);
    }
    {
        Launch L(c.prof, st, "chunk_union");
        k_ch_flatten<<<(unsigned)ceil_div(std::max<int32_t>(n_cores, 1), 256), 256, 0, st>>>(c.ch_parent.as<int32_t>(),
                                                                                            n_cores);
    }
    // C: border search against the global cluster ids, owned labels kept
    for (const Slab &sl : c.h_slabs) {
        Ctx &s = slab_grid(c, sl, false);
        slab_inject(c, s);
        c.ch_ext.ensure((size_t)s.n * 4);
        {
            Launch L(c.prof, st, "chunk_ext");
            k_ch_ext<<<(unsigned)ceil_div(s.n, 256), 256, 0, st>>>(s.core.as<uint8_t>(), c.ch_gid.as<int32_t>(),
                                                                   c.ch_grank.as<int32_t>(), c.ch_parent.as<int32_t>(),
                                                                   s.n, c.ch_ext.as<int32_t>());
        }
        s.ext_label = c.ch_ext.as<int32_t>();
        grid_label(s);
        s.ext_label = nullptr;
        const int64_t o0 = sl.own_lo - sl.sub_lo, o1 = sl.own_hi - sl.sub_lo;
        {
            Launch L(c.prof, st, "chunk_scatter");
            k_ch_scatter<<<(unsigned)ceil_div(std::max<int64_t>(o1 - o0, 1), 256), 256, 0, st>>>(
                s.labels.as<int32_t>(), c.ch_gid.as<int32_t>(), o0, o1, c.labels.as<int32_t>());
        }
        s.sync();
        merge_prof(c.prof, s.prof);
    }
    fill_i32(c.meta.as<int32_t>(), 0, 3, st, c.prof);
    {
        Launch L(c.prof, st, "meta");
        k_ch_meta<<<(unsigned)std::min<int64_t>(ceil_div(n, 256), 1024), 256, 0, st>>>(
            c.core.as<uint8_t>(), c.labels.as<int32_t>(), n, c.ch_parent.as<int32_t>(), n_cores,
            c.meta.as<int32_t>());
    }
}

}  // namespace ndbx
