// All-pairs path: any d, and inputs the cell grid does not take (coordinates
// that are non-finite or FLT_MAX in dimension 0, eps^2 not a normal float,
// extents beyond 2^21 cells).  It keeps the reference's exact semantics,
// including the padding rule: a candidate whose coordinate 0 is FLT_MAX never
// counts and is never a core to attach to (data_process.cpp:440, :194).
//
// Counting: one wave per 64 query points, a wave-uniform loop over all
// candidates (scalar loads), the float32 predicate per lane.
// Labels: union-find over point ids (root = lowest id of the set).
#include <cfloat>
#include <climits>

#include "ndbx_internal.h"

namespace ndbx {

__device__ __forceinline__ int32_t ld_rx(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int32_t bf_find(int32_t *parent, int32_t x) {
    for (;;) {
        const int32_t p = ld_rx(parent + x);
        if (p == x) return x;
        const int32_t gp = ld_rx(parent + p);
        if (gp != p) atomicCAS(parent + x, p, gp);
        x = gp;
    }
}

__device__ void bf_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = bf_find(parent, a);
        b = bf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + b, b, a) == b) return;
    }
}

// Query coordinates of one lane.  D > 0: exactly D dims in registers.
template <int D>
struct Query {
    float v[D];
    __device__ void load(const float *__restrict__ x, int64_t q, bool valid, int, float *) {
        const int64_t qq = valid ? q : 0;  // inactive lanes read row 0
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = x[qq * D + k];
    }
    __device__ float len(const float *__restrict__ p, int) const {
        float l = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float df = v[k] - p[k];
            l = __fadd_rn(l, __fmul_rn(df, df));
        }
        return l;
    }
};

template <>
struct Query<0> {  // runtime d: the query row is re-read from global memory (cached)
    const float *qp;
    __device__ void load(const float *__restrict__ x, int64_t q, bool valid, int d, float *) {
        qp = x + (valid ? q : 0) * d;
    }
    __device__ float len(const float *__restrict__ p, int d) const {
        float l = 0.0f;
        for (int k = 0; k < d; ++k) {
            const float df = qp[k] - p[k];
            l = __fadd_rn(l, __fmul_rn(df, df));
        }
        return l;
    }
};

template <int D>
__global__ __launch_bounds__(256) void k_brute_count(const float *__restrict__ x, int64_t n, int d, float e2,
                                                     int32_t *__restrict__ nn) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t q = ((int64_t)blockIdx.x * 4 + w) * 64 + lane;
    const bool valid = q < n;
    Query<D> Q;
    Q.load(x, q, valid, d, nullptr);
    const int dd = D > 0 ? D : d;
    int32_t cnt = 0;
    for (int64_t j = 0; j < n; ++j) {
        const float *p = x + j * dd;
        if (p[0] == FLT_MAX) continue;  // padding row (data_process.cpp:440)
        cnt += Q.len(p, dd) <= e2 ? 1 : 0;
    }
    if (valid) nn[q] = cnt;
}

__global__ __launch_bounds__(256) void k_brute_core(const int32_t *__restrict__ nn, int64_t n, int32_t m,
                                                    uint8_t *__restrict__ core, int32_t *__restrict__ core_i32,
                                                    int32_t *__restrict__ parent) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool c = nn[i] >= m;
    core[i] = c;
    core_i32[i] = c;
    parent[i] = (int32_t)i;
}

// Core-core edges (q < j).  An edge needs the predicate and at least one end
// that is a valid source: point i against core jj is skipped only when jj is
// padding (data_process.cpp:194), and either core can play either role.
template <int D>
__global__ __launch_bounds__(256) void k_brute_union(const float *__restrict__ x, int64_t n, int d, float e2,
                                                     const uint8_t *__restrict__ core, int32_t *__restrict__ parent) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + w) * 64;
    const int64_t q = q0 + lane;
    const bool active = q < n && core[q];
    Query<D> Q;
    Q.load(x, q, active, d, nullptr);
    if (__ballot(active) == 0) return;
    const int dd = D > 0 ? D : d;
    const bool q_src = active && x[q * dd] != FLT_MAX;
    for (int64_t j = q0 + 1; j < n; ++j) {
        if (!core[j]) continue;
        const float *p = x + j * dd;
        const bool j_src = p[0] != FLT_MAX;
        const bool hit = active && j > q && (q_src || j_src) && Q.len(p, dd) <= e2;
        if (hit) bf_unite(parent, (int32_t)q, (int32_t)j);
    }
}

__global__ __launch_bounds__(256) void k_brute_core_label(const uint8_t *__restrict__ core, int64_t n,
                                                          int32_t *__restrict__ parent,
                                                          const int32_t *__restrict__ rank,
                                                          int32_t *__restrict__ labels, int32_t *__restrict__ meta) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !core[i]) return;
    const int32_t r = bf_find(parent, (int32_t)i);
    labels[i] = rank[r];
    if (r == (int32_t)i) atomicAdd(meta + 2, 1);
}

template <int D>
__global__ __launch_bounds__(256) void k_brute_border(const float *__restrict__ x, int64_t n, int d, float e2,
                                                      const uint8_t *__restrict__ core,
                                                      int32_t *__restrict__ labels) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t q = ((int64_t)blockIdx.x * 4 + w) * 64 + lane;
    const bool valid = q < n;
    bool need = valid && !core[q];
    Query<D> Q;
    Q.load(x, q, need, d, nullptr);
    if (__ballot(need) == 0) return;
    const int dd = D > 0 ? D : d;
    int32_t best = NDBX_NO_CLUSTER;
    for (int64_t j = 0; j < n; ++j) {
        if (!core[j]) continue;
        const float *p = x + j * dd;
        if (p[0] == FLT_MAX) continue;  // padding core (data_process.cpp:194)
        const int32_t lj = labels[j];
        const bool want = need && lj < best;
        if (__ballot(want) == 0) continue;
        if (want && Q.len(p, dd) <= e2) best = lj;
    }
    if (need) labels[q] = best;
}

__global__ __launch_bounds__(256) void k_brute_meta(const uint8_t *__restrict__ core,
                                                    const int32_t *__restrict__ labels, int64_t n,
                                                    int32_t *__restrict__ meta) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = i < n;
    int a = v && core[i], b = v && labels[i] == NDBX_NO_CLUSTER;
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(meta + 0, a);
        if (b) atomicAdd(meta + 1, b);
    }
}

template <int D>
static void brute_classify_d(Ctx &c, float e2) {
    const unsigned g = (unsigned)ceil_div(c.n, 256);
    {
        Launch L(c.prof, c.stream, "brute_count");
        k_brute_count<D><<<g, 256, 0, c.stream>>>(c.coords, c.n, c.d, e2, c.nn.as<int32_t>());
    }
    c.b_parent.ensure(c.n * 4);
    c.label_s.ensure(c.n * 4);
    Launch L(c.prof, c.stream, "brute_core");
    k_brute_core<<<g, 256, 0, c.stream>>>(c.nn.as<int32_t>(), c.n, c.h_m[0], c.core.as<uint8_t>(),
                                          c.label_s.as<int32_t>(), c.b_parent.as<int32_t>());
}

template <int D>
static void brute_label_d(Ctx &c, float e2) {
    const unsigned g = (unsigned)ceil_div(c.n, 256);
    {
        Launch L(c.prof, c.stream, "brute_union");
        k_brute_union<D><<<g, 256, 0, c.stream>>>(c.coords, c.n, c.d, e2, c.core.as<uint8_t>(),
                                                   c.b_parent.as<int32_t>());
    }
    {
        Launch L(c.prof, c.stream, "brute_core_label");
        k_brute_core_label<<<g, 256, 0, c.stream>>>(c.core.as<uint8_t>(), c.n, c.b_parent.as<int32_t>(),
                                                    c.rank.as<int32_t>(), c.labels.as<int32_t>(),
                                                    c.meta.as<int32_t>());
    }
    {
        Launch L(c.prof, c.stream, "brute_border");
        k_brute_border<D><<<g, 256, 0, c.stream>>>(c.coords, c.n, c.d, e2, c.core.as<uint8_t>(),
                                                    c.labels.as<int32_t>());
    }
}

#define NDBX_BRUTE_DISPATCH(fn, ...)                  \
    switch (c.d) {                                    \
        case 1: fn<1>(__VA_ARGS__); break;            \
        case 2: fn<2>(__VA_ARGS__); break;            \
        case 3: fn<3>(__VA_ARGS__); break;            \
        case 4: fn<4>(__VA_ARGS__); break;            \
        case 5: fn<5>(__VA_ARGS__); break;            \
        case 6: fn<6>(__VA_ARGS__); break;            \
        case 7: fn<7>(__VA_ARGS__); break;            \
        case 8: fn<8>(__VA_ARGS__); break;            \
        case 16: fn<16>(__VA_ARGS__); break;          \
        case 32: fn<32>(__VA_ARGS__); break;          \
        case 64: fn<64>(__VA_ARGS__); break;          \
        default: fn<0>(__VA_ARGS__); break;           \
    }

void brute_classify(Ctx &c) {
    if (c.n_inst != 1) throw Error(NDBX_EUNSUPPORTED, "batched instances need the cell grid (d <= 3, finite input)");
    const float e = c.h_eps[0];
    const float e2 = e * e;  // data_process.h:73
    if (c.n == 0) return;
    NDBX_BRUTE_DISPATCH(brute_classify_d, c, e2);
}

void brute_label(Ctx &c) {
    const float e = c.h_eps[0];
    const float e2 = e * e;
    const int64_t n = c.n;
    const hipStream_t s = c.stream;
    fill_i32(c.meta.as<int32_t>(), 0, 3, s, c.prof);
    if (n == 0) return;
    const unsigned g = (unsigned)ceil_div(n, 256);
    c.rank.ensure((n + 1) * 4);
    c.misc.ensure(64);
    scan_exclusive_i32(c.label_s.as<int32_t>(), c.rank.as<int32_t>(), n, c.misc.as<int32_t>(), c.scratch, s, c.prof);
    fill_i32(c.labels.as<int32_t>(), NDBX_NO_CLUSTER, n, s, c.prof);
    NDBX_BRUTE_DISPATCH(brute_label_d, c, e2);
    Launch L(c.prof, s, "meta");
    k_brute_meta<<<g, 256, 0, s>>>(c.core.as<uint8_t>(), c.labels.as<int32_t>(), n, c.meta.as<int32_t>());
}

}  // namespace ndbx
