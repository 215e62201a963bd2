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
#include "ndbx_tile.h"

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

// Core flags as int32 in input order, for the core-rank scan; per cell, the
// lowest input index among its cores.  The points of a cell are contiguous
// in sorted order, so a wave whose points share one cell reduces first and
// issues a single atomicMin.
__global__ __launch_bounds__(256) void k_core_points(const uint8_t *__restrict__ core_s,
                                                     const uint32_t *__restrict__ perm,
                                                     const int32_t *__restrict__ cell_of, int64_t n,
                                                     const int32_t *__restrict__ ncp,
                                                     const uint8_t *__restrict__ core_in,
                                                     int32_t *__restrict__ core_i32,
                                                     int32_t *__restrict__ cell_core_min) {
    // cells at or past *ncp exist only in a captured run over its capacity
    // (flagged for an eager rerun): they are left out
    const int32_t nc = *ncp;
    const int64_t wave_base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    if (wave_base >= n) return;  // whole wave past the end
    const int64_t i = wave_base + (threadIdx.x & 63);
    const bool v = i < n;
    const int64_t iv = v ? i : wave_base;
    const bool is_core = v && core_s[iv];
    const uint32_t p = perm[iv];
    if (v) core_i32[iv] = core_in[iv];  // input order i, coalesced
    const int32_t c = cell_of[iv];
    int32_t val = is_core ? (int32_t)p : INT_MAX;
    // minimum over each run of lanes of one cell (points are sorted by cell):
    // a segmented scan, then the run's last lane makes the one atomic
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t vo = __shfl_up(val, o, 64), co = __shfl_up(c, o, 64);
        if (lane >= o && co == c) val = min(val, vo);
    }
    const int32_t c_next = __shfl_down(c, 1, 64);
    const bool last = lane == 63 || c_next != c || !(i + 1 < n);
    if (v && last && val != INT_MAX && c < nc) atomicMin(cell_core_min + c, val);
}

// Core cells as int32 flags and as a bitmask (bit c & 63 of word c >> 6): the
// union passes test the bits of a row's cells with one or two loads, and the
// mask (nc / 8 bytes) stays in L2 where the flags would not.
__global__ __launch_bounds__(256) void k_has_core(const int32_t *__restrict__ cell_core_min, const int32_t *__restrict__ ncp,
                                                  int32_t *__restrict__ cell_has_core, uint64_t *__restrict__ core_bits) {
    const int64_t nc = *ncp;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // a wave starts at a multiple of 64
    const bool h = c < nc && cell_core_min[c] != INT_MAX;
    if (c < nc) cell_has_core[c] = h;
    const uint64_t m = __ballot(h);
    if ((threadIdx.x & 63) == 0 && c < nc) core_bits[c >> 6] = m;
}

__device__ __forceinline__ bool core_cell(const uint64_t *__restrict__ core_bits, int c) {
    return (core_bits[c >> 6] >> (c & 63)) & 1;
}

// Entries of a global list reserved for the whole block with one atomic
// (atomics on one counter serialise, so list kernels take one per block, not
// per wave): `cnt` is the calling wave's count (wave-uniform); returns the
// wave's first entry.  Every thread of the block calls it; `sh` is NW + 1
// ints of LDS.
template <int NW>
__device__ __forceinline__ int32_t block_reserve(int32_t cnt, int32_t *counter, int32_t *sh) {
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
        for (int w = 0; w < NW; ++w) {
            const int32_t t = sh[w];
            sh[w] = tot;
            tot += t;
        }
        sh[NW] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int32_t r = sh[NW] + sh[wave];
    __syncthreads();
    return r;
}

// Every listed cell hung directly under its root (between the union
// launches: the second one's finds then take one step).
__global__ __launch_bounds__(256) void k_compress(const int32_t *__restrict__ list, const int32_t *__restrict__ n_list,
                                                  int32_t *__restrict__ parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_list) return;
    const int32_t c = list[i];
    const int32_t r = uf_find(parent, c);
    if (r != c) atomicMin(parent + c, r);  // parents only decrease
}

__global__ __launch_bounds__(256) void k_iota(int32_t *__restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// Core-only copy of the sorted coordinates, one point after another
// (kCoreStride floats each: a 16-byte slot in 3-D), so the few points of a
// small cell are one or two cache lines: a non-core point becomes NaN in
// every dimension, so the float32 predicate rejects it without a flag test.
template <int D>
constexpr int kCoreStride = D == 3 ? 4 : D;

template <int D>
__device__ __forceinline__ void load_core(const float *__restrict__ pts, int64_t i, float (&p)[D]) {
    if constexpr (D == 3) {
        const float4 v = reinterpret_cast<const float4 *>(pts)[i];
        p[0] = v.x, p[1] = v.y, p[2] = v.z;
    } else if constexpr (D == 2) {
        const float2 v = reinterpret_cast<const float2 *>(pts)[i];
        p[0] = v.x, p[1] = v.y;
    } else {
        p[0] = pts[i];
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_core_pts(const float *__restrict__ soa, const uint8_t *__restrict__ core_s,
                                                  int64_t n, int64_t npad, float *__restrict__ core_pts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool c = core_s[i];
    float p[kCoreStride<D>];
#pragma unroll
    for (int k = 0; k < kCoreStride<D>; ++k) p[k] = c && k < D ? soa[k * npad + i] : __builtin_nanf("");
    if constexpr (D == 3) reinterpret_cast<float4 *>(core_pts)[i] = make_float4(p[0], p[1], p[2], p[3]);
    else if constexpr (D == 2) reinterpret_cast<float2 *>(core_pts)[i] = make_float2(p[0], p[1]);
    else core_pts[i] = p[0];
}

// Any of U candidates [j, j+U) within eps of this lane's query.
template <int D, int U>
__device__ __forceinline__ bool any_block(const float (&qx)[D], const float *__restrict__ pts, int j, float e2) {
    float c[D][U];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) c[k][u] = pts[(int64_t)(j + u) * kCoreStride<D> + k];
    bool hit = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float len = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float df = qx[k] - c[k][u];
            len = __fadd_rn(len, __fmul_rn(df, df));
        }
        hit |= len <= e2;
    }
    return hit;
}

// True when some lane's query has a candidate in [b, e) within eps.
template <int D>
__device__ __forceinline__ bool wave_any_within(const float (&qx)[D], const float *__restrict__ pts, int b, int e,
                                                float e2) {
    int j = b;
    for (; j + 8 <= e; j += 8)
        if (__ballot(any_block<D, 8>(qx, pts, j, e2)) != 0) return true;
    for (; j < e; ++j)
        if (__ballot(any_block<D, 1>(qx, pts, j, e2)) != 0) return true;
    return false;
}

// Union of core cells, in two launches over all cells (one wave per cell A
// with a core).
//   phase 0: the face neighbours B = A+e_k (k < D) that hold a core, tested
//            with early exit; this joins the bulk of every dense cluster.
//   phase 1: every stencil neighbour B > A with a core, minus those already
//            in A's set (finds run one lane per candidate), tested nearest
//            first, and skipped when the box of A's current 64 cores is
//            farther than the search radius from B's cell.
// A test of (A, B) stops at the first core pair within eps; on a hit A and B
// are united.  Either phase alone is correct; phase 0 makes phase 1 cheap.
constexpr int kMaxStencilCells = 128;  // 25 rows x 5 cells for d <= 3

// One big core cell A = list[li] for one wave (the wave's LDS rows passed in).
template <int D>
__device__ __forceinline__ void union_cell(int64_t li, const int32_t *__restrict__ list, int64_t nc, int phase,
                                           const int32_t *__restrict__ cell_start,
                                           const uint64_t *__restrict__ cell_key, const InstGrid *__restrict__ ig,
                                           int lin_bits, const int2 *__restrict__ ranges, int rows_per_cell,
                                           const int32_t *__restrict__ cell_c, const float *__restrict__ core_pts,
                                           const uint64_t *__restrict__ core_bits, int32_t *__restrict__ parent,
                                           int wave, int32_t (*cand)[kMaxStencilCells],
                                           int32_t (*cand_root)[kMaxStencilCells],
                                           int32_t (*cand_off)[kMaxStencilCells], int2 (*cand_pts)[kMaxStencilCells]) {
    const int A = list[li];
    const int lane = threadIdx.x & 63;
    const InstGrid &g = ig[cell_key[A] >> lin_bits];
    const float e2 = g.e2;
    const int a0 = cell_start[A], a1 = cell_start[A + 1];
    const int ac[3] = {cell_c[A], cell_c[nc + A], cell_c[2 * nc + A]};
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int center = rows_per_cell / 2;  // slot of (dz, dy) = (0, 0)
    // the slot's row offsets (the stencil reaches 2 cells)
    const int dy = D >= 2 ? (D == 3 ? lane % 5 : lane) - 2 : 0, dz = D == 3 ? lane / 5 - 2 : 0;

    // ---- candidate list
    int n_cand = 0;
    if (phase == 0) {
        // lane k < D looks for the cell at +1 along dimension k
        int B = -1;
        if (lane < D) {
            const int slot = lane == 0 ? center : (lane == 1 ? center + 1 : center + 5);
            const int2 w = ranges[(int64_t)A * rows_per_cell + slot];
            const int want0 = lane == 0 ? ac[0] + 1 : ac[0];
            for (int b = w.x; b < w.y; ++b)
                if (cell_c[b] == want0) B = b;
            if (B >= 0 && !core_cell(core_bits, B)) B = -1;
        }
        const uint64_t mask = __ballot(B >= 0);
        if (B >= 0) {
            const int i = __popcll(mask & lt);
            cand[wave][i] = B;
            cand_off[wave][i] = 62 + (lane == 0 ? 1 : lane == 1 ? 5 : 25);  // (2, 2, 2) + e_lane
            cand_pts[wave][i] = make_int2(cell_start[B], cell_start[B + 1]);
        }
        n_cand = __popcll(mask);
    } else {
        // rows before the centre hold only cells below A
        int2 my = make_int2(0, 0);
        if (lane >= center && lane < rows_per_cell) my = ranges[(int64_t)A * rows_per_cell + lane];
        // gather B > A with a core, grouped nearest first by the number of
        // dimensions in which B is more than one cell from A
        int Bs[5], cls[5], offs[5];
        const int cl12 = (abs(dy) > 1) + (abs(dz) > 1);
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            const int B = my.x + it;
            const bool keep = B < my.y && B > A && core_cell(core_bits, B);
            Bs[it] = keep ? B : -1;
            int cl = 0, off = 0;
            if (keep) {
                const int d0 = cell_c[B] - ac[0];
                cl = (abs(d0) > 1) + cl12;
                off = (d0 + 2) + 5 * (dy + 2) + 25 * (dz + 2);
            }
            cls[it] = cl;
            offs[it] = off;
        }
        for (int want = 0; want <= D; ++want) {
#pragma unroll
            for (int it = 0; it < 5; ++it) {
                const bool take = Bs[it] >= 0 && cls[it] == want;
                const uint64_t mask = __ballot(take);
                if (take) {
                    cand[wave][n_cand + __popcll(mask & lt)] = Bs[it];
                    cand_off[wave][n_cand + __popcll(mask & lt)] = offs[it];
                }
                n_cand += __popcll(mask);
            }
        }
        __builtin_amdgcn_wave_barrier();
        // drop candidates already in A's set (one lane per candidate); the
        // kept ones take their point ranges along
        const int32_t root_a = uf_find(parent, A);
        int kept = 0;
        for (int base = 0; base < n_cand; base += 64) {
            const int ci = base + lane;
            int B = -1, rb = -1, off = 0;
            if (ci < n_cand) {
                B = cand[wave][ci];
                off = cand_off[wave][ci];
                rb = uf_find(parent, B);
                if (rb == root_a) B = -1;
            }
            const int2 pts = B >= 0 ? make_int2(cell_start[B], cell_start[B + 1]) : make_int2(0, 0);
            __builtin_amdgcn_wave_barrier();
            const uint64_t mask = __ballot(B >= 0);
            if (B >= 0) {
                const int i = kept + __popcll(mask & lt);
                cand[wave][i] = B;
                cand_root[wave][i] = rb;
                cand_off[wave][i] = off;
                cand_pts[wave][i] = pts;
            }
            kept += __popcll(mask);
        }
        n_cand = kept;
    }
    __builtin_amdgcn_wave_barrier();
    if (n_cand == 0) return;

    // ---- tests
    const double s = 1.0 / g.inv_s, slack = s * 9.5367431640625e-07;
    const double r2 = (double)e2 * (1.0 + 3.814697265625e-06);
    int32_t joined_root[4] = {-1, -1, -1, -1};  // roots met earlier in this wave (phase 1)
    int n_joined = 0;
    int left = n_cand;  // candidates not yet joined (the walk over A's cores ends at 0)
    for (int q0 = a0; q0 < a1 && left > 0; q0 += 64) {
        const int q = q0 + lane;
        float qx[D], lo[D], hi[D];
        load_core<D>(core_pts, q < a1 ? q : a0, qx);
        if (q >= a1) qx[0] = __builtin_nanf("");
        const bool is_core = qx[0] == qx[0];
        if (__ballot(is_core) == 0) continue;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            lo[k] = wave_min(is_core ? qx[k] : __builtin_inff());
            hi[k] = wave_max(is_core ? qx[k] : -__builtin_inff());
        }
        for (int ci = 0; ci < n_cand; ++ci) {
            const int B = __builtin_amdgcn_readfirstlane(cand[wave][ci]);
            if (B < 0) continue;  // joined already
            if (phase == 1) {
                const int rb = __builtin_amdgcn_readfirstlane(cand_root[wave][ci]);
                bool seen = false;
                for (int u = 0; u < n_joined && u < 4; ++u) seen |= joined_root[u] == rb;
                if (seen) {  // B's set (as listed) has joined A's: done
                    if (lane == 0) cand[wave][ci] = -1;
                    --left;
                    __builtin_amdgcn_wave_barrier();
                    continue;
                }
            }
            // box of this chunk's cores against B's cell box
            const int off = __builtin_amdgcn_readfirstlane(cand_off[wave][ci]);
            const int bc[3] = {ac[0] + off % 5 - 2, ac[1] + off / 5 % 5 - 2, ac[2] + off / 25 - 2};
            double gap2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double blo = g.origin[k] + (double)bc[k] * s - slack, bhi = blo + s + 2.0 * slack;
                const double gk = fmax(0.0, fmax(blo - (double)hi[k], (double)lo[k] - bhi));
                gap2 += gk * gk;
            }
            if (__builtin_amdgcn_readfirstlane(gap2 > r2)) continue;
            const int b0 = __builtin_amdgcn_readfirstlane(cand_pts[wave][ci].x);
            const int b1 = __builtin_amdgcn_readfirstlane(cand_pts[wave][ci].y);
            if (wave_any_within<D>(qx, core_pts, b0, b1, e2)) {
                if (lane == 0) {
                    cand[wave][ci] = -1;
                    uf_unite(parent, A, B);
                }
                --left;
                if (phase == 1 && n_joined < 4) joined_root[n_joined++] = __builtin_amdgcn_readfirstlane(cand_root[wave][ci]);
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

// Waves take the listed cells in a grid-stride loop (a fixed grid: launching a
// wave per grid cell, most of them past the list, cost more than the work on
// sparse grids).
template <int D>
__global__ __launch_bounds__(256) void k_union(const int32_t *__restrict__ list, const int32_t *__restrict__ n_list,
                                               const int32_t *__restrict__ ncp, int phase, const int32_t *__restrict__ cell_start,
                                               const uint64_t *__restrict__ cell_key,
                                               const InstGrid *__restrict__ ig, int lin_bits,
                                               const int2 *__restrict__ ranges, int rows_per_cell,
                                               const int32_t *__restrict__ cell_c,
                                               const float *__restrict__ core_pts,
                                               const uint64_t *__restrict__ core_bits,
                                               int32_t *__restrict__ parent) {
    const int64_t nc = *ncp;
    // per wave: candidate cells, their roots, their offsets from A (packed
    // (dc0 + 2) + 5 (dc1 + 2) + 25 (dc2 + 2)) and point ranges
    __shared__ int32_t cand[4][kMaxStencilCells];
    __shared__ int32_t cand_root[4][kMaxStencilCells];
    __shared__ int32_t cand_off[4][kMaxStencilCells];
    __shared__ int2 cand_pts[4][kMaxStencilCells];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t nl = *n_list;
    for (int64_t li = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; li < nl;
         li += (int64_t)gridDim.x * (blockDim.x >> 6)) {
        union_cell<D>(li, list, nc, phase, cell_start, cell_key, ig, lin_bits, ranges, rows_per_cell, cell_c,
                      core_pts, core_bits, parent, wave, cand, cand_root, cand_off, cand_pts);
        __builtin_amdgcn_wave_barrier();  // the wave's LDS rows are reused by its next cell
    }
}

// Core cells split by size: cells with more than kSmallCell points go to the
// wave-per-cell union, the rest to the lane-per-cell union (measured on the
// 4,096-instance batch: unions 11.9 ms at 8, 10.1 at 16, 9.6 at 32; the
// 10M 3-D set 2.29 / 2.23 / 2.20 ms).  A lane keeps kCoreCache of its cell's
// cores in registers.
constexpr int kSmallCell = 32;
constexpr int kCoreCache = 8;

constexpr int kListPer = 8;  // cells per thread of the list kernel

__global__ __launch_bounds__(256) void k_cell_lists(const int32_t *__restrict__ cell_has_core,
                                                    const int32_t *__restrict__ cell_start, const int32_t *__restrict__ ncp,
                                                    int32_t *__restrict__ big, int32_t *__restrict__ small,
                                                    int32_t *__restrict__ counts) {
    const int64_t nc = *ncp;
    __shared__ int32_t shb[5], shs[5];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // wave w takes 64 kListPer consecutive cells, so the lists stay in cell order
    const int64_t c0 = (int64_t)blockIdx.x * (256 * kListPer) + (threadIdx.x >> 6) * (64 * kListPer) + lane;
    uint64_t mb[kListPer], ms[kListPer];
    int32_t nb = 0, ns = 0;
#pragma unroll
    for (int it = 0; it < kListPer; ++it) {
        const int64_t c = c0 + it * 64;
        const bool v = c < nc && cell_has_core[c];
        const bool is_big = v && cell_start[c + 1] - cell_start[c] > kSmallCell;
        mb[it] = __ballot(is_big);
        ms[it] = __ballot(v && !is_big);
        nb += __popcll(mb[it]);
        ns += __popcll(ms[it]);
    }
    int32_t ob = block_reserve<4>(nb, counts + 0, shb);
    int32_t os = block_reserve<4>(ns, counts + 1, shs);
#pragma unroll
    for (int it = 0; it < kListPer; ++it) {
        const int32_t c = (int32_t)(c0 + it * 64);
        if ((mb[it] >> lane) & 1) big[ob + __popcll(mb[it] & lt)] = c;
        if ((ms[it] >> lane) & 1) small[os + __popcll(ms[it] & lt)] = c;
        ob += __popcll(mb[it]);
        os += __popcll(ms[it]);
    }
}

// Core bits of the cells [b0, e) of one window (at most 2 reach + 1 <= 5
// cells), bit i for cell b0 + i.
__device__ __forceinline__ uint32_t window_bits(const uint64_t *__restrict__ core_bits, int b0, int e) {
    const int sh = b0 & 63;
    uint64_t bits = core_bits[b0 >> 6] >> sh;
    if (sh != 0 && ((e - 1) >> 6) != (b0 >> 6)) bits |= core_bits[(b0 >> 6) + 1] << (64 - sh);
    return (uint32_t)(bits & 31u);
}

// One lane per small core cell A: every stencil neighbour B > A with a core
// and a different root is tested pair by pair (its cores against A's cores,
// non-cores are NaN in core_pts) until a pair within eps unites them.  Rows
// before A's own row (slots below the centre) hold only cells below A.  The
// cells of a row window are consecutive, so their core bits come from one or
// two mask words.  A's root is kept between candidates and found again when
// a candidate's root differs from it (sets only merge, so equal roots at any
// time mean one set).
template <int D, bool PF>
__global__ __launch_bounds__(256) void k_union_small(const int32_t *__restrict__ list,
                                                     const int32_t *__restrict__ n_list,
                                                     const int32_t *__restrict__ cell_start,
                                                     const uint64_t *__restrict__ cell_key,
                                                     const InstGrid *__restrict__ ig, int lin_bits,
                                                     const int2 *__restrict__ ranges, int rows_per_cell,
                                                     const float *__restrict__ core_pts,
                                                     const uint64_t *__restrict__ core_bits,
                                                     int32_t *__restrict__ parent) {
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= *n_list) return;
    const int A = list[li];
    const float e2 = ig[cell_key[A] >> lin_bits].e2;
    const int a0 = cell_start[A], a1 = cell_start[A + 1];
    // A's first kCoreCache cores in registers; points from `rest` on (cores
    // past those, and non-cores as NaN, which no test accepts) read again
    float ax[kCoreCache][D];
    int na = 0, rest = a1;
    for (int i = a0; i < a1; ++i) {
        float p[D];
        load_core<D>(core_pts, i, p);
        if (p[0] != p[0]) continue;  // not a core
        if (na == kCoreCache) {
            rest = i;
            break;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) ax[na][k] = p[k];
        ++na;
    }
    int32_t ra = uf_find(parent, A);
    // the cells [b0, bend) of one window (<= 5), bit i of bits: cell b0 + i
    // has a core
    auto window = [&](int b0, int bend, uint32_t bits) {
        for (int B = b0; B < bend; ++B, bits >>= 1) {
        if (!(bits & 1)) continue;
        const int32_t rb = uf_find(parent, B);
        if (rb != ra) ra = uf_find(parent, A);  // A's set may have been hung under another since
        if (rb == ra) continue;
        const int b1 = cell_start[B + 1];
        bool hit = false;
        for (int j = cell_start[B]; j < b1 && !hit; ++j) {
            float p[D];
            load_core<D>(core_pts, j, p);
            for (int i = 0; i < na && !hit; ++i) {
                float len = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float df = ax[i][k] - p[k];
                    len = __fadd_rn(len, __fmul_rn(df, df));
                }
                hit = len <= e2;
            }
            for (int i = rest; i < a1 && !hit; ++i) {
                float qa[D];
                load_core<D>(core_pts, i, qa);
                float len = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float df = qa[k] - p[k];
                    len = __fadd_rn(len, __fmul_rn(df, df));
                }
                hit = len <= e2;
            }
        }
        if (hit) uf_unite(parent, A, B);
        }
    };
    if constexpr (PF) {
        // the windows from the centre row on, and their cells' core bits,
        // loaded before any union (none of them depends on one): a lane's
        // load chain is then two levels deep instead of two per window.  Per
        // window its first cell above A, and (its length <= 5) << 5 | (its
        // <= 5 core bits) in one byte, four to a register.
        constexpr int NW = kStencilRows<D> - kStencilRows<D> / 2;
        int32_t wb0[NW];
        uint32_t wmeta[(NW + 3) / 4];
#pragma unroll
        for (int q = 0; q < (NW + 3) / 4; ++q) wmeta[q] = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const int2 w = ranges[(int64_t)A * rows_per_cell + kStencilRows<D> / 2 + q];
            const int b0 = w.x > A ? w.x : A + 1;
            wb0[q] = b0;
            if (b0 < w.y)
                wmeta[q >> 2] |= (window_bits(core_bits, b0, w.y) | (uint32_t)(w.y - b0) << 5) << (8 * (q & 3));
        }
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const uint32_t mq = (wmeta[q >> 2] >> (8 * (q & 3))) & 255u;
            if ((mq & 31u) != 0) window(wb0[q], wb0[q] + (int)(mq >> 5), mq & 31u);
        }
    } else {
        for (int slot = rows_per_cell / 2; slot < rows_per_cell; ++slot) {
            const int2 w = ranges[(int64_t)A * rows_per_cell + slot];
            const int b0 = w.x > A ? w.x : A + 1;
            if (b0 < w.y) window(b0, w.y, window_bits(core_bits, b0, w.y));
        }
    }
}

// Every core cell's root and the lowest core id of each root's set.  A few
// roots take all the cells, so neighbouring lanes with one root (cells of a
// row mostly share one) take their minimum first: a segmented scan over runs
// of equal roots, one atomic per run.
__global__ __launch_bounds__(256) void k_flatten(const int32_t *__restrict__ cell_has_core,
                                                 const int32_t *__restrict__ cell_core_min, const int32_t *__restrict__ ncp,
                                                 int32_t *__restrict__ parent, int32_t *__restrict__ root_min) {
    const int64_t nc = *ncp;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = c < nc && cell_has_core[c];
    const int32_t r = v ? uf_find(parent, (int32_t)c) : -1;
    int32_t val = v ? cell_core_min[c] : INT_MAX;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        // (a hop to an equal root outside the run only adds values of the
        // same root; within a run the scan is complete)
        const int32_t vo = __shfl_up(val, o, 64), ro = __shfl_up(r, o, 64);
        if (lane >= o && ro == r) val = min(val, vo);
    }
    const int32_t r_next = __shfl_down(r, 1, 64);
    if (v && (lane == 63 || r_next != r)) atomicMin(root_min + r, val);
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
                                                    const InstGrid *__restrict__ ig, int lin_bits, const int32_t *__restrict__ ncp,
                                                    const int32_t *__restrict__ parent,
                                                    const int32_t *__restrict__ root_min,
                                                    const int32_t *__restrict__ rank, int32_t *__restrict__ cell_label) {
    const int64_t nc = *ncp;
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

// Cluster id of each core cell from ids given per point (input order) for
// the cores: the id of the cell's lowest-index core (every core of a cell is
// in one cluster).  Used in place of the union-find when the clusters were
// found elsewhere (chunked runs).
__global__ __launch_bounds__(256) void k_cell_label_ext(const int32_t *__restrict__ cell_has_core,
                                                        const int32_t *__restrict__ cell_core_min,
                                                        const int32_t *__restrict__ ncp,
                                                        const int32_t *__restrict__ ext,
                                                        int32_t *__restrict__ cell_label) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *ncp) return;
    cell_label[c] = cell_has_core[c] ? ext[cell_core_min[c]] : NDBX_NO_CLUSTER;
}

// Labels per point: cores take their cell's id; a non-core point takes the
// smallest id among core cells holding a core within eps (its own cell's
// cores are within eps by construction).  Three launches:
//   k_border_own   one lane per point: its own cell's id, in input order;
//   k_border_list  one lane per count tile: non-core points with a
//                  neighbour besides themselves (nn > 1; with nn = 1 a point
//                  is within eps of no core) search; tiles of more than
//                  kBorderLaneTile points that hold one are listed for a wave
//                  each, the searching points of smaller tiles for a lane
//                  each;
//   k_border       waves over the listed tiles: the searching points look
//                  for a core cell with a smaller id holding a core within
//                  eps, the candidates broadcast to the wave;
//   k_border_pts   one lane per listed point, the same search alone: sparse
//                  tiles would leave most of a wave idle, and a wave has many
//                  more lane searches than tile searches in flight.
constexpr int kBorderLaneTile = 16;
// Cells at or past *ncp exist only in a captured run over its cell capacity
// (cell_label holds the capacity's cells; the run is flagged for an eager
// rerun): their points get NO_CLUSTER.
__global__ __launch_bounds__(256) void k_border_own(const int32_t *__restrict__ cell_of,
                                                    const int32_t *__restrict__ ncp,
                                                    const int32_t *__restrict__ cell_label,
                                                    const uint32_t *__restrict__ perm, int64_t n,
                                                    int32_t *__restrict__ labels) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cell_of[i];
    labels[perm[i]] = c < *ncp ? cell_label[c] : NDBX_NO_CLUSTER;  // NO_CLUSTER also when the cell has no core
}

constexpr int kBorderListThreads = 1024;

__global__ __launch_bounds__(kBorderListThreads) void k_border_list(const int4 *__restrict__ ctiles, const int32_t *__restrict__ ntp,
                                                     const uint8_t *__restrict__ core_s,
                                                     const int32_t *__restrict__ nn_s, int32_t *__restrict__ list,
                                                     int32_t *__restrict__ n_list, int32_t *__restrict__ plist,
                                                     int32_t *__restrict__ n_plist) {
    constexpr int NW = kBorderListThreads / 64;
    __shared__ int32_t sh_t[NW + 1], sh_p[NW + 1];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int4 tile = t < *ntp ? ctiles[t] : make_int4(0, 0, 0, 0);
    // a large tile: its lane walks the points until one searches
    bool wave_tile = false;
    if (tile.w > kBorderLaneTile)
        for (int q = tile.z; q < tile.z + tile.w && !wave_tile; ++q) wave_tile = !core_s[q] && nn_s[q] > 1;
    const uint64_t m = __ballot(wave_tile);
    const int32_t tbase = block_reserve<NW>(__popcll(m), n_list, sh_t);
    if (wave_tile) list[tbase + __popcll(m & lt)] = (int32_t)t;
    // the small tiles' points flattened over the wave (consecutive tiles hold
    // consecutive points, so the loads coalesce), one lane each; the wave's
    // searching points are reserved with one atomic, so the list keeps
    // neighbouring points together
    const int len = tile.w <= kBorderLaneTile ? tile.w : 0;
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t2 = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t2;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);  // <= 64 kBorderLaneTile
    constexpr int kSteps = kBorderLaneTile;
    uint64_t pm[kSteps];
    int qs[kSteps];
    int n_need = 0;
#pragma unroll
    for (int it = 0; it < kSteps; ++it) {
        pm[it] = 0;
        qs[it] = 0;
        const int k = it * 64 + lane;
        if (it * 64 >= total) continue;  // wave-uniform
        int lo_w = 0, hi_w = 63;  // lowest lane with incl > k
#pragma unroll
        for (int st = 0; st < 6; ++st) {
            const int mid = (lo_w + hi_w) >> 1;
            if (__shfl(incl, mid, 64) > k) hi_w = mid;
            else lo_w = mid + 1;
        }
        const int wz = __shfl(tile.z, lo_w, 64), wi = __shfl(incl, lo_w, 64), wl = __shfl(len, lo_w, 64);
        const int q = wz + k - (wi - wl);
        const bool need = k < total && !core_s[q] && nn_s[q] > 1;
        pm[it] = __ballot(need);
        qs[it] = q;
        n_need += __popcll(pm[it]);
    }
    int32_t base = block_reserve<NW>(n_need, n_plist, sh_p);
#pragma unroll
    for (int it = 0; it < kSteps; ++it) {
        if ((pm[it] >> lane) & 1) plist[base + __popcll(pm[it] & lt)] = qs[it];
        base += __popcll(pm[it]);
    }
}

// One lane per listed point q (sorted order): the rows of its cell's stencil
// that some point within the search radius can lie in, and in them the core
// cells with an id below the point's current one, tested point by point.
template <int D>
__global__ __launch_bounds__(256) void k_border_pts(const int32_t *__restrict__ plist, const int32_t *__restrict__ n_plist,
                                                    const int32_t *__restrict__ cell_start,
                                                    const uint64_t *__restrict__ cell_key,
                                                    const InstGrid *__restrict__ ig, int lin_bits,
                                                    const int2 *__restrict__ ranges, int rows_per_cell,
                                                    const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                                    const int32_t *__restrict__ cell_of,
                                                    const float *__restrict__ soa, const float *__restrict__ core_pts,
                                                    int64_t npad, const int32_t *__restrict__ cell_label,
                                                    const uint32_t *__restrict__ perm, int32_t *__restrict__ labels) {
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= *n_plist) return;
    const int64_t nc = *ncp;
    const int q = plist[li];
    const int A = cell_of[q];
    const int32_t own = cell_label[A];
    int32_t best = own;
    const InstGrid &g = ig[cell_key[A] >> lin_bits];
    const float e2 = g.e2;
    float qx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) qx[k] = soa[k * npad + q];
    const int c1 = D >= 2 ? cell_c[nc + A] : 0, c2 = D >= 3 ? cell_c[2 * nc + A] : 0;
    for (int slot = 0; slot < rows_per_cell; ++slot) {
        int xl, xh;
        if (!row_reach<D>(slot, qx, qx, g, c1, c2, xl, xh)) continue;
        const int2 w = ranges[(int64_t)A * rows_per_cell + slot];
        for (int B = w.x; B < w.y; ++B) {
            const int32_t LB = cell_label[B];
            if (LB >= best) continue;  // NO_CLUSTER (no core) included
            const int b1 = cell_start[B + 1];
            bool hit = false;
            // four points per round trip (a hit is any core within eps, so
            // the order of the tests does not matter)
            for (int j = cell_start[B]; j < b1 && !hit; j += 4) {
                float p[4][D];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j + u < b1) load_core<D>(core_pts, j + u, p[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j + u < b1) {
                        float len = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) {
                            const float df = qx[k] - p[u][k];
                            len = __fadd_rn(len, __fmul_rn(df, df));
                        }
                        hit = hit || len <= e2;
                    }
            }
            if (hit) best = LB;
        }
    }
    if (best != own) labels[perm[q]] = best;
}

// A non-core point of a listed tile starts from its own cell's id and takes
// the id of any core cell B with a smaller id that holds a core within eps.
// Windows are the count windows of the tile, over the box of its searching
// points.  Each wave takes listed tiles in a grid-stride loop.
template <int D>
// 8 waves per SIMD (SGPRs capped at 8 waves' share).
__global__ __launch_bounds__(256, 8) void k_border(const int4 *__restrict__ ctiles, const int32_t *__restrict__ list,
                                                const int32_t *__restrict__ n_list,
                                                const int32_t *__restrict__ cell_start,
                                                const uint64_t *__restrict__ cell_key,
                                                const InstGrid *__restrict__ ig, int lin_bits,
                                                const int2 *__restrict__ ranges, int rows_per_cell,
                                                const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                                const int32_t *__restrict__ cell_of,
                                                const float *__restrict__ soa, const float *__restrict__ core_pts,
                                                int64_t npad, const uint8_t *__restrict__ core_s,
                                                const int32_t *__restrict__ nn_s,
                                                const int32_t *__restrict__ cell_label,
                                                const uint32_t *__restrict__ perm, int32_t *__restrict__ labels) {
    const int64_t nc = *ncp;
    const int64_t nl = *n_list;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t li = wave0; li < nl; li += (int64_t)gridDim.x * (blockDim.x >> 6)) {
        const int4 tile = ctiles[list[li]];
        const bool valid = lane < tile.w;
        const int q = tile.z + (valid ? lane : 0);
        const bool is_core = valid && core_s[q];
        const int32_t own = cell_label[cell_of[q]];
        int32_t best = own;
        const bool need = valid && !is_core && nn_s[q] > 1;
        if (__ballot(need) != 0) {
            const InstGrid &g = ig[cell_key[tile.x] >> lin_bits];
            const float e2 = g.e2;
            float qx[D], lo[D], hi[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                qx[k] = soa[k * npad + q];
                lo[k] = wave_min(need ? qx[k] : __builtin_inff());
                hi[k] = wave_max(need ? qx[k] : -__builtin_inff());
            }
            const int2 w = tile_window<D>(lane, rows_per_cell, tile.x, tile.y, lo, hi, g, cell_c[nc + tile.x],
                                          cell_c[2 * nc + tile.x], ranges, cell_c, nc);
            // the windows' cells flattened over the lanes (a scan of their
            // lengths): each lane loads one cell's id and point range, so the
            // serial walk below visits only cells that can lower some
            // searching point's id, with no load before its tests
            const int len = w.y > w.x ? w.y - w.x : 0;
            int incl = len;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t2 = __shfl_up(incl, o, 64);
                if (lane >= o) incl += t2;
            }
            const int total = __builtin_amdgcn_readlane(incl, 63);
            int32_t best_max = need ? best : INT_MIN;
            for (int o = 32; o >= 1; o >>= 1) best_max = max(best_max, __shfl_xor(best_max, o, 64));
            for (int base = 0; base < total; base += 64) {
                const int k = base + lane;
                // the window holding flattened cell k: the lane r with incl[r] > k, lowest such
                int lo_w = 0, hi_w = 63;
                while (lo_w < hi_w) {
                    const int mid = (lo_w + hi_w) >> 1;
                    if (__shfl(incl, mid, 64) > k) hi_w = mid;
                    else lo_w = mid + 1;
                }
                const int wx = __shfl(w.x, lo_w, 64), wi = __shfl(incl, lo_w, 64), wl = __shfl(len, lo_w, 64);
                const int B = k < total ? wx + k - (wi - wl) : -1;
                const int32_t LB = B >= 0 ? cell_label[B] : NDBX_NO_CLUSTER;
                const bool useful = LB != NDBX_NO_CLUSTER && LB < best_max;
                const int j0 = useful ? cell_start[B] : 0, j1 = useful ? cell_start[B + 1] : 0;
                uint64_t um = __ballot(useful);
                while (um) {
                    const int src = __builtin_ctzll(um);
                    um &= um - 1;
                    const int32_t LBs = __builtin_amdgcn_readlane(LB, src);
                    // only lanes that would improve take part; others see NaN
                    const bool want = need && LBs < best;
                    if (__ballot(want) == 0) continue;
                    float qw[D];
#pragma unroll
                    for (int kk = 0; kk < D; ++kk) qw[kk] = want ? qx[kk] : __builtin_nanf("");
                    const int b1 = __builtin_amdgcn_readlane(j1, src);
                    bool got = false;
                    int j = __builtin_amdgcn_readlane(j0, src);
                    bool open = true;  // some wanting lane has no hit yet (wave-uniform)
                    for (; j + 8 <= b1 && open; j += 8) {
                        got |= any_block<D, 8>(qw, core_pts, j, e2);
                        open = __ballot(want && !got) != 0;
                    }
                    for (; j < b1 && open; ++j) {
                        got |= any_block<D, 1>(qw, core_pts, j, e2);
                        open = __ballot(want && !got) != 0;
                    }
                    if (got) best = LBs;
                }
            }
        }
        if (need && best != own) labels[perm[q]] = best;
    }
}

// Per-instance (cores, noise, clusters) from input-order flags and labels;
// clusters = core cells that are roots.  A wave whose points all lie in one
// instance adds its sums with one atomic per counter.
__global__ __launch_bounds__(256) void k_meta_points(const uint8_t *__restrict__ core, const int32_t *__restrict__ labels,
                                                     const int64_t *__restrict__ off, int n_inst, int64_t n,
                                                     int32_t *__restrict__ meta) {
    // each wave takes one contiguous span of points, 64 at a time; the
    // instance of the span's first point is searched once, later ones are
    // reached by stepping over instance boundaries
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t span = (n + n_waves * 64 - 1) / (n_waves * 64) * 64;
    const int64_t s0 = wave * span, s1 = s0 + span < n ? s0 + span : n;
    if (s0 >= s1) return;
    int inst = 0;
    {
        int lo = 0, hi = n_inst - 1;  // last instance starting at or before s0
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= s0) lo = mid;
            else hi = mid - 1;
        }
        inst = lo;
    }
    int64_t next = off[inst + 1];  // start of the next instance
    int a_acc = 0, b_acc = 0;      // counts of instance `inst` so far
    auto flush = [&]() {
        for (int o = 32; o >= 1; o >>= 1) {
            a_acc += __shfl_xor(a_acc, o, 64);
            b_acc += __shfl_xor(b_acc, o, 64);
        }
        if (lane == 0) {
            if (a_acc) atomicAdd(meta + 3 * inst + 0, a_acc);
            if (b_acc) atomicAdd(meta + 3 * inst + 1, b_acc);
        }
        a_acc = b_acc = 0;
    };
    for (int64_t base = s0; base < s1; base += 64) {
        const int64_t i = base + lane;
        const bool v = i < s1;
        const int c = v && core[i];
        const int z = v && labels[i] == NDBX_NO_CLUSTER;
        while (next <= base) {  // the wave has left instance `inst`
            flush();
            ++inst;
            next = off[inst + 1];
        }
        const int64_t last = base + 63 < s1 ? base + 63 : s1 - 1;
        if (next > last) {
            a_acc += c;
            b_acc += z;
        } else if (v) {  // the 64 points cross into later instances: per lane
            int li = inst;
            while (off[li + 1] <= i) ++li;
            if (li == inst) {
                a_acc += c;
                b_acc += z;
            } else {
                if (c) atomicAdd(meta + 3 * li + 0, 1);
                if (z) atomicAdd(meta + 3 * li + 1, 1);
            }
        }
    }
    flush();
}
__global__ __launch_bounds__(256) void k_meta_cells(const int32_t *__restrict__ cell_has_core,
                                                    const uint64_t *__restrict__ cell_key, int lin_bits,
                                                    const int32_t *__restrict__ parent, const int32_t *__restrict__ ncp,
                                                    int32_t *__restrict__ meta) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (cell_has_core[c] && parent[c] == (int32_t)c) atomicAdd(meta + 3 * (cell_key[c] >> lin_bits) + 2, 1);
}

template <int D>
static void grid_label_d(Ctx &c) {
    // capacity mode (captured runs): cell and tile counts on the device,
    // launches over the capacities
    const int64_t n = c.n, nc = c.capacity_mode ? c.cap_cells : c.n_cells;
    const int32_t *ncp = c.dcount.as<int32_t>(), *ntp = ncp + 1;
    const hipStream_t s = c.stream;
    const int nb = (int)ceil_div(n, 256), ncb = (int)ceil_div(nc, 256);
    InstGrid *ig = c.inst_d.as<InstGrid>();
    const int rows = c.rows_per_cell;
    const int64_t npad = ceil_div(n, 64) * 64;

    c.cell_core_min.ensure(nc * 4);
    c.cell_has_core.ensure(nc * 4);
    c.core_bits.ensure((size_t)ceil_div(nc, 64) * 8);
    c.rank.ensure((n + 1) * 4);
    c.label_s.ensure(n * 4);  // core flags as int32, input order
    fill_i32(c.cell_core_min.as<int32_t>(), INT_MAX, nc, s, c.prof);
    {
        Launch L(c.prof, s, "core");
        k_core_points<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.vals.as<uint32_t>(), c.cell_of.as<int32_t>(), n,
                                         ncp, c.core.as<uint8_t>(), c.label_s.as<int32_t>(),
                                         c.cell_core_min.as<int32_t>());
        k_has_core<<<ncb, 256, 0, s>>>(c.cell_core_min.as<int32_t>(), ncp, c.cell_has_core.as<int32_t>(),
                                       c.core_bits.as<uint64_t>());
    }
    // clusters given per point by the caller (chunked runs): no union-find,
    // the border search runs against those ids
    const int32_t *ext = c.ext_label;
    int32_t *d_total = c.misc.as<int32_t>();
    scan_exclusive_i32(c.label_s.as<int32_t>(), c.rank.as<int32_t>(), n, d_total, c.scratch, s, c.prof);
    if (!ext) {
        Launch L(c.prof, s, "core_base");
        k_core_base<<<(unsigned)ceil_div(c.n_inst, 256), 256, 0, s>>>(c.rank.as<int32_t>(), d_total, n, ig, c.n_inst);
    }
    c.parent.ensure(nc * 4);
    c.comp_min.ensure(nc * 4);
    c.cell_label.ensure(nc * 4);
    if (!ext) {
        Launch L(c.prof, s, "uf_init");
        k_iota<<<ncb, 256, 0, s>>>(c.parent.as<int32_t>(), nc);
    }
    if (!ext) fill_i32(c.comp_min.as<int32_t>(), INT_MAX, nc, s, c.prof);
    c.soa_core.ensure((size_t)n * kCoreStride<D> * 4);
    {
        Launch L(c.prof, s, "core_pts");
        k_core_pts<D><<<nb, 256, 0, s>>>(c.soa.as<float>(), c.core_s.as<uint8_t>(), n, npad, c.soa_core.as<float>());
    }
    if (ext) {
        Launch L(c.prof, s, "cell_label");
        k_cell_label_ext<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_core_min.as<int32_t>(), ncp, ext,
                                             c.cell_label.as<int32_t>());
    }
    // compacted core-cell lists (big: wave per cell; small: lane per cell)
    c.gflag.ensure(nc * 4);
    c.ctile_off.ensure(nc * 4);
    int32_t *big = c.gflag.as<int32_t>(), *small = c.ctile_off.as<int32_t>();
    int32_t *counts = c.misc.as<int32_t>() + 8;
    fill_i32(counts, 0, 4, s, c.prof);  // big list, small list, border tiles, border points
    if (!ext) {
    {
        Launch L(c.prof, s, "cell_lists");
        k_cell_lists<<<(unsigned)ceil_div(nc, 256 * kListPer), 256, 0, s>>>(
            c.cell_has_core.as<int32_t>(), c.cell_start.as<int32_t>(), ncp, big, small, counts);
    }
    for (int phase = 0; phase < 2; ++phase) {
        if (phase == 1) {
            Launch L(c.prof, s, "compress");
            k_compress<<<(unsigned)ceil_div(nc, 256), 256, 0, s>>>(big, counts, c.parent.as<int32_t>());
        }
        Launch L(c.prof, s, phase == 0 ? "union_faces" : "union");
        k_union<D><<<(unsigned)std::min<int64_t>(ceil_div(nc, 4), 4096), 256, 0, s>>>(big, counts, ncp, phase, c.cell_start.as<int32_t>(),
                                                             c.cell_key.as<uint64_t>(), ig, c.lin_bits,
                                                             c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(),
                                                             c.soa_core.as<float>(), c.core_bits.as<uint64_t>(),
                                                             c.parent.as<int32_t>());
    }
    {
        // windows loaded ahead (more registers, fewer waves) where cells hold
        // enough points that each lane has work to overlap: on the 10M 3-D
        // set (8 points per cell) 1.8 -> 1.1 ms; on the 4,096-instance batch
        // (2.3 per cell, mostly empty windows) 6.8 -> 8.4 ms
        Launch L(c.prof, s, "union_small");
        if (n >= 4 * nc)
            k_union_small<D, true><<<ncb, 256, 0, s>>>(small, counts + 1, c.cell_start.as<int32_t>(),
                                                       c.cell_key.as<uint64_t>(), ig, c.lin_bits, c.ranges.as<int2>(),
                                                       rows, c.soa_core.as<float>(), c.core_bits.as<uint64_t>(),
                                                       c.parent.as<int32_t>());
        else
            k_union_small<D, false><<<ncb, 256, 0, s>>>(small, counts + 1, c.cell_start.as<int32_t>(),
                                                        c.cell_key.as<uint64_t>(), ig, c.lin_bits, c.ranges.as<int2>(),
                                                        rows, c.soa_core.as<float>(), c.core_bits.as<uint64_t>(),
                                                        c.parent.as<int32_t>());
    }
    {
        Launch L(c.prof, s, "flatten");
        k_flatten<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_core_min.as<int32_t>(), ncp,
                                      c.parent.as<int32_t>(), c.comp_min.as<int32_t>());
    }
    {
        Launch L(c.prof, s, "cell_label");
        k_cell_label<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_key.as<uint64_t>(), ig, c.lin_bits, ncp,
                                         c.parent.as<int32_t>(), c.comp_min.as<int32_t>(), c.rank.as<int32_t>(),
                                         c.cell_label.as<int32_t>());
    }
    }  // !ext
    {
        Launch L(c.prof, s, "border");
        const int64_t nct = c.capacity_mode ? c.cap_tiles : c.n_tiles;
        c.border_tiles.ensure((size_t)(std::max<int64_t>(nct, 1) + n) * 4);
        int32_t *blist = c.border_tiles.as<int32_t>(), *n_blist = counts + 2;
        int32_t *plist = blist + std::max<int64_t>(nct, 1), *n_plist = counts + 3;
        k_border_own<<<nb, 256, 0, s>>>(c.cell_of.as<int32_t>(), ncp, c.cell_label.as<int32_t>(), c.vals.as<uint32_t>(),
                                        n, c.labels.as<int32_t>());
        k_border_list<<<(unsigned)ceil_div(std::max<int64_t>(nct, 1), kBorderListThreads), kBorderListThreads, 0, s>>>(
            c.ctiles.as<int4>(), ntp, c.core_s.as<uint8_t>(), c.nn_s.as<int32_t>(), blist, n_blist, plist, n_plist);
        stencil_back(c, blist, n_blist, plist, n_plist);
        k_border_pts<D><<<nb, 256, 0, s>>>(plist, n_plist, c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), ig,
                                           c.lin_bits, c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(), ncp,
                                           c.cell_of.as<int32_t>(), c.soa.as<float>(), c.soa_core.as<float>(), npad,
                                           c.cell_label.as<int32_t>(), c.vals.as<uint32_t>(), c.labels.as<int32_t>());
        // waves loop over the list (its length is on the device)
        k_border<D><<<(unsigned)std::min<int64_t>(ceil_div(std::max<int64_t>(nct, 1), 4), 4096), 256, 0, s>>>(
            c.ctiles.as<int4>(), blist, n_blist, c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), ig,
            c.lin_bits, c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(), ncp, c.cell_of.as<int32_t>(),
            c.soa.as<float>(), c.soa_core.as<float>(), npad, c.core_s.as<uint8_t>(), c.nn_s.as<int32_t>(),
            c.cell_label.as<int32_t>(), c.vals.as<uint32_t>(), c.labels.as<int32_t>());
    }
    if (ext) return;  // the caller counts clusters
    fill_i32(c.meta.as<int32_t>(), 0, 3 * c.n_inst, s, c.prof);
    {
        Launch L(c.prof, s, "meta");
        k_meta_points<<<(unsigned)std::min<int64_t>(ceil_div(n, 256), 2048), 256, 0, s>>>(c.core.as<uint8_t>(), c.labels.as<int32_t>(), c.offsets_d.as<int64_t>(),
                                         c.n_inst, n, c.meta.as<int32_t>());
        k_meta_cells<<<ncb, 256, 0, s>>>(c.cell_has_core.as<int32_t>(), c.cell_key.as<uint64_t>(), c.lin_bits,
                                         c.parent.as<int32_t>(), ncp, c.meta.as<int32_t>());
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
