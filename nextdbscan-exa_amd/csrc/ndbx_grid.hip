// Cell grid for d <= 3: bounds, plan, cell keys, radix sort, cell table, row
// hash, neighbour-cell stencil, counting tiles and the neighbour-count kernel.
//
// The grid has no counterpart in the reference, whose counting is all-pairs
// (data_process.cpp:433-454).  It is a filter only: every candidate pair
// still goes through the reference's float32 predicate, so nn is the same
// integer.  Cell side s = sqrt(e2/d)*(1-2^-18): two points of one cell are
// always within eps under that predicate (the label pass and the symmetric
// count take a cell's own points without tests), and a stencil of +-reach
// cells per dimension contains every point the predicate can accept.
#include <cfloat>
#include <cmath>

#include "ndbx_internal.h"
#include "ndbx_dense.h"
#include "ndbx_tile.h"

namespace ndbx {

constexpr int64_t kBoundsChunk = 1 << 13;  // points per bounds workgroup (enough workgroups to fill the chip)
constexpr int kFineBits = 6;  // 2-D: points of a cell in Morton order of a 64 x 64 sub-grid
constexpr int kPartMin = 64;  // count kernel: cells with fewer points are taken whole, not by parts
constexpr int kRunGap = 8;    // symmetric count: candidate runs closer than this are walked as one
// symmetric count: cells of 64+ points are taken by their parts in reach of
// tile boxes up to this many cells wide (10M 3-D count: 12.07 ms at 1, 11.72
// at 1.5, 11.59 at 2, 11.90 unbounded; the batch: 21.1, 21.1, 21.2, 23.2 ms)
constexpr double kPartBoxCells = 2.0;

// ------------------------------------------------------------ bounds --
__global__ __launch_bounds__(256) void k_bounds(const float *__restrict__ x, int d, int dg,
                                                const int64_t *__restrict__ chunk_begin,
                                                const int64_t *__restrict__ chunk_end,
                                                BoundsPart *__restrict__ out) {
    __shared__ float sh[6][256 / 64];
    __shared__ int32_t shs[256 / 64];
    const int64_t b = chunk_begin[blockIdx.x], e = chunk_end[blockIdx.x];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    int32_t special = 0;
    // rows of 4k coordinates at a 16-byte aligned base: 32 coordinates of a
    // row per round trip
    const bool v4 = (d & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const float *p = x + i * d;
        bool sp = p[0] == FLT_MAX;
        if (v4) {
            const float4 *p4 = reinterpret_cast<const float4 *>(p);
            const int n4 = d >> 2;
            for (int k0 = 0; k0 < n4; k0 += 8) {
                float4 a[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4) a[j] = p4[k0 + j];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4)
                        sp |= !isfinite(a[j].x) || !isfinite(a[j].y) || !isfinite(a[j].z) || !isfinite(a[j].w);
            }
        } else {
            for (int k = 0; k < d; ++k) sp |= !isfinite(p[k]);
        }
        if (sp) {
            ++special;
            continue;
        }
        for (int k = 0; k < dg; ++k) {
            lo[k] = fminf(lo[k], p[k]);
            hi[k] = fmaxf(hi[k], p[k]);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off >= 1; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off, 64));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off, 64));
        }
    }
    for (int off = 32; off >= 1; off >>= 1) special += __shfl_xor(special, off, 64);
    if (lane == 0) {
        for (int k = 0; k < 3; ++k) {
            sh[k][w] = lo[k];
            sh[3 + k][w] = hi[k];
        }
        shs[w] = special;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        BoundsPart r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = sh[k][0];
            r.hi[k] = sh[3 + k][0];
        }
        r.special = shs[0];
        for (int q = 1; q < 4; ++q) {
            for (int k = 0; k < 3; ++k) {
                r.lo[k] = fminf(r.lo[k], sh[k][q]);
                r.hi[k] = fmaxf(r.hi[k], sh[3 + k][q]);
            }
            r.special += shs[q];
        }
        r.count = (int32_t)(e - b);
        out[blockIdx.x] = r;
    }
}

static int bits_for(uint64_t v) {  // bits needed to hold values 0..v-1
    int b = 0;
    while (b < 64 && (v > (1ull << b))) ++b;
    return b;
}

// Each instance's cell geometry on the device, for captured runs: the same
// arithmetic as grid_plan below.  An instance the fixed key layout cannot
// hold (special values, reach > 2, too many cells) sets bit 2 of *status.
__global__ __launch_bounds__(64) void k_plan(const BoundsPart *__restrict__ parts, const int32_t *__restrict__ first,
                                             const float *__restrict__ eps, const int32_t *__restrict__ min_pts,
                                             const int64_t *__restrict__ off, int n_inst, int d, int dg, int lin_bits,
                                             InstGrid *__restrict__ ig, int32_t *__restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    InstGrid g;
    g.row_base = 0;  // set by k_row_base where the run takes a direct row table
    const float e = eps[i];
    const float e2 = __fmul_rn(e, e);  // data_process.h:73
    g.e2 = e2;
    g.m = min_pts[i];
    g.point_begin = off[i];
    g.point_end = off[i + 1];
    g.core_base = 0;
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = 0.0;
        g.G[k] = 1;
    }
    g.inv_s = 1.0;
    g.reach = 2;
    bool fail = false;
    if (g.point_end != g.point_begin) {
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        int64_t special = 0;
        for (int q = first[i]; q < first[i + 1]; ++q) {
            special += parts[q].special;
            for (int k = 0; k < dg; ++k) {
                lo[k] = fminf(lo[k], parts[q].lo[k]);
                hi[k] = fmaxf(hi[k], parts[q].hi[k]);
            }
        }
        if (special > 0 || !isfinite(e2) || !(e2 >= 1e-30f)) {
            fail = true;
        } else {
            const double s = sqrt((double)e2 / (double)d) * (1.0 - 3.814697265625e-06);
            g.inv_s = 1.0 / s;
            const double rmax = sqrt((double)e2 * (1.0 + 9.5367431640625e-07)) * (1.0 + 9.313225746154785e-10);
            g.reach = (int32_t)ceil(rmax * g.inv_s);
            uint64_t total = 1;
            bool too_big = g.reach > 2;
            for (int k = 0; k < dg && !too_big; ++k) {
                g.origin[k] = (double)lo[k];
                const double span = ((double)hi[k] - g.origin[k]) * g.inv_s;
                if (!(span < 2097152.0)) {
                    too_big = true;
                    break;
                }
                g.G[k] = (int64_t)floor(span) + 1;
                total *= (uint64_t)g.G[k];
            }
            fail = too_big || total > (1ull << lin_bits);
        }
    }
    if (fail) {  // one cell for the whole instance: every later index stays in range
        for (int k = 0; k < 3; ++k) {
            g.origin[k] = 0.0;
            g.G[k] = 1;
        }
        g.inv_s = 1.0;
        g.reach = 2;
    }
    ig[i] = g;
    if (fail) atomicOr(status, 2);
}

// Captured runs with a direct row table: each instance's first row (an
// exclusive scan of G1 G2 over the instances, as grid_plan), in one block.
// Rows beyond the table's capacity set bit 3 of *status and shrink every
// instance to one cell, so that every later index stays in range (the run is
// then done again eagerly).
__global__ __launch_bounds__(1024) void k_row_base(InstGrid *__restrict__ ig, int n_inst, int64_t cap_rows,
                                                   int32_t *__restrict__ status) {
    __shared__ int64_t sh[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < n_inst; b += 1024) {
        const int i = b + (int)threadIdx.x;
        const int64_t rows = i < n_inst ? ig[i].G[1] * ig[i].G[2] : 0;
        sh[threadIdx.x] = rows;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int64_t v = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < n_inst) ig[i].row_base = carry + sh[threadIdx.x] - rows;
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[1023];
        __syncthreads();
    }
    if (carry > cap_rows) {
        for (int i = (int)threadIdx.x; i < n_inst; i += 1024) {
            InstGrid &g = ig[i];
            for (int k = 0; k < 3; ++k) {
                g.origin[k] = 0.0;
                g.G[k] = 1;
            }
            g.inv_s = 1.0;
            g.row_base = 0;
        }
        if (threadIdx.x == 0) atomicOr(status, 8);
    }
}

// Decides the path and, for the grid, each instance's cell geometry.
void grid_plan(Ctx &c) {
    if (c.capacity_mode) {
        // captured runs keep the path and key layout of the eager run; the
        // geometry comes from the device
        const int dg = c.d < 3 ? c.d : 3;
        const int64_t nch = c.n_chunks;
        int64_t *d_cb = c.plan_in.as<int64_t>();
        int32_t *d_first = reinterpret_cast<int32_t *>(d_cb + 2 * nch);
        c.dcount.ensure(64);
        {
            Launch L(c.prof, c.stream, "bounds");
            k_bounds<<<(unsigned)nch, 256, 0, c.stream>>>(c.coords, c.d, dg, d_cb, d_cb + nch,
                                                          c.bounds_part.as<BoundsPart>());
            k_plan<<<(unsigned)ceil_div(c.n_inst, 64), 64, 0, c.stream>>>(
                c.bounds_part.as<BoundsPart>(), d_first, reinterpret_cast<const float *>(d_first + c.n_inst + 1),
                d_first + 2 * c.n_inst + 1, c.offsets_d.as<int64_t>(), c.n_inst, c.d, dg, c.lin_bits,
                c.inst_d.as<InstGrid>(), c.dcount.as<int32_t>() + 2);
            if (c.cap_rows > 0)
                k_row_base<<<1, 1024, 0, c.stream>>>(c.inst_d.as<InstGrid>(), c.n_inst, c.cap_rows,
                                                     c.dcount.as<int32_t>() + 2);
        }
        return;
    }
    c.path = PATH_GRID;
    if (c.n == 0) {
        c.h_inst.clear();
        return;
    }
    const int dg = c.d < 3 ? c.d : 3;
    // chunk table (each chunk inside one instance)
    std::vector<int64_t> cb, ce;
    std::vector<int> ci;
    for (int i = 0; i < c.n_inst; ++i) {
        for (int64_t b = c.h_offsets[i]; b < c.h_offsets[i + 1]; b += kBoundsChunk) {
            cb.push_back(b);
            ce.push_back(std::min(b + kBoundsChunk, c.h_offsets[i + 1]));
            ci.push_back(i);
        }
    }
    const int64_t nch = (int64_t)cb.size();
    // chunk table, first chunk of each instance, eps and min_pts: kept on the
    // device for the plan kernel of captured runs
    std::vector<int32_t> first(c.n_inst + 1, 0);
    for (int64_t q = nch - 1; q >= 0; --q) first[ci[q]] = (int32_t)q;
    first[c.n_inst] = (int32_t)nch;
    for (int i = c.n_inst - 1; i >= 0; --i)
        if (c.h_offsets[i + 1] == c.h_offsets[i]) first[i] = first[i + 1];
    const size_t plan_bytes = (size_t)nch * 16 + (size_t)(c.n_inst + 1) * 4 + (size_t)c.n_inst * 8;
    c.plan_in.ensure(plan_bytes);
    c.n_chunks = nch;
    int64_t *d_cb = c.plan_in.as<int64_t>();
    int32_t *d_first = reinterpret_cast<int32_t *>(d_cb + 2 * nch);
    NDBX_HIP(hipMemcpyAsync(d_cb, cb.data(), nch * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    NDBX_HIP(hipMemcpyAsync(d_cb + nch, ce.data(), nch * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    NDBX_HIP(hipMemcpyAsync(d_first, first.data(), (c.n_inst + 1) * 4, hipMemcpyHostToDevice, c.stream));
    NDBX_HIP(hipMemcpyAsync(d_first + c.n_inst + 1, c.h_eps.data(), c.n_inst * 4, hipMemcpyHostToDevice, c.stream));
    NDBX_HIP(hipMemcpyAsync(d_first + 2 * c.n_inst + 1, c.h_m.data(), c.n_inst * 4, hipMemcpyHostToDevice,
                            c.stream));
    c.bounds_part.ensure((size_t)nch * sizeof(BoundsPart));
    {
        Launch L(c.prof, c.stream, "bounds");
        k_bounds<<<(unsigned)nch, 256, 0, c.stream>>>(c.coords, c.d, dg, d_cb, d_cb + nch,
                                                      c.bounds_part.as<BoundsPart>());
    }
    std::vector<BoundsPart> parts(nch);
    NDBX_HIP(hipMemcpyAsync(parts.data(), c.bounds_part.p, nch * sizeof(BoundsPart), hipMemcpyDeviceToHost,
                            c.stream));
    c.sync();

    c.h_inst.assign(c.n_inst, InstGrid{});
    std::vector<float> lo(c.n_inst * 3, FLT_MAX), hi(c.n_inst * 3, -FLT_MAX);
    std::vector<int64_t> special(c.n_inst, 0);
    for (int64_t q = 0; q < nch; ++q) {
        int i = ci[q];
        special[i] += parts[q].special;
        for (int k = 0; k < dg; ++k) {
            lo[i * 3 + k] = std::min(lo[i * 3 + k], parts[q].lo[k]);
            hi[i * 3 + k] = std::max(hi[i * 3 + k], parts[q].hi[k]);
        }
    }
    int lin_bits = 1;
    for (int i = 0; i < c.n_inst; ++i) {
        InstGrid &g = c.h_inst[i];
        const float e = c.h_eps[i];
        const float e2 = e * e;  // data_process.h:73
        g.e2 = e2;
        g.m = c.h_m[i];
        g.point_begin = c.h_offsets[i];
        g.point_end = c.h_offsets[i + 1];
        g.core_base = 0;
        g.row_base = 0;
        for (int k = 0; k < 3; ++k) {
            g.origin[k] = 0.0;
            g.G[k] = 1;
        }
        g.inv_s = 1.0;
        g.reach = 2;
        if (g.point_end == g.point_begin) continue;
        if (special[i] > 0 || !std::isfinite(e2) || !(e2 >= 1e-30f)) {
            c.path = PATH_BRUTE;
            continue;
        }
        const double s = std::sqrt((double)e2 / (double)c.d) * (1.0 - std::ldexp(1.0, -18));
        g.inv_s = 1.0 / s;
        const double rmax = std::sqrt((double)e2 * (1.0 + std::ldexp(1.0, -20))) * (1.0 + std::ldexp(1.0, -30));
        g.reach = (int32_t)std::ceil(rmax * g.inv_s);
        uint64_t total = 1;
        bool too_big = g.reach > 2;
        for (int k = 0; k < dg; ++k) {
            g.origin[k] = (double)lo[i * 3 + k];
            const double span = ((double)hi[i * 3 + k] - g.origin[k]) * g.inv_s;
            if (!(span < 2097152.0)) {  // 2^21 cells per dimension
                too_big = true;
                break;
            }
            g.G[k] = (int64_t)std::floor(span) + 1;
            total *= (uint64_t)g.G[k];
        }
        if (too_big || total >= (1ull << 52)) {
            c.path = PATH_BRUTE;
            continue;
        }
        lin_bits = std::max(lin_bits, bits_for(total));
    }
    // d > 3: the MFMA filter where it applies (finite input, normal e2,
    // d <= 128, one instance), else all pairs
    if (c.d > 3) {
        bool ok = c.n_inst == 1 && mfma_supported(c.d);
        for (int i = 0; i < c.n_inst && ok; ++i) {
            const float e2 = c.h_eps[i] * c.h_eps[i];
            ok = special[i] == 0 && std::isfinite(e2) && e2 >= 1e-30f;
        }
        c.path = ok ? PATH_MFMA : PATH_BRUTE;
    }
    if (c.flags & NDBX_FORCE_BRUTE) c.path = PATH_BRUTE;
    // rows (c2, c1) of every instance's grid, one after another: a direct
    // row table is indexed by row_base + c2 G1 + c1
    c.n_rows_direct = 0;
    for (int i = 0; i < c.n_inst; ++i) {
        InstGrid &g = c.h_inst[i];
        g.row_base = c.n_rows_direct;
        const int64_t rows = g.G[1] * g.G[2];
        c.n_rows_direct = rows > ((int64_t)1 << 40) - c.n_rows_direct ? ((int64_t)1 << 40) : c.n_rows_direct + rows;
    }
    const int inst_bits = bits_for((uint64_t)c.n_inst);
    c.lin_bits = lin_bits;
    c.cell_bits = lin_bits + inst_bits;
    // 2 bits per dimension of sub-cell position below the cell key, so that
    // within a cell the points are in Morton order of a 4^d sub-grid
    const int want_sub = dg == 2 ? 2 * kFineBits : 2 * dg;
    c.sub_bits = c.cell_bits + want_sub <= 62 ? want_sub : 0;
    c.key_bits = c.cell_bits + c.sub_bits;
    c.reach_max = 2;
    if (c.cell_bits > 62) c.path = PATH_BRUTE;
}

// -------------------------------------------------------------- keys --
__device__ __forceinline__ int find_inst(const int64_t *__restrict__ off, int n_inst, int64_t i) {
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {  // last instance with off[inst] <= i
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int D>
__global__ __launch_bounds__(256) void k_keys(const float *__restrict__ x, int64_t n,
                                              const InstGrid *__restrict__ ig,
                                              const int64_t *__restrict__ off, int n_inst, int lin_bits,
                                              int sub_bits, uint64_t *__restrict__ keys,
                                              uint32_t *__restrict__ vals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the instance of the wave's first point (one wave-uniform search), then
    // per lane only where the wave's points cross into later instances
    int inst = 0;
    if (n_inst > 1) {
        const int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
        inst = find_inst(off, n_inst, base);
        while (off[inst + 1] <= i) ++inst;
    }
    const InstGrid &g = ig[inst];
    uint64_t lin = 0;
    uint32_t sub = 0;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
        const double t = ((double)x[i * D + k] - g.origin[k]) * g.inv_s;
        int64_t ck = (int64_t)floor(t);
        ck = ck < 0 ? 0 : (ck >= g.G[k] ? g.G[k] - 1 : ck);
        lin = lin * (uint64_t)g.G[k] + (uint64_t)ck;
        if (D == 2) {
            // 64 x 64 sub-grid, bits interleaved (dimension 0 in the odd bits)
            int f = (int)floor((t - (double)ck) * (double)(1 << kFineBits));
            f = f < 0 ? 0 : (f >= (1 << kFineBits) ? (1 << kFineBits) - 1 : f);
#pragma unroll
            for (int b = 0; b < kFineBits; ++b) sub |= (uint32_t)((f >> b) & 1) << (2 * b + (k == 0 ? 1 : 0));
        } else {
            int sc = (int)floor((t - (double)ck) * 4.0);
            sc = sc < 0 ? 0 : (sc > 3 ? 3 : sc);
            sub |= (uint32_t)(sc & 1) << k;          // bit 0 of each dimension
            sub |= (uint32_t)(sc >> 1) << (D + k);   // bit 1 of each dimension
        }
    }
    const uint64_t cell = ((uint64_t)inst << lin_bits) | lin;
    keys[i] = sub_bits ? (cell << sub_bits) | sub : cell;
    vals[i] = (uint32_t)i;
}

// --------------------------------------------------------- cell table --
__global__ __launch_bounds__(256) void k_cell_flags(const uint64_t *__restrict__ keys, int64_t n, int sub_bits,
                                                    int32_t *__restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || (keys[i] >> sub_bits) != (keys[i - 1] >> sub_bits)) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_cell_fill(const uint64_t *__restrict__ keys, int sub_bits,
                                                   const int32_t *__restrict__ flag,
                                                   const int32_t *__restrict__ excl, int64_t n,
                                                   int32_t *__restrict__ cell_of, int32_t *__restrict__ cell_start,
                                                   uint64_t *__restrict__ cell_key) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t f = flag[i];
    const int32_t cid = excl[i] + f - 1;
    cell_of[i] = cid;
    if (f) {
        cell_start[cid] = (int32_t)i;
        cell_key[cid] = keys[i] >> sub_bits;
    }
    if (i == n - 1) cell_start[cid + 1] = (int32_t)n;
}

template <int D>
__global__ __launch_bounds__(256) void k_gather_soa(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                    int64_t n, int64_t npad, float *__restrict__ soa) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
#pragma unroll
    for (int k = 0; k < D; ++k) soa[k * npad + i] = x[p * D + k];
}

// ------------------------------------------------------------ row hash --
__device__ __forceinline__ uint64_t mix64(uint64_t v) {
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdull;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ull;
    v ^= v >> 33;
    return v;
}

constexpr uint64_t kEmpty = ~0ull;

// Cell coordinates (c0, c1, c2) of every cell from its key, as three planes
// of cell_c, decoded once per cell (32-bit division when the key's linear
// index and the grid sides fit, as they do unless an instance has 2^32 cells).
__global__ __launch_bounds__(256) void k_cell_coords(const uint64_t *__restrict__ cell_key, const int32_t *__restrict__ ncp,
                                                     const InstGrid *__restrict__ ig, int lin_bits,
                                                     int32_t *__restrict__ cell_c) {
    const int64_t nc = *ncp;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t key = cell_key[c];
    const InstGrid &g = ig[key >> lin_bits];
    const uint64_t lin = key & ((1ull << lin_bits) - 1);
    const uint64_t G0 = (uint64_t)g.G[0], G1 = (uint64_t)g.G[1];
    uint64_t c0, c1, c2;
    if (((lin | G0 | G1) >> 32) == 0) {
        const uint32_t l = (uint32_t)lin, g0 = (uint32_t)G0, g1 = (uint32_t)G1;
        const uint32_t r = l / g0;
        c0 = l - r * g0;
        c1 = r % g1;
        c2 = r / g1;
    } else {
        const uint64_t r = lin / G0;
        c0 = lin - r * G0;
        c1 = r % G1;
        c2 = r / G1;
    }
    cell_c[c] = (int32_t)c0;
    cell_c[nc + c] = (int32_t)c1;
    cell_c[2 * nc + c] = (int32_t)c2;
}

// Row table slot: the row key and the row's cell range [first, end), 16
// bytes, so a lookup that finds its row reads one line.
struct RowSlot {
    uint64_t key;
    int32_t first, end;
};

// A cell's row key is its key minus its c0.
__global__ __launch_bounds__(256) void k_row_insert(const uint64_t *__restrict__ cell_key, const int32_t *__restrict__ ncp,
                                                    const int32_t *__restrict__ cell_c,
                                                    RowSlot *__restrict__ table, uint64_t mask) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t rk = cell_key[c] - (uint64_t)cell_c[c];
    const bool first = (c == 0) || cell_key[c - 1] - (uint64_t)cell_c[c - 1] != rk;
    const bool last = (c == nc - 1) || cell_key[c + 1] - (uint64_t)cell_c[c + 1] != rk;
    if (!first && !last) return;
    uint64_t slot = mix64(rk) & mask;
    for (;;) {
        unsigned long long prev = atomicCAS((unsigned long long *)&table[slot].key, (unsigned long long)kEmpty,
                                            (unsigned long long)rk);
        if (prev == kEmpty || prev == rk) break;
        slot = (slot + 1) & mask;
    }
    if (first) table[slot].first = (int32_t)c;
    if (last) table[slot].end = (int32_t)(c + 1);
}

__device__ __forceinline__ bool row_lookup(const RowSlot *__restrict__ table, uint64_t mask, uint64_t rk, int2 *out) {
    uint64_t slot = mix64(rk) & mask;
    for (;;) {
        const uint4 e = reinterpret_cast<const uint4 *>(table)[slot];  // one 16-byte load
        const uint64_t k = (uint64_t)e.y << 32 | e.x;
        if (k == rk) {
            *out = make_int2((int32_t)e.z, (int32_t)e.w);
            return true;
        }
        if (k == kEmpty) return false;
        slot = (slot + 1) & mask;
    }
}

// ------------------------------------------------------------- stencil --
// For every cell, one slot per stencil row (dz, dy) in [-reach, reach]^(D-1):
// the neighbour cells of that row with c0' in [c0-reach, c0+reach], as a cell
// index range [a, e).  Cells contiguous along dimension 0 are contiguous in
// sorted order, and so are their points, [cell_start[a], cell_start[e]).  An
// empty window keeps its insertion position (a == e), so the window of a run
// of cells [cf, cl] of one row is [slot(cf).a, slot(cl).e).
// Direct row table (eager runs whose grids have few enough rows): entry
// row_base + c2 G1 + c1 of the cell's instance holds its row's [first, end)
// cells, (0, 0) when the row is empty.  A cell's 5 rows of one c2 are
// adjacent entries, and cells of one row read the same ones.
__global__ __launch_bounds__(256) void k_row_insert_direct(const uint64_t *__restrict__ cell_key,
                                                           const int32_t *__restrict__ ncp,
                                                           const InstGrid *__restrict__ ig, int lin_bits,
                                                           const int32_t *__restrict__ cell_c,
                                                           int2 *__restrict__ table) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t rk = cell_key[c] - (uint64_t)cell_c[c];
    const bool first = c == 0 || cell_key[c - 1] - (uint64_t)cell_c[c - 1] != rk;
    const bool last = c == nc - 1 || cell_key[c + 1] - (uint64_t)cell_c[c + 1] != rk;
    if (!first && !last) return;
    const InstGrid &g = ig[cell_key[c] >> lin_bits];
    const int64_t row = g.row_base + (int64_t)cell_c[2 * nc + c] * g.G[1] + cell_c[nc + c];
    if (first) table[row].x = (int32_t)c;
    if (last) table[row].y = (int32_t)(c + 1);
}

// The rows of cell c with one dz (plane dz + 2 of its 2k+1; one plane in 2-D
// and 1-D): slots plane * P .. plane * P + P - 1.
template <int D, bool DIRECT>
__device__ __forceinline__ void stencil_plane(int64_t c, int plane, int64_t nc, const uint64_t *__restrict__ cell_key,
                                              const InstGrid *__restrict__ ig, int lin_bits,
                                              const RowSlot *__restrict__ table, uint64_t mask,
                                              const int2 *__restrict__ dtable, int2 *__restrict__ ranges,
                                              const int32_t *__restrict__ cell_c) {
    constexpr int R = kStencilRows<D>, k = 2;
    constexpr int P = D >= 2 ? 2 * k + 1 : 1;
    const uint64_t key = cell_key[c];
    const uint64_t inst = key >> lin_bits;
    const InstGrid &g = ig[inst];
    const int64_t G0 = g.G[0], G1 = g.G[1], G2 = g.G[2];
    const int64_t c0 = cell_c[c], c1 = cell_c[nc + c], c2 = cell_c[2 * nc + c];
    const int dz = D >= 3 ? plane - k : 0;
    const int64_t lo0 = c0 - k < 0 ? 0 : c0 - k;
    const int64_t hi0 = c0 + k >= G0 ? G0 - 1 : c0 + k;
    const int64_t z = c2 + dz;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int dy = D >= 2 ? r - k : 0;
        const int64_t y = c1 + dy;
        int2 out = make_int2(0, 0);
        if (z >= 0 && z < G2 && y >= 0 && y < G1) {
            const uint64_t rk = (inst << lin_bits) | (uint64_t)((z * G1 + y) * G0);
            int2 rv;
            bool found;
            if (DIRECT) {
                rv = dtable[g.row_base + z * G1 + y];
                found = rv.x < rv.y;
            } else {
                found = row_lookup(table, mask, rk, &rv);
            }
            if (found) {
                int a = rv.x, b = rv.y;  // first cell with key >= rk+lo0
                const uint64_t klo = rk + (uint64_t)lo0, khi = rk + (uint64_t)hi0;
                while (a < b) {
                    int mid = (a + b) >> 1;
                    if (cell_key[mid] < klo) a = mid + 1;
                    else b = mid;
                }
                int e = a, f = rv.y;
                while (e < f) {  // first cell with key > khi
                    int mid = (e + f) >> 1;
                    if (cell_key[mid] <= khi) e = mid + 1;
                    else f = mid;
                }
                out = make_int2(a, e);
            }
        }
        ranges[c * R + plane * P + r] = out;
    }
}

// One thread per (cell, plane) for planes [plane_lo, plane_lo + n_planes):
// a plane's rows are consecutive in the direct table and in the ranges.
template <int D, bool DIRECT>
__global__ __launch_bounds__(256) void k_stencil(const uint64_t *__restrict__ cell_key, const int32_t *__restrict__ ncp,
                                                 const InstGrid *__restrict__ ig, int lin_bits,
                                                 const RowSlot *__restrict__ table, uint64_t mask,
                                                 const int2 *__restrict__ dtable, int2 *__restrict__ ranges,
                                                 const int32_t *__restrict__ cell_c, int plane_lo, int n_planes) {
    const int64_t nc = *ncp;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nc * n_planes) return;
    const int64_t c = tid / n_planes;
    const int plane = plane_lo + (int)(tid - c * n_planes);
    stencil_plane<D, DIRECT>(c, plane, nc, cell_key, ig, lin_bits, table, mask, dtable, ranges, cell_c);
}

// The planes below the centre (dz < 0) of the cells the border search reads
// them for: the first and last cell of each listed tile and the cell of each
// listed point.  A cell met more than once is written with the same values.
template <int D, bool DIRECT>
__global__ __launch_bounds__(256) void k_stencil_back(const int4 *__restrict__ ctiles,
                                                      const int32_t *__restrict__ blist,
                                                      const int32_t *__restrict__ n_blist,
                                                      const int32_t *__restrict__ plist,
                                                      const int32_t *__restrict__ n_plist,
                                                      const int32_t *__restrict__ cell_of,
                                                      const uint64_t *__restrict__ cell_key,
                                                      const int32_t *__restrict__ ncp, const InstGrid *__restrict__ ig,
                                                      int lin_bits, const RowSlot *__restrict__ table, uint64_t mask,
                                                      const int2 *__restrict__ dtable, int2 *__restrict__ ranges,
                                                      const int32_t *__restrict__ cell_c) {
    const int64_t nc = *ncp, nbt = *n_blist, npt = *n_plist;
    const int64_t total = 4 * nbt + 2 * npt;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t cell;
        if (i < 4 * nbt) {
            const int4 tile = ctiles[blist[i >> 2]];
            cell = (i & 2) ? tile.y : tile.x;
        } else {
            cell = cell_of[plist[(i - 4 * nbt) >> 1]];
        }
        if (cell < nc) stencil_plane<D, DIRECT>(cell, (int)(i & 1), nc, cell_key, ig, lin_bits, table, mask, dtable,
                                                ranges, cell_c);
    }
}

void stencil_back(Ctx &c, const int32_t *blist, const int32_t *n_blist, const int32_t *plist,
                  const int32_t *n_plist) {
    if (!c.back_rows_pending) return;
    c.back_rows_pending = false;
    const hipStream_t s = c.stream;
    const int32_t *ncp = c.dcount.as<int32_t>();
    const InstGrid *ig = c.inst_d.as<InstGrid>();
    Launch L(c.prof, s, "stencil");
    if (c.row_direct)
        k_stencil_back<3, true><<<2048, 256, 0, s>>>(c.ctiles.as<int4>(), blist, n_blist, plist, n_plist,
                                                     c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(), ncp, ig,
                                                     c.lin_bits, nullptr, 0, c.row_keys.as<int2>(),
                                                     c.ranges.as<int2>(), c.cell_c.as<int32_t>());
    else
        k_stencil_back<3, false><<<2048, 256, 0, s>>>(c.ctiles.as<int4>(), blist, n_blist, plist, n_plist,
                                                      c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(), ncp, ig,
                                                      c.lin_bits, c.row_keys.as<RowSlot>(), c.row_mask, nullptr,
                                                      c.ranges.as<int2>(), c.cell_c.as<int32_t>());
}

// Counting groups: the cells of one row whose c0 share c0 / group_cells; the
// group's points are contiguous in sorted order and are cut into tiles of at
// most 64 points (one wave each).

// A dense cell is a group of its own, cut into tiles strip by strip.
__global__ __launch_bounds__(256) void k_group_flags(const uint64_t *__restrict__ cell_key,
                                                     const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                                     int group_cells, const int32_t *__restrict__ dense_id,
                                                     int32_t *__restrict__ gflag) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    bool f = c == 0;
    if (!f) {
        const uint64_t rk = cell_key[c] - (uint64_t)cell_c[c], rk_prev = cell_key[c - 1] - (uint64_t)cell_c[c - 1];
        f = rk != rk_prev || cell_c[c] / group_cells != cell_c[c - 1] / group_cells;
        if (dense_id) f = f || dense_id[c] >= 0 || dense_id[c - 1] >= 0;
    }
    gflag[c] = f;
}

// Tiles of the strips of dense cell dc.
__device__ __forceinline__ int32_t dense_strip_tiles(const DenseCell &dc, const int32_t *__restrict__ tbl, int col) {
    const int32_t *t = tbl + dc.tbl;
    const int32_t a = t[col << dc.R], b = t[(col + 1) << dc.R];
    return (b - a + kDenseTile - 1) / kDenseTile;
}

// Per cell that opens a group: number of tiles of the group.
// Entries [nc, ncap) are written as 0, so a scan over ncap entries totals
// the tiles.
__global__ __launch_bounds__(256) void k_group_tiles(const int32_t *__restrict__ gflag,
                                                     const int32_t *__restrict__ cell_start, const int32_t *__restrict__ ncp,
                                                     int64_t ncap, const DenseView dv, int32_t *__restrict__ ntiles) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) {
        if (c < ncap) ntiles[c] = 0;
        return;
    }
    int32_t t = 0;
    if (gflag[c]) {
        const int32_t di = dv.id ? dv.id[c] : -1;
        if (di >= 0) {
            const DenseCell dc = dv.info[di];
            for (int col = 0; col < (1 << dc.L); ++col) t += dense_strip_tiles(dc, dv.tbl, col);
        } else {
            int64_t e = c + 1;
            while (e < nc && !gflag[e]) ++e;
            t = (cell_start[e] - cell_start[c] + kTile - 1) / kTile;
        }
    }
    ntiles[c] = t;
}

// Tiles of dense cell blockIdx.x.  The strips are taken in groups of 32;
// inside a group the tiles are ordered by (tile index in its strip, strip),
// so consecutive tiles -- which run at the same time on one XCD -- cover a
// compact patch (32 strips wide, a few tiles tall) and share most of their
// candidates in that XCD's L2.
constexpr int kStripGroup = 32;

__global__ __launch_bounds__(256) void k_dense_tile_fill(const DenseView dv, const int32_t *__restrict__ toff,
                                                         int4 *__restrict__ ctiles) {
    __shared__ int32_t nt[1 << kDenseYBits];
    __shared__ int32_t gbase[(1 << kDenseYBits) / kStripGroup + 1];
    const DenseCell dc = dv.info[blockIdx.x];
    const int ncol = 1 << dc.L;
    const int G = ncol < kStripGroup ? ncol : kStripGroup, ng = ncol / G;
    for (int col = threadIdx.x; col < ncol; col += blockDim.x) nt[col] = dense_strip_tiles(dc, dv.tbl, col);
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = toff[dc.cell];
        for (int g = 0; g < ng; ++g) {
            gbase[g] = run;
            for (int k = g * G; k < g * G + G; ++k) run += nt[k];
        }
    }
    __syncthreads();
    const int32_t *tb = dv.tbl + dc.tbl;
    for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
        const int g0 = col - col % G;
        const int32_t a = tb[col << dc.R], b = tb[(col + 1) << dc.R];
        for (int32_t j = 0; j < nt[col]; ++j) {
            // tiles before (j, col) in the group: all tiles of index < j, and
            // tile j of the strips left of col
            int32_t pos = gbase[g0 / G];
            for (int k = g0; k < g0 + G; ++k) pos += (nt[k] < j ? nt[k] : j) + (k < col && nt[k] > j ? 1 : 0);
            const int32_t p = a + j * kDenseTile;
            ctiles[pos] = make_int4(dc.cell, dc.cell, p, b - p < kDenseTile ? b - p : kDenseTile);
        }
    }
}

// Count tiles: (first cell, last cell, first point, points).
__global__ __launch_bounds__(256) void k_group_tile_fill(const int32_t *__restrict__ gflag,
                                                         const int32_t *__restrict__ cell_start,
                                                         const int32_t *__restrict__ cell_of,
                                                         const int32_t *__restrict__ toff, const int32_t *__restrict__ ncp,
                                                         int64_t tcap, const int32_t *__restrict__ dense_id,
                                                         int4 *__restrict__ ctiles) {
    const int64_t nc = *ncp;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc || !gflag[c] || (dense_id && dense_id[c] >= 0)) return;
    int64_t e = c + 1;
    while (e < nc && !gflag[e]) ++e;
    const int32_t p1 = cell_start[e];
    int32_t t = toff[c];
    for (int32_t p = cell_start[c]; p < p1 && t < tcap; p += kTile) {
        const int32_t len = p1 - p < kTile ? p1 - p : kTile;
        ctiles[t++] = make_int4(cell_of[p], cell_of[p + len - 1], p, len);
    }
}

// -------------------------------------------------------------- count --
// One wave per tile of <= 64 points of one cell.  The candidate loop is
// wave-uniform, so candidate coordinates come through scalar loads and each
// lane tests its own query point with the reference's arithmetic
// (data_process.cpp:443-447): ascending d, no contraction, '<='.
// Unrolled candidate block: U candidates [j, j+U) of a wave-uniform range.
template <int D, int U>
__device__ __forceinline__ int32_t count_block(const float (&qx)[D], const float *__restrict__ soa, int64_t npad,
                                               int j, float e2, int32_t cnt) {
    float c[D][U];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) c[k][u] = soa[k * npad + j + u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float len = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float df = qx[k] - c[k][u];
            len = __fadd_rn(len, __fmul_rn(df, df));
        }
        cnt = add_if(cnt, len <= e2);
    }
    return cnt;
}

// One wave per count tile (<= 64 points of a run of cells of one row).  The
// lanes first compute the candidate window of every stencil row in parallel
// (tile_window: rows and cells out of reach of the tile's bounding box are
// dropped), then the wave walks the non-empty windows.  The candidate loop is
// wave-uniform (scalar loads) and each lane tests its own point with the
// reference's arithmetic (data_process.cpp:443-447): ascending d, no
// contraction, '<='.  2-D tiles whose windows reach a dense cell take that
// cell through its strip table.
template <int D>
// 6 waves per SIMD (VGPRs capped at 80, a few spilled): measured on the 100M
// 2-D workload 705 ms at 5 (the unbounded allocation), 640 at 6, 658 at 7,
// 669 at 8.
__global__ __launch_bounds__(256, 6) void k_count(const int4 *__restrict__ ctiles, const int32_t *__restrict__ ntp,
                                               const int32_t *__restrict__ cell_start,
                                               const uint64_t *__restrict__ cell_key,
                                               const InstGrid *__restrict__ ig, int lin_bits,
                                               const int2 *__restrict__ ranges, int rows_per_cell,
                                               const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                               const float *__restrict__ soa, int64_t npad,
                                               int32_t *__restrict__ nn_s,
                                               int64_t *__restrict__ tests,
                                               const DenseView dv, const int32_t *__restrict__ part_off) {
    const int64_t nc = *ncp;
    const int64_t n_tiles = *ntp;
    __shared__ __align__(16) float stage[4][D][kStage];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (t >= n_tiles) return;
    const int lane = threadIdx.x & 63;
    const int4 tile = ctiles[t];
    const int cf = tile.x, cl = tile.y;
    const bool valid = lane < tile.w;
    // tiles of dense 2-D cells hold <= 16 points: lane = (query lane & 15, slice lane >> 4)
    const bool dtile = D == 2 && dv.id && dv.id[cf] >= 0;
    const int q = tile.z + (dtile ? ((lane & 15) < tile.w ? (lane & 15) : 0) : (valid ? lane : 0));
    float qx[D], lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        qx[k] = soa[k * npad + q];
        lo[k] = wave_min(qx[k]);
        hi[k] = wave_max(qx[k]);
    }
    const InstGrid &g = ig[cell_key[cf] >> lin_bits];
    const float e2 = g.e2;
    const int2 w = tile_window<D>(lane, rows_per_cell, cf, cl, lo, hi, g, cell_c[nc + cf], cell_c[2 * nc + cf],
                                  ranges, cell_c, nc);
    const int ps = w.x < w.y ? cell_start[w.x] : 0, pe = w.x < w.y ? cell_start[w.y] : 0;
    uint64_t rows = __ballot(pe > ps);
    int32_t cnt = 0;
    int64_t issued = 0;
    if constexpr (D == 2) {
        if (dtile) {
            int32_t bulk = 0;
            while (rows) {
                const int slot = __builtin_ctzll(rows);
                rows &= rows - 1;
                const int a = __builtin_amdgcn_readlane(w.x, slot), e = __builtin_amdgcn_readlane(w.y, slot);
                int run = cell_start[a];  // start of the pending run of ordinary cells
                for (int B = a; B <= e; ++B) {
                    const int di = B < e ? dv.id[B] : -1;
                    if (B < e && di < 0) continue;
                    cnt += staged_run<D, 4>(run, cell_start[B], qx, soa, npad, e2, &stage[wave][0][0], lane, issued);
                    if (B == e) break;
                    cnt += dense_cell_count<4>(qx, lo, hi, g, cell_c[B], cell_c[nc + B], dv.info[di], dv.tbl, soa,
                                               npad, e2, &stage[wave][0][0], lane, bulk, issued);
                    run = cell_start[B + 1];
                }
            }
            cnt += __shfl_xor(cnt, 16, 64);
            cnt += __shfl_xor(cnt, 32, 64);
            if (valid) nn_s[q] = cnt + bulk;
            if (tests && lane == 0) tests[t] = issued * tile.w;
            return;
        }
        if (dv.id) {
            bool near = false;
            for (int B = w.x; B < w.y; ++B) near = near || dv.id[B] >= 0;
            if (tile.w > 16 || __ballot(near) != 0) {
                // as below, with the dense cells of each window taken by strips
                int32_t bulk = 0;
                while (rows) {
                    const int slot = __builtin_ctzll(rows);
                    rows &= rows - 1;
                    const int a = __builtin_amdgcn_readlane(w.x, slot), e = __builtin_amdgcn_readlane(w.y, slot);
                    int run = cell_start[a];  // start of the pending run of ordinary cells
                    for (int B = a; B <= e; ++B) {
                        const int di = B < e ? dv.id[B] : -1;
                        if (B < e && di < 0) continue;
                        const int end = cell_start[B];
                        int j = run;
                        issued += end - j;
                        for (; j + 8 <= end; j += 8) cnt = count_block<D, 8>(qx, soa, npad, j, e2, cnt);
                        for (; j < end; ++j) cnt = count_block<D, 1>(qx, soa, npad, j, e2, cnt);
                        if (B == e) break;
                        cnt += dense_cell_count<1>(qx, lo, hi, g, cell_c[B], cell_c[nc + B], dv.info[di], dv.tbl,
                                                   soa, npad, e2, &stage[wave][0][0], lane, bulk, issued);
                        run = cell_start[B + 1];
                    }
                }
                if (valid) nn_s[q] = cnt + bulk;
                if (tests && lane == 0) tests[t] = issued * tile.w;
                return;
            }
        }
    }
    if (tile.w > 16) {
        // wave-uniform candidates: every lane tests its own query.  With a
        // part table, each cell of a window contributes only the runs of its
        // halves-per-dimension parts within reach of the tile box.
        const double r2o = search_r2(e2);
        // parts pay off when the tile box is within about one cell
        bool small_box = part_off != nullptr;
#pragma unroll
        for (int k = 0; k < D; ++k) small_box = small_box && (double)(hi[k] - lo[k]) * g.inv_s <= 1.0;
        while (rows) {
            const int slot = __builtin_ctzll(rows);
            rows &= rows - 1;
            if (small_box) {
                const int a = __builtin_amdgcn_readlane(w.x, slot), e = __builtin_amdgcn_readlane(w.y, slot);
                // the row's cells share dimensions 1..D-1: their half gaps once
                const double s_ = 1.0 / g.inv_s, h = 0.5 * s_, slack = s_ * 9.5367431640625e-07;
                double gy[2] = {0.0, 0.0}, gz[2] = {0.0, 0.0};
                if (D >= 2) {
#pragma unroll
                    for (int k = 1; k < D; ++k) {
                        const double X = g.origin[k] + (double)cell_c[k * nc + a] * s_;
#pragma unroll
                        for (int b = 0; b < 2; ++b) {
                            const double a0 = X + (double)b * h - slack, a1 = a0 + h + 2.0 * slack;
                            const double gp = fmax(0.0, fmax(a0 - (double)hi[k], (double)lo[k] - a1));
                            (k == 1 ? gy : gz)[b] = gp * gp;
                        }
                    }
                }
                int run = cell_start[a];  // pending run of whole cells
                for (int B = a; B <= e; ++B) {
                    const int32_t b0 = cell_start[B];
                    // cells with few points are taken whole (merged into the run)
                    if (B < e && cell_start[B + 1] - b0 < kPartMin) continue;
                    int j = run;
                    issued += b0 - j;
                    for (; j + 8 <= b0; j += 8) cnt = count_block<D, 8>(qx, soa, npad, j, e2, cnt);
                    for (; j < b0; ++j) cnt = count_block<D, 1>(qx, soa, npad, j, e2, cnt);
                    if (B == e) break;
                    const double X = g.origin[0] + (double)cell_c[B] * s_;
                    double gx[2];
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const double a0 = X + (double)b * h - slack, a1 = a0 + h + 2.0 * slack;
                        const double gp = fmax(0.0, fmax(a0 - (double)hi[0], (double)lo[0] - a1));
                        gx[b] = gp * gp;
                    }
                    uint32_t mask = 0;
#pragma unroll
                    for (int o = 0; o < (1 << D); ++o) {
                        double sum = gx[(o >> part_bit<D>(0)) & 1];
                        if (D >= 2) sum += gy[(o >> part_bit<D>(1)) & 1];
                        if (D >= 3) sum += gz[(o >> part_bit<D>(2)) & 1];
                        if (sum <= r2o) mask |= 1u << o;
                    }
                    const int32_t *po = part_off + (int64_t)B * 8;
                    const int32_t b_end = cell_start[B + 1];
                    while (mask) {
                        const int o1 = __builtin_ctz(mask);
                        const uint32_t above = ~((1u << o1) - 1u);
                        const int o2 = (~mask & above) ? __builtin_ctz(~mask & above) : (1 << D);
                        mask &= o2 < 32 ? ~((1u << o2) - 1u) : 0u;
                        int jj = po[o1];
                        const int end = o2 < (1 << D) ? po[o2] : b_end;
                        issued += end - jj;
                        for (; jj + 8 <= end; jj += 8) cnt = count_block<D, 8>(qx, soa, npad, jj, e2, cnt);
                        for (; jj < end; ++jj) cnt = count_block<D, 1>(qx, soa, npad, jj, e2, cnt);
                    }
                    run = b_end;
                }
                continue;
            }
            const int end = __builtin_amdgcn_readlane(pe, slot);
            int j = __builtin_amdgcn_readlane(ps, slot);
            issued += end - j;
            for (; j + 8 <= end; j += 8) cnt = count_block<D, 8>(qx, soa, npad, j, e2, cnt);
            for (; j < end; ++j) cnt = count_block<D, 1>(qx, soa, npad, j, e2, cnt);
        }
        if (valid) nn_s[q] = cnt;
        if (tests && lane == 0) tests[t] = issued * tile.w;
        return;
    }
    // Few queries (sparse tiles): the windows are flattened into one candidate
    // sequence (wave scan of their lengths), staged through LDS 256 at a time
    // with independent loads, and split over lanes = (query qi, slice sub).
    const int len = pe - ps;
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t2 = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t2;
    }
    const int pre = incl - len;
    const int total = __builtin_amdgcn_readlane(incl, 63);
    const int qb = tile.w <= 1 ? 0 : tile.w <= 2 ? 1 : tile.w <= 4 ? 2 : tile.w <= 8 ? 3 : 4;
    const int qi = lane & ((1 << qb) - 1), sub = lane >> qb;
    const bool qvalid = qi < tile.w;
    float my[D];
#pragma unroll
    for (int k = 0; k < D; ++k) my[k] = __shfl(qx[k], qvalid ? qi : 0, 64);
    for (int cb = 0; cb < total; cb += kStage) {
        const int n_here = total - cb < kStage ? total - cb : kStage;
#pragma unroll
        for (int u = 0; u < kStage / 64; ++u) {
            const int gi = cb + u * 64 + lane;
            const int gc = gi < total ? gi : total - 1;  // all lanes take part in the shuffles
            int lo_w = 0, hi_w = 63;  // last lane whose window starts at or before gc
#pragma unroll
            for (int it = 0; it < 6; ++it) {
                const int mid = (lo_w + hi_w + 1) >> 1;
                if (__shfl(pre, mid, 64) <= gc) lo_w = mid;
                else hi_w = mid - 1;
            }
            const int j = __shfl(ps, lo_w, 64) + gc - __shfl(pre, lo_w, 64);
#pragma unroll
            for (int k = 0; k < D; ++k)
                stage[wave][k][u * 64 + lane] = gi < total ? soa[k * npad + j] : __builtin_inff();
        }
        __builtin_amdgcn_wave_barrier();
        cnt += stage_test_sliced<D>(my, &stage[wave][0][0], n_here, e2, lane, qb);
        __builtin_amdgcn_wave_barrier();
    }
    issued = total;
    // sum the slices of each query (lanes qi, qi + 2^qb, ...)
    for (int o = 32; o >= (1 << qb); o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (sub == 0 && qvalid) nn_s[tile.z + qi] = cnt;
    if (tests && lane == 0) tests[t] = issued * tile.w;
}

// ------------------------------------------------------ symmetric count --
// Quarter-cell parts: the top 2D bits of a point's sub-cell code, v in
// [0, 4^D), in sort order.  Quarter coordinate of part v in dimension k:
// d = 1, 3: bit k is the low bit and bit D + k the high bit of it (the 4^d
// Morton code of k_keys); d = 2: the top 4 bits of the 64 x 64 code
// interleave dimension 0 in the odd bits.
template <int D>
__device__ __forceinline__ int fine_q(int v, int k) {
    if (D == 2) return k == 0 ? (((v >> 3) & 1) << 1) | ((v >> 1) & 1) : (((v >> 2) & 1) << 1) | (v & 1);
    return (((v >> (D + k)) & 1) << 1) | ((v >> k) & 1);
}
// Grids without dense cells test every pair of distinct cells once.  Cells
// are numbered in key order, (c2, c1, c0)-major, so a query's candidates of
// higher cell index are exactly those of the stencil rows after the centre
// row and of the centre row's cells after the query's own cell.  A query is
// tested against those only, and each tested candidate receives the number
// of queries that accepted it (a popcount of the wave's compare mask).  The
// pair's decision is the same from either side: (a-b)^2 and (b-a)^2 are the
// same float.  A query's own cell is counted without tests: the cell side
// sqrt(e2/d)(1-2^-18) puts every pair of one cell within eps under the
// float32 predicate (self included, as data_process.cpp:443-447 counts it).
// Every count is an atomic add into nn_s, which starts at zero.

// a | b << 16 for wave-uniform a, b < 2^16 in one scalar instruction
__device__ __forceinline__ uint32_t pack16(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("s_pack_ll_b32_b16 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}

// Candidates [j, j+U) of a wave-uniform range against every lane's query
// (lanes without a query hold +inf, which accepts nothing); returns the
// lane's count and puts each candidate's count (a popcount of the compare
// mask, <= 64) into lane j+u-jb of cacc.  The U = 8 counts are packed a byte
// each into one wave-uniform 64-bit word, candidate u at byte bpos(u) =
// (u & 4) | (u & 1) << 1 | (u & 2) >> 1 (pairs are joined at 16 bits, then
// pairs of pairs at 8); j - jb is a multiple of 8 and lane l unpacks byte
// bpos(l & 7) (shift `sh`).  MASKED: a lane accepts only candidates at or
// after my_end (the end of its own cell); the others meet a threshold of -1.
template <int D, int U, bool MASKED>
__device__ __forceinline__ int32_t count_block_sym(const float (&qx)[D], const float *__restrict__ soa, int64_t npad,
                                                   int j, float e2, int my_end, int jb, int lane, uint32_t sh,
                                                   int32_t cnt, int32_t &cacc) {
    float c[D][U];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) c[k][u] = soa[k * npad + j + u];
    uint32_t bc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float len = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float df = qx[k] - c[k][u];
            len = __fadd_rn(len, __fmul_rn(df, df));
        }
        const float lim = MASKED ? (j + u >= my_end ? e2 : -1.0f) : e2;
        const bool ok = len <= lim;
        cnt = add_if(cnt, ok);
        bc[u] = (uint32_t)__builtin_popcountll(__ballot(ok));
    }
    const int base = j - jb;
    if constexpr (U == 8) {
        const uint32_t w0 = pack16(bc[0], bc[1]) | pack16(bc[2], bc[3]) << 8;
        const uint32_t w1 = pack16(bc[4], bc[5]) | pack16(bc[6], bc[7]) << 8;
        const uint64_t pk = (uint64_t)w1 << 32 | w0;
        const int32_t v = (int32_t)((pk >> sh) & 0xffu);
        cacc = (lane >> 3) == (base >> 3) ? v : cacc;
    } else {
        cacc = lane == base ? (int32_t)bc[0] : cacc;
    }
    return cnt;
}

// Candidates [j0, end) in chunks of 64: the chunk's candidate counts are one
// coalesced atomic add.
template <int D, bool MASKED>
__device__ __forceinline__ int32_t count_run_sym(const float (&qx)[D], const float *__restrict__ soa, int64_t npad,
                                                 int j0, int end, float e2, int my_end, int lane,
                                                 int32_t *__restrict__ nn_s, int32_t cnt) {
    const uint32_t sh = 8u * ((lane & 4) | (lane & 1) << 1 | (lane & 2) >> 1);
    for (int jb = j0; jb < end; jb += 64) {
        const int je = end - jb < 64 ? end : jb + 64;
        int32_t cacc = 0;
        int j = jb;
        for (; j + 8 <= je; j += 8) cnt = count_block_sym<D, 8, MASKED>(qx, soa, npad, j, e2, my_end, jb, lane, sh, cnt, cacc);
        for (; j < je; ++j) cnt = count_block_sym<D, 1, MASKED>(qx, soa, npad, j, e2, my_end, jb, lane, sh, cnt, cacc);
        if (lane < je - jb && cacc != 0) atomicAdd(nn_s + jb + lane, cacc);
    }
    return cnt;
}

// One wave per count tile, as k_count, over the higher-index candidates only.
// Tiles of more than 16 queries walk the windows wave-uniformly (scalar
// candidate loads, one query per lane; cells of 64+ points by the parts in
// reach of a tile box at most kPartBoxCells wide).  Smaller tiles put one candidate
// on each lane and loop over their queries: the candidate's count stays in
// its lane and each query's count is a popcount of the compare mask.
template <int D>
__device__ __forceinline__ void count_tile_sym(int64_t t, const int4 *__restrict__ ctiles,
                                               const int32_t *__restrict__ cell_start,
                                               const int32_t *__restrict__ cell_of,
                                               const uint64_t *__restrict__ cell_key,
                                               const InstGrid *__restrict__ ig, int lin_bits,
                                               const int2 *__restrict__ ranges, int rows_per_cell,
                                               const int32_t *__restrict__ cell_c, int64_t nc,
                                               const float *__restrict__ soa, int64_t npad,
                                               int32_t *__restrict__ nn_s, int64_t *__restrict__ tests,
                                               const int32_t *__restrict__ part_fine,
                                               const int32_t *__restrict__ big_id, float4 *__restrict__ qsh) {
    const int lane = threadIdx.x & 63;
    const int4 tile = ctiles[t];
    const int cf = tile.x, cl = tile.y;
    const bool valid = lane < tile.w;
    const int q = tile.z + (valid ? lane : 0);
    float qx[D], lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        qx[k] = soa[k * npad + q];
        lo[k] = wave_min(qx[k]);
        hi[k] = wave_max(qx[k]);
    }
    const int my_cell = cell_of[q];
    const int my_beg = cell_start[my_cell], my_end = cell_start[my_cell + 1];
    const InstGrid &g = ig[cell_key[cf] >> lin_bits];
    const float e2 = g.e2;
    // rows before the centre row hold lower cells only (and are not read:
    // 3-D grids compute them for the border search only); the centre row
    // keeps the cells after the tile's last cell
    const int centre = rows_per_cell >> 1;
    int2 w = tile_window<D>(lane < centre ? rows_per_cell : lane, rows_per_cell, cf, cl, lo, hi, g, cell_c[nc + cf],
                            cell_c[2 * nc + cf], ranges, cell_c, nc);
    if (lane == centre && w.x < cl + 1) w.x = cl + 1;
    if (w.x >= w.y) w = make_int2(0, 0);
    int ps = w.x < w.y ? cell_start[w.x] : 0, pe = w.x < w.y ? cell_start[w.y] : 0;
    // the tile's own cells after its first: tested per lane (candidates at
    // or after the end of the lane's own cell)
    const int m_beg = cell_start[cf + 1], m_end = cl > cf ? cell_start[cl + 1] : m_beg;
    int32_t cnt = 0;
    int64_t issued = 0;
    if (tile.w > 16) {
        float qw[D];  // lanes without a query: +inf
#pragma unroll
        for (int k = 0; k < D; ++k) qw[k] = valid ? qx[k] : __builtin_inff();
        issued += m_end - m_beg;
        cnt = count_run_sym<D, true>(qw, soa, npad, m_beg, m_end, e2, my_end, lane, nn_s, cnt);
        uint64_t rows = __ballot(pe > ps);
        const double r2o = search_r2(e2), r2in = (double)e2 * (1.0 - 9.5367431640625e-07);
        bool small_box = part_fine != nullptr;
#pragma unroll
        for (int k = 0; k < D; ++k) small_box = small_box && (double)(hi[k] - lo[k]) * g.inv_s <= kPartBoxCells;
        while (rows) {
            const int slot = __builtin_ctzll(rows);
            rows &= rows - 1;
            if (small_box) {
                const int a = __builtin_amdgcn_readlane(w.x, slot), e = __builtin_amdgcn_readlane(w.y, slot);
                // Cells of kPartMin+ points are taken by their quarter-cell
                // parts (4^D, Morton order) within reach of the tile box:
                // lane v tests part v.  Dimensions 1..D-1 are the row's.
                // A part whose farthest point from the box is within
                // sqrt(e2 (1 - 2^-20)) holds only neighbours of every query:
                // the float32 sum of squares is within (1 + 5.1 2^-24) of the
                // exact square distance, so it cannot exceed e2.  Such parts
                // are counted without tests.
                const double s_ = 1.0 / g.inv_s, qs = 0.25 * s_, slack = s_ * 9.5367431640625e-07;
                double gyz = 0.0, fyz = 0.0;
#pragma unroll
                for (int k = 1; k < D; ++k) {
                    const double a0 = g.origin[k] + (double)cell_c[k * nc + a] * s_ + (double)fine_q<D>(lane, k) * qs -
                                      slack;
                    const double a1 = a0 + qs + 2.0 * slack;
                    const double gp = fmax(0.0, fmax(a0 - (double)hi[k], (double)lo[k] - a1));
                    const double fp = fmax((double)hi[k] - a0, a1 - (double)lo[k]);
                    gyz += gp * gp;
                    fyz += fp * fp;
                }
                int run = cell_start[a];  // pending run of whole cells
                for (int B = a; B <= e; ++B) {
                    const int32_t b0 = cell_start[B];
                    // cells with few points are taken whole (merged into the run)
                    if (B < e && cell_start[B + 1] - b0 < kPartMin) continue;
                    issued += b0 - run;
                    cnt = count_run_sym<D, false>(qw, soa, npad, run, b0, e2, my_end, lane, nn_s, cnt);
                    if (B == e) break;
                    const double a0 = g.origin[0] + (double)cell_c[B] * s_ + (double)fine_q<D>(lane, 0) * qs - slack;
                    const double a1 = a0 + qs + 2.0 * slack;
                    const double gp = fmax(0.0, fmax(a0 - (double)hi[0], (double)lo[0] - a1));
                    const double fp = fmax((double)hi[0] - a0, a1 - (double)lo[0]);
                    const bool is_part = lane < (1 << (2 * D));
                    uint64_t mask = __ballot(is_part && gyz + gp * gp <= r2o);
                    const uint64_t inside = __ballot(is_part && fyz + fp * fp <= r2in) & mask;
                    const int32_t *po = part_fine + (int64_t)big_id[B] * 64;
                    const int32_t b_end = cell_start[B + 1];
                    // runs of parts wholly within eps: each query counts the
                    // run, each point of it gets the tile's query count
                    for (uint64_t im = inside; im;) {
                        const int o1 = __builtin_ctzll(im);
                        const uint64_t gaps = ~im & ~((1ull << o1) - 1ull);
                        const int o2 = gaps ? __builtin_ctzll(gaps) : 64;
                        im &= o2 < 64 ? ~((1ull << o2) - 1ull) : 0ull;
                        const int jj = po[o1];
                        const int end = o2 < (1 << (2 * D)) ? po[o2] : b_end;
                        cnt += end - jj;
                        for (int jb = jj; jb < end; jb += 64)
                            if (jb + lane < end) atomicAdd(nn_s + jb + lane, tile.w);
                    }
                    mask &= ~inside;
                    // runs of the other parts in reach; a gap of at most
                    // kRunGap points that holds no counted part is tested
                    // rather than split the run (short runs cost more per
                    // candidate than the few extra tests)
                    int pj = -1, pe = -1;
                    uint64_t seen = 0;  // parts before the end of the pending run
                    while (mask) {
                        const int o1 = __builtin_ctzll(mask);
                        const uint64_t gaps = ~mask & ~((1ull << o1) - 1ull);  // parts >= o1 not in reach
                        const int o2 = gaps ? __builtin_ctzll(gaps) : 64;
                        mask &= o2 < 64 ? ~((1ull << o2) - 1ull) : 0ull;
                        const int jj = po[o1];
                        const int end = o2 < (1 << (2 * D)) ? po[o2] : b_end;
                        const uint64_t upto = o2 < 64 ? (1ull << o2) - 1ull : ~0ull;
                        if (pj >= 0 && jj - pe <= kRunGap && (inside & ((1ull << o1) - 1ull) & ~seen) == 0) {
                            pe = end;
                            seen = upto;
                            continue;
                        }
                        if (pj >= 0) {
                            issued += pe - pj;
                            cnt = count_run_sym<D, false>(qw, soa, npad, pj, pe, e2, my_end, lane, nn_s, cnt);
                        }
                        pj = jj;
                        pe = end;
                        seen = upto;
                    }
                    if (pj >= 0) {
                        issued += pe - pj;
                        cnt = count_run_sym<D, false>(qw, soa, npad, pj, pe, e2, my_end, lane, nn_s, cnt);
                    }
                    run = b_end;
                }
                continue;
            }
            const int end = __builtin_amdgcn_readlane(pe, slot);
            const int j = __builtin_amdgcn_readlane(ps, slot);
            issued += end - j;
            cnt = count_run_sym<D, false>(qw, soa, npad, j, end, e2, my_end, lane, nn_s, cnt);
        }
        issued *= tile.w;
    } else {
        // one candidate per lane, the tile's queries in turn; the tile's own
        // cells after its first are the window of lane R (the lane after the
        // stencil rows, which have none), so the windows sit in lanes
        // [centre, R] and a lane finds its window in log2(kSearch) steps
        constexpr int R = kStencilRows<D>, C0 = R / 2;
        constexpr int kSearch = D == 3 ? 16 : D == 2 ? 4 : 2;  // power of two >= R - C0 + 1
        if (lane == R) {
            ps = m_beg;
            pe = m_end;
        }
        const int len = pe - ps;
        int incl = len;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t2 = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t2;
        }
        const int pre = incl - len;
        const int total = __builtin_amdgcn_readlane(incl, 63);
        // the queries (coordinates, end of the query's own cell) for broadcast
        // reads: one 16-byte LDS read per query and round
        if (lane < 16) qsh[lane] = make_float4(qx[0], D >= 2 ? qx[D >= 2 ? 1 : 0] : 0.0f,
                                                     D >= 3 ? qx[D >= 3 ? 2 : 0] : 0.0f, __int_as_float(my_end));
        __builtin_amdgcn_wave_barrier();
        int32_t qcnt[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) qcnt[i] = 0;
        // candidate gc of the flattened windows: its point index
        auto cand_of = [&](int cb0) {
            const int gc = cb0 + lane < total ? cb0 + lane : total - 1;  // all lanes take part in the shuffles
            int lo_w = C0, hi_w = C0 + kSearch - 1;  // last lane whose window starts at or before gc
#pragma unroll
            for (int it = 1; it < kSearch; it <<= 1) {
                const int mid = (lo_w + hi_w + 1) >> 1;
                if (__shfl(pre, mid, 64) <= gc) lo_w = mid;
                else hi_w = mid - 1;
            }
            return __shfl(ps, lo_w, 64) + gc - __shfl(pre, lo_w, 64);
        };
        // rounds of 64 candidates; the next round's point loads are issued
        // before this round's tests
        int jn = total > 0 ? cand_of(0) : 0;
        float cn[D];
#pragma unroll
        for (int k = 0; k < D; ++k) cn[k] = total > 0 ? soa[k * npad + jn] : 0.0f;
        for (int cb = 0; cb < total; cb += 64) {
            const bool okc = cb + lane < total;
            const int j = jn;
            float c[D];
#pragma unroll
            for (int k = 0; k < D; ++k) c[k] = cn[k];
            if (cb + 64 < total) {
                jn = cand_of(cb + 64);
#pragma unroll
                for (int k = 0; k < D; ++k) cn[k] = soa[k * npad + jn];
            }
            int32_t ccount = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                if (i < tile.w) {
                    const float4 qv = qsh[i];
                    const float qq[3] = {qv.x, qv.y, qv.z};
                    float l2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const float df = qq[k] - c[k];
                        l2 = __fadd_rn(l2, __fmul_rn(df, df));
                    }
                    const bool ok = okc & (l2 <= e2) & (j >= __float_as_int(qv.w));
                    ccount = add_if(ccount, ok);
                    qcnt[i] += __builtin_popcountll(__ballot(ok));
                }
            }
            if (okc && ccount != 0) atomicAdd(nn_s + j, ccount);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (lane == i) cnt = qcnt[i];
        issued = (int64_t)total * tile.w;
    }
    if (valid) atomicAdd(nn_s + q, cnt + (my_end - my_beg));
    if (tests && lane == 0) tests[t] = issued;
}

// Four tiles of at most 16 queries in one wave: each 16-lane group takes a
// tile.  The lanes of a group compute its windows (the 13 stencil rows from
// the centre row on in 3-D, and the tile's own cells after its first), then
// take 16 candidates at a time and test them against the group's queries
// in turn; each lane keeps per-query partial counts, summed over the group
// at the end.  On sparse grids (a few points per tile) the time of a tile
// is its chain of dependent loads (tile, points, instance, windows, cell
// ranges, candidates); four tiles per wave run four such chains at once.
template <int D>
__device__ __forceinline__ void count_quad_sym(int64_t t0, int64_t n_tiles, const int4 *__restrict__ ctiles,
                                               const int32_t *__restrict__ cell_start,
                                               const int32_t *__restrict__ cell_of,
                                               const uint64_t *__restrict__ cell_key,
                                               const InstGrid *__restrict__ ig, int lin_bits,
                                               const int2 *__restrict__ ranges, int rows_per_cell,
                                               const int32_t *__restrict__ cell_c, int64_t nc,
                                               const float *__restrict__ soa, int64_t npad,
                                               int32_t *__restrict__ nn_s, int64_t *__restrict__ tests,
                                               float4 *__restrict__ qsh) {
    constexpr int R = kStencilRows<D>, C0 = R / 2, OWN = R - C0;  // lanes [0, OWN): rows C0.., lane OWN: own cells
    const int lane = threadIdx.x & 63, g = lane >> 4, gl = lane & 15;
    const int64_t t = t0 + g;
    const bool tv = t < n_tiles;
    const int4 tile = tv ? ctiles[t] : make_int4(0, 0, 0, 0);
    const int cf = tile.x, cl = tile.y, wq = tile.w;
    const bool valid = gl < wq;
    const int q = tile.z + (valid ? gl : 0);
    float qx[D], lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        qx[k] = soa[k * npad + q];
        float a = valid ? qx[k] : __builtin_inff(), b = valid ? qx[k] : -__builtin_inff();
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) {
            a = fminf(a, __shfl_xor(a, o, 16));
            b = fmaxf(b, __shfl_xor(b, o, 16));
        }
        lo[k] = a;
        hi[k] = b;
    }
    const int my_cell = cell_of[q];
    const int my_beg = cell_start[my_cell], my_end = cell_start[my_cell + 1];
    const InstGrid &gg = ig[cell_key[cf] >> lin_bits];
    const float e2 = gg.e2;
    int2 w = make_int2(0, 0);
    if (wq > 0 && gl < OWN) {
        w = tile_window<D>(C0 + gl, rows_per_cell, cf, cl, lo, hi, gg, D >= 2 ? cell_c[nc + cf] : 0,
                           D >= 3 ? cell_c[2 * nc + cf] : 0, ranges, cell_c, nc);
        if (gl == 0 && w.x < cl + 1) w.x = cl + 1;  // centre row: the cells after the tile's last
        if (w.x >= w.y) w = make_int2(0, 0);
    }
    int ps = w.x < w.y ? cell_start[w.x] : 0, pe = w.x < w.y ? cell_start[w.y] : 0;
    if (gl == OWN && wq > 0) {  // the tile's own cells after its first
        ps = cell_start[cf + 1];
        pe = cl > cf ? cell_start[cl + 1] : ps;
    }
    const int len = pe - ps;
    int incl = len;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        const int t2 = __shfl_up(incl, o, 16);
        if (gl >= o) incl += t2;
    }
    const int pre = incl - len;
    const int total = __shfl(incl, 15, 16);
    int tmax = total;
    for (int o = 32; o >= 16; o >>= 1) tmax = max(tmax, __shfl_xor(tmax, o, 64));
    tmax = __builtin_amdgcn_readfirstlane(tmax);
    if (valid) qsh[lane] = make_float4(qx[0], D >= 2 ? qx[D >= 2 ? 1 : 0] : 0.0f, D >= 3 ? qx[D >= 3 ? 2 : 0] : 0.0f,
                                       __int_as_float(my_end));
    __builtin_amdgcn_wave_barrier();
    // candidate gc of the group's flattened windows: its point index (the
    // last lane whose window starts at or before gc holds it)
    auto cand_of = [&](int cb0) {
        const int gc = cb0 + gl < total ? cb0 + gl : (total > 0 ? total - 1 : 0);
        int lo_w = 0, hi_w = 15;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int mid = (lo_w + hi_w + 1) >> 1;
            if (__shfl(pre, mid, 16) <= gc) lo_w = mid;
            else hi_w = mid - 1;
        }
        return __shfl(ps, lo_w, 16) + gc - __shfl(pre, lo_w, 16);
    };
    int32_t qp[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) qp[i] = 0;
    int jn = cand_of(0);
    float cn[D];
#pragma unroll
    for (int k = 0; k < D; ++k) cn[k] = soa[k * npad + jn];
    for (int cb = 0; cb < tmax; cb += 16) {
        const bool okc = cb + gl < total;
        const int j = jn;
        float c[D];
#pragma unroll
        for (int k = 0; k < D; ++k) c[k] = cn[k];
        if (cb + 16 < tmax) {  // the next round's candidates load during this round's tests
            jn = cand_of(cb + 16);
#pragma unroll
            for (int k = 0; k < D; ++k) cn[k] = soa[k * npad + jn];
        }
        int32_t cc = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < wq) {
                const float4 qv = qsh[g * 16 + i];
                const float qq[3] = {qv.x, qv.y, qv.z};
                float l2 = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float df = qq[k] - c[k];
                    l2 = __fadd_rn(l2, __fmul_rn(df, df));
                }
                const bool ok = okc & (l2 <= e2) & (j >= __float_as_int(qv.w));
                cc = add_if(cc, ok);
                qp[i] = add_if(qp[i], ok);
            }
        }
        if (okc && cc != 0) atomicAdd(nn_s + j, cc);
    }
    // each query's count: its partials summed over the group's lanes
    int32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int32_t v = qp[i];
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 16);
        if (gl == i) cnt = v;
    }
    if (valid) atomicAdd(nn_s + q, cnt + (my_end - my_beg));
    if (tests && tv && gl == 0) tests[t] = (int64_t)total * wq;
}

// Whether the four tiles [t0, t0 + 4) all hold at most 16 queries (tiles
// past the end count as such): those run side by side in k_count_quad.
__device__ __forceinline__ bool quad_small(const int4 *__restrict__ ctiles, int64_t t0, int64_t n_tiles) {
    const int lane = threadIdx.x & 63;
    const int w4 = lane < 4 && t0 + lane < n_tiles ? ctiles[t0 + lane].w : 0;
    return __ballot(w4 > 16) == 0;
}

// One wave per count tile, over the higher-index candidates only; with
// `quads`, tiles of quads of small tiles are left to k_count_quad.  Tiles of more than 16
// queries walk the windows wave-uniformly (scalar candidate loads, one query
// per lane; cells of 64+ points by the parts in reach of a tile box at most
// kPartBoxCells wide).  Smaller tiles put one candidate on each lane and loop
// over their queries: the candidate's count stays in its lane and each
// query's count is a popcount of the compare mask.
template <int D>
// 8 waves per SIMD (SGPRs capped, the rest spilled to VGPR lanes): 7 waves
// left the tiles' load chains exposed (10M 3-D 12.9 -> 12.0 ms, batch 23.9 ->
// 21.0 ms).
__global__ __launch_bounds__(256, 8) void k_count_sym(const int4 *__restrict__ ctiles, const int32_t *__restrict__ ntp,
                                                   const int32_t *__restrict__ cell_start,
                                                   const int32_t *__restrict__ cell_of,
                                                   const uint64_t *__restrict__ cell_key,
                                                   const InstGrid *__restrict__ ig, int lin_bits,
                                                   const int2 *__restrict__ ranges, int rows_per_cell,
                                                   const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                                   const float *__restrict__ soa, int64_t npad,
                                                   int32_t *__restrict__ nn_s, int64_t *__restrict__ tests,
                                                   const int32_t *__restrict__ part_fine,
                                                   const int32_t *__restrict__ big_id, int quads) {
    __shared__ float4 qsh[4][16];  // per wave: the queries of a small tile
    const int64_t nc = *ncp;
    const int64_t n_tiles = *ntp;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (t >= n_tiles || (quads && quad_small(ctiles, t & ~(int64_t)3, n_tiles))) return;
    count_tile_sym<D>(t, ctiles, cell_start, cell_of, cell_key, ig, lin_bits, ranges, rows_per_cell, cell_c, nc, soa,
                      npad, nn_s, tests, part_fine, big_id, qsh[wave]);
}

// One wave per four consecutive count tiles that all hold at most 16
// queries (count_quad_sym); other quads are k_count_sym's.
template <int D>
__global__ __launch_bounds__(256, 8) void k_count_quad(const int4 *__restrict__ ctiles, const int32_t *__restrict__ ntp,
                                                    const int32_t *__restrict__ cell_start,
                                                    const int32_t *__restrict__ cell_of,
                                                    const uint64_t *__restrict__ cell_key,
                                                    const InstGrid *__restrict__ ig, int lin_bits,
                                                    const int2 *__restrict__ ranges, int rows_per_cell,
                                                    const int32_t *__restrict__ cell_c, const int32_t *__restrict__ ncp,
                                                    const float *__restrict__ soa, int64_t npad,
                                                    int32_t *__restrict__ nn_s, int64_t *__restrict__ tests) {
    __shared__ float4 qsh[4][64];  // per wave: the queries of its four tiles
    const int64_t nc = *ncp;
    const int64_t n_tiles = *ntp;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * 4;
    if (t0 >= n_tiles || !quad_small(ctiles, t0, n_tiles)) return;
    count_quad_sym<D>(t0, n_tiles, ctiles, cell_start, cell_of, cell_key, ig, lin_bits, ranges, rows_per_cell, cell_c,
                      nc, soa, npad, nn_s, tests, qsh[wave]);
}

// Sum of int64 values (pair-test statistics): one atomic per block.
__global__ __launch_bounds__(256) void k_sum_i64(const int64_t *__restrict__ v, int64_t n,
                                                 unsigned long long *__restrict__ out) {
    __shared__ int64_t sh[4];
    int64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += v[i];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(sh[0] + sh[1] + sh[2] + sh[3]));
}

// Cell table from sorted keys: flags, exclusive scan, cell of each point,
// first point and key of each cell; *total (device) = number of cells.
void cell_table(const uint64_t *keys, int64_t n, int sub_bits, int32_t *flag, int32_t *excl, int32_t *cell_of,
                int32_t *cell_start, uint64_t *cell_key, int32_t *total, DevBuf &scratch, hipStream_t s,
                Profiler &prof) {
    const int nb = (int)ceil_div(n, 256);
    {
        Launch L(prof, s, "cell_flags");
        k_cell_flags<<<nb, 256, 0, s>>>(keys, n, sub_bits, flag);
    }
    scan_exclusive_i32(flag, excl, n, total, scratch, s, prof);
    Launch L(prof, s, "cell_fill");
    k_cell_fill<<<nb, 256, 0, s>>>(keys, sub_bits, flag, excl, n, cell_of, cell_start, cell_key);
}

// nn and core flags in input order (v_coord_nn, data_process.h:57), and core
// flags in sorted order for the label pass.
__global__ __launch_bounds__(256) void k_classify_out(const int32_t *__restrict__ nn_s,
                                                      const uint32_t *__restrict__ perm,
                                                      const int32_t *__restrict__ cell_of,
                                                      const uint64_t *__restrict__ cell_key,
                                                      const InstGrid *__restrict__ ig, int lin_bits, int64_t n,
                                                      int32_t *__restrict__ nn, uint8_t *__restrict__ core_s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = nn_s[i];
    const bool c = v >= ig[cell_key[cell_of[i]] >> lin_bits].m;  // data_process.cpp:282
    nn[perm[i]] = v;
    core_s[i] = c;
}

// Core flags in input order from the input-order counts (coalesced; a byte
// scatter through the permutation costs a partial-line write per point).
// Instances are contiguous input ranges [off[k], off[k+1]): a wave finds the
// instance of its first point, and its lanes search only when the wave
// crosses into the next instance.
__global__ __launch_bounds__(256) void k_core_in(const int32_t *__restrict__ nn, const int64_t *__restrict__ off,
                                                 const InstGrid *__restrict__ ig, int n_inst, int64_t n,
                                                 uint8_t *__restrict__ core) {
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    if (base >= n) return;
    const int64_t i = base + (threadIdx.x & 63);
    int inst = 0;
    if (n_inst > 1) {
        int lo = 0, hi = n_inst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= base) lo = mid;
            else hi = mid - 1;
        }
        inst = lo;
        const int64_t last = base + 63 < n ? base + 63 : n - 1;
        if (off[inst + 1] <= last) {
            const int64_t iv = i < n ? i : n - 1;
            hi = n_inst - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (off[mid] <= iv) lo = mid;
                else hi = mid - 1;
            }
            inst = lo;
        }
    }
    if (i < n) core[i] = nn[i] >= ig[inst].m;  // data_process.cpp:282
}

// Part table: part_off[8 c + o] = first point of cell c whose part (top D
// bits of its sub-cell code) is >= o, for o < 2^D.
template <int D>
__global__ __launch_bounds__(256) void k_part_off(const uint64_t *__restrict__ keys, int sub_bits,
                                                  const int32_t *__restrict__ cell_start,
                                                  const int32_t *__restrict__ ncp, int32_t *__restrict__ part_off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = i >> D;
    const int o = (int)(i & ((1 << D) - 1));
    if (c >= *ncp) return;
    int lo = cell_start[c], hi = cell_start[c + 1];
    if (o > 0) {
        const uint64_t mask = (1ull << sub_bits) - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)((keys[mid] & mask) >> (sub_bits - D)) < o) lo = mid + 1;
            else hi = mid;
        }
    }
    part_off[c * 8 + o] = lo;
}

// Cells of kPartMin+ points get a table of the first point of each of their
// 4^D quarter-cell parts (part_fine[64 b + v], b the cell's index among such
// cells, an exclusive scan of these flags).  At most n / kPartMin cells
// qualify, so the table never exceeds 4 n bytes.
__global__ __launch_bounds__(256) void k_big_flag(const int32_t *__restrict__ cell_start,
                                                  const int32_t *__restrict__ ncp, int64_t ncap,
                                                  int32_t *__restrict__ flag) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncap) return;
    flag[c] = c < *ncp && cell_start[c + 1] - cell_start[c] >= kPartMin;
}

// One thread per point of a qualifying cell: the parts after the previous
// point's, up to its own, start at it (the last point also closes the
// parts after its own at the cell end).  Cells at or past *ncp exist only in
// a captured run over its cell capacity (big_id holds the capacity's cells;
// the run is flagged for an eager rerun): their points are left out.
template <int D>
__global__ __launch_bounds__(256) void k_part_fine(const uint64_t *__restrict__ keys, int sub_bits,
                                                   const int32_t *__restrict__ cell_of,
                                                   const int32_t *__restrict__ cell_start,
                                                   const int32_t *__restrict__ ncp,
                                                   const int32_t *__restrict__ big_id, int64_t n,
                                                   int32_t *__restrict__ part_fine) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t c = cell_of[p];
    if (c >= *ncp) return;
    const int32_t b = cell_start[c], e = cell_start[c + 1];
    if (e - b < kPartMin) return;
    const int shift = sub_bits - 2 * D;
    const uint64_t mask = (1ull << sub_bits) - 1;
    const int vp = (int)((keys[p] & mask) >> shift);
    const int vprev = p == b ? -1 : (int)((keys[p - 1] & mask) >> shift);
    int32_t *t = part_fine + (int64_t)big_id[c] * 64;
    for (int v = vprev + 1; v <= vp; ++v) t[v] = (int32_t)p;
    if (p == e - 1)
        for (int v = vp + 1; v < (1 << (2 * D)); ++v) t[v] = e;
}

// A captured run's count above its capacity: clamp it (so every later
// launch stays in bounds) and flag the run for an eager rerun.
__global__ void k_clamp_count(int32_t *__restrict__ count, int32_t cap, int32_t *__restrict__ status) {
    if (threadIdx.x == 0 && *count > cap) {
        *count = cap;
        atomicOr(status, 1);
    }
}

// ------------------------------------------------------------ pipeline --
static void upload_inst(Ctx &c) {
    c.inst_d.ensure(sizeof(InstGrid) * c.n_inst);
    NDBX_HIP(hipMemcpyAsync(c.inst_d.p, c.h_inst.data(), sizeof(InstGrid) * c.n_inst, hipMemcpyHostToDevice,
                            c.stream));
    c.offsets_d.ensure(sizeof(int64_t) * (c.n_inst + 1));
    NDBX_HIP(hipMemcpyAsync(c.offsets_d.p, c.h_offsets.data(), sizeof(int64_t) * (c.n_inst + 1),
                            hipMemcpyHostToDevice, c.stream));
}

template <int D>
static void grid_classify_d(Ctx &c) {
    const int64_t n = c.n;
    const hipStream_t s = c.stream;
    const int nb = (int)ceil_div(n, 256);
    // capture: no host reads; cell and tile counts stay on the device and
    // every cell / tile launch covers the capacities of the eager run
    const bool cap_mode = c.capacity_mode;
    if (!cap_mode) upload_inst(c);
    const InstGrid *ig = c.inst_d.as<InstGrid>();
    c.dcount.ensure(64);
    int32_t *ncp = c.dcount.as<int32_t>(), *ntp = ncp + 1, *status = ncp + 2;

    c.keys.ensure(n * 8);
    c.vals.ensure(n * 4);
    c.keys2.ensure(n * 8);
    c.vals2.ensure(n * 4);
    {
        Launch L(c.prof, s, "keys");
        k_keys<D><<<nb, 256, 0, s>>>(c.coords, n, ig, c.offsets_d.as<int64_t>(), c.n_inst, c.lin_bits, c.sub_bits,
                                    c.keys.as<uint64_t>(), c.vals.as<uint32_t>());
    }
    bool alt = false;
    radix_sort_u64(c.keys.as<uint64_t>(), c.vals.as<uint32_t>(), c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>(), n,
                   c.key_bits, c.scratch, s, c.prof, &alt);
    if (alt) {
        std::swap(c.keys, c.keys2);
        std::swap(c.vals, c.vals2);
    }
    const uint64_t *keys = c.keys.as<uint64_t>();

    c.flag.ensure(n * 4);
    c.cid.ensure(n * 4);
    c.cell_of.ensure(n * 4);
    c.cell_start.ensure((n + 1) * 4);
    c.cell_key.ensure(n * 8);
    c.misc.ensure(64);
    int32_t *d_total = c.misc.as<int32_t>();
    unsigned long long *d_tests = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 4);
    cell_table(keys, n, c.sub_bits, c.flag.as<int32_t>(), c.cid.as<int32_t>(), c.cell_of.as<int32_t>(),
               c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), ncp, c.scratch, s, c.prof);
    int32_t h_total = 0;
    int64_t nc;  // cells (eager) or cell capacity (capture)
    if (cap_mode) {
        nc = c.cap_cells;
        k_clamp_count<<<1, 64, 0, s>>>(ncp, (int32_t)nc, status);
    } else {
        NDBX_HIP(hipMemcpyAsync(&h_total, ncp, 4, hipMemcpyDeviceToHost, s));
        c.sync();
        nc = h_total;
        c.n_cells = nc;
    }
    const int ncb = (int)ceil_div(nc, 256);
    // part table (before the dense re-sort: ordinary cells keep this order)
    const int32_t *part_off = nullptr;
    if (D == 2 && c.sub_bits > 0) {  // read by the dense 2-D count only
        c.part_off.ensure((size_t)nc * 8 * 4);
        Launch L(c.prof, s, "part_off");
        k_part_off<D><<<(unsigned)ceil_div(nc << D, 256), 256, 0, s>>>(keys, c.sub_bits, c.cell_start.as<int32_t>(),
                                                                        ncp, c.part_off.as<int32_t>());
        part_off = c.part_off.as<int32_t>();
    }
    // quarter-cell part tables of cells with kPartMin+ points (before the
    // dense re-sort, like part_off)
    const int32_t *part_fine = nullptr, *big_id = nullptr;
    if (c.sub_bits >= 2 * D) {
        c.big_id.ensure((size_t)(nc + 1) * 4);
        c.gflag.ensure((size_t)(nc + 1) * 4);
        c.part_fine.ensure((size_t)(n / kPartMin + 1) * 64 * 4);
        {
            Launch L(c.prof, s, "part_off");
            k_big_flag<<<(unsigned)ceil_div(nc, 256), 256, 0, s>>>(c.cell_start.as<int32_t>(), ncp, nc,
                                                                   c.gflag.as<int32_t>());
        }
        scan_exclusive_i32(c.gflag.as<int32_t>(), c.big_id.as<int32_t>(), nc, nullptr, c.scratch, s, c.prof);
        {
            Launch L(c.prof, s, "part_off");
            k_part_fine<D><<<nb, 256, 0, s>>>(keys, c.sub_bits, c.cell_of.as<int32_t>(), c.cell_start.as<int32_t>(),
                                             ncp, c.big_id.as<int32_t>(), n, c.part_fine.as<int32_t>());
        }
        part_fine = c.part_fine.as<int32_t>();
        big_id = c.big_id.as<int32_t>();
    }
    // dense 2-D cells: strip order and tables (may reorder points inside
    // cells); eager runs only, a captured run counts them as ordinary cells
    DenseView dv{nullptr, nullptr, nullptr};
    if (!cap_mode) {
        c.n_dense = 0;
        if (D == 2 && c.sub_bits == 2 * kFineBits) dv = dense_prepare(c, ig, n, nc, d_total, s);
    }
    const uint32_t *perm = c.vals.as<uint32_t>();
    const int64_t npad = ceil_div(n, 64) * 64;
    c.soa.ensure((size_t)npad * D * 4);
    {
        Launch L(c.prof, s, "gather_soa");
        k_gather_soa<D><<<nb, 256, 0, s>>>(c.coords, perm, n, npad, c.soa.as<float>());
    }

    // row hash: power of two >= 2 * cells
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * nc)) cap <<= 1;
    c.cell_c.ensure((size_t)nc * 3 * 4);
    {
        Launch L(c.prof, s, "cell_coords");
        k_cell_coords<<<ncb, 256, 0, s>>>(c.cell_key.as<uint64_t>(), ncp, ig, c.lin_bits, c.cell_c.as<int32_t>());
    }
    // eager runs take a direct row table when it is no larger than the hash
    // (8-byte entries, at most 2 per hash slot); captured runs the hash
    // (captured runs: when the eager run took one, sized by cap_rows)
    const bool direct = cap_mode ? c.cap_rows > 0 : c.n_rows_direct <= 2 * (int64_t)cap;
    if (direct) {
        const int64_t n_rows = cap_mode ? c.cap_rows : c.n_rows_direct;
        c.row_keys.ensure((size_t)n_rows * sizeof(int2));
        fill_u64(c.row_keys.as<uint64_t>(), 0, n_rows, s, c.prof);  // (0, 0): empty row
        Launch L(c.prof, s, "row_insert");
        k_row_insert_direct<<<ncb, 256, 0, s>>>(c.cell_key.as<uint64_t>(), ncp, ig, c.lin_bits,
                                                c.cell_c.as<int32_t>(), c.row_keys.as<int2>());
    } else {
        c.row_keys.ensure(cap * sizeof(RowSlot));
        fill_u64(c.row_keys.as<uint64_t>(), kEmpty, (int64_t)cap * 2, s, c.prof);  // keys (and ranges) all ones
        Launch L(c.prof, s, "row_insert");
        k_row_insert<<<ncb, 256, 0, s>>>(c.cell_key.as<uint64_t>(), ncp, c.cell_c.as<int32_t>(),
                                         c.row_keys.as<RowSlot>(), cap - 1);
    }
    const int rows = c.rows_per_cell = kStencilRows<D>;
    c.ranges.ensure((size_t)nc * rows * 8);
    // 3-D: the planes from the centre on (dz >= 0) here, the ones below
    // for the cells the border search reads them for (stencil_back); the
    // count and the union read rows from the centre row on only
    c.back_rows_pending = D == 3;
    c.row_direct = direct;
    c.row_mask = cap - 1;
    {
        const int plane_lo = D == 3 ? 2 : 0, n_planes = D == 3 ? 3 : 1;
        const unsigned nbs = (unsigned)ceil_div(nc * n_planes, 256);
        Launch L(c.prof, s, "stencil");
        if (direct)
            k_stencil<D, true><<<nbs, 256, 0, s>>>(c.cell_key.as<uint64_t>(), ncp, ig, c.lin_bits, nullptr, 0,
                                                  c.row_keys.as<int2>(), c.ranges.as<int2>(), c.cell_c.as<int32_t>(),
                                                  plane_lo, n_planes);
        else
            k_stencil<D, false><<<nbs, 256, 0, s>>>(c.cell_key.as<uint64_t>(), ncp, ig, c.lin_bits,
                                                   c.row_keys.as<RowSlot>(), cap - 1, nullptr, c.ranges.as<int2>(),
                                                   c.cell_c.as<int32_t>(), plane_lo, n_planes);
    }
    // count tiles over groups of cells of one row
    c.gflag.ensure(nc * 4);
    c.ctile_off.ensure((nc + 1) * 4);
    {
        Launch L(c.prof, s, "group_tiles");
        k_group_flags<<<ncb, 256, 0, s>>>(c.cell_key.as<uint64_t>(), c.cell_c.as<int32_t>(), ncp, c.group_cells,
                                          dv.id, c.gflag.as<int32_t>());
        k_group_tiles<<<ncb, 256, 0, s>>>(c.gflag.as<int32_t>(), c.cell_start.as<int32_t>(), ncp, nc, dv,
                                          c.flag.as<int32_t>());
    }
    scan_exclusive_i32(c.flag.as<int32_t>(), c.ctile_off.as<int32_t>(), nc, ntp, c.scratch, s, c.prof);
    int64_t nct;  // tiles (eager) or tile capacity (capture)
    if (cap_mode) {
        nct = c.cap_tiles;
        k_clamp_count<<<1, 64, 0, s>>>(ntp, (int32_t)nct, status);
    } else {
        NDBX_HIP(hipMemcpyAsync(&h_total, ntp, 4, hipMemcpyDeviceToHost, s));
        c.sync();
        nct = h_total;
        c.n_tiles = nct;
    }
    c.ctiles.ensure(nct * 16);
    {
        Launch L(c.prof, s, "tiles");
        k_group_tile_fill<<<ncb, 256, 0, s>>>(c.gflag.as<int32_t>(), c.cell_start.as<int32_t>(),
                                              c.cell_of.as<int32_t>(), c.ctile_off.as<int32_t>(), ncp, nct, dv.id,
                                              c.ctiles.as<int4>());
        if (dv.id)
            k_dense_tile_fill<<<(unsigned)c.n_dense, 256, 0, s>>>(dv, c.ctile_off.as<int32_t>(), c.ctiles.as<int4>());
    }
    c.nn_s.ensure(n * 4);
    if (c.no_count) {  // grid and tiles only: counts and core flags are written by the caller
        c.core_s.ensure(n);
        return;
    }
    const bool stats = c.prof.on && c.prof.only.empty() && !cap_mode;  // pair tests when every launch is evented
    if (stats) {
        c.tile_tests.ensure(nct * 8);
        fill_u64(reinterpret_cast<uint64_t *>(d_tests), 0, 1, s, c.prof);
    }
    if (dv.id) {
        // dense 2-D cells: every tile counts its own queries against all candidates
        if constexpr (D == 2) {
            Launch L(c.prof, s, "count");
            k_count<D><<<(unsigned)ceil_div(nct, 4), 256, 0, s>>>(
                c.ctiles.as<int4>(), ntp, c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), ig, c.lin_bits,
                c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(), ncp, c.soa.as<float>(), npad, c.nn_s.as<int32_t>(),
                stats ? c.tile_tests.as<int64_t>() : nullptr, dv, part_off);
        }
    } else {
        // each pair of distinct cells tested once, counts added to both points
        fill_i32(c.nn_s.as<int32_t>(), 0, n, s, c.prof);
        Launch L(c.prof, s, "count");
        // sparse grids (fewer than 8 points per tile on average): quads of
        // small tiles share a wave (4-instance batch of 50M 3-D points: count
        // 21.2 -> 18.4 ms); on denser grids the second launch only costs
        // (10M 3-D: 11.4 -> 11.9 ms), so it is left out
        const int quads = c.quad_mode >= 0 ? c.quad_mode : c.n_tiles > 0 && c.n < 8 * c.n_tiles;
        k_count_sym<D><<<(unsigned)ceil_div(nct, 4), 256, 0, s>>>(
            c.ctiles.as<int4>(), ntp, c.cell_start.as<int32_t>(), c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(),
            ig, c.lin_bits, c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(), ncp, c.soa.as<float>(), npad,
            c.nn_s.as<int32_t>(), stats ? c.tile_tests.as<int64_t>() : nullptr, part_fine, big_id, quads);
        if (quads) k_count_quad<D><<<(unsigned)ceil_div(nct, 16), 256, 0, s>>>(
            c.ctiles.as<int4>(), ntp, c.cell_start.as<int32_t>(), c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(),
            ig, c.lin_bits, c.ranges.as<int2>(), rows, c.cell_c.as<int32_t>(), ncp, c.soa.as<float>(), npad,
            c.nn_s.as<int32_t>(), stats ? c.tile_tests.as<int64_t>() : nullptr);
    }
    if (stats) {
        Launch L(c.prof, s, "stats");
        k_sum_i64<<<1024, 256, 0, s>>>(c.tile_tests.as<int64_t>(), nct, d_tests);
    }
    c.core_s.ensure(n);
    {
        Launch L(c.prof, s, "classify_out");
        k_classify_out<<<nb, 256, 0, s>>>(c.nn_s.as<int32_t>(), perm, c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(),
                                          ig, c.lin_bits, n, c.nn.as<int32_t>(), c.core_s.as<uint8_t>());
        k_core_in<<<nb, 256, 0, s>>>(c.nn.as<int32_t>(), c.offsets_d.as<int64_t>(), ig, c.n_inst, n,
                                     c.core.as<uint8_t>());
    }
    if (stats) {
        unsigned long long h_tests = 0;
        NDBX_HIP(hipMemcpyAsync(&h_tests, d_tests, 8, hipMemcpyDeviceToHost, s));
        c.sync();
        c.pair_tests = (int64_t)h_tests;
    }
}

void grid_classify(Ctx &c) {
    switch (c.d) {
        case 1: grid_classify_d<1>(c); break;
        case 2: grid_classify_d<2>(c); break;
        case 3: grid_classify_d<3>(c); break;
        default: throw Error(NDBX_EUNSUPPORTED, "grid path needs d <= 3");
    }
}

}  // namespace ndbx
