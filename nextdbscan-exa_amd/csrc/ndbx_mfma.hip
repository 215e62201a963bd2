// MFMA distance filter for 4 <= d <= 128: neighbour counts, union-find
// and border attachment on 32x32 tiles of bf16 matrix products.
//
// 1. Partition: K sampled points are cell centres; every point joins its
//    nearest centre (bf16 matrix-product estimates, lowest index on ties); a
//    radix sort by cell makes each cell a contiguous range.  A cell's radius R_C (float64,
//    max member distance to its centre) gives a rigorous skip: cells C, C'
//    whose centres are farther apart than R_C + R_C' + r can hold no pair
//    within r = sqrt(e2 (1 + 2^-18)), so their tiles are never computed.
// 2. Estimate: coordinates centred on the data mean, a' = fl(x - mu), are
//    split into bf16 hi + lo (|a' - hi - lo| <= 2^-16 |a'|); three MFMAs per
//    16-wide k step (hi*hi + hi*lo + lo*hi, float32 accumulation) give a.b';
//    est = |a'|^2 + |b'|^2 - 2 a.b'.
// 3. Guard band: |est - len| <= band, len being the reference's float32
//    direct-difference sum (data_process.cpp:443-447), with
//       band = 2 * ( B1 * R^2 + 2^-21 * sqrt(e2) * R + (d+3) 2^-24 e2 ),
//       B1 = 2 (3.1 2^-16 + 3 Kp 2^-24) + (2 Kp + 16) 2^-24,
//    R = largest |a'| of the tile's rows and the candidate cell.  Pairs with
//    est <= e2 - band are neighbours, est > e2 + band are not, and the rest
//    are decided by the reference's own float32 arithmetic on the original
//    coordinates.  Every decision therefore equals the float32 predicate.
// 4. Block pruning: within a cell, points are ordered by their distance to
//    the cell's centroid m_C, so a 32-point tile's radius (largest distance)
//    is small except for the cell's outliers.  For every tile t and cell C,
//    T[t][C] bounds from below the distance of t's points to m_C; a row group
//    G of cell A and a column tile J of cell B hold no pair within r when
//    max(T[G][B] - rad(J), T[J][A] - rad(G)) > r (triangle inequality), and
//    only neighbours when rad(G) + |m_A - m_B| + rad(J) is small enough; the
//    count pass computes neither kind (the second is counted wholesale).
#include <cfloat>
#include <climits>

#include "ndbx_internal.h"

namespace ndbx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// A load of read-only data through the constant address space: at a
// wave-uniform address it is a scalar load, which waits on its own counter
// instead of behind the vector loads in flight.
template <class T>
__device__ __forceinline__ T ld_const(const T *p) {
    return *(const __attribute__((address_space(4))) T *)(p);
}
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// LDS reads of 16 bytes, waited for in the same statement: one at p, or
// four at p + k * STRIDE bytes.  Used where the LDS is filled by loads
// straight from global memory (global_load_lds): the compiler cannot tell
// which buffer such a load writes and would otherwise wait for every load in
// flight before each read; the caller orders the fills with explicit waits
// and barriers instead.
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
__device__ __forceinline__ bf16x8_t lds_read16(const bf16x8_t *p) {
    bf16x8_t v;
    asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_addr(p)));
    return v;
}
template <int STRIDE>
__device__ __forceinline__ void lds_read16x4(const bf16x8_t *p, bf16x8_t &v0, bf16x8_t &v1, bf16x8_t &v2,
                                             bf16x8_t &v3) {
    asm volatile(
        "ds_read_b128 %0, %4\n\t"
        "ds_read_b128 %1, %4 offset:%5\n\t"
        "ds_read_b128 %2, %4 offset:%6\n\t"
        "ds_read_b128 %3, %4 offset:%7\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
        : "v"(lds_addr(p)), "i"(STRIDE), "i"(2 * STRIDE), "i"(3 * STRIDE));
}

constexpr int kCellTarget = 1024;  // points per partition cell, on average
// largest size of the block pruning's distance bounds (two tiles x cells
// float tables); past it the count pass computes every candidate block
constexpr size_t kPruneBytes = (size_t)16 << 30;

// ---------------------------------------------------------------- prep --
// Column sums in two fixed-order levels (the result is deterministic):
// workgroup g sums rows [g rows_per, (g + 1) rows_per) in order, one thread
// per coordinate, into part[g][k]; k_colsum_total sums the workgroups'
// parts in order.
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ x, int64_t n, int d,
                                                double *__restrict__ part) {
    __shared__ double sp[256];
    const int64_t rows_per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = blockIdx.x * rows_per, e = b + rows_per < n ? b + rows_per : n;
    if (d > 256) {
        for (int k = threadIdx.x; k < d; k += blockDim.x) {
            double s = 0.0;
            for (int64_t i = b; i < e; ++i) s += x[i * d + k];
            part[(int64_t)blockIdx.x * d + k] = s;
        }
        return;
    }
    // P interleaved row streams of the group's rows, each summed in order,
    // then the streams' sums added in stream order
    const int P = 256 / d, k = threadIdx.x % d, st = threadIdx.x / d;
    double s = 0.0;
    if (st < P) {
        int64_t i = b + st;
        for (; i + 3 * P < e; i += 4 * P) {
            const float v0 = x[i * d + k], v1 = x[(i + P) * d + k], v2 = x[(i + 2 * P) * d + k],
                        v3 = x[(i + 3 * P) * d + k];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; i < e; i += P) s += x[i * d + k];
    }
    sp[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < d) {
        double t = 0.0;
        for (int q = 0; q < P; ++q) t += sp[q * d + threadIdx.x];
        part[(int64_t)blockIdx.x * d + threadIdx.x] = t;
    }
}
// One workgroup per coordinate: thread t adds parts t, t + 256, ... in
// order, then the threads' sums are added in a fixed tree order.
__global__ __launch_bounds__(256) void k_colsum_total(const double *__restrict__ part, int groups, int d,
                                                      double *__restrict__ sum) {
    __shared__ double sp[256];
    const int k = blockIdx.x;
    double t = 0.0;
    for (int g = threadIdx.x; g < groups; g += 256) t += part[(int64_t)g * d + k];
    sp[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) sp[threadIdx.x] += sp[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sum[k] = sp[0];
}

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }  // round to nearest even

// Partition centres: row j is point j n / k_cent (one launch; a copy per
// centre queued ~1,000 copies a run).
__global__ __launch_bounds__(256) void k_pick_centres(const float *__restrict__ x, int64_t n, int d, int64_t k_cent,
                                                      float *__restrict__ cent) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k_cent * d) return;
    const int64_t j = t / d, k = t - j * d;
    cent[t] = x[(j * n / k_cent) * d + k];
}

// The centres' bf16 rows (centred on mu, padded to KP, the padding centres
// up to a multiple of 32 all zero) and their squared norms (+inf for the
// padding, which is then never the nearest).
template <int KP>
__global__ __launch_bounds__(256) void k_centres_bf16(const float *__restrict__ cent, int k_cent, int k_pad, int d,
                                                      const float *__restrict__ mu, __bf16 *__restrict__ chi,
                                                      float *__restrict__ cn) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k_pad) return;
    float s = 0.0f;
    for (int k = 0; k < KP; ++k) {
        const __bf16 h = to_bf16(c < k_cent && k < d ? cent[(int64_t)c * d + k] - mu[k] : 0.0f);
        chi[(int64_t)c * KP + k] = h;
        s += (float)h * (float)h;
    }
    cn[c] = c < k_cent ? s : __builtin_inff();
}

// Nearest centre, approximately: the smallest |c|^2 - 2 a.c over the
// centres, a and c centred and rounded to bf16, from matrix products (lowest
// index on ties).  Any assignment is a valid partition (cell radii, pair
// lists and bounds are computed from the points themselves); the nearest
// centre only keeps the cells compact.  One wave per 32 points: lane l holds
// row l % 32 (k elements [16 ks + 8 (l / 32), +8) of each 16-wide k step)
// and takes column l % 32 of every 32 centres.
template <int KP>
__global__ __launch_bounds__(256) void k_assign(const float *__restrict__ x, int64_t n, int d,
                                                const float *__restrict__ mu, const __bf16 *__restrict__ chi,
                                                const float *__restrict__ cn, int k_pad,
                                                uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    constexpr int KS = KP / 16;
    const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
    const int64_t row0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 32;
    if (row0 >= n) return;
    const int64_t r = row0 + col;
    bf16x8_t a[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = 16 * ks + 8 * half + u;
            a[ks][u] = to_bf16(r < n && k < d ? x[r * d + k] - mu[k] : 0.0f);
        }
    const bf16x8_t *c8 = reinterpret_cast<const bf16x8_t *>(chi);
    float best[16];
    int arg[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) best[q] = __builtin_inff(), arg[q] = 0;
    for (int c0 = 0; c0 < k_pad; c0 += 32) {
        const int cc = c0 + col;
        f32x16_t acc = {};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], c8[(int64_t)cc * (KP / 8) + 2 * ks + half], acc, 0, 0, 0);
        const float nc = cn[cc];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const float est = nc - 2.0f * acc[q];
            if (est < best[q]) best[q] = est, arg[q] = cc;
        }
    }
    // accumulator row of register q: 8 (q / 4) + 4 half + q % 4; the smallest
    // over the half's 32 lanes (columns), lowest index on ties
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        float b = best[q];
        int ix = arg[q];
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) {
            const float ob = __shfl_xor(b, o, 64);
            const int oi = __shfl_xor(ix, o, 64);
            if (ob < b || (ob == b && oi < ix)) b = ob, ix = oi;
        }
        const int64_t row = row0 + 8 * (q >> 2) + 4 * half + (q & 3);
        if (col == 0 && row < n) {
            keys[row] = (uint64_t)ix;
            vals[row] = (uint32_t)row;
        }
    }
}

// Sorted float32 rows, centred bf16 hi/lo rows padded to KP, |a'|^2 in
// float32 (ascending k) and |a'|.  One thread per row, held in registers:
// the bf16 rows are written 16 bytes at a time, and with vec (d % 4 == 0
// and x 16-byte aligned) so are the float32 rows.
// One row's centred coordinates a = v - mu (float32, zero past d) as bf16
// hi = bf16(a) and lo = bf16(a - hi), its |a|^2 (float32 products and sums in
// order) and |a| (that norm's root, rounded up by 1e-6).
template <int KP>
__device__ __forceinline__ void split_row(const float *v, const float *__restrict__ mu, int d, __bf16 *hi,
                                          __bf16 *lo, float &na, float &rho) {
    float s = 0.0f;
    bf16x8_t h8[KP / 8], l8[KP / 8];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const float a = k < d ? v[k] - mu[k] : 0.0f;
        const __bf16 h = to_bf16(a);
        h8[k / 8][k % 8] = h;
        l8[k / 8][k % 8] = to_bf16(a - (float)h);
        s = __fadd_rn(s, __fmul_rn(a, a));
    }
    bf16x8_t *hd = reinterpret_cast<bf16x8_t *>(hi), *ld = reinterpret_cast<bf16x8_t *>(lo);
#pragma unroll
    for (int u = 0; u < KP / 8; ++u) {
        hd[u] = h8[u];
        ld[u] = l8[u];
    }
    na = s;
    rho = sqrtf(s) * (1.0f + 1e-6f);
}

template <int KP>
__global__ __launch_bounds__(256) void k_mfma_prep(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                   int64_t n, int d, bool vec, const float *__restrict__ mu,
                                                   float *__restrict__ xs, __bf16 *__restrict__ hi,
                                                   __bf16 *__restrict__ lo, float *__restrict__ na,
                                                   float *__restrict__ rho) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    float v[KP];
    if (vec) {
        const float4 *src = reinterpret_cast<const float4 *>(x + p * d);
        float4 *dst = reinterpret_cast<float4 *>(xs + i * d);
#pragma unroll
        for (int k4 = 0; k4 < KP / 4; ++k4) {
            float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (4 * k4 < d) {
                q = src[k4];
                dst[k4] = q;
            }
            v[4 * k4] = q.x, v[4 * k4 + 1] = q.y, v[4 * k4 + 2] = q.z, v[4 * k4 + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            v[k] = 0.0f;
            if (k < d) {
                v[k] = x[p * d + k];
                xs[i * d + k] = v[k];
            }
        }
    }
    split_row<KP>(v, mu, d, hi + i * KP, lo + i * KP, na[i], rho[i]);
}

// Per cell centroid: its centred bf16 hi / lo halves, |m'|^2 and |m'| as the
// points' (k_mfma_prep), for the matrix-core tile bounds (k_tile_dist).
template <int KP>
__global__ __launch_bounds__(256) void k_centroid_prep(const float *__restrict__ mc, int64_t nc, int d,
                                                       const float *__restrict__ mu, __bf16 *__restrict__ hi,
                                                       __bf16 *__restrict__ lo, float *__restrict__ na,
                                                       float *__restrict__ rho) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    float v[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) v[k] = k < d ? mc[c * d + k] : 0.0f;
    split_row<KP>(v, mu, d, hi + c * KP, lo + c * KP, na[c], rho[c]);
}

// Square distance of row p to row m, summed by a 16-lane group (gl: the
// lane in the group): each lane takes 4 consecutive coordinates in turn
// (16-byte loads; vec: d % 4 == 0 and both rows 16-byte aligned), or every
// 16th coordinate, then the lanes' sums are added over the group in a fixed
// tree order.  Every lane of the group must call it; each gets the total.
__device__ __forceinline__ double grp_sqdist64(const float *p, const float *m, int d, bool vec, int gl) {
    double s = 0.0;
    if (vec) {
        for (int k = 4 * gl; k < d; k += 64) {
            const float4 a = *reinterpret_cast<const float4 *>(p + k), b = *reinterpret_cast<const float4 *>(m + k);
            double t = (double)a.x - (double)b.x;
            s += t * t;
            t = (double)a.y - (double)b.y;
            s += t * t;
            t = (double)a.z - (double)b.z;
            s += t * t;
            t = (double)a.w - (double)b.w;
            s += t * t;
        }
    } else {
        for (int k = gl; k < d; k += 16) {
            const double t = (double)p[k] - (double)m[k];
            s += t * t;
        }
    }
    for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 16);
    return s;
}
// The same in float32 (differences, products and sums rounded to float).
__device__ __forceinline__ float grp_sqdist32(const float *p, const float *m, int d, bool vec, int gl) {
    float s = 0.0f;
    if (vec) {
        for (int k = 4 * gl; k < d; k += 64) {
            const float4 a = *reinterpret_cast<const float4 *>(p + k), b = *reinterpret_cast<const float4 *>(m + k);
            float t = a.x - b.x;
            s += t * t;
            t = a.y - b.y;
            s += t * t;
            t = a.z - b.z;
            s += t * t;
            t = a.w - b.w;
            s += t * t;
        }
    } else {
        for (int k = gl; k < d; k += 16) {
            const float t = p[k] - m[k];
            s += t * t;
        }
    }
    for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 16);
    return s;
}

// Per cell: radius (float64, max member distance to the centre point) and
// the largest |a'| of its members.
__global__ __launch_bounds__(256) void k_cell_radius(const float *__restrict__ xs, const float *__restrict__ rho,
                                                     const int32_t *__restrict__ cell_start,
                                                     const uint64_t *__restrict__ cell_key, int64_t nc, int d,
                                                     const float *__restrict__ cent, double *__restrict__ rad,
                                                     float *__restrict__ rmax) {
    const int64_t c = (int64_t)blockIdx.x;
    if (c >= nc) return;
    const int32_t b = cell_start[c], e = cell_start[c + 1];
    const float *cc = cent + (int64_t)cell_key[c] * d;
    double r = 0.0;
    float m = 0.0f;
    const bool v4 = (d & 3) == 0;  // rows and the centroid 16-byte aligned
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        double s2 = 0.0;
        if (v4) {
            // 32 coordinates of the row per round trip, summed in ascending k
            const float4 *x4 = reinterpret_cast<const float4 *>(xs + (int64_t)i * d);
            const float4 *c4 = reinterpret_cast<const float4 *>(cc);
            const int n4 = d >> 2;
            for (int k0 = 0; k0 < n4; k0 += 8) {
                float4 a[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4) a[j] = x4[k0 + j];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4) {
                        const float4 cv = c4[k0 + j];
                        double df = (double)a[j].x - (double)cv.x;
                        s2 += df * df;
                        df = (double)a[j].y - (double)cv.y;
                        s2 += df * df;
                        df = (double)a[j].z - (double)cv.z;
                        s2 += df * df;
                        df = (double)a[j].w - (double)cv.w;
                        s2 += df * df;
                    }
            }
        } else {
            for (int k = 0; k < d; ++k) {
                const double df = (double)xs[(int64_t)i * d + k] - (double)cc[k];
                s2 += df * df;
            }
        }
        r = fmax(r, s2);
        m = fmaxf(m, rho[i]);
    }
    __shared__ double sr[256];
    __shared__ float sm[256];
    sr[threadIdx.x] = r;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rad[c] = sqrt(sr[0]) * (1.0 + 1e-12);
        rmax[c] = sm[0];
    }
}

// Each unordered pair of candidate cells is computed once: cell c takes
// candidate o when o == c, or when (c < o) equals (c + o even) -- about half
// of every cell's candidates, whatever the cell order.  A pair (c, o != c)
// adds its decisions to the rows' counts and to the columns' counts; the
// diagonal pair (c, c) adds to the rows only (every row of c meets every
// column of c, itself included).
__device__ __forceinline__ bool mf_takes(int c, int o) { return o == c || ((c < o) == (((c + o) & 1) == 0)); }

__device__ __forceinline__ int32_t mf_ld(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ int32_t mf_find(int32_t *parent, int32_t x) {
    for (;;) {
        const int32_t p = mf_ld(parent + x);
        if (p == x) return x;
        const int32_t gp = mf_ld(parent + p);
        if (gp != p) atomicCAS(parent + x, p, gp);
        x = gp;
    }
}
// Returns the root of the united set as this call left it (a later union
// may hang it lower); a may be any member of its set, such as a root
// returned before, which shortens its find.
__device__ int32_t mf_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = mf_find(parent, a);
        b = mf_find(parent, b);
        if (a == b) return a;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + b, b, a) == b) return a;
    }
}

// ------------------------------------------------------ block pruning --
// Per cell: the centroid of its points (float64 sums, rounded to float32);
// the float32 centroid is the reference point of every radius and distance
// bound of the block pruning.
// One workgroup per cell: P = 1024 / d point streams (point b + s + P j to
// stream s), each summing its points per coordinate, then the streams
// summed in order (a fixed order: the result is deterministic).
__global__ __launch_bounds__(1024) void k_cell_centroid(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                        const int32_t *__restrict__ cell_start, int64_t nc, int d,
                                                        float *__restrict__ mc) {
    __shared__ double part[1024];
    const int64_t c = blockIdx.x;
    if (c >= nc) return;
    const int32_t b = cell_start[c], e = cell_start[c + 1];
    for (int k0 = 0; k0 < d; k0 += 1024) {
        const int dk = d - k0 < 1024 ? d - k0 : 1024;  // coordinates of this round
        const int P = 1024 / dk;
        const int s = threadIdx.x / dk, k = k0 + (threadIdx.x - s * dk);
        double acc = 0.0;
        if (s < P) {
            int32_t i = b + s;
            double a1 = 0.0;
            for (; i + P < e; i += 2 * P) {
                acc += (double)x[(int64_t)perm[i] * d + k];
                a1 += (double)x[(int64_t)perm[i + P] * d + k];
            }
            if (i < e) acc += (double)x[(int64_t)perm[i] * d + k];
            acc += a1;
        }
        part[threadIdx.x] = acc;
        __syncthreads();
        if (threadIdx.x < dk) {
            double t = 0.0;
            for (int q = 0; q < P; ++q) t += part[q * dk + threadIdx.x];
            mc[c * d + k0 + threadIdx.x] = e > b ? (float)(t / (double)(e - b)) : 0.0f;
        }
        __syncthreads();
    }
}

// Sort keys of the order within cells: (centre index, distance to the cell
// centroid as float bits, which order like the distances).
__global__ __launch_bounds__(256) void k_radius_keys(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                     const uint64_t *__restrict__ keys,
                                                     const int32_t *__restrict__ cell_of,
                                                     const float *__restrict__ mc, int64_t n, int d, bool vec,
                                                     uint64_t *__restrict__ keys2, uint32_t *__restrict__ vals2) {
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;  // one 16-lane group per point
    const int gl = threadIdx.x & 15;
    const int64_t i = i0 < n ? i0 : n - 1;
    const uint32_t p = perm[i];
    const float s = grp_sqdist32(x + (int64_t)p * d, mc + (int64_t)cell_of[i] * d, d, vec, gl);
    if (gl == 0 && i0 < n) {
        keys2[i] = keys[i] << 32 | (uint64_t)__float_as_uint(sqrtf(s));
        vals2[i] = p;
    }
}

// Per sorted point: an upper bound of its distance to its cell's centroid
// (float64, rounded up to float32).
__global__ __launch_bounds__(256) void k_point_radius(const float *__restrict__ xs, const int32_t *__restrict__ cell_of,
                                                      const float *__restrict__ mc, int64_t n, int d, bool vec,
                                                      float *__restrict__ rad) {
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;  // one 16-lane group per point
    const int gl = threadIdx.x & 15;
    const int64_t i = i0 < n ? i0 : n - 1;
    const double s = grp_sqdist64(xs + i * d, mc + (int64_t)cell_of[i] * d, d, vec, gl);
    if (gl != 0 || i0 >= n) return;
    const double r = sqrt(s) * (1.0 + 1e-12);
    float f = (float)r;
    if ((double)f < r) f = __uint_as_float(__float_as_uint(f) + 1u);  // the next float up (f >= 0, finite)
    rad[i] = f;
}

// Per row tile t (up to 32 sorted points of one cell) and cell C: tab[t][C],
// a lower bound of the smallest distance of t's points to C's centroid;
// rad_t[t], the largest radius of t's points, and rho_t[t], their largest
// centred norm |a'| (the guard band's R for a block of t's columns).  One
// wave per tile, 32 centroids per step: est = (|a'|^2 + |m'|^2) - 2 a'.m',
// the product from bf16 hi / lo halves (hi*hi + hi*lo + lo*hi, float32
// accumulation).  With R >= |a'|, |m'|:
//  - centring (a' = fl(a - mu), m' likewise) moves the distance by at most
//    2 u R (u = 2^-24), its square by at most 8.01 u R^2 (distances <= 2R);
//  - each float32 norm is within (d + 1) u R^2, the product's accumulation
//    within 3.03 Kp u R^2 and its dropped terms (lo*lo and the lo halves'
//    own rounding, |lo - (a - hi)| <= 2^-16 |a|) within 3.03 2^-16 R^2, both
//    doubled in est, and est's two roundings within 6.1 u R^2;
// so |est - |a - m|^2| <= (2d + 6.06 Kp + 16.3) u R^2 + 6.06 2^-16 R^2, below
// 1.6e-4 R^2 for d, Kp <= 128, and est - 2^-12 R^2 (1.0001) is a lower bound
// of the square distance; its root times (1 - 2^-20), of the distance.
template <int KP>
__global__ __launch_bounds__(256) void k_tile_dist(const __bf16 *__restrict__ hi, const __bf16 *__restrict__ lo,
                                                   const float *__restrict__ na, const float *__restrict__ rho,
                                                   const int4 *__restrict__ tiles, int64_t nt,
                                                   const __bf16 *__restrict__ chi, const __bf16 *__restrict__ clo,
                                                   const float *__restrict__ cna, const float *__restrict__ crho,
                                                   int64_t nc, const float *__restrict__ rad,
                                                   float *__restrict__ tab, float *__restrict__ rad_t,
                                                   float *__restrict__ rho_t) {
    constexpr int KS = KP / 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + wave;
    if (t >= nt) return;
    const int half = lane >> 5, col = lane & 31;
    const int4 tile = tiles[t];
    const int rows = tile.z;
    const bf16x8_t *hi8 = reinterpret_cast<const bf16x8_t *>(hi), *lo8 = reinterpret_cast<const bf16x8_t *>(lo);
    const bf16x8_t *chi8 = reinterpret_cast<const bf16x8_t *>(chi), *clo8 = reinterpret_cast<const bf16x8_t *>(clo);
    const bool arow_ok = col < rows;
    const int64_t arow = tile.y + (arow_ok ? col : 0);
    bf16x8_t ahi[KS], alo[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        ahi[ks] = arow_ok ? hi8[arow * (KP / 8) + ks * 2 + half] : bf16x8_t{};
        alo[ks] = arow_ok ? lo8[arow * (KP / 8) + ks * 2 + half] : bf16x8_t{};
    }
    float m = arow_ok && half == 0 ? rad[arow] : 0.0f;
    float h = arow_ok ? rho[arow] : 0.0f;
    for (int o = 32; o >= 1; o >>= 1) {
        m = fmaxf(m, __shfl_xor(m, o, 64));
        h = fmaxf(h, __shfl_xor(h, o, 64));
    }
    if (lane == 0) {
        rad_t[t] = m;
        rho_t[t] = h;
    }
    // rows held in the accumulator: r -> tile row 8 (r/4) + 4 half + (r%4)
    float na_r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
        na_r[r] = tr < rows ? na[tile.y + tr] : __builtin_inff();
    }
    for (int64_t cb = 0; cb < nc; cb += 32) {
        const int64_t C = cb + col < nc ? cb + col : nc - 1;
        f32x16_t acc = {};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const bf16x8_t bh = chi8[C * (KP / 8) + ks * 2 + half];
            const bf16x8_t bl = clo8[C * (KP / 8) + ks * 2 + half];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[ks], bh, acc, 0, 0, 0);
        }
        const float nb = cna[C];
        float e = __builtin_inff();
#pragma unroll
        for (int r = 0; r < 16; ++r) e = fminf(e, __fsub_rn(__fadd_rn(na_r[r], nb), 2.0f * acc[r]));
        e = fminf(e, __shfl_xor(e, 32, 64));
        const float R = fmaxf(h, crho[C]);
        const float lb = e - 2.44140625e-4f * (R * R) * 1.0001f;
        if (half == 0 && cb + col < nc)
            tab[t * nc + cb + col] = lb > 0.0f ? sqrtf(lb) * (1.0f - 9.5367431640625e-07f) : 0.0f;
    }
}

// tab (tiles x cells) to cell-major (cells x tiles), 32 x 32 tiles via LDS.
__global__ __launch_bounds__(256) void k_transpose_f32(const float *__restrict__ in, int64_t rows, int64_t cols,
                                                       float *__restrict__ out) {
    __shared__ float tl[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int y = ty; y < 32; y += 8)
        if (r0 + y < rows && c0 + tx < cols) tl[y][tx] = in[(r0 + y) * cols + c0 + tx];
    __syncthreads();
    for (int y = ty; y < 32; y += 8)
        if (c0 + y < cols && r0 + tx < rows) out[(c0 + y) * rows + r0 + tx] = tl[tx][y];
}

// Per count group g (row tiles [grp.y, grp.y + grp.z) of cell A) and
// candidate entry i of A that A takes (candidate cell B), column tile j of B
// (j < 64) is
//  - pruned when no point of it is within r of a row of the group:
//    max(T[G][B] - rad(j), T[j][A] - rad(G)) > r, with T[G][B] the smallest
//    of the group's tiles' bounds and rad(G) their largest radius;
//  - wholly within eps when every pair is a neighbour: the pair distance is
//    at most U = rad(G) + |m_A - m_B| + rad(j), and the reference's float32
//    sum is within (d + 3) 2^-24 of the exact square distance, so
//    U^2 (1 + 1.01 (d + 3) 2^-24) <= e2 makes the predicate accept them all.
//    Such a tile is counted here: each row of the group gains its columns,
//    each of its columns the group's rows (not in A's own pair, where the
//    rows' counts take every pair), through per-tile sums (tacc) that
//    k_tile_acc adds to every point of the tile; the pair is flagged for the
//    label passes;
//  - otherwise computed by the count pass: bit j of masks[moff[g] + i].
// Bit j of amasks[moff[g] + i] marks the tiles wholly within eps (the union
// of their cores needs no products either, k_allin_union).  Entries A does
// not take are 0 in both.  Adds the point pairs of the computed
// tiles (rows x columns) to *pairs.  tcm: the bounds cell-major; dab_e: per
// candidate entry the upper bound of |m_A - m_B| (k_pair_cdist).
__global__ __launch_bounds__(256) void k_block_masks(const int4 *__restrict__ groups, const int32_t *__restrict__ cand_off,
                                                     const int32_t *__restrict__ cand,
                                                     const int32_t *__restrict__ toff, const int4 *__restrict__ tiles,
                                                     const float *__restrict__ tcm, const float *__restrict__ rad_t,
                                                     int64_t nt, const double *__restrict__ dab_e, int d, double r,
                                                     double e2_in, const int32_t *__restrict__ moff,
                                                     uint64_t *__restrict__ masks, uint64_t *__restrict__ amasks,
                                                     int32_t *__restrict__ tacc,
                                                     uint8_t *__restrict__ pflag,
                                                     unsigned long long *__restrict__ pairs) {
    __shared__ int32_t s_rowadd;
    const int64_t g = blockIdx.x;
    const int4 grp = groups[g];
    const int A = grp.x;
    if (threadIdx.x == 0) s_rowadd = 0;
    float rg = 0.0f;
    int rows = 0;
    for (int u = 0; u < grp.z; ++u) {
        rg = fmaxf(rg, rad_t[grp.y + u]);
        rows += tiles[grp.y + u].z;
    }
    __syncthreads();
    const int32_t cb = cand_off[A], ce = cand_off[A + 1];
    const float *tA = tcm + (int64_t)A * nt;  // tiles to A's centroid
    unsigned long long kept = 0;
    int32_t rowadd = 0;
    for (int32_t i = cb + threadIdx.x; i < ce; i += blockDim.x) {
        const int B = cand[i];
        uint64_t m = 0, am = 0;
        if (mf_takes(A, B)) {
            const float *tB = tcm + (int64_t)B * nt;  // tiles to B's centroid
            float tg = __builtin_inff();
            for (int u = 0; u < grp.z; ++u) tg = fminf(tg, tB[grp.y + u]);
            const double dab = dab_e[i];  // above |m_A - m_B| (k_pair_cdist)
            const int32_t j0 = toff[B], j1 = toff[B + 1];
            const int nj = j1 - j0 < 64 ? j1 - j0 : 64;
            bool any_in = false;
            for (int j = 0; j < nj; ++j) {
                const double rj = (double)rad_t[j0 + j];
                const double lb = fmax((double)tg - rj, (double)tA[j0 + j] - (double)rg);
                if (lb > r) continue;  // pruned
                const double ub = (double)rg + dab + rj;
                const int cols = tiles[j0 + j].z;
                if (ub * ub * (1.0 + 1.01 * (d + 3) * 5.9604644775390625e-08) * (1.0 + 1e-12) <= e2_in) {
                    any_in = true;
                    am |= 1ull << j;
                    rowadd += cols;
                    if (B != A) atomicAdd(tacc + j0 + j, rows);
                } else {
                    m |= 1ull << j;
                    kept += (unsigned long long)cols;
                }
            }
            for (int j = j0 + 64; j < j1; ++j) kept += (unsigned long long)tiles[j].z;
            if (any_in) pflag[i] = 1;
        }
        masks[moff[g] + (i - cb)] = m;
        amasks[moff[g] + (i - cb)] = am;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        kept += __shfl_xor(kept, o, 64);
        rowadd += __shfl_xor(rowadd, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (kept) atomicAdd(pairs, kept * (unsigned long long)rows);
        if (rowadd) atomicAdd(&s_rowadd, rowadd);
    }
    __syncthreads();
    if (threadIdx.x < grp.z && s_rowadd) atomicAdd(tacc + grp.y + threadIdx.x, s_rowadd);
}

// Per candidate entry i of cell A (candidate cell B = cand[i]): an upper
// bound of the centroids' distance |m_A - m_B| (float64 sum of squares in
// ascending k, root times 1 + 1e-12), for k_block_masks.  One workgroup per
// cell, threads over its entries.
__global__ __launch_bounds__(256) void k_pair_cdist(const int32_t *__restrict__ cand_off,
                                                    const int32_t *__restrict__ cand, int64_t nc,
                                                    const float *__restrict__ mc, int d, double *__restrict__ dab) {
    const int64_t A = blockIdx.x;
    if (A >= nc) return;
    const int32_t cb = cand_off[A], ce = cand_off[A + 1];
    const float *mA = mc + A * d;
    for (int32_t i = cb + threadIdx.x; i < ce; i += blockDim.x) {
        const float *mB = mc + (int64_t)cand[i] * d;
        double dd = 0.0;
        if ((d & 3) == 0) {  // 32 coordinates of B per round trip, ascending k
            const float4 *a4 = reinterpret_cast<const float4 *>(mA), *b4 = reinterpret_cast<const float4 *>(mB);
            const int n4 = d >> 2;
            for (int k0 = 0; k0 < n4; k0 += 8) {
                float4 bv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4) bv[j] = b4[k0 + j];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (k0 + j < n4) {
                        const float4 av = a4[k0 + j];
                        double df = (double)av.x - (double)bv[j].x;
                        dd += df * df;
                        df = (double)av.y - (double)bv[j].y;
                        dd += df * df;
                        df = (double)av.z - (double)bv[j].z;
                        dd += df * df;
                        df = (double)av.w - (double)bv[j].w;
                        dd += df * df;
                    }
            }
        } else {
            for (int k = 0; k < d; ++k) {
                const double df = (double)mA[k] - (double)mB[k];
                dd += df * df;
            }
        }
        dab[i] = sqrt(dd) * (1.0 + 1e-12);
    }
}

// Per sorted point: its row tile's sum from k_block_masks added to its count.
__global__ __launch_bounds__(256) void k_tile_acc(const int32_t *__restrict__ tacc, const int32_t *__restrict__ cell_of,
                                                  const int32_t *__restrict__ cell_start,
                                                  const int32_t *__restrict__ toff, int64_t n,
                                                  int32_t *__restrict__ nn_s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_of[i];
    const int32_t v = tacc[toff[c] + (int32_t)((i - cell_start[c]) >> 5)];
    if (v) nn_s[i] += v;
}

// Per row tile: its first core (sorted index), or -1; its flag cleared.
__global__ __launch_bounds__(256) void k_tile_rep(const int4 *__restrict__ tiles, int64_t nt,
                                                  const uint8_t *__restrict__ core_s, int32_t *__restrict__ rep,
                                                  int32_t *__restrict__ tflag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int4 tile = tiles[t];
    int32_t r = -1;
    for (int32_t i = tile.y; i < tile.y + tile.z; ++i)
        if (core_s[i]) {
            r = i;
            break;
        }
    rep[t] = r;
    tflag[t] = 0;
}

// The union of the tiles k_block_masks found wholly within eps: per count
// group g and candidate entry, every such column tile J with a core has all
// its cores within eps of every core of the group's tiles, so all of them
// are one set.  The group's first core (r0) is united with J's first core
// and, once per group, with its other tiles' first cores; the tiles are
// flagged (k_tile_unite then joins each of their cores to its tile's
// first).
__global__ __launch_bounds__(256) void k_allin_union(const int4 *__restrict__ groups, const int32_t *__restrict__ cand_off,
                                                     const int32_t *__restrict__ cand, const int32_t *__restrict__ toff,
                                                     const int32_t *__restrict__ moff,
                                                     const uint64_t *__restrict__ amasks,
                                                     const int32_t *__restrict__ rep, int32_t *__restrict__ tflag,
                                                     int32_t *__restrict__ parent) {
    __shared__ int32_t s_any;
    const int64_t g = blockIdx.x;
    const int4 grp = groups[g];
    int32_t r0 = -1;  // the group's first core
    for (int u = 0; u < grp.z && r0 < 0; ++u) r0 = rep[grp.y + u];
    if (r0 < 0) return;  // no core in the group: nothing to unite (the same for every thread)
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    const int32_t cb = cand_off[grp.x], ce = cand_off[grp.x + 1];
    bool any = false;
    int32_t ra = r0;  // a member of the group's set: the last root seen
    for (int32_t i = cb + threadIdx.x; i < ce; i += blockDim.x) {
        uint64_t m = amasks[moff[g] + (i - cb)];
        if (!m) continue;
        const int32_t j0 = toff[cand[i]];
        while (m) {
            const int32_t J = j0 + __builtin_ctzll(m);
            m &= m - 1;
            const int32_t rj = rep[J];
            if (rj < 0) continue;
            ra = mf_unite(parent, ra, rj);
            tflag[J] = 1;
            any = true;
        }
    }
    if (any) s_any = 1;
    __syncthreads();
    if (s_any && threadIdx.x < grp.z) {
        const int32_t rt = rep[grp.y + threadIdx.x];
        if (rt >= 0) {
            mf_unite(parent, r0, rt);
            tflag[grp.y + threadIdx.x] = 1;
        }
    }
}

// Per sorted point: a core of a flagged tile joins its tile's first core.
// Only first cores (and their roots) have been united so far, so every
// other core is still its own root, and the first core has the smaller
// index: it is hung under it directly (the union-find's roots are the
// smallest index of their set).
__global__ __launch_bounds__(256) void k_tile_unite(const uint8_t *__restrict__ core_s,
                                                    const int32_t *__restrict__ cell_of,
                                                    const int32_t *__restrict__ cell_start,
                                                    const int32_t *__restrict__ toff, int64_t n,
                                                    const int32_t *__restrict__ rep,
                                                    const int32_t *__restrict__ tflag, int32_t *__restrict__ parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !core_s[i]) return;
    const int c = cell_of[i];
    const int32_t t = toff[c] + (int32_t)((i - cell_start[c]) >> 5);
    if (tflag[t] && rep[t] != (int32_t)i) parent[i] = rep[t];
}

// Per count group: its cell's candidate entries (mask words).
__global__ __launch_bounds__(256) void k_group_entries(const int4 *__restrict__ groups, int64_t ng,
                                                       const int32_t *__restrict__ cand_off,
                                                       int32_t *__restrict__ cnt) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) cnt[g] = cand_off[groups[g].x + 1] - cand_off[groups[g].x];
}

// Candidate cell pairs: |centre_C - centre_C'| <= R_C + R_C' + r (+ slack),
// a symmetric relation (the same float64 sums from either side).  Pass 0
// counts per cell; pass 1 writes each cell's list in ascending cell order
// (a block scan per 256 candidates), so the entry of C in the list of C' can
// be found by binary search.  Pass 1 also adds |C| * (points of C's
// candidate cells) to *pts: the point pairs the filter covers.
__global__ __launch_bounds__(256) void k_cell_pairs(const float *__restrict__ cent, const uint64_t *__restrict__ cell_key,
                                                    const double *__restrict__ rad, int64_t nc, int d, double r,
                                                    const int32_t *__restrict__ off, int32_t *__restrict__ cnt_or_list,
                                                    int pass, const int32_t *__restrict__ cell_start,
                                                    unsigned long long *__restrict__ pts) {
    const int64_t c = (int64_t)blockIdx.x;
    if (c >= nc) return;
    const float *a = cent + (int64_t)cell_key[c] * d;
    __shared__ int32_t wsum[4];
    __shared__ unsigned long long cand_pts;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) cand_pts = 0;
    int32_t base = 0;  // entries written so far (pass 1) or counted (pass 0)
    unsigned long long mine = 0;
    for (int64_t o0 = 0; o0 < nc; o0 += blockDim.x) {
        const int64_t o = o0 + threadIdx.x;
        bool in = false;
        if (o < nc) {
            const float *b = cent + (int64_t)cell_key[o] * d;
            double s = 0.0;
            for (int k = 0; k < d; ++k) {
                const double df = (double)a[k] - (double)b[k];
                s += df * df;
            }
            const double lim = rad[c] + rad[o] + r;
            in = s <= lim * lim * (1.0 + 1e-9);
        }
        const uint64_t m = __ballot(in);
        __syncthreads();  // wsum of the previous round has been read
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int32_t before = base;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const int32_t pos = before + (int32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (in && pass == 1) {
            cnt_or_list[off[c] + pos] = (int32_t)o;
            mine += (unsigned long long)(cell_start[o + 1] - cell_start[o]);
        }
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    }
    if (pass == 1 && mine) atomicAdd(&cand_pts, mine);
    __syncthreads();
    if (pass == 0 && threadIdx.x == 0) cnt_or_list[c] = base;
    if (pass == 1 && threadIdx.x == 0 && cand_pts)
        atomicAdd(pts, cand_pts * (unsigned long long)(cell_start[c + 1] - cell_start[c]));
}

// Row tiles: (cell, first row, rows) with <= 32 rows of one cell.
__global__ __launch_bounds__(256) void k_row_tiles(const int32_t *__restrict__ cell_start,
                                                   const int32_t *__restrict__ toff, int64_t nc,
                                                   int4 *__restrict__ tiles) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t b = cell_start[c], e = cell_start[c + 1];
    int32_t t = toff[c];
    for (int32_t q = b; q < e; q += 32) tiles[t++] = make_int4((int32_t)c, q, e - q < 32 ? e - q : 32, 0);
}

__global__ __launch_bounds__(256) void k_row_tile_count(const int32_t *__restrict__ cell_start, int64_t nc,
                                                        int32_t *__restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) cnt[c] = (cell_start[c + 1] - cell_start[c] + 31) / 32;
}

// ------------------------------------------------------------ MFMA pass --
// The reference's predicate on original coordinates (data_process.cpp:443-447).
__device__ __forceinline__ bool exact_pred(const float *__restrict__ p, const float *__restrict__ q, int d, float e2) {
    float len = 0.0f;
    for (int k = 0; k < d; ++k) {
        const float df = p[k] - q[k];
        len = __fadd_rn(len, __fmul_rn(df, df));
    }
    return len <= e2;
}

// The same predicate, coordinates read 16 bytes at a time (d % 4 == 0, rows
// 16-byte aligned), 32 of each row per round trip: the loads of a chunk are
// all issued before its sum, which is still taken in ascending k.
__device__ __forceinline__ bool exact_pred4(const float *__restrict__ p, const float *__restrict__ q, int d,
                                            float e2) {
    const float4 *p4 = reinterpret_cast<const float4 *>(p), *q4 = reinterpret_cast<const float4 *>(q);
    const int n4 = d >> 2;
    float len = 0.0f;
    for (int k0 = 0; k0 < n4; k0 += 8) {
        float4 a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k0 + j < n4) {
                a[j] = p4[k0 + j];
                b[j] = q4[k0 + j];
            }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k0 + j < n4) {
                float df = a[j].x - b[j].x;
                len = __fadd_rn(len, __fmul_rn(df, df));
                df = a[j].y - b[j].y;
                len = __fadd_rn(len, __fmul_rn(df, df));
                df = a[j].z - b[j].z;
                len = __fadd_rn(len, __fmul_rn(df, df));
                df = a[j].w - b[j].w;
                len = __fadd_rn(len, __fmul_rn(df, df));
            }
    }
    return len <= e2;
}

struct MfmaArgs {
    const int4 *tiles;
    int64_t n_tiles;
    const int4 *groups;
    int64_t n_groups;
    const int32_t *rows;  // tile row -> point (null: identity)
    const int32_t *cand_off, *cand, *cell_start;
    const uint64_t *bmask;     // count pass: kept column tiles per (group, candidate entry), or null
    const int32_t *bmask_off;  // per group: its first mask word
    uint64_t *emask;           // count pass: per mask word, the computed tiles holding a pair within eps or a
                               // guard-band element (zero on entry; the union passes walk these)
    const float *tile_rho;     // count pass: per row tile its largest |a'| (null: per cell, cell_rmax)
    const int32_t *ctoff;      // cell -> its first row tile
    const float2 *tile_band;   // count pass: per column tile (hb, hskip) of its R (below)
    const __bf16 *hi, *lo;
    const float *na, *rho, *cell_rmax, *xs;
    int64_t n;
    int d;
    float e2;
    double b1, b2, b3;  // band = 2 (b1 R^2 + b2 R + b3)
    int32_t *nn_s;                    // pass 0
    const uint8_t *core_s;            // pass 1, 2
    int32_t *parent;                  // pass 1
    const int32_t *label_s;           // pass 2
    int32_t *best_s;                  // pass 2
    uint8_t *pflag;                   // per candidate entry: the pair holds a pair within eps
    unsigned long long *audit;        // [band pairs, fp64 flips, filter flips], pass 3
    float *audit_err;                 // max |est - fp64 len|, pass 3
    int audit_stride, audit_offset;  // pass 3: groups audit_offset + audit_stride j
    int union_phase;          // pass 1: 0 the cells' own pairs, 1 the other flagged pairs
    const int32_t *cell_rep;  // pass 1, phase 1: a core of the cell when all its cores share a root, else -1
    int4 *bq;          // count pass: guard-band pairs (row, column, candidate entry, own pair)
    unsigned long long *bq_count;
    int64_t bq_cap;
};

// Work groups: up to `per` consecutive row tiles of one cell, one workgroup
// (one wave per tile: 4 for the label passes, 8 for the count pass).  The
// waves share every 32-column B block through LDS.
__global__ __launch_bounds__(256) void k_groups(const int32_t *__restrict__ toff, int64_t nc,
                                                const int32_t *__restrict__ goff, int per,
                                                int4 *__restrict__ groups) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t t0 = toff[c], t1 = toff[c + 1];
    int32_t g = goff[c];
    for (int32_t t = t0; t < t1; t += per) groups[g++] = make_int4((int32_t)c, t, t1 - t < per ? t1 - t : per, 0);
}

__global__ __launch_bounds__(256) void k_group_count(const int32_t *__restrict__ toff, int64_t nc, int per,
                                                     int32_t *__restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) cnt[c] = (toff[c + 1] - toff[c] + per - 1) / per;
}


// PASS 0: neighbour counts.  PASS 1: unite core pairs, over the cell pairs
// this cell takes whose count pass found a pair within eps (col > row in
// the cell's own pair).  PASS 2: smallest core cluster id within eps for
// non-core rows, over the candidate cells whose pair was flagged (the entry
// of the side that takes it).  PASS 3: audit of every estimate of every
// audit_stride-th group against float32 and float64.
// Rows and columns that take no part are masked by an infinite norm, so the
// estimate never passes; the few estimates inside the guard band are decided
// afterwards with the exact predicate, under a ballot.
// Candidate entries per cell whose same-set state the union's second phase
// takes at workgroup start (later entries are tested per block only).
constexpr int kSameSetCap = 4096;

template <int KP, int PASS>
__global__ __launch_bounds__(256, (PASS == 1 && KP < 128) ? 3 : 1) void k_mfma_pass(MfmaArgs A) {
    constexpr int KS = KP / 16;
    constexpr int UNITS = 32 * (KP / 8);  // bf16x8 units of one 32-column block, per hi / lo
    constexpr int ROWU = KP / 8 + 1;      // LDS row stride in units: one pad unit breaks bank conflicts
    __shared__ bf16x8_t sb_hi[2][32 * ROWU], sb_lo[2][32 * ROWU];  // double buffer
    __shared__ float sb_na[2][32];
    __shared__ int32_t sb_aux[2][32];  // PASS 1: column root; PASS 2: column cluster id
    __shared__ int32_t any_work;
    __shared__ uint32_t ss_bits[kSameSetCap / 32];
    __shared__ int32_t ss_pub[2];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t gidx = blockIdx.x;
    if (PASS == 3) gidx = gidx * A.audit_stride + A.audit_offset;
    if (gidx >= A.n_groups) return;
    const int4 grp = A.groups[gidx];
    const bool wave_on = wave < grp.z;
    const int4 tile = A.tiles[grp.y + (wave_on ? wave : 0)];
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int cell = grp.x, row0 = tile.y, nrows = wave_on ? tile.z : 0;
    const bf16x8_t *hi8 = reinterpret_cast<const bf16x8_t *>(A.hi);
    const bf16x8_t *lo8 = reinterpret_cast<const bf16x8_t *>(A.lo);
    const float INF = __builtin_inff();
    // tile row t is point rows[row0 + t] (or row0 + t)
    auto row_point = [&](int t) -> int64_t { return A.rows ? (int64_t)A.rows[row0 + t] : (int64_t)(row0 + t); };
    // A operand: row `col` of the tile, k elements [16 ks + 8 half, +8)
    const bool arow_ok = col < nrows;
    const int64_t arow = row_point(arow_ok ? col : 0);
    bf16x8_t ahi[KS], alo[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        ahi[ks] = hi8[arow * (KP / 8) + ks * 2 + half];
        alo[ks] = lo8[arow * (KP / 8) + ks * 2 + half];
        if (!arow_ok) {
            ahi[ks] = bf16x8_t{};
            alo[ks] = bf16x8_t{};
        }
    }
    // rows held in the accumulator: r -> tile row 8 (r/4) + 4 half + (r%4)
    float na_r[16];
    int32_t acc_i[16], rp_r[16], root_r[16];
    uint32_t rlive = 0;
    // PASS 1: each live row's root found once, by lane `row` (the 32 lanes
    // of a half hold the same rows: finding them in every lane repeated each
    // path-halving write 32 times), then handed to the lanes that hold it
    int32_t my_root = -1;
    if (PASS == 1 && half == 0 && col < nrows) {
        const int64_t p = row_point(col);
        if (A.core_s[p]) my_root = mf_find(A.parent, (int32_t)p);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
        const bool ok = tr < nrows;
        const int64_t p = row_point(ok ? tr : 0);
        rp_r[r] = (int32_t)p;
        bool live = ok;
        if (PASS == 1) live = ok && A.core_s[p];
        if (PASS == 2) live = ok && !A.core_s[p];
        rlive |= (uint32_t)live << r;
        na_r[r] = live ? A.na[p] : INF;
        acc_i[r] = PASS == 2 ? NDBX_NO_CLUSTER : 0;
        if (PASS == 1) {
            const int32_t rt = __shfl(my_root, tr, 64);
            root_r[r] = live ? rt : -1;
        } else {
            root_r[r] = -1;
        }
    }
    if (PASS == 1 || PASS == 2) {  // whole workgroup leaves when no wave has work
        if (threadIdx.x == 0) any_work = 0;
        __syncthreads();
        if (__ballot(rlive != 0) != 0 && lane == 0) any_work = 1;
        __syncthreads();
        if (!any_work) return;
    }
    float rmax_row = arow_ok ? A.rho[arow] : 0.0f;
    for (int o = 32; o >= 1; o >>= 1) rmax_row = fmaxf(rmax_row, __shfl_xor(rmax_row, o, 64));
    rmax_row = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rmax_row)));
    const int32_t min_row_point = A.rows ? 0 : A.tiles[grp.y].y;  // PASS 1 (identity rows only)

    float amax = 0.0f;  // audit: largest |est - float64 distance^2|
    // The column blocks of all candidate cells form one sequence; the B data
    // of the next block is fetched into registers while the current block is
    // computed from LDS.
    constexpr int UPT = (UNITS + 255) / 256;  // units per thread
    const int c_beg = A.cand_off[cell], c_end = A.cand_off[cell + 1];
    int ci = c_beg, cb = 0, ce = 0, cs = 0;
    // Union, phase 1: a pair whose cells each have all their cores in one
    // set, the same set, holds no union to make (read-only root walks).
    // Other workgroups unite sets while this one runs, so the test can turn
    // true at any moment; every decision that shapes the block sequence is
    // therefore taken once for the workgroup and read from LDS, never by
    // each wave for itself (waves that disagreed would walk different block
    // sequences and fill one LDS block from two).
    const bool set_skip = PASS == 1 && A.union_phase == 1;
    auto root_of = [&](int32_t x) -> int32_t {
        for (int32_t p = mf_ld(A.parent + x); p != x; p = mf_ld(A.parent + x)) x = p;
        return x;
    };
    const int32_t rep_c = set_skip ? A.cell_rep[cell] : -1;
    auto same_set_now = [&](int o) -> bool {
        if (!set_skip || rep_c < 0) return false;
        const int32_t rep_o = A.cell_rep[o];
        return rep_o >= 0 && root_of(rep_c) == root_of(rep_o);
    };
    // ss_bits: candidate entries (the first kSameSetCap of the cell) whose
    // pair was one set when the workgroup started; ss_pub[b]: the entry whose
    // pair was found to be one set while its block was fetched into buffer b,
    // or -1.
    if (set_skip) {
        const int ncand = min(c_end - c_beg, kSameSetCap);
        for (int base = wave * 64; base < ncand; base += 256) {
            const int idx = base + lane;
            const uint64_t m = __ballot(idx < ncand && same_set_now(A.cand[c_beg + idx]));
            if (lane == 0) {
                ss_bits[base >> 5] = (uint32_t)m;
                ss_bits[(base >> 5) + 1] = (uint32_t)(m >> 32);
            }
        }
        if (threadIdx.x < 2) ss_pub[threadIdx.x] = -1;
        __syncthreads();
    }
    auto same_set_at_start = [&](int ci_) -> bool {
        const int idx = ci_ - c_beg;
        return set_skip && idx < kSameSetCap && ((ss_bits[idx >> 5] >> (idx & 31)) & 1u);
    };
    int buf = 0;  // LDS buffer of the block being computed
    // PASS 1 with block masks: the column tiles of each pair that the count
    // pass computed and found to hold a pair within eps or a guard-band
    // element (the others hold no pair within eps, or only such pairs, whose
    // cores k_allin_union has united); bm: those of the current cell among
    // its first 64 not yet passed
    const uint64_t *gm = (PASS == 1 && A.bmask) ? A.bmask + (A.bmask_off[gidx] - c_beg) : nullptr;
    uint64_t bm = ~0ull;
    auto to_kept = [&](int &cb_, int cs_) {  // cb_ to the first kept tile at or after it
        if (!gm || cb_ >= cs_ + 64 * 32) return;
        bm &= ~((1ull << ((cb_ - cs_) >> 5)) - 1ull);
        cb_ = bm ? cs_ + 32 * __builtin_ctzll(bm) : cs_ + 64 * 32;
    };
    // the candidate cells a pass visits (an empty range for the others)
    auto wanted = [&](int ci_, int o) -> bool {
        if (PASS == 0 || PASS == 3) return true;
        if (PASS == 1 && (o == cell) != (A.union_phase == 0)) return false;
        if (PASS == 1 && same_set_at_start(ci_)) return false;
        if (mf_takes(cell, o)) return A.pflag[ci_] != 0;
        if (PASS == 1) return false;
        int lo_ = A.cand_off[o], hi_ = A.cand_off[o + 1];  // entry of `cell` in o's list
        const int end_ = hi_;
        while (lo_ < hi_) {
            const int mid = (lo_ + hi_) >> 1;
            if (A.cand[mid] < cell) lo_ = mid + 1;
            else hi_ = mid;
        }
        return lo_ < end_ && A.cand[lo_] == cell && A.pflag[lo_] != 0;
    };
    auto load_cell = [&](int ci_, int &cb_, int &cs_, int &ce_) {
        const int cc_ = A.cand[ci_];
        cs_ = A.cell_start[cc_];
        ce_ = A.cell_start[cc_ + 1];
        cb_ = wanted(ci_, cc_) ? cs_ : ce_;
        if (gm && cb_ < ce_) bm = gm[ci_];  // PASS 1 visits only pairs this cell takes
    };
    // The candidate entries worth loading (wanted, and with a tile to visit
    // under the masks), 64 at a time, one per lane: win holds those of the
    // window [win_base, win_base + 64).  Every input is fixed while the pass
    // runs (pair flags, masks, cell sizes, the same-set bits taken at start),
    // so every wave finds the same entries.
    int win_base = -1;
    uint64_t win = 0;
    auto next_worth = [&](int from) -> int {  // the first such entry >= from, or c_end
        while (from < c_end) {
            const int wb = c_beg + ((from - c_beg) & ~63);
            if (wb != win_base) {
                const int e = wb + lane;
                bool w = false;
                if (e < c_end) {
                    const int o = A.cand[e];
                    w = wanted(e, o);
                    if (w && gm) w = gm[e] != 0 || A.cell_start[o + 1] - A.cell_start[o] > 64 * 32;
                }
                win = __ballot(w);
                win_base = wb;
            }
            const uint64_t m = win & (~0ull << (from - wb));
            if (m) return wb + __builtin_ctzll(m);
            from = wb + 64;
        }
        return c_end;
    };
    // move (ci, cb) to the first block at or after it that has work; ci ==
    // c_end when none is left.  PASS 1 jumps over blocks of the cell's own
    // pair that lie entirely below every row of the group.
    auto settle = [&](int &ci_, int &cb_, int &cs_, int &ce_) {
        while (ci_ < c_end) {
            if (PASS == 1 && cb_ + 31 < min_row_point && A.cand[ci_] == cell)
                cb_ += 32 * ((min_row_point - cb_) / 32);
            to_kept(cb_, cs_);
            // a pair whose cells have come to share one set since its first block
            if (set_skip && cb_ > cs_ && cb_ < ce_ && ss_pub[buf] == ci_) cb_ = ce_;
            if (cb_ < ce_) return;
            ci_ = next_worth(ci_ + 1);
            if (ci_ < c_end) load_cell(ci_, cb_, cs_, ce_);
        }
    };
    auto advance = [&](int &ci_, int &cb_, int &cs_, int &ce_, bool first) {
        if (first) {
            ci_ = next_worth(ci_);
            if (ci_ < c_end) load_cell(ci_, cb_, cs_, ce_);
        } else {
            cb_ += 32;
        }
        settle(ci_, cb_, cs_, ce_);
    };
    bf16x8_t pf_hi[UPT], pf_lo[UPT];
    float pf_na = INF;
    int32_t pf_aux = 0;
    auto fetch = [&](int cb_, int cs_, int ce_) {
#pragma unroll
        for (int q = 0; q < UPT; ++q) {
            const int u = threadIdx.x + q * 256;
            const int bc = u / (KP / 8), kk = u - bc * (KP / 8);
            const bool ok = u < UNITS && cb_ + bc < ce_;
            const int64_t src = (int64_t)(ok ? cb_ + bc : cs_) * (KP / 8) + kk;
            pf_hi[q] = ok ? hi8[src] : bf16x8_t{};
            pf_lo[q] = ok ? lo8[src] : bf16x8_t{};
        }
        if (threadIdx.x < 32) {
            const int j = cb_ + (int)threadIdx.x;
            bool live = j < ce_;
            pf_aux = 0;
            if (PASS == 1 || PASS == 2) live = live && A.core_s[j];
            if (PASS == 1 && live) pf_aux = mf_find(A.parent, j);
            if (PASS == 2 && live) pf_aux = A.label_s[j];
            pf_na = live ? A.na[j] : INF;
        }
    };
    auto publish = [&](int buf) {
#pragma unroll
        for (int q = 0; q < UPT; ++q) {
            const int u = threadIdx.x + q * 256;
            if (u < UNITS) {
                const int bc = u / (KP / 8), kk = u - bc * (KP / 8);
                sb_hi[buf][bc * ROWU + kk] = pf_hi[q];
                sb_lo[buf][bc * ROWU + kk] = pf_lo[q];
            }
        }
        if (threadIdx.x < 32) {
            sb_na[buf][threadIdx.x] = pf_na;
            sb_aux[buf][threadIdx.x] = pf_aux;
        }
    };
    advance(ci, cb, cs, ce, true);
    if (ci < c_end) {
        fetch(cb, cs, ce);
        publish(0);
    }
    __syncthreads();
    float mid_r[16];  // PASS 0: (|a|^2 - e2)/2 per row (+inf for rows out of range)
#pragma unroll
    for (int r = 0; r < 16; ++r) mid_r[r] = 0.5f * (na_r[r] - A.e2);
    while (ci < c_end) {
        const int cc = A.cand[ci];
        const int cur_cb = cb;
        // fetch the next block while this one is computed; one barrier per block
        advance(ci, cb, cs, ce, false);
        const bool more = ci < c_end;
        if (more) fetch(cb, cs, ce);
        const int cur = buf;
        buf ^= 1;
        do {  // compute this block ("break" ends it; the barrier below always runs)
            if (!wave_on) break;
            // R: the largest centred norm of the rows and of the block's
            // columns (its tile's, or its cell's without tile bounds)
            const float cr = A.tile_rho ? A.tile_rho[A.ctoff[cc] + ((cur_cb - A.cell_start[cc]) >> 5)]
                                        : A.cell_rmax[cc];
            const double R = fmax((double)rmax_row, (double)cr);
            const float band = (float)(2.0 * (A.b1 * R * R + A.b2 * R + A.b3));
            const float thr_in = A.e2 - band, thr_out = A.e2 + band;

            const float nb = sb_na[cur][col];
            const int32_t aux = sb_aux[cur][col];
            if (PASS == 1) {
                // skip the block when every live row and column is in one set
                int32_t lo_r = INT_MAX, hi_r = -1;
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    if ((rlive >> r) & 1) {
                        lo_r = min(lo_r, root_r[r]);
                        hi_r = max(hi_r, root_r[r]);
                    }
                if (nb != INF) {
                    lo_r = min(lo_r, aux);
                    hi_r = max(hi_r, aux);
                }
                for (int o = 32; o >= 1; o >>= 1) {
                    lo_r = min(lo_r, __shfl_xor(lo_r, o, 64));
                    hi_r = max(hi_r, __shfl_xor(hi_r, o, 64));
                }
                if (lo_r == hi_r || hi_r < 0 || __ballot(nb != INF) == 0) break;
            }
            if (PASS == 2 && __ballot(nb != INF) == 0) break;
            f32x16_t acc = {};
            if (PASS == 0) {  // count: the accumulator starts at -|b|^2/2 (this lane's column)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[r] = -0.5f * nb;
            }
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const bf16x8_t bhi = sb_hi[cur][col * ROWU + ks * 2 + half];
                const bf16x8_t blo = sb_lo[cur][col * ROWU + ks * 2 + half];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], bhi, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], blo, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[ks], bhi, acc, 0, 0, 0);
            }
            const int32_t bcol = cur_cb + col;
            if (PASS == 0) {
                // With dv = acc - (|a|^2 - e2)/2 = (e2 - est)/2 (the accumulator
                // starts at -|b|^2/2) and hb = band/2: dv >= hb is a neighbour,
                // dv < -hb is not, and |dv| <= hb flags the guard band, which the
                // exact predicate decides.  An element costs a subtract, two
                // compares (|.| is free) and a count.
                const float hb = 0.5f * band;
                bool inband = false;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float dv = acc[r] - mid_r[r];
                    acc_i[r] += dv >= hb ? 1 : 0;
                    inband = inband || fabsf(dv) <= hb;
                }
                if (__ballot(inband) != 0) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const float dv = acc[r] - mid_r[r];
                        if (dv >= hb || dv < -hb) continue;
                        acc_i[r] += exact_pred(A.xs + (int64_t)rp_r[r] * A.d, A.xs + (int64_t)bcol * A.d, A.d,
                                               A.e2) ? 1 : 0;
                    }
                }
                break;
            }
            uint32_t in_m = 0, band_m = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float est = __builtin_fmaf(-2.0f, acc[r], na_r[r] + nb);
                const bool in = est <= thr_in;
                in_m |= (uint32_t)in << r;
                band_m |= (uint32_t)(!in && est <= thr_out) << r;
            }            if (PASS == 3) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if (!((rlive >> r) & 1) || nb == INF) continue;
                    const float est = __builtin_fmaf(-2.0f, acc[r], na_r[r] + nb);
                    const int64_t rp = rp_r[r];
                    const bool ex = exact_pred(A.xs + rp * A.d, A.xs + (int64_t)bcol * A.d, A.d, A.e2);
                    double d64 = 0.0;
                    for (int k = 0; k < A.d; ++k) {
                        const double df = (double)A.xs[rp * A.d + k] - (double)A.xs[(int64_t)bcol * A.d + k];
                        d64 += df * df;
                    }
                    const bool in = (in_m >> r) & 1, band_pair = (band_m >> r) & 1;
                    if (band_pair) atomicAdd(A.audit + 0, 1ull);
                    if (ex != (d64 <= (double)A.e2)) atomicAdd(A.audit + 1, 1ull);
                    if (!band_pair && in != ex) atomicAdd(A.audit + 2, 1ull);
                    amax = fmaxf(amax, (float)fabs((double)est - d64));
                }
                break;
            }
            // guard-band estimates: the reference's own arithmetic decides
            if (__ballot(band_m != 0) != 0 && band_m) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    if ((band_m >> r) & 1) {
                        const bool ex = exact_pred(A.xs + (int64_t)rp_r[r] * A.d, A.xs + (int64_t)bcol * A.d,
                                                   A.d, A.e2);
                        in_m |= (uint32_t)ex << r;
                        if (PASS == 0) acc_i[r] += ex ? 1 : 0;
                    }
            }
            if (PASS == 1 && in_m) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if (!((in_m >> r) & 1) || (cc == cell && bcol <= rp_r[r]) || root_r[r] == aux) continue;
                    root_r[r] = mf_unite(A.parent, root_r[r] >= 0 ? root_r[r] : rp_r[r], bcol);
                }
            }
            if (PASS == 2 && in_m) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    if (((in_m >> r) & 1) && aux < acc_i[r]) acc_i[r] = aux;
            }
        } while (0);
        if (more) publish(buf);
        if (set_skip && more && threadIdx.x == 0) ss_pub[buf] = same_set_now(A.cand[ci]) ? ci : -1;
        __syncthreads();
    }
    if (!wave_on) return;
    if (PASS == 3) {
        for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
        if (lane == 0) atomicMax(reinterpret_cast<int *>(A.audit_err), __float_as_int(amax));
    }
    if (PASS == 0 || PASS == 2) {
        // reduce each row over the 32 lanes (columns) of its half
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int32_t v = acc_i[r];
            for (int o = 16; o >= 1; o >>= 1) {
                const int32_t w = __shfl_xor(v, o, 64);
                v = PASS == 0 ? v + w : (w < v ? w : v);
            }
            acc_i[r] = v;
        }
        if (col == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (!((rlive >> r) & 1)) continue;
                if (PASS == 0) A.nn_s[rp_r[r]] = acc_i[r];
                if (PASS == 2) A.best_s[rp_r[r]] = acc_i[r];
            }
        }
    }
}

// Half the guard band hb and the hi*hi skip threshold hskip of the count
// pass for R = the largest centred norm of a column tile (its own, or its
// cell's without tile bounds), in the arithmetic the count pass used per
// block.  Both are monotone in R (every operation and rounding is), so the
// count pass's value for R = max(rows' R, columns' R) is the larger hb and
// the smaller hskip of the row side's and the column side's, exactly.
__global__ __launch_bounds__(256) void k_tile_band(const int4 *__restrict__ tiles, int64_t nt,
                                                   const float *__restrict__ tile_rho,
                                                   const float *__restrict__ cell_rmax, double b1, double b2,
                                                   double b3, float2 *__restrict__ band) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const double R = (double)(tile_rho ? tile_rho[t] : cell_rmax[tiles[t].x]);
    const float hb = (float)(b1 * R * R + b2 * R + b3);
    band[t] = make_float2(hb, -(float)(3.0 * (double)hb + 0.008 * R * R));
}

// larger / smaller of two floats of one sign, on their bits (scalar unit)
__device__ __forceinline__ float max_nonneg(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}
__device__ __forceinline__ float min_nonpos(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// ------------------------------------------------------ symmetric count --
// Groups are taken in chunks of `chunk` consecutive groups (a few cells),
// chunk q on XCD q % 8 (workgroups go to XCDs round-robin by index), so the
// groups of one cell run on one XCD at about the same time and read their
// candidate blocks from its L2.
__device__ __forceinline__ int64_t mf_group_of(int64_t b, int chunk) {
    const int64_t x = b & 7, i = b >> 3;
    return ((i / chunk) * 8 + x) * chunk + (i % chunk);
}

// PASS 0 as k_mfma_pass, over the cell pairs this cell takes.  nn_s starts at
// zero; row and column counts are atomic adds.  Column counts are summed in
// an LDS window of kColWin consecutive columns of one candidate cell (all
// waves, all blocks of the window) and added to nn_s, one atomic per
// column, when the block sequence leaves the window; so is the pair flag.
// The blocks themselves write nothing to global memory (a store or atomic
// there would hold up the next block's loads, which wait on the same
// counter).
constexpr int kColWin = 2048;
constexpr int kBandQ = 32;       // guard-band pairs queued per wave in LDS
// Row tiles (waves) per count workgroup.  Measured on the 1M 64-D workload:
// 4 (3 waves per SIMD) 220 ms with 1.0 TB of L2 misses; 8 (2 per SIMD) 284 ms
// with 0.40 TB; 12 (3 per SIMD) 234 ms with 0.30 TB: the pass is bound by
// issue and latency, not by its candidate-block traffic.
constexpr int kCountWaves = 4;
// Waves per SIMD the count pass is built for (d <= 64).
#ifndef NDBX_MF_COUNT_OCC
#define NDBX_MF_COUNT_OCC 4
#endif
constexpr int kCountOcc = NDBX_MF_COUNT_OCC;
// LDS buffers of column blocks per count workgroup: the block in the
// products and the kCountBufs - 1 after it, in flight.
#ifndef NDBX_MF_COUNT_BUFS
#define NDBX_MF_COUNT_BUFS 3
#endif
constexpr int kCountBufs = NDBX_MF_COUNT_BUFS;
static_assert(kCountBufs >= 3 && kCountBufs <= 6, "count pass LDS buffers");

template <int KP, int WAVES>
__global__ __launch_bounds__(64 * WAVES, (WAVES == 4 && KP < 128) ? kCountOcc : 2) void k_mf_count(MfmaArgs A, int chunk) {
    static_assert(WAVES == 4, "a block's loads are split over four waves");
    constexpr int KS = KP / 16;
    constexpr int UNITS = 32 * (KP / 8);  // bf16x8 units of a block, per hi / lo
    // Only the hi halves are staged (the lo halves are read from global
    // memory by the few blocks that need the full products), unit-major:
    // unit u of column c at u * 32 + c, so the 32 columns a wave's half
    // reads for one k step are 32 consecutive slots.  kCountBufs buffers:
    // the block in the products and the next ones, loaded straight into LDS.
    __shared__ bf16x8_t sb_hi[kCountBufs][UNITS];
    __shared__ float sb_na[kCountBufs][32];
    __shared__ int32_t sb_cacc[kColWin];  // column counts of the current window
    __shared__ int32_t sb_edge;
    __shared__ unsigned long long sb_emask;  // edge tiles of the current candidate among its first 64
    __shared__ int4 sb_q[WAVES][kBandQ];  // per wave: guard-band pairs not yet in the global queue
    __shared__ bf16x8_t sa_lo[WAVES][UNITS];  // per wave: its rows' lo halves, unit-major as the blocks
    const int64_t gidx = mf_group_of(blockIdx.x, chunk);
    if (gidx >= A.n_groups) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int4 grp = A.groups[gidx];
    const bool wave_on = wave < grp.z;
    const int4 tile = A.tiles[grp.y + (wave_on ? wave : 0)];
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int cell = grp.x, row0 = tile.y, nrows = wave_on ? tile.z : 0;
    const bf16x8_t *hi8 = reinterpret_cast<const bf16x8_t *>(A.hi);
    const bf16x8_t *lo8 = reinterpret_cast<const bf16x8_t *>(A.lo);
    const float INF = __builtin_inff();
    // A operand: row `col` of the tile, k elements [16 ks + 8 half, +8);
    // the lo halves wait in LDS (zero past the tile's rows) for the blocks
    // that take the full products
    const bool arow_ok = col < nrows;
    const int64_t arow = row0 + (arow_ok ? col : 0);
    bf16x8_t ahi[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) ahi[ks] = arow_ok ? hi8[arow * (KP / 8) + ks * 2 + half] : bf16x8_t{};
    for (int u = lane; u < UNITS; u += 64) {
        const int r = u & 31;
        sa_lo[wave][u] = r < nrows ? lo8[(int64_t)(row0 + r) * (KP / 8) + (u >> 5)] : bf16x8_t{};
    }
    // accumulator rows: r -> tile row 8 (r/4) + 4 half + (r%4); the start
    // value of every block is -(|a|^2 - e2)/2 per row (-inf for rows past
    // the tile, so dv = -inf: never counted, never in the band)
    f32x16_t acc0;
    int32_t rcnt[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
        acc0[r] = tr < nrows ? -(0.5f * (A.na[row0 + tr] - A.e2)) : -INF;
        rcnt[r] = 0;
    }
    float rmax_row = arow_ok ? A.rho[arow] : 0.0f;
    for (int o = 32; o >= 1; o >>= 1) rmax_row = fmaxf(rmax_row, __shfl_xor(rmax_row, o, 64));
    rmax_row = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rmax_row)));
    // the rows' side of hb and hskip (k_tile_band)
    const double Rr = (double)rmax_row;
    const float hb_row = (float)(A.b1 * Rr * Rr + A.b2 * Rr + A.b3);
    const float hs_row = -(float)(3.0 * (double)hb_row + 0.008 * Rr * Rr);

    // This wave's queued guard-band pairs to the global queue (one atomic);
    // past its capacity they are decided here.
    int qn = 0;
    auto band_flush = [&]() {
        if (qn == 0) return;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(A.bq_count, (unsigned long long)qn);
        base = __shfl(base, 0, 64);
        for (int i = lane; i < qn; i += 64) {
            const int4 e = sb_q[wave][i];
            if (base + (unsigned long long)i < (unsigned long long)A.bq_cap) {
                A.bq[base + i] = e;
            } else if (exact_pred(A.xs + (int64_t)e.x * A.d, A.xs + (int64_t)e.y * A.d, A.d, A.e2)) {
                atomicAdd(A.nn_s + e.x, 1);
                if (!e.w) atomicAdd(A.nn_s + e.y, 1);
                A.pflag[e.z] = 1;
            }
        }
        qn = 0;
    };
    // block sequence: the 32-column blocks of the cells this cell takes,
    // the first 64 of a cell those its mask keeps (all without masks)
    const int c_beg = ld_const(A.cand_off + cell), c_end = ld_const(A.cand_off + cell + 1);
    const uint64_t *gmask = A.bmask ? A.bmask + (ld_const(A.bmask_off + gidx) - c_beg) : nullptr;
    uint64_t *emask_g = A.bmask && A.emask ? A.emask + (ld_const(A.bmask_off + gidx) - c_beg) : nullptr;
    // cursor: cell co (candidate ci), columns [cb, cb + 32) of [cs, ce), ct
    // the cell's first row tile; bm: the current cell's kept tiles among its
    // first 64 not yet visited
    int ci = c_beg - 1, cb = 0, cs = 0, ce = 0, co = -1, ct = 0;
    uint64_t bm = 0;
    // The candidate entries with blocks to compute, 64 at a time, one per
    // lane (win: those of [win_base, win_base + 64)); the entry found is
    // then read again with scalar loads.
    int win_base = -1;
    uint64_t win = 0;
    auto next_cell = [&]() {  // the next taken cell with blocks to compute (ci >= c_end: none)
        for (++ci; ci < c_end; ++ci) {
            const int wb = c_beg + ((ci - c_beg) & ~63);
            if (wb != win_base) {
                const int e = wb + lane;
                bool w = false;
                if (e < c_end) {
                    const int o = A.cand[e];
                    if (mf_takes(cell, o)) {
                        const int s0 = A.cell_start[o], s1 = A.cell_start[o + 1];
                        w = s0 < s1 && (s1 - s0 > 64 * 32 || !gmask || gmask[e] != 0);
                    }
                }
                win = __ballot(w);
                win_base = wb;
            }
            const uint64_t wm = win & (~0ull << (ci - wb));
            if (wm == 0) {
                ci = wb + 63;
                continue;
            }
            ci = wb + __builtin_ctzll(wm);
            const int o = ld_const(A.cand + ci);
            if (!mf_takes(cell, o)) continue;
            const int s0 = ld_const(A.cell_start + o), s1 = ld_const(A.cell_start + o + 1);
            if (s0 >= s1) continue;
            const uint64_t m = gmask ? ld_const(gmask + ci) : ~0ull;
            if (m == 0 && s1 - s0 <= 64 * 32) continue;  // every block pruned
            co = o, cs = s0, ce = s1;
            bm = m;
            cb = bm ? s0 + 32 * __builtin_ctzll(bm) : s0 + 64 * 32;
            if (cb >= ce) continue;
            ct = ld_const(A.ctoff + o);
            return;
        }
    };
    auto advance = [&]() {
        if (cb < cs + 64 * 32) {
            bm &= bm - 1;
            cb = bm ? cs + 32 * __builtin_ctzll(bm) : cs + 64 * 32;
        } else {
            cb += 32;
        }
        if (cb >= ce) next_cell();
    };
    // A block's loads go straight to LDS (no registers, so kCountBufs - 1
    // blocks can be in flight behind the one in the products): each wave
    // loads KP of its UNITS hi units, 64 per instruction, and 8 of its 32
    // norms, so every wave issues kLoads loads per block, and waiting until
    // at most (kCountBufs - 2) kLoads are outstanding leaves the oldest block
    // in flight landed (loads complete in issue order for the counter).  Rows are read 32 at a
    // time from the block's first (from row n past the sequence's end): the
    // hi rows and the norms hold 32 rows of padding.  Columns past their
    // cell's end are given norm +inf (accumulator -inf) where the norm is
    // read.
    constexpr int kOps = (KP + 63) / 64;
    constexpr int kLoads = kOps + 1;
    static_assert((kCountBufs - 2) * kLoads < 16, "vmcnt field");
    const int64_t n_last = A.n - 1;
    // per-lane byte offsets of the loads within a block
    uint32_t hoff[kOps];
#pragma unroll
    for (int q = 0; q < kOps; ++q) {
        const int slot = wave * KP + 64 * q + lane;
        hoff[q] = (uint32_t)((slot & 31) * (KP * 2) + (slot >> 5) * 16);
    }
    const uint32_t noff = (uint32_t)(8 * wave + lane) * 4;
    auto fetch = [&](int b) {
        const int fb = cb < (int)A.n ? cb : (int)A.n;  // past the sequence's end: any rows in range
        const char *hb8 = reinterpret_cast<const char *>(hi8) + (int64_t)fb * (KP * 2);
        const char *nb8 = reinterpret_cast<const char *>(A.na + fb);
#pragma unroll
        for (int q = 0; q < kOps; ++q) {
            if (lane < KP - 64 * q)
                __builtin_amdgcn_global_load_lds((const void *)(hb8 + hoff[q]),
                                                 (__attribute__((address_space(3))) void *)&sb_hi[b][wave * KP + 64 * q],
                                                 16, 0, 0);
        }
        if (lane < 8)
            __builtin_amdgcn_global_load_lds((const void *)(nb8 + noff),
                                             (__attribute__((address_space(3))) void *)&sb_na[b][8 * wave], 4, 0, 0);
    };
    // every wave's loads of the older block landed and its LDS operations
    // done, then the workgroup barrier; the empty asm statements keep the
    // compiler from moving memory operations across it
    auto landed_barrier = [&]() {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0x0070 | (kCountBufs - 2) * kLoads);  // vmcnt(...) lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };
    // The blocks fetched and not yet computed, oldest first (q[0] is in the
    // products): place in the sequence, the column tile's (hb, hskip), read
    // when the block is fetched so that the scalar load is long done when it
    // is computed; m: the block exists.
    struct Blk {
        int o, cb, ci, cs, ce;
        float2 band;
        bool m;
    };
    auto at_cursor = [&]() {
        Blk b;
        b.o = co, b.cb = cb, b.ci = ci, b.cs = cs, b.ce = ce;
        b.m = ci < c_end;
        const uint64_t bw = !b.m ? 0ull : ld_const(reinterpret_cast<const uint64_t *>(A.tile_band) + (ct + ((cb - cs) >> 5)));
        b.band = make_float2(__uint_as_float((uint32_t)bw), __uint_as_float((uint32_t)(bw >> 32)));
        return b;
    };
    next_cell();
    if (ci >= c_end) return;  // cannot happen: a cell always takes itself
    for (int t = threadIdx.x; t < kColWin; t += blockDim.x) sb_cacc[t] = 0;
    if (threadIdx.x == 0) {
        sb_edge = 0;
        sb_emask = 0;
    }
    Blk q[kCountBufs - 1];
    q[0] = at_cursor();
    fetch(0);
#pragma unroll
    for (int j = 1; j < kCountBufs - 1; ++j) {
        advance();
        q[j] = at_cursor();
        fetch(j);
    }
    landed_barrier();
    int buf = 0;
    for (;;) {
        // this block's place in the sequence (wave-uniform, in scalar
        // registers), then the loads of the block kCountBufs - 1 places on,
        // which overlap this block's products and the next ones'
        const int o = q[0].o, bcb = q[0].cb, bci = q[0].ci, bcs = q[0].cs, bce = q[0].ce;
        const float2 band = q[0].band;
        advance();
        const Blk nb = at_cursor();
        fetch(buf == 0 ? kCountBufs - 1 : buf - 1);
        const bool diag = o == cell;
        if (wave_on) {
            // R = the largest centred norm of the rows and of the block's
            // columns; hb = (float)(b1 R^2 + b2 R + b3), half the band.
            // hskip, the skip threshold of the hi*hi partial sum: the two
            // products it leaves out are each at most 2^-8 (1 + 2^-8)^2 |a|
            // |b| (|lo| <= 2^-8 |x| componentwise), 0.00787 R^2 together; its
            // own float32 accumulation error and the full sum's are each
            // within hb, so a partial below -(3 hb + 0.00787 R^2) leaves the
            // full dv below -hb.  The 0.008 R^2 taken also covers the one
            // rounding of the per-column threshold below, at most 2^-24
            // (|hskip| + R^2 / 2).  Both from the rows' and the column tile's
            // values (k_tile_band).
            const float hb = max_nonneg(hb_row, band.x);
            const float hskip = min_nonpos(hs_row, band.y);
            // dv = (e2 - est)/2 = -(|a|^2 - e2)/2 + a.b - |b|^2/2 (-inf for
            // rows and columns out of range).  The hi*hi partial is taken
            // without the column term -|b|^2/2, which moves to the skip
            // threshold instead (+inf for columns out of range), and is added
            // only in the blocks that go on to the full products.
            const float nav = sb_na[buf][col];
            const float nbh = bcb + col < bce ? -0.5f * nav : -INF;
            const float hcol = hskip - nbh;
            // hi*hi first: a block whose partial largest dv is below hskip
            // holds no neighbour and no band element under the full
            // products, and costs 2 MFMAs per 16-wide k step less
            f32x16_t acc = acc0;
            const bf16x8_t *bp = &sb_hi[buf][half * 32 + col];
#pragma unroll
            for (int k4 = 0; k4 < KS; k4 += 4) {
                // B operands of four k steps (slots 64 apart), read together
                bf16x8_t b4[4];
                if (KS - k4 >= 4) {
                    lds_read16x4<64 * 16>(bp + k4 * 64, b4[0], b4[1], b4[2], b4[3]);
                } else {
                    for (int k = 0; k < KS - k4; ++k) b4[k] = lds_read16(bp + (k4 + k) * 64);
                }
#pragma unroll
                for (int k = 0; k < 4 && k4 + k < KS; ++k)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[k4 + k], b4[k], acc, 0, 0, 0);
            }
            float hmax = acc[0];
#pragma unroll
            for (int r = 1; r < 16; ++r) hmax = fmaxf(hmax, acc[r]);
            int32_t ccnt = 0;
            if (__ballot(hmax >= hcol) != 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] += nbh;
            const int64_t bpt = bcb + col < n_last ? bcb + col : n_last;  // clamped to the last row
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const bf16x8_t blo = lo8[bpt * (KP / 8) + ks * 2 + half];
                const bf16x8_t alo = lds_read16(&sa_lo[wave][(ks * 2 + half) * 32 + col]);
                const bf16x8_t bhi = lds_read16(&sb_hi[buf][(ks * 2 + half) * 32 + col]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], blo, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
            }
            // dv >= hb is a neighbour, dv < -hb is not, the rest (the guard
            // band) the exact predicate decides.  A block whose largest dv is
            // below -hb (no neighbour, no band element) needs nothing more.
            float amax = acc[0];
#pragma unroll
            for (int r = 1; r < 16; ++r) amax = fmaxf(amax, acc[r]);
            if (__ballot(amax >= -hb) != 0) {
            float amin = INF;
#pragma unroll
            for (int r = 0; r < 16; r += 2) {
                const float dv0 = acc[r], dv1 = acc[r + 1];
                const int32_t in0 = dv0 >= hb ? 1 : 0, in1 = dv1 >= hb ? 1 : 0;
                rcnt[r] += in0;
                rcnt[r + 1] += in1;
                ccnt += in0 + in1;
                amin = fminf(amin, fminf(fabsf(dv0), fabsf(dv1)));
            }
            if (__ballot(amin <= hb) != 0) {
                // guard-band elements (-hb <= dv < hb) go to this wave's queue
                uint32_t bm = 0;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float dv = acc[r];
                    bm |= (uint32_t)(dv < hb && dv >= -hb) << r;
                }
                const int cnt = __builtin_popcount(bm);
                int incl = cnt;
#pragma unroll
                for (int o2 = 1; o2 < 64; o2 <<= 1) {
                    const int t2 = __shfl_up(incl, o2, 64);
                    if (lane >= o2) incl += t2;
                }
                const int total = __builtin_amdgcn_readlane(incl, 63);
                if (qn + total > kBandQ) band_flush();
                if (total > kBandQ) {
                    // too many for the queue: decided here, one lane per column
                    uint32_t hit = 0;
                    const float *pc = A.xs + (int64_t)(bcb + col) * A.d;
                    while (bm) {
                        const int r = __builtin_ctz(bm);
                        bm &= bm - 1;
                        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
                        if (exact_pred(A.xs + (int64_t)(row0 + tr) * A.d, pc, A.d, A.e2)) hit |= 1u << r;
                    }
                    ccnt += __builtin_popcount(hit);
#pragma unroll
                    for (int r = 0; r < 16; ++r) rcnt[r] += (hit >> r) & 1;
                } else {
                    int pos = qn + incl - cnt;
                    while (bm) {
                        const int r = __builtin_ctz(bm);
                        bm &= bm - 1;
                        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
                        sb_q[wave][pos++] = make_int4(row0 + tr, bcb + col, bci, diag ? 1 : 0);
                    }
                    qn += total;
                }
            }
            // the pair holds a pair within eps: the label passes visit it;
            // the tile holds one, or a guard-band element: the union does
            const bool edge = __ballot(ccnt != 0) != 0;
            if (edge && lane == 0) sb_edge = 1;
            if ((edge || __ballot(amin <= hb) != 0) && lane == 0 && bcb - bcs < 64 * 32)
                atomicOr(&sb_emask, 1ull << ((bcb - bcs) >> 5));
            if (!diag) {
                ccnt += __shfl_xor(ccnt, 32, 64);
                if (half == 0 && ccnt != 0) atomicAdd(&sb_cacc[(bcb - bcs) % kColWin + col], ccnt);
            }
            }
            }
        }
        landed_barrier();
        // leaving the window (another cell, the next kColWin columns, or the
        // end): its column counts and pair flag go to global memory
        const bool m1 = q[1].m;
        const int o1 = q[1].o, cb1 = q[1].cb, cs1 = q[1].cs;
        if (!m1 || o1 != o || (uint32_t)(cb1 - cs1) / kColWin != (uint32_t)(bcb - bcs) / kColWin) {
            if (!diag) {
                const int32_t wb = bcs + (int32_t)((uint32_t)(bcb - bcs) / kColWin * kColWin);
                for (int t = threadIdx.x; t < kColWin; t += blockDim.x) {
                    const int32_t v = sb_cacc[t];
                    if (v != 0) {
                        atomicAdd(A.nn_s + wb + t, v);
                        sb_cacc[t] = 0;
                    }
                }
            }
            if (threadIdx.x == 0 && sb_edge) {
                A.pflag[bci] = 1;
                sb_edge = 0;
            }
            if (threadIdx.x == 0 && sb_emask) {
                if (emask_g) atomicOr(reinterpret_cast<unsigned long long *>(emask_g + bci), sb_emask);
                sb_emask = 0;
            }
            __syncthreads();
        }
        if (!m1) break;
#pragma unroll
        for (int j = 0; j + 1 < kCountBufs - 1; ++j) q[j] = q[j + 1];
        q[kCountBufs - 2] = nb;
        buf = __builtin_amdgcn_readfirstlane(buf == kCountBufs - 1 ? 0 : buf + 1);
    }
    // no load into LDS outlives the workgroup
    __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
    if (!wave_on) return;
    band_flush();
    // each row over the 32 lanes (columns) of its half
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int32_t v = rcnt[r];
        for (int o = 16; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        rcnt[r] = v;
    }
    if (col == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
            if (tr < nrows && rcnt[r] != 0) atomicAdd(A.nn_s + row0 + tr, rcnt[r]);
        }
    }
}

// Guard-band pairs of the count pass, decided by the reference's float32
// arithmetic on the original coordinates (data_process.cpp:443-447): a
// neighbour adds to both counts (the row's only, in a cell's own pair) and
// flags its cell pair.
__global__ __launch_bounds__(256) void k_mf_band(const int4 *__restrict__ bq,
                                                 const unsigned long long *__restrict__ bq_count, int64_t cap,
                                                 const float *__restrict__ xs, int d, float e2,
                                                 int32_t *__restrict__ nn_s, uint8_t *__restrict__ pflag) {
    const int64_t nq = *bq_count < (unsigned long long)cap ? (int64_t)*bq_count : cap;
    const bool v4 = (d & 3) == 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x) {
        const int4 e = bq[i];
        const float *p = xs + (int64_t)e.x * d, *q = xs + (int64_t)e.y * d;
        if (v4 ? exact_pred4(p, q, d, e2) : exact_pred(p, q, d, e2)) {
            atomicAdd(nn_s + e.x, 1);
            if (!e.w) atomicAdd(nn_s + e.y, 1);
            pflag[e.z] = 1;
        }
    }
}

// Non-core rows for the border pass: compacted list (sorted order, so grouped
// by cell), per-cell tile counts.
// A non-core point whose only neighbour is itself (nn = 1) is within eps of
// no core: it keeps NO_CLUSTER without a border row.
__global__ __launch_bounds__(256) void k_noncore_flag(const uint8_t *__restrict__ core_s,
                                                      const int32_t *__restrict__ nn_s, int64_t n,
                                                      int32_t *__restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = !core_s[i] && nn_s[i] > 1;
}
__global__ __launch_bounds__(256) void k_noncore_list(const int32_t *__restrict__ f, const int32_t *__restrict__ pos,
                                                      int64_t n, int32_t *__restrict__ list) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && f[i]) list[pos[i]] = (int32_t)i;
}
// Border row tiles of cell c: list entries [pos[cell_start[c]], pos[cell_start[c+1]]).
__global__ __launch_bounds__(256) void k_btile_count(const int32_t *__restrict__ pos,
                                                     const int32_t *__restrict__ cell_start, int64_t nc,
                                                     int32_t *__restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) cnt[c] = (pos[cell_start[c + 1]] - pos[cell_start[c]] + 31) / 32;
}
__global__ __launch_bounds__(256) void k_btiles(const int32_t *__restrict__ pos,
                                                const int32_t *__restrict__ cell_start,
                                                const int32_t *__restrict__ toff, int64_t nc,
                                                int4 *__restrict__ tiles) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t b = pos[cell_start[c]], e = pos[cell_start[c + 1]];
    int32_t t = toff[c];
    for (int32_t q = b; q < e; q += 32) tiles[t++] = make_int4((int32_t)c, q, e - q < 32 ? e - q : 32, 0);
}

// ------------------------------------------------------------- labels --
// Per cell: a core of the cell when all its cores have one root (after the
// union of the cells' own pairs), else -1 (no core, or several sets).
__global__ __launch_bounds__(256) void k_mf_cell_rep(const uint8_t *__restrict__ core_s,
                                                     const int32_t *__restrict__ cell_start, int64_t nc,
                                                     const int32_t *__restrict__ parent, int32_t *__restrict__ rep) {
    const int64_t c = blockIdx.x;
    if (c >= nc) return;
    __shared__ int32_t first, root0, mixed;
    const int32_t b = cell_start[c], e = cell_start[c + 1];
    if (threadIdx.x == 0) {
        first = INT_MAX;
        mixed = 0;
    }
    __syncthreads();
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x)
        if (core_s[i]) atomicMin(&first, i);
    __syncthreads();
    if (first == INT_MAX) {
        if (threadIdx.x == 0) rep[c] = -1;
        return;
    }
    auto root = [&](int32_t x) {
        while (parent[x] != x) x = parent[x];
        return x;
    };
    if (threadIdx.x == 0) root0 = root(first);
    __syncthreads();
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x)
        if (core_s[i] && root(i) != root0) mixed = 1;
    __syncthreads();
    if (threadIdx.x == 0) rep[c] = mixed ? -1 : first;
}

__global__ __launch_bounds__(256) void k_mf_core(const int32_t *__restrict__ nn_s, const uint32_t *__restrict__ perm,
                                                 int64_t n, int32_t m, int32_t *__restrict__ nn,
                                                 uint8_t *__restrict__ core, uint8_t *__restrict__ core_s,
                                                 int32_t *__restrict__ core_i32, int32_t *__restrict__ parent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = nn_s[i];
    const bool c = v >= m;
    const uint32_t p = perm[i];
    nn[p] = v;
    core[p] = c;
    core_s[i] = c;
    core_i32[p] = c;
    parent[i] = (int32_t)i;
}

// Lowest input index of each set's cores, at its root.  Consecutive points
// mostly share a root: when every core lane of a wave has the same one, the
// wave takes the minimum first and makes one atomic (a few sets take all the
// cores, so per-lane atomics would queue on a few addresses).
__global__ __launch_bounds__(256) void k_mf_rootmin(const uint8_t *__restrict__ core_s,
                                                    const uint32_t *__restrict__ perm, int64_t n,
                                                    int32_t *__restrict__ parent, int32_t *__restrict__ root_min) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool core = i < n && core_s[i];
    const int32_t r = core ? mf_find(parent, (int32_t)i) : -1;
    const int32_t v = core ? (int32_t)perm[i] : INT_MAX;
    const uint64_t cores = __ballot(core);
    if (cores == 0) return;
    const int32_t r0 = __shfl(r, __builtin_ctzll(cores), 64);
    if (__ballot(core && r != r0) == 0) {
        int32_t m = v;
        for (int o = 32; o >= 1; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) atomicMin(root_min + r0, m);
    } else if (core) {
        atomicMin(root_min + r, v);
    }
}

__global__ __launch_bounds__(256) void k_mf_label(const uint8_t *__restrict__ core_s, int64_t n,
                                                  const int32_t *__restrict__ parent,
                                                  const int32_t *__restrict__ root_min,
                                                  const int32_t *__restrict__ rank, int32_t *__restrict__ label_s,
                                                  int32_t *__restrict__ meta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!core_s[i]) {
        label_s[i] = NDBX_NO_CLUSTER;
        return;
    }
    int32_t r = (int32_t)i;
    while (parent[r] != r) r = parent[r];
    label_s[i] = rank[root_min[r]];
    if (r == (int32_t)i) atomicAdd(meta + 2, 1);
}

__global__ __launch_bounds__(256) void k_mf_out(const uint8_t *__restrict__ core_s, const int32_t *__restrict__ label_s,
                                                const int32_t *__restrict__ best_s, const uint32_t *__restrict__ perm,
                                                int64_t n, int32_t *__restrict__ labels, int32_t *__restrict__ meta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = i < n;
    int a = 0, z = 0;
    if (v) {
        const bool c = core_s[i];
        const int32_t l = c ? label_s[i] : best_s[i];
        labels[perm[i]] = l;
        a = c;
        z = l == NDBX_NO_CLUSTER;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        z += __shfl_xor(z, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(meta + 0, a);
        if (z) atomicAdd(meta + 1, z);
    }
}

// ------------------------------------------------------------ pipeline --
static int kp_for(int d) { return d <= 16 ? 16 : d <= 32 ? 32 : d <= 64 ? 64 : 128; }

bool mfma_supported(int d) { return d >= 4 && d <= 128; }

template <int PASS>
static void launch_pass(Ctx &c, const MfmaArgs &a, int kp, const char *name) {
    int64_t blocks = a.n_groups;
    if (PASS == 3) blocks = ceil_div(a.n_groups - a.audit_offset, a.audit_stride);
    const unsigned g = (unsigned)blocks;
    if (g == 0) return;
    Launch L(c.prof, c.stream, name);
    switch (kp) {
        case 16: k_mfma_pass<16, PASS><<<g, 256, 0, c.stream>>>(a); break;
        case 32: k_mfma_pass<32, PASS><<<g, 256, 0, c.stream>>>(a); break;
        case 64: k_mfma_pass<64, PASS><<<g, 256, 0, c.stream>>>(a); break;
        default: k_mfma_pass<128, PASS><<<g, 256, 0, c.stream>>>(a); break;
    }
}

// Symmetric count pass: groups of kCountWaves row tiles, in XCD chunks of
// kCountChunk groups.
constexpr int kCountChunk = 16;
static void launch_count(Ctx &c, const MfmaArgs &a, int kp) {
    if (a.n_groups == 0) return;
    const int64_t per = 8 * kCountChunk;
    const unsigned g = (unsigned)(ceil_div(a.n_groups, per) * per);
    constexpr int NT = 64 * kCountWaves;
    Launch L(c.prof, c.stream, "mf_count");
    switch (kp) {
        case 16: k_mf_count<16, kCountWaves><<<g, NT, 0, c.stream>>>(a, kCountChunk); break;
        case 32: k_mf_count<32, kCountWaves><<<g, NT, 0, c.stream>>>(a, kCountChunk); break;
        case 64: k_mf_count<64, kCountWaves><<<g, NT, 0, c.stream>>>(a, kCountChunk); break;
        default: k_mf_count<128, kCountWaves><<<g, NT, 0, c.stream>>>(a, kCountChunk); break;
    }
}

static MfmaArgs mfma_args(Ctx &c) {
    MfmaArgs a{};
    a.tiles = c.ctiles.as<int4>();
    a.n_tiles = c.n_tiles;
    a.groups = c.tiles.as<int4>();
    a.n_groups = c.n_label_tiles;
    a.cand_off = c.tile_off.as<int32_t>();
    a.cand = c.ranges.as<int32_t>();
    a.cell_start = c.cell_start.as<int32_t>();
    a.hi = c.keys2.as<__bf16>();
    a.lo = reinterpret_cast<const __bf16 *>(c.vals2.p);
    a.na = c.soa_core.as<float>();
    a.rho = c.soa_core.as<float>() + c.n;
    a.cell_rmax = c.cell_c.as<float>();
    a.xs = c.soa.as<float>();
    a.pflag = c.pflag.as<uint8_t>();
    a.n = c.n;
    a.d = c.d;
    const float e = c.h_eps[0];
    a.e2 = e * e;  // data_process.h:73
    const int kp = kp_for(c.d);
    const double u = std::ldexp(1.0, -24);
    // the last terms cover the count pass: its accumulator starts at
    // -(|a|^2 - e2)/2 (the threshold folded in), takes -|b|^2/2 after the
    // hi*hi products, and every value it takes is at most
    // |a||b| + (|a|^2 + |b|^2)/2 + e2/2 <= 2 R^2 + e2/2 in magnitude: one
    // rounding of that per MFMA (3 kp / 16 of them) and four more for
    // (|a|^2 - e2)/2, |b|^2/2 and the addition of the column term
    const double cnt_r = (3.0 * kp / 16.0 + 4.0) * u;
    a.b1 = 2.0 * (3.1 * std::ldexp(1.0, -16) + 3.0 * kp * u) + (2.0 * kp + 16.0) * u + 2.0 * cnt_r;
    a.b2 = std::ldexp(1.0, -21) * std::sqrt((double)a.e2 * 1.01);
    a.b3 = (c.d + 3) * u * (double)a.e2 + 0.5 * cnt_r * (double)a.e2;
    return a;
}

// Partition, sort, prep, candidate lists, row tiles, then the count pass.
void mfma_classify(Ctx &c) {
    const int64_t n = c.n;
    const int d = c.d, kp = kp_for(d);
    const hipStream_t s = c.stream;
    const int nb = (int)ceil_div(n, 256);
    // centre of the data (float64 sums)
    c.misc.ensure(4096);
    double *sum = reinterpret_cast<double *>(c.misc.p);
    constexpr int kColsumGroups = 1024;
    c.mfp_colpart.ensure((size_t)kColsumGroups * d * 8);
    {
        Launch L(c.prof, s, "mf_colsum");
        k_colsum<<<kColsumGroups, 256, 0, s>>>(c.coords, n, d, c.mfp_colpart.as<double>());
        k_colsum_total<<<(unsigned)d, 256, 0, s>>>(c.mfp_colpart.as<double>(), kColsumGroups, d, sum);
    }
    std::vector<double> h_sum(d);
    NDBX_HIP(hipMemcpyAsync(h_sum.data(), sum, d * 8, hipMemcpyDeviceToHost, s));
    c.sync();
    std::vector<float> h_mu(d);
    for (int k = 0; k < d; ++k) h_mu[k] = (float)(h_sum[k] / (double)n);
    // centres: every (n/K)-th point
    const int64_t k_cent = std::max<int64_t>(1, std::min<int64_t>(n, std::max<int64_t>(1, n / kCellTarget)));
    c.flag.ensure((size_t)k_cent * d * 4 + (size_t)d * 4);
    float *cent = c.flag.as<float>();
    float *mu = cent + k_cent * d;
    NDBX_HIP(hipMemcpyAsync(mu, h_mu.data(), d * 4, hipMemcpyHostToDevice, s));
    {
        Launch L(c.prof, s, "mf_assign");
        k_pick_centres<<<(unsigned)ceil_div(k_cent * d, 256), 256, 0, s>>>(c.coords, n, d, k_cent, cent);
    }
    c.keys.ensure(n * 8);
    c.vals.ensure(n * 4);
    // the hi rows are read 32 at a time from any row up to n: 32 rows of
    // padding past the last
    c.keys2.ensure(std::max<size_t>(n * 8, (size_t)(n + 32) * kp * 2));
    c.vals2.ensure(std::max<size_t>(n * 4, (size_t)n * kp * 2));
    {
        const int k_pad = (int)ceil_div(k_cent, 32) * 32;
        c.mfp_chi.ensure((size_t)k_pad * kp * 2 + (size_t)k_pad * 4);
        __bf16 *chi = c.mfp_chi.as<__bf16>();
        float *cn = reinterpret_cast<float *>(chi + (size_t)k_pad * kp);
        const unsigned na = (unsigned)ceil_div(n, 128);  // four waves of 32 points per block
        Launch L(c.prof, s, "mf_assign");
        switch (kp) {
            case 16:
                k_centres_bf16<16><<<(unsigned)ceil_div(k_pad, 256), 256, 0, s>>>(cent, (int)k_cent, k_pad, d, mu, chi, cn);
                k_assign<16><<<na, 256, 0, s>>>(c.coords, n, d, mu, chi, cn, k_pad, c.keys.as<uint64_t>(), c.vals.as<uint32_t>());
                break;
            case 32:
                k_centres_bf16<32><<<(unsigned)ceil_div(k_pad, 256), 256, 0, s>>>(cent, (int)k_cent, k_pad, d, mu, chi, cn);
                k_assign<32><<<na, 256, 0, s>>>(c.coords, n, d, mu, chi, cn, k_pad, c.keys.as<uint64_t>(), c.vals.as<uint32_t>());
                break;
            case 64:
                k_centres_bf16<64><<<(unsigned)ceil_div(k_pad, 256), 256, 0, s>>>(cent, (int)k_cent, k_pad, d, mu, chi, cn);
                k_assign<64><<<na, 256, 0, s>>>(c.coords, n, d, mu, chi, cn, k_pad, c.keys.as<uint64_t>(), c.vals.as<uint32_t>());
                break;
            default:
                k_centres_bf16<128><<<(unsigned)ceil_div(k_pad, 256), 256, 0, s>>>(cent, (int)k_cent, k_pad, d, mu, chi, cn);
                k_assign<128><<<na, 256, 0, s>>>(c.coords, n, d, mu, chi, cn, k_pad, c.keys.as<uint64_t>(), c.vals.as<uint32_t>());
                break;
        }
    }
    int bits = 1;
    while ((1ll << bits) < k_cent) ++bits;
    bool alt = false;
    radix_sort_u64(c.keys.as<uint64_t>(), c.vals.as<uint32_t>(), c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>(), n,
                   bits, c.scratch, s, c.prof, &alt);
    if (alt) {
        std::swap(c.keys, c.keys2);
        std::swap(c.vals, c.vals2);
        c.keys2.ensure((size_t)(n + 32) * kp * 2);
        c.vals2.ensure((size_t)n * kp * 2);
    }
    const uint32_t *perm = c.vals.as<uint32_t>();
    // cells = runs of equal keys
    c.cid.ensure(n * 4);
    c.cell_of.ensure(n * 4);
    c.cell_start.ensure((n + 1) * 4);
    c.cell_key.ensure(n * 8);
    c.gflag.ensure(n * 4);
    int32_t *d_total = c.misc.as<int32_t>() + 600;
    cell_table(c.keys.as<uint64_t>(), n, 0, c.gflag.as<int32_t>(), c.cid.as<int32_t>(), c.cell_of.as<int32_t>(),
               c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), d_total, c.scratch, s, c.prof);
    int32_t h_total = 0;
    NDBX_HIP(hipMemcpyAsync(&h_total, d_total, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    const int64_t nc = h_total;
    // within each cell, points in ascending distance to the cell's centroid
    // (the block pruning's tiles: small radii except for a cell's outliers)
    c.mfp_cent.ensure((size_t)nc * d * 4);
    {
        Launch L(c.prof, s, "mf_prune");
        k_cell_centroid<<<(unsigned)nc, 1024, 0, s>>>(c.coords, perm, c.cell_start.as<int32_t>(), nc, d,
                                                     c.mfp_cent.as<float>());
        const bool vin = (d & 3) == 0 && ((uintptr_t)c.coords & 15) == 0;  // 16-byte rows of the input
        k_radius_keys<<<(unsigned)ceil_div(n * 16, 256), 256, 0, s>>>(c.coords, perm, c.keys.as<uint64_t>(),
                                                                       c.cell_of.as<int32_t>(), c.mfp_cent.as<float>(),
                                                                       n, d, vin, c.keys2.as<uint64_t>(),
                                                                       c.vals2.as<uint32_t>());
    }
    bool alt2 = false;
    radix_sort_u64(c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>(), c.keys.as<uint64_t>(), c.vals.as<uint32_t>(), n,
                   bits + 32, c.scratch, s, c.prof, &alt2);
    if (!alt2) {  // the order lands in keys / vals, the hi / lo rows in keys2 / vals2
        std::swap(c.keys, c.keys2);
        std::swap(c.vals, c.vals2);
        c.keys2.ensure(std::max<size_t>(n * 8, (size_t)(n + 32) * kp * 2));
        c.vals2.ensure(std::max<size_t>(n * 4, (size_t)n * kp * 2));
    }
    perm = c.vals.as<uint32_t>();
    // sorted rows, hi/lo, norms
    c.soa.ensure((size_t)n * d * 4);
    c.soa_core.ensure(((size_t)n * 2 + 32) * 4);  // norms, radii, 32 floats read past them
    {
        Launch L(c.prof, s, "mf_prep");
        __bf16 *hi = c.keys2.as<__bf16>(), *lo = reinterpret_cast<__bf16 *>(c.vals2.p);
        float *na = c.soa_core.as<float>(), *rho = c.soa_core.as<float>() + n;
        const bool vec = (d & 3) == 0 && ((uintptr_t)c.coords & 15) == 0;
        switch (kp) {
            case 16: k_mfma_prep<16><<<nb, 256, 0, s>>>(c.coords, perm, n, d, vec, mu, c.soa.as<float>(), hi, lo, na, rho); break;
            case 32: k_mfma_prep<32><<<nb, 256, 0, s>>>(c.coords, perm, n, d, vec, mu, c.soa.as<float>(), hi, lo, na, rho); break;
            case 64: k_mfma_prep<64><<<nb, 256, 0, s>>>(c.coords, perm, n, d, vec, mu, c.soa.as<float>(), hi, lo, na, rho); break;
            default: k_mfma_prep<128><<<nb, 256, 0, s>>>(c.coords, perm, n, d, vec, mu, c.soa.as<float>(), hi, lo, na, rho); break;
        }
    }
    // the padding rows hold zeros (their columns are out of range wherever
    // they are read)
    fill_u64(reinterpret_cast<uint64_t *>(c.keys2.as<__bf16>() + (size_t)n * kp), 0, 8 * kp, s, c.prof);
    fill_u64(reinterpret_cast<uint64_t *>(c.soa_core.as<float>() + 2 * n), 0, 16, s, c.prof);
    c.sync();
    c.n_cells = nc;
    const int ncb = (int)ceil_div(nc, 256);
    // cell radii and largest |a'|
    c.b_rootmin.ensure(nc * 8);
    c.cell_c.ensure(nc * 4);
    {
        Launch L(c.prof, s, "mf_cell_radius");
        k_cell_radius<<<(unsigned)nc, 256, 0, s>>>(c.soa.as<float>(), c.soa_core.as<float>() + n,
                                                   c.cell_start.as<int32_t>(), c.cell_key.as<uint64_t>(), nc, d, cent,
                                                   reinterpret_cast<double *>(c.b_rootmin.p), c.cell_c.as<float>());
    }
    // candidate cell pairs (CSR)
    const float e = c.h_eps[0];
    const float e2 = e * e;
    const double r = std::sqrt((double)e2 * (1.0 + std::ldexp(1.0, -18)));
    c.nranges.ensure((nc + 1) * 4);
    c.tile_off.ensure((nc + 1) * 4);
    {
        Launch L(c.prof, s, "mf_cell_pairs");
        k_cell_pairs<<<(unsigned)nc, 256, 0, s>>>(cent, c.cell_key.as<uint64_t>(),
                                                  reinterpret_cast<double *>(c.b_rootmin.p), nc, d, r, nullptr,
                                                  c.nranges.as<int32_t>(), 0, nullptr, nullptr);
    }
    // the pair list has int32 offsets: past mfma_safe_cells() cells the total
    // can exceed them, so it is summed in 64 bits and checked first
    if (nc > mfma_safe_cells()) {
        unsigned long long *d_sum = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 624);
        sum_i32_u64(c.nranges.as<int32_t>(), nc, d_sum, s, c.prof);
        unsigned long long h_sum = 0;
        NDBX_HIP(hipMemcpyAsync(&h_sum, d_sum, 8, hipMemcpyDeviceToHost, s));
        c.sync();
        if (!mfma_pairs_fit((int64_t)h_sum))
            throw Error(NDBX_EUNSUPPORTED, "MFMA filter: " + std::to_string(h_sum) + " candidate cell pairs exceed "
                                           "the int32 pair list (" + std::to_string(kMfmaMaxPairs) +
                                           "); use a smaller eps or fewer points per run");
    }
    scan_exclusive_i32(c.nranges.as<int32_t>(), c.tile_off.as<int32_t>(), nc, c.tile_off.as<int32_t>() + nc,
                       c.scratch, s, c.prof);
    int32_t h_pairs = 0;
    NDBX_HIP(hipMemcpyAsync(&h_pairs, c.tile_off.as<int32_t>() + nc, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.ranges.ensure((size_t)h_pairs * 4 + 4);
    c.pflag.ensure((size_t)h_pairs + 4);
    fill_i32(c.pflag.as<int32_t>(), 0, ceil_div(h_pairs, 4), s, c.prof);
    unsigned long long *d_pts = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 620);
    fill_u64(reinterpret_cast<uint64_t *>(d_pts), 0, 1, s, c.prof);
    {
        Launch L(c.prof, s, "mf_cell_pairs");
        k_cell_pairs<<<(unsigned)nc, 256, 0, s>>>(cent, c.cell_key.as<uint64_t>(),
                                                  reinterpret_cast<double *>(c.b_rootmin.p), nc, d, r,
                                                  c.tile_off.as<int32_t>(), c.ranges.as<int32_t>(), 1,
                                                  c.cell_start.as<int32_t>(), d_pts);
    }
    unsigned long long h_pts = 0;
    NDBX_HIP(hipMemcpyAsync(&h_pts, d_pts, 8, hipMemcpyDeviceToHost, s));
    // row tiles
    c.ctile_off.ensure((nc + 1) * 4);
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_row_tile_count<<<ncb, 256, 0, s>>>(c.cell_start.as<int32_t>(), nc, c.nranges.as<int32_t>());
    }
    // the total lands in ctile_off[nc], so a cell's tiles are [toff[c], toff[c+1])
    scan_exclusive_i32(c.nranges.as<int32_t>(), c.ctile_off.as<int32_t>(), nc, c.ctile_off.as<int32_t>() + nc,
                       c.scratch, s, c.prof);
    NDBX_HIP(hipMemcpyAsync(&h_total, c.ctile_off.as<int32_t>() + nc, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.n_tiles = h_total;
    c.ctiles.ensure(c.n_tiles * 16 + 16);
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_row_tiles<<<ncb, 256, 0, s>>>(c.cell_start.as<int32_t>(), c.ctile_off.as<int32_t>(), nc, c.ctiles.as<int4>());
    }
    // work groups of up to 4 row tiles of one cell
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_group_count<<<ncb, 256, 0, s>>>(c.ctile_off.as<int32_t>(), nc, 4, c.nranges.as<int32_t>());
    }
    c.cid.ensure((nc + 1) * 4);
    scan_exclusive_i32(c.nranges.as<int32_t>(), c.cid.as<int32_t>(), nc, d_total, c.scratch, s, c.prof);
    NDBX_HIP(hipMemcpyAsync(&h_total, d_total, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.n_label_tiles = h_total;
    c.tiles.ensure(c.n_label_tiles * 16 + 16);
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_groups<<<ncb, 256, 0, s>>>(c.ctile_off.as<int32_t>(), nc, c.cid.as<int32_t>(), 4, c.tiles.as<int4>());
    }
    // count-pass groups: up to kCountWaves row tiles of one cell (the label
    // passes' groups when that is 4)
    int64_t n_count_groups = c.n_label_tiles;
    const int4 *count_groups = c.tiles.as<int4>();
    if (kCountWaves != 4) {
        {
            Launch L(c.prof, s, "mf_row_tiles");
            k_group_count<<<ncb, 256, 0, s>>>(c.ctile_off.as<int32_t>(), nc, kCountWaves, c.nranges.as<int32_t>());
        }
        scan_exclusive_i32(c.nranges.as<int32_t>(), c.cid.as<int32_t>(), nc, d_total, c.scratch, s, c.prof);
        NDBX_HIP(hipMemcpyAsync(&h_total, d_total, 4, hipMemcpyDeviceToHost, s));
        c.sync();
        n_count_groups = h_total;
        c.cgroups.ensure(n_count_groups * 16 + 16);
        {
            Launch L(c.prof, s, "mf_row_tiles");
            k_groups<<<ncb, 256, 0, s>>>(c.ctile_off.as<int32_t>(), nc, c.cid.as<int32_t>(), kCountWaves,
                                         c.cgroups.as<int4>());
        }
        count_groups = c.cgroups.as<int4>();
    }
    c.pair_tests = (int64_t)h_pts;  // point pairs of the candidate cell pairs (read after the syncs above)
    // block pruning: per tile, its largest radius and the bounds of its
    // distance to every cell's centroid; per (count group, candidate entry),
    // the column tiles not pruned.  Left out when the bounds would take more
    // than kPruneBytes or the mask words overflow int32 offsets.
    const int64_t nt = c.n_tiles;
    bool prune = nt > 0 && (size_t)nt * (size_t)nc * 8 <= kPruneBytes;
    if (prune) {
        c.mfp_rad.ensure(n * 4);
        c.mfp_trad.ensure(nt * 4);
        c.mfp_trho.ensure(nt * 4);
        c.mfp_tab.ensure((size_t)nt * nc * 4);
        c.mfp_tcm.ensure((size_t)nt * nc * 4);
        c.mfp_moff.ensure((n_count_groups + 1) * 4);
        Launch L(c.prof, s, "mf_prune");
        k_point_radius<<<(unsigned)ceil_div(n * 16, 256), 256, 0, s>>>(
            c.soa.as<float>(), c.cell_of.as<int32_t>(), c.mfp_cent.as<float>(), n, d, (d & 3) == 0,
            c.mfp_rad.as<float>());
        // the centroids centred and split as the points, then the bounds
        c.mfp_mhi.ensure((size_t)nc * kp * 2);
        c.mfp_mlo.ensure((size_t)nc * kp * 2);
        c.mfp_mna.ensure((size_t)nc * 8);
        __bf16 *mhi = c.mfp_mhi.as<__bf16>(), *mlo = c.mfp_mlo.as<__bf16>();
        float *cna = c.mfp_mna.as<float>(), *crho = cna + nc;
        const __bf16 *phi = c.keys2.as<__bf16>(), *plo = reinterpret_cast<const __bf16 *>(c.vals2.p);
        const float *pna = c.soa_core.as<float>(), *prho = c.soa_core.as<float>() + n;
        const unsigned gc = (unsigned)ceil_div(nc, 256), gt = (unsigned)ceil_div(nt, 4);
#define NDBX_TILE_DIST(KPV)                                                                                    \
    k_centroid_prep<KPV><<<gc, 256, 0, s>>>(c.mfp_cent.as<float>(), nc, d, mu, mhi, mlo, cna, crho);           \
    k_tile_dist<KPV><<<gt, 256, 0, s>>>(phi, plo, pna, prho, c.ctiles.as<int4>(), nt, mhi, mlo, cna, crho, nc, \
                                        c.mfp_rad.as<float>(), c.mfp_tab.as<float>(), c.mfp_trad.as<float>(),  \
                                        c.mfp_trho.as<float>())
        switch (kp) {
            case 16: NDBX_TILE_DIST(16); break;
            case 32: NDBX_TILE_DIST(32); break;
            case 64: NDBX_TILE_DIST(64); break;
            default: NDBX_TILE_DIST(128); break;
        }
#undef NDBX_TILE_DIST
        k_transpose_f32<<<dim3((unsigned)ceil_div(nc, 32), (unsigned)ceil_div(nt, 32)), 256, 0, s>>>(
            c.mfp_tab.as<float>(), nt, nc, c.mfp_tcm.as<float>());
        k_group_entries<<<(unsigned)ceil_div(n_count_groups, 256), 256, 0, s>>>(
            count_groups, n_count_groups, c.tile_off.as<int32_t>(), c.mfp_moff.as<int32_t>());
    }
    int32_t h_words = 0;
    if (prune) {
        int32_t *d_words = c.misc.as<int32_t>() + 604;
        // (mask words per group are at most nc, so the scan of int32 counts
        // cannot wrap before the total is checked: its sum is bounded by
        // n_count_groups * nc, checked in 64 bits first)
        if ((int64_t)n_count_groups * nc >= INT32_MAX) prune = false;
        else {
            scan_exclusive_i32(c.mfp_moff.as<int32_t>(), c.mfp_moff.as<int32_t>(), n_count_groups, d_words, c.scratch,
                               s, c.prof);
            NDBX_HIP(hipMemcpyAsync(&h_words, d_words, 4, hipMemcpyDeviceToHost, s));
            c.sync();
        }
    }
    if (prune) {
        c.mfp_mask.ensure((size_t)h_words * 8 + 8);
        c.mfp_amask.ensure((size_t)h_words * 8 + 8);
        c.mfp_emask.ensure((size_t)h_words * 8 + 8);
        fill_u64(c.mfp_emask.as<uint64_t>(), 0, h_words, s, c.prof);
        c.mfp_tacc.ensure((size_t)nt * 4);
        fill_i32(c.mfp_tacc.as<int32_t>(), 0, nt, s, c.prof);
        unsigned long long *d_kept = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 608);
        fill_u64(reinterpret_cast<uint64_t *>(d_kept), 0, 1, s, c.prof);
        const double rr = std::sqrt((double)e2 * (1.0 + std::ldexp(1.0, -18)));
        c.mfp_dab.ensure((size_t)h_pairs * 8 + 8);
        {
            Launch L(c.prof, s, "mf_prune");
            k_pair_cdist<<<(unsigned)nc, 256, 0, s>>>(c.tile_off.as<int32_t>(), c.ranges.as<int32_t>(), nc,
                                                      c.mfp_cent.as<float>(), d, c.mfp_dab.as<double>());
            k_block_masks<<<(unsigned)n_count_groups, 256, 0, s>>>(
                count_groups, c.tile_off.as<int32_t>(), c.ranges.as<int32_t>(), c.ctile_off.as<int32_t>(),
                c.ctiles.as<int4>(), c.mfp_tcm.as<float>(), c.mfp_trad.as<float>(), nt, c.mfp_dab.as<double>(), d,
                rr, (double)e2, c.mfp_moff.as<int32_t>(), c.mfp_mask.as<uint64_t>(), c.mfp_amask.as<uint64_t>(),
                c.mfp_tacc.as<int32_t>(),
                c.pflag.as<uint8_t>(), d_kept);
        }
        unsigned long long h_kept = 0;
        NDBX_HIP(hipMemcpyAsync(&h_kept, d_kept, 8, hipMemcpyDeviceToHost, s));
        c.sync();
        c.pair_tests = (int64_t)h_kept;  // point pairs of the blocks the count pass computes
    }
    c.pruned = prune;
    // count
    c.nn_s.ensure(n * 4);
    MfmaArgs a = mfma_args(c);
    a.nn_s = c.nn_s.as<int32_t>();
    fill_i32(a.nn_s, 0, n, s, c.prof);
    // guard-band queue: sized by the last run's count (pairs past the
    // capacity are decided in the count pass itself)
    c.bq_last = (int64_t)c.bq_last64;  // copied back after the previous run's count pass
    c.bq_cap = std::max<int64_t>(c.bq_cap, std::max<int64_t>(1 << 20, 8 * n));
    if (c.bq_last > c.bq_cap) c.bq_cap = c.bq_last + c.bq_last / 4;
    c.bq_cap = std::min<int64_t>(c.bq_cap, 1ll << 30);
    if (c.bq_fixed > 0) c.bq_cap = c.bq_fixed;
    c.bq.ensure((size_t)c.bq_cap * 16);
    a.bq = c.bq.as<int4>();
    a.bq_count = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 660);
    a.bq_cap = c.bq_cap;
    fill_u64(reinterpret_cast<uint64_t *>(a.bq_count), 0, 1, s, c.prof);
    {
        MfmaArgs ac = a;
        ac.groups = count_groups;
        ac.n_groups = n_count_groups;
        if (prune) {
            ac.bmask = c.mfp_mask.as<uint64_t>();
            ac.bmask_off = c.mfp_moff.as<int32_t>();
            ac.emask = c.mfp_emask.as<uint64_t>();
            ac.tile_rho = c.mfp_trho.as<float>();
        }
        ac.ctoff = c.ctile_off.as<int32_t>();
        c.mfp_tband.ensure((size_t)std::max<int64_t>(nt, 1) * 8);
        {
            Launch L(c.prof, s, "mf_prune");
            k_tile_band<<<(unsigned)ceil_div(std::max<int64_t>(nt, 1), 256), 256, 0, s>>>(
                c.ctiles.as<int4>(), nt, prune ? c.mfp_trho.as<float>() : nullptr, a.cell_rmax, a.b1, a.b2, a.b3,
                c.mfp_tband.as<float2>());
        }
        ac.tile_band = c.mfp_tband.as<float2>();
        launch_count(c, ac, kp);
    }
    if (prune) {  // the tiles counted wholly within eps
        Launch L(c.prof, s, "mf_prune");
        k_tile_acc<<<nb, 256, 0, s>>>(c.mfp_tacc.as<int32_t>(), c.cell_of.as<int32_t>(), c.cell_start.as<int32_t>(),
                                      c.ctile_off.as<int32_t>(), n, a.nn_s);
    }
    {
        Launch L(c.prof, s, "mf_band");
        k_mf_band<<<2048, 256, 0, s>>>(a.bq, a.bq_count, a.bq_cap, a.xs, a.d, a.e2, a.nn_s, a.pflag);
    }
    NDBX_HIP(hipMemcpyAsync(&c.bq_last64, a.bq_count, 8, hipMemcpyDeviceToHost, s));
    c.core_s.ensure(n);
    c.label_s.ensure(n * 4);
    c.parent.ensure(n * 4);
    {
        Launch L(c.prof, s, "mf_core");
        k_mf_core<<<nb, 256, 0, s>>>(c.nn_s.as<int32_t>(), perm, n, c.h_m[0], c.nn.as<int32_t>(), c.core.as<uint8_t>(),
                                     c.core_s.as<uint8_t>(), c.label_s.as<int32_t>(), c.parent.as<int32_t>());
    }
    if (c.flags & NDBX_FP64_AUDIT) {
        // stats: band pairs, fp64 flips, filter flips, max |est - fp64|
        unsigned long long *au = reinterpret_cast<unsigned long long *>(c.misc.as<int32_t>() + 640);
        fill_u64(reinterpret_cast<uint64_t *>(au), 0, 4, s, c.prof);
        a.audit = au;
        a.audit_err = reinterpret_cast<float *>(au + 3);
        // every audit_stride-th group (about 512 groups, or every
        // NDBX_AUDIT_STRIDE-th from NDBX_AUDIT_OFFSET), with the count pass's band: R from the
        // column tile's own largest norm when the tile bounds were built
        a.audit_stride = c.audit_stride > 0 ? c.audit_stride : (int)std::max<int64_t>(1, c.n_label_tiles / 512);
        a.audit_offset = c.audit_offset < a.audit_stride ? c.audit_offset : 0;
        if (c.pruned) {
            a.tile_rho = c.mfp_trho.as<float>();
            a.ctoff = c.ctile_off.as<int32_t>();
        }
        launch_pass<3>(c, a, kp, "mf_audit");
        a.tile_rho = nullptr;
        a.ctoff = nullptr;
        unsigned long long h[4];
        NDBX_HIP(hipMemcpyAsync(h, au, 32, hipMemcpyDeviceToHost, s));
        c.sync();
        c.audit_band = (int64_t)h[0];
        c.audit_fp64_flips = (int64_t)h[1];
        c.audit_filter_flips = (int64_t)h[2];
        float f;
        std::memcpy(&f, &h[3], 4);
        c.audit_max_err = f;
    }
}

void mfma_label(Ctx &c) {
    const int64_t n = c.n;
    const int kp = kp_for(c.d);
    const hipStream_t s = c.stream;
    const int nb = (int)ceil_div(n, 256);
    fill_i32(c.meta.as<int32_t>(), 0, 3, s, c.prof);
    // core rank in input order (label_s holds core flags as int32 here)
    c.rank.ensure((n + 1) * 4);
    scan_exclusive_i32(c.label_s.as<int32_t>(), c.rank.as<int32_t>(), n, c.misc.as<int32_t>() + 604, c.scratch, s,
                       c.prof);
    MfmaArgs a = mfma_args(c);
    a.core_s = c.core_s.as<uint8_t>();
    a.parent = c.parent.as<int32_t>();
    if (c.pruned) {
        // the tiles counted wholly within eps: their cores united without
        // products; the union passes then visit only the count pass's tiles
        const int64_t nt = c.n_tiles, ng = c.n_label_tiles;
        c.mfp_trep.ensure((size_t)nt * 8);
        int32_t *rep = c.mfp_trep.as<int32_t>(), *tflag = rep + nt;
        Launch L(c.prof, s, "mf_allin_union");
        k_tile_rep<<<(unsigned)ceil_div(nt, 256), 256, 0, s>>>(c.ctiles.as<int4>(), nt, c.core_s.as<uint8_t>(), rep,
                                                               tflag);
        k_allin_union<<<(unsigned)ng, 256, 0, s>>>(c.tiles.as<int4>(), c.tile_off.as<int32_t>(), c.ranges.as<int32_t>(),
                                                   c.ctile_off.as<int32_t>(), c.mfp_moff.as<int32_t>(),
                                                   c.mfp_amask.as<uint64_t>(), rep, tflag, c.parent.as<int32_t>());
        k_tile_unite<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.cell_of.as<int32_t>(), c.cell_start.as<int32_t>(),
                                        c.ctile_off.as<int32_t>(), n, rep, tflag, c.parent.as<int32_t>());
        a.bmask = c.mfp_emask.as<uint64_t>();  // the tiles holding a pair within eps (or a band element)
        a.bmask_off = c.mfp_moff.as<int32_t>();
        a.tile_rho = c.mfp_trho.as<float>();
        a.ctoff = c.ctile_off.as<int32_t>();
    }
    // union: the cells' own pairs, then the other flagged pairs, skipped
    // while both cells' cores already share one set
    a.union_phase = 0;
    launch_pass<1>(c, a, kp, "mf_union_own");
    const int64_t nc = c.n_cells;
    c.b_parent.ensure((size_t)nc * 4 + 4);
    {
        Launch L(c.prof, s, "mf_cell_rep");
        k_mf_cell_rep<<<(unsigned)nc, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.cell_start.as<int32_t>(), nc,
                                                   c.parent.as<int32_t>(), c.b_parent.as<int32_t>());
    }
    a.union_phase = 1;
    a.cell_rep = c.b_parent.as<int32_t>();
    launch_pass<1>(c, a, kp, "mf_union");
    // the border pass's groups are its own, and ctile_off is reused below
    a.bmask = nullptr;
    a.bmask_off = nullptr;
    a.tile_rho = nullptr;
    a.ctoff = nullptr;
    c.comp_min.ensure(n * 4);
    fill_i32(c.comp_min.as<int32_t>(), INT_MAX, n, s, c.prof);
    {
        Launch L(c.prof, s, "mf_label");
        k_mf_rootmin<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.vals.as<uint32_t>(), n, c.parent.as<int32_t>(),
                                        c.comp_min.as<int32_t>());
        k_mf_label<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), n, c.parent.as<int32_t>(), c.comp_min.as<int32_t>(),
                                      c.rank.as<int32_t>(), c.label_s.as<int32_t>(), c.meta.as<int32_t>());
    }
    c.cell_label.ensure(n * 4);
    fill_i32(c.cell_label.as<int32_t>(), NDBX_NO_CLUSTER, n, s, c.prof);
    // border rows: the non-core points, compacted (still grouped by cell)
    const int ncb = (int)ceil_div(nc, 256);
    c.flag.ensure((n + 1) * 4);
    c.cid.ensure((n + 1) * 4);
    c.cell_of.ensure(n * 4);
    int32_t *bpos = c.cid.as<int32_t>();
    {
        Launch L(c.prof, s, "mf_border_rows");
        k_noncore_flag<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.nn_s.as<int32_t>(), n, c.flag.as<int32_t>());
    }
    scan_exclusive_i32(c.flag.as<int32_t>(), bpos, n, bpos + n, c.scratch, s, c.prof);
    int32_t *btoff = c.ctile_off.as<int32_t>(), *bgoff = c.nranges.as<int32_t>();
    int32_t *bcnt = c.gflag.as<int32_t>();
    {
        Launch L(c.prof, s, "mf_border_rows");
        k_noncore_list<<<nb, 256, 0, s>>>(c.flag.as<int32_t>(), bpos, n, c.cell_of.as<int32_t>());
        k_btile_count<<<ncb, 256, 0, s>>>(bpos, c.cell_start.as<int32_t>(), nc, bcnt);
    }
    scan_exclusive_i32(bcnt, btoff, nc, btoff + nc, c.scratch, s, c.prof);
    {
        Launch L(c.prof, s, "mf_border_rows");
        k_group_count<<<ncb, 256, 0, s>>>(btoff, nc, 4, bcnt);
    }
    scan_exclusive_i32(bcnt, bgoff, nc, bgoff + nc, c.scratch, s, c.prof);
    int32_t h_t[2] = {0, 0};
    NDBX_HIP(hipMemcpyAsync(&h_t[0], btoff + nc, 4, hipMemcpyDeviceToHost, s));
    NDBX_HIP(hipMemcpyAsync(&h_t[1], bgoff + nc, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.keys.ensure((size_t)(h_t[0] + h_t[1]) * 16 + 32);
    int4 *btiles = reinterpret_cast<int4 *>(c.keys.p), *bgroups = btiles + h_t[0];
    {
        Launch L(c.prof, s, "mf_border_rows");
        k_btiles<<<ncb, 256, 0, s>>>(bpos, c.cell_start.as<int32_t>(), btoff, nc, btiles);
        k_groups<<<ncb, 256, 0, s>>>(btoff, nc, bgoff, 4, bgroups);
    }
    a.label_s = c.label_s.as<int32_t>();
    a.best_s = c.cell_label.as<int32_t>();
    a.rows = c.cell_of.as<int32_t>();
    a.tiles = btiles;
    a.n_tiles = h_t[0];
    a.groups = bgroups;
    a.n_groups = h_t[1];
    launch_pass<2>(c, a, kp, "mf_border");
    Launch L(c.prof, s, "meta");
    k_mf_out<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.label_s.as<int32_t>(), c.cell_label.as<int32_t>(),
                                c.vals.as<uint32_t>(), n, c.labels.as<int32_t>(), c.meta.as<int32_t>());
}

}  // namespace ndbx
