// MFMA distance filter for d >= 16 (d <= 128): neighbour counts, union-find
// and border attachment on 32x32 tiles of bf16 matrix products.
//
// 1. Partition: K sampled points are cell centres; every point joins its
//    nearest centre (float32 distances, lowest index on ties); a radix sort by
//    cell makes each cell a contiguous range.  A cell's radius R_C (float64,
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
//       B1 = 2 (3.1 2^-16 + 3 Kp 2^-24) + (2 Kp + 8) 2^-24,
//    R = largest |a'| of the tile's rows and the candidate cell.  Pairs with
//    est <= e2 - band are neighbours, est > e2 + band are not, and the rest
//    are decided by the reference's own float32 arithmetic on the original
//    coordinates.  Every decision therefore equals the float32 predicate.
#include <cfloat>
#include <climits>

#include "ndbx_internal.h"

namespace ndbx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kCellTarget = 1024;  // points per partition cell, on average

// ---------------------------------------------------------------- prep --
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ x, int64_t n, int d,
                                                double *__restrict__ sum) {
    const int64_t rows_per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = blockIdx.x * rows_per, e = b + rows_per < n ? b + rows_per : n;
    for (int k = threadIdx.x; k < d; k += blockDim.x) {
        double s = 0.0;
        for (int64_t i = b; i < e; ++i) s += x[i * d + k];
        if (e > b) atomicAdd(sum + k, s);
    }
}

// Nearest centre (float32 distance, lowest index on ties), the point held in
// registers (zero-padded to KP); centres staged through LDS 64 at a time.
template <int KP>
__global__ __launch_bounds__(256) void k_assign(const float *__restrict__ x, int64_t n, int d,
                                                const float *__restrict__ cent, int k_cent,
                                                uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    __shared__ float sc[64 * KP];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = i < n;
    float p[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) p[k] = (v && k < d) ? x[i * d + k] : 0.0f;
    float best = __builtin_inff();
    int arg = 0;
    for (int c0 = 0; c0 < k_cent; c0 += 64) {
        const int cn = k_cent - c0 < 64 ? k_cent - c0 : 64;
        __syncthreads();
        for (int t = threadIdx.x; t < 64 * KP; t += blockDim.x) {
            const int cc = t / KP, k = t - cc * KP;
            sc[t] = (cc < cn && k < d) ? cent[(int64_t)(c0 + cc) * d + k] : 0.0f;
        }
        __syncthreads();
        for (int c = 0; c < cn; ++c) {
            float sum = 0.0f;
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const float df = p[k] - sc[c * KP + k];
                sum = __fadd_rn(sum, __fmul_rn(df, df));
            }
            if (sum < best) {
                best = sum;
                arg = c0 + c;
            }
        }
    }
    if (v) {
        keys[i] = (uint64_t)arg;
        vals[i] = (uint32_t)i;
    }
}

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }  // round to nearest even

// Sorted float32 rows, centred bf16 hi/lo rows padded to kp, |a'|^2 in
// float32 (ascending k) and |a'|.
__global__ __launch_bounds__(256) void k_mfma_prep(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                   int64_t n, int d, int kp, const float *__restrict__ mu,
                                                   float *__restrict__ xs, __bf16 *__restrict__ hi,
                                                   __bf16 *__restrict__ lo, float *__restrict__ na,
                                                   float *__restrict__ rho) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    float s = 0.0f;
    for (int k = 0; k < kp; ++k) {
        float a = 0.0f;
        if (k < d) {
            const float v = x[p * d + k];
            xs[i * d + k] = v;
            a = v - mu[k];
        }
        const __bf16 h = to_bf16(a);
        const __bf16 l = to_bf16(a - (float)h);
        hi[i * kp + k] = h;
        lo[i * kp + k] = l;
        s = __fadd_rn(s, __fmul_rn(a, a));
    }
    na[i] = s;
    rho[i] = sqrtf(s) * (1.0f + 1e-6f);
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
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < d; ++k) {
            const double df = (double)xs[(int64_t)i * d + k] - (double)cc[k];
            s += df * df;
        }
        r = fmax(r, s);
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

// Candidate cell pairs: |centre_C - centre_C'| <= R_C + R_C' + r (+ slack).
// Pass 0 counts per cell, pass 1 writes (in no particular order: every use is
// order-independent).
__global__ __launch_bounds__(256) void k_cell_pairs(const float *__restrict__ cent, const uint64_t *__restrict__ cell_key,
                                                    const double *__restrict__ rad, int64_t nc, int d, double r,
                                                    const int32_t *__restrict__ off, int32_t *__restrict__ cnt_or_list,
                                                    int pass) {
    const int64_t c = (int64_t)blockIdx.x;
    if (c >= nc) return;
    const float *a = cent + (int64_t)cell_key[c] * d;
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int64_t o = threadIdx.x; o < nc; o += blockDim.x) {
        const float *b = cent + (int64_t)cell_key[o] * d;
        double s = 0.0;
        for (int k = 0; k < d; ++k) {
            const double df = (double)a[k] - (double)b[k];
            s += df * df;
        }
        const double lim = rad[c] + rad[o] + r;
        if (s <= lim * lim * (1.0 + 1e-9)) {
            if (pass == 0) atomicAdd(&cnt, 1);
            else cnt_or_list[off[c] + atomicAdd(&cnt, 1)] = (int32_t)o;
        }
    }
    __syncthreads();
    if (pass == 0 && threadIdx.x == 0) cnt_or_list[c] = cnt;
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
__device__ void mf_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = mf_find(parent, a);
        b = mf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + b, b, a) == b) return;
    }
}

// The reference's predicate on original coordinates (data_process.cpp:443-447).
__device__ __forceinline__ bool exact_pred(const float *__restrict__ p, const float *__restrict__ q, int d, float e2) {
    float len = 0.0f;
    for (int k = 0; k < d; ++k) {
        const float df = p[k] - q[k];
        len = __fadd_rn(len, __fmul_rn(df, df));
    }
    return len <= e2;
}

struct MfmaArgs {
    const int4 *tiles;
    int64_t n_tiles;
    const int32_t *cand_off, *cand, *cell_start;
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
    unsigned long long *audit;        // [band pairs, fp64 flips, filter flips], pass 3
    float *audit_err;                 // max |est - fp64 len|, pass 3
    int audit_stride;
};

// PASS 0: neighbour counts.  PASS 1: unite core pairs (col > row).  PASS 2:
// smallest core cluster id within eps for non-core rows.  PASS 3: audit of
// every estimate of every audit_stride-th tile against float32 and float64.
template <int KP, int PASS>
__global__ __launch_bounds__(256) void k_mfma_pass(MfmaArgs A) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (PASS == 3) t *= A.audit_stride;
    if (t >= A.n_tiles) return;
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int4 tile = A.tiles[t];
    const int cell = tile.x, row0 = tile.y, nrows = tile.z;
    constexpr int KS = KP / 16;
    const bf16x8_t *hi8 = reinterpret_cast<const bf16x8_t *>(A.hi);
    const bf16x8_t *lo8 = reinterpret_cast<const bf16x8_t *>(A.lo);
    // A operand: row `col` of the tile, k elements [16 ks + 8 half, +8)
    const bool arow_ok = col < nrows;
    const int64_t arow = row0 + (arow_ok ? col : 0);
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
    int32_t acc_i[16];
    uint32_t rvalid = 0, rcore = 0;
    int32_t root_r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int tr = 8 * (r >> 2) + 4 * half + (r & 3);
        const bool ok = tr < nrows;
        const int64_t p = row0 + (ok ? tr : 0);
        na_r[r] = A.na[p];
        rvalid |= (uint32_t)ok << r;
        acc_i[r] = PASS == 2 ? NDBX_NO_CLUSTER : 0;
        root_r[r] = 0;
        if (PASS == 1 || PASS == 2) {
            const bool c = ok && A.core_s[p];
            rcore |= (uint32_t)c << r;
            if (PASS == 1 && c) root_r[r] = mf_find(A.parent, (int32_t)p);
        }
    }
    if (PASS == 2 && (__ballot(rvalid & ~rcore) == 0)) return;  // no non-core row
    if (PASS == 1 && (__ballot(rcore) == 0)) return;
    float rmax_row = A.rho[arow];
    for (int o = 32; o >= 1; o >>= 1) rmax_row = fmaxf(rmax_row, __shfl_xor(rmax_row, o, 64));
    rmax_row = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(rmax_row)));

    float amax = 0.0f;  // audit: largest |est - float64 distance^2|
    const int c_beg = A.cand_off[cell], c_end = A.cand_off[cell + 1];
    for (int ci = c_beg; ci < c_end; ++ci) {
        const int cc = A.cand[ci];
        const double R = fmax((double)rmax_row, (double)A.cell_rmax[cc]);
        const float band = (float)(2.0 * (A.b1 * R * R + A.b2 * R + A.b3));
        const float thr_in = A.e2 - band, thr_out = A.e2 + band;
        const int cs = A.cell_start[cc], ce = A.cell_start[cc + 1];
        for (int cb = cs; cb < ce; cb += 32) {
            const bool bcol_ok = cb + col < ce;
            const int64_t bcol = bcol_ok ? cb + col : cs;
            if (PASS == 1 && cb + 31 < row0) continue;  // all columns below every row
            f32x16_t acc = {};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                bf16x8_t bhi = hi8[bcol * (KP / 8) + ks * 2 + half];
                bf16x8_t blo = lo8[bcol * (KP / 8) + ks * 2 + half];
                if (!bcol_ok) {
                    bhi = bf16x8_t{};
                    blo = bf16x8_t{};
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], bhi, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[ks], blo, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[ks], bhi, acc, 0, 0, 0);
            }
            const float nb = A.na[bcol];
            bool col_core = false;
            int32_t col_label = 0, col_root = 0;
            if (PASS == 1 || PASS == 2) col_core = bcol_ok && A.core_s[bcol];
            if (PASS == 2 && col_core) col_label = A.label_s[bcol];
            if (PASS == 1 && col_core) col_root = mf_find(A.parent, (int32_t)bcol);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (!bcol_ok || !((rvalid >> r) & 1)) continue;
                const int64_t rp = row0 + 8 * (r >> 2) + 4 * half + (r & 3);
                if (PASS == 1 && (!col_core || !((rcore >> r) & 1) || bcol <= rp || root_r[r] == col_root))
                    continue;
                if (PASS == 2 && (!col_core || ((rcore >> r) & 1) || col_label >= acc_i[r])) continue;
                const float est = (na_r[r] + nb) - 2.0f * acc[r];
                bool in = est <= thr_in;
                if (PASS == 3) {
                    const bool ex = exact_pred(A.xs + rp * A.d, A.xs + bcol * A.d, A.d, A.e2);
                    double d64 = 0.0;
                    for (int k = 0; k < A.d; ++k) {
                        const double df = (double)A.xs[rp * A.d + k] - (double)A.xs[bcol * A.d + k];
                        d64 += df * df;
                    }
                    const bool band_pair = !in && est <= thr_out;
                    const bool filter_wrong = (!band_pair) && (in != ex);
                    const bool f64_flip = ex != (d64 <= (double)A.e2);
                    if (band_pair) atomicAdd(A.audit + 0, 1ull);
                    if (f64_flip) atomicAdd(A.audit + 1, 1ull);
                    if (filter_wrong) atomicAdd(A.audit + 2, 1ull);
                    amax = fmaxf(amax, (float)fabs((double)est - d64));
                    continue;
                }
                if (!in && est <= thr_out) in = exact_pred(A.xs + rp * A.d, A.xs + bcol * A.d, A.d, A.e2);
                if (!in) continue;
                if (PASS == 0) acc_i[r] += 1;
                if (PASS == 1) {
                    mf_unite(A.parent, (int32_t)rp, (int32_t)bcol);
                    root_r[r] = mf_find(A.parent, (int32_t)rp);
                    col_root = mf_find(A.parent, (int32_t)bcol);
                }
                if (PASS == 2) acc_i[r] = col_label;
            }
        }
    }
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
                if (!((rvalid >> r) & 1)) continue;
                const int64_t rp = row0 + 8 * (r >> 2) + 4 * half + (r & 3);
                if (PASS == 0) A.nn_s[rp] = acc_i[r];
                if (PASS == 2 && !((rcore >> r) & 1)) A.best_s[rp] = acc_i[r];
            }
        }
    }
}

// ------------------------------------------------------------- labels --
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

__global__ __launch_bounds__(256) void k_mf_rootmin(const uint8_t *__restrict__ core_s,
                                                    const uint32_t *__restrict__ perm, int64_t n,
                                                    int32_t *__restrict__ parent, int32_t *__restrict__ root_min) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !core_s[i]) return;
    atomicMin(root_min + mf_find(parent, (int32_t)i), (int32_t)perm[i]);
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
    int64_t waves = a.n_tiles;
    if (PASS == 3) waves = ceil_div(a.n_tiles, a.audit_stride);
    const unsigned g = (unsigned)ceil_div(waves, 4);
    if (g == 0) return;
    Launch L(c.prof, c.stream, name);
    switch (kp) {
        case 16: k_mfma_pass<16, PASS><<<g, 256, 0, c.stream>>>(a); break;
        case 32: k_mfma_pass<32, PASS><<<g, 256, 0, c.stream>>>(a); break;
        case 64: k_mfma_pass<64, PASS><<<g, 256, 0, c.stream>>>(a); break;
        default: k_mfma_pass<128, PASS><<<g, 256, 0, c.stream>>>(a); break;
    }
}

static MfmaArgs mfma_args(Ctx &c) {
    MfmaArgs a{};
    a.tiles = c.ctiles.as<int4>();
    a.n_tiles = c.n_tiles;
    a.cand_off = c.tile_off.as<int32_t>();
    a.cand = c.ranges.as<int32_t>();
    a.cell_start = c.cell_start.as<int32_t>();
    a.hi = c.keys2.as<__bf16>();
    a.lo = reinterpret_cast<const __bf16 *>(c.vals2.p);
    a.na = c.soa_core.as<float>();
    a.rho = c.soa_core.as<float>() + c.n;
    a.cell_rmax = c.cell_c.as<float>();
    a.xs = c.soa.as<float>();
    a.n = c.n;
    a.d = c.d;
    const float e = c.h_eps[0];
    a.e2 = e * e;  // data_process.h:73
    const int kp = kp_for(c.d);
    const double u = std::ldexp(1.0, -24);
    a.b1 = 2.0 * (3.1 * std::ldexp(1.0, -16) + 3.0 * kp * u) + (2.0 * kp + 8.0) * u;
    a.b2 = std::ldexp(1.0, -21) * std::sqrt((double)a.e2 * 1.01);
    a.b3 = (c.d + 3) * u * (double)a.e2;
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
    fill_u64(reinterpret_cast<uint64_t *>(sum), 0, d, s, c.prof);
    {
        Launch L(c.prof, s, "mf_colsum");
        k_colsum<<<1024, 256, 0, s>>>(c.coords, n, d, sum);
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
    for (int64_t j = 0; j < k_cent; ++j)
        NDBX_HIP(hipMemcpyAsync(cent + j * d, c.coords + (j * n / k_cent) * d, d * 4, hipMemcpyDeviceToDevice, s));
    c.keys.ensure(n * 8);
    c.vals.ensure(n * 4);
    c.keys2.ensure(std::max<size_t>(n * 8, (size_t)n * kp * 2));
    c.vals2.ensure(std::max<size_t>(n * 4, (size_t)n * kp * 2));
    {
        Launch L(c.prof, s, "mf_assign");
        switch (kp) {
            case 16: k_assign<16><<<nb, 256, 0, s>>>(c.coords, n, d, cent, (int)k_cent, c.keys.as<uint64_t>(), c.vals.as<uint32_t>()); break;
            case 32: k_assign<32><<<nb, 256, 0, s>>>(c.coords, n, d, cent, (int)k_cent, c.keys.as<uint64_t>(), c.vals.as<uint32_t>()); break;
            case 64: k_assign<64><<<nb, 256, 0, s>>>(c.coords, n, d, cent, (int)k_cent, c.keys.as<uint64_t>(), c.vals.as<uint32_t>()); break;
            default: k_assign<128><<<nb, 256, 0, s>>>(c.coords, n, d, cent, (int)k_cent, c.keys.as<uint64_t>(), c.vals.as<uint32_t>()); break;
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
        c.keys2.ensure((size_t)n * kp * 2);
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
    // sorted rows, hi/lo, norms
    c.soa.ensure((size_t)n * d * 4);
    c.soa_core.ensure((size_t)n * 2 * 4);
    {
        Launch L(c.prof, s, "mf_prep");
        k_mfma_prep<<<nb, 256, 0, s>>>(c.coords, perm, n, d, kp, mu, c.soa.as<float>(), c.keys2.as<__bf16>(),
                                       reinterpret_cast<__bf16 *>(c.vals2.p), c.soa_core.as<float>(),
                                       c.soa_core.as<float>() + n);
    }
    c.sync();
    const int64_t nc = h_total;
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
                                                  c.nranges.as<int32_t>(), 0);
    }
    scan_exclusive_i32(c.nranges.as<int32_t>(), c.tile_off.as<int32_t>(), nc, c.tile_off.as<int32_t>() + nc,
                       c.scratch, s, c.prof);
    int32_t h_pairs = 0;
    NDBX_HIP(hipMemcpyAsync(&h_pairs, c.tile_off.as<int32_t>() + nc, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.ranges.ensure((size_t)h_pairs * 4 + 4);
    {
        Launch L(c.prof, s, "mf_cell_pairs");
        k_cell_pairs<<<(unsigned)nc, 256, 0, s>>>(cent, c.cell_key.as<uint64_t>(),
                                                  reinterpret_cast<double *>(c.b_rootmin.p), nc, d, r,
                                                  c.tile_off.as<int32_t>(), c.ranges.as<int32_t>(), 1);
    }
    // row tiles
    c.ctile_off.ensure((nc + 1) * 4);
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_row_tile_count<<<ncb, 256, 0, s>>>(c.cell_start.as<int32_t>(), nc, c.nranges.as<int32_t>());
    }
    scan_exclusive_i32(c.nranges.as<int32_t>(), c.ctile_off.as<int32_t>(), nc, d_total, c.scratch, s, c.prof);
    NDBX_HIP(hipMemcpyAsync(&h_total, d_total, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    c.n_tiles = h_total;
    c.ctiles.ensure(c.n_tiles * 16 + 16);
    {
        Launch L(c.prof, s, "mf_row_tiles");
        k_row_tiles<<<ncb, 256, 0, s>>>(c.cell_start.as<int32_t>(), c.ctile_off.as<int32_t>(), nc, c.ctiles.as<int4>());
    }
    c.pair_tests = (int64_t)h_pairs;  // candidate cell pairs
    // count
    c.nn_s.ensure(n * 4);
    MfmaArgs a = mfma_args(c);
    a.nn_s = c.nn_s.as<int32_t>();
    launch_pass<0>(c, a, kp, "mf_count");
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
        a.audit_stride = (int)std::max<int64_t>(1, c.n_tiles / 2048);
        launch_pass<3>(c, a, kp, "mf_audit");
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
    launch_pass<1>(c, a, kp, "mf_union");
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
    a.label_s = c.label_s.as<int32_t>();
    a.best_s = c.cell_label.as<int32_t>();
    launch_pass<2>(c, a, kp, "mf_border");
    Launch L(c.prof, s, "meta");
    k_mf_out<<<nb, 256, 0, s>>>(c.core_s.as<uint8_t>(), c.label_s.as<int32_t>(), c.cell_label.as<int32_t>(),
                                c.vals.as<uint32_t>(), n, c.labels.as<int32_t>(), c.meta.as<int32_t>());
}

}  // namespace ndbx
