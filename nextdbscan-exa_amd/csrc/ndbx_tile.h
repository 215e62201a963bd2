// Device helpers shared by the tile kernels (count, border): wave reductions
// and the per-row candidate windows of a tile.
#pragma once

#include "ndbx_internal.h"

namespace ndbx {

// Stencil rows per cell, (2 reach + 1)^(D-1) at reach 2.
template <int D>
constexpr int kStencilRows = D == 3 ? 25 : D == 2 ? 5 : 1;

constexpr int kStage = 256;  // candidates staged in LDS per wave

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// cnt + (ok ? 1 : 0) as one v_addc_co_u32 whose carry-in is the compare's
// lane mask (the mask the ballot reads).  Written plainly, each compare is
// widened to 0/1 by a v_cndmask before the adds: 1.5 VALU per test instead
// of 1.
__device__ __forceinline__ int32_t add_if(int32_t cnt, bool ok) {
    const uint64_t m = __ballot(ok);
    int32_t r;
    uint64_t carry;
    asm("v_addc_co_u32_e64 %0, %1, 0, %2, %3" : "=v"(r), "=s"(carry) : "v"(cnt), "s"(m));
    return r;
}

// Tests the staged candidates [0, n) of a wave's LDS stage (kStage slots per
// dimension; slots past n hold +inf, which no query accepts) for lanes split
// as (query lane & (2^qb - 1), slice lane >> qb): slice s takes the groups of
// four candidates i with (i / 4) % (64 >> qb) == s, one 16-byte read per
// dimension each.  Returns the lane's count for its slice.
template <int D>
__device__ __forceinline__ int32_t stage_test_sliced(const float (&qx)[D], const float *stg, int n, float e2,
                                                     int lane, int qb) {
    const int stride = 4 * (64 >> qb);
    const int npad = __builtin_amdgcn_readfirstlane((n + stride - 1) / stride * stride);
    const int off = 4 * (lane >> qb);
    int32_t cnt = 0;
    for (int i = 0; i < npad; i += stride) {
        float c[D][4];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float4 a = *reinterpret_cast<const float4 *>(stg + k * kStage + i + off);
            c[k][0] = a.x, c[k][1] = a.y, c[k][2] = a.z, c[k][3] = a.w;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float l2 = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float df = qx[k] - c[k][u];
                l2 = __fadd_rn(l2, __fmul_rn(df, df));
            }
            cnt = add_if(cnt, l2 <= e2);
        }
    }
    return cnt;
}

// Search radius^2 above anything the float32 predicate can accept: e2 (1 + 2^-18).
__device__ __forceinline__ double search_r2(float e2) { return (double)e2 * (1.0 + 3.814697265625e-06); }

// Halves of a cell per dimension: the points of a cell are sorted by a
// Morton code of their sub-cell whose top D bits say, per dimension, whether
// the point lies in the upper half of the cell, so each of the 2^D parts is a
// contiguous run.  Part bit of dimension k: k for D = 1, 3; 1 - k for D = 2.
template <int D>
__device__ __forceinline__ int part_bit(int k) { return D == 2 ? 1 - k : k; }

// Parts of cell (cc) that can hold a point within the search radius of the
// tile box [lo, hi] (bounds widened by 2^-20 of a cell), as a 2^D-bit mask.
template <int D>
__device__ __forceinline__ uint32_t parts_in_reach(const float (&lo)[D], const float (&hi)[D], const InstGrid &g,
                                                   const int (&cc)[D], double r2o) {
    const double s = 1.0 / g.inv_s, h = 0.5 * s, slack = s * 9.5367431640625e-07;
    double gap2[D][2];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double X = g.origin[k] + (double)cc[k] * s;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double a0 = X + (double)b * h - slack, a1 = a0 + h + 2.0 * slack;
            const double gp = fmax(0.0, fmax(a0 - (double)hi[k], (double)lo[k] - a1));
            gap2[k][b] = gp * gp;
        }
    }
    uint32_t mask = 0;
#pragma unroll
    for (int o = 0; o < (1 << D); ++o) {
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) sum += gap2[k][(o >> part_bit<D>(k)) & 1];
        if (sum <= r2o) mask |= 1u << o;
    }
    return mask;
}

// Stencil row `slot` (reach 2) around a cell of row (c1, c2): whether a
// point of the box [lo, hi] can have a point within the search radius in it,
// and if so the cells along dimension 0 that can hold one, [xl, xh].  Cell
// bounds are widened by 2^-20 of a cell, so no point the predicate accepts is
// dropped.
template <int D>
__device__ __forceinline__ bool row_reach(int slot, const float (&lo)[D], const float (&hi)[D], const InstGrid &g,
                                          int c1, int c2, int &xl, int &xh) {
    const double inv_s = g.inv_s, s = 1.0 / inv_s, slack = s * 9.5367431640625e-07;  // 2^-20 cell
    double rem = search_r2(g.e2);
    if (D >= 2) {
        const int dy = (D == 3 ? slot % 5 : slot) - 2;
        const double ylo = g.origin[1] + (double)(c1 + dy) * s - slack, yhi = ylo + s + 2.0 * slack;
        const double gy = fmax(0.0, fmax(ylo - (double)hi[1], (double)lo[1] - yhi));
        rem -= gy * gy;
    }
    if (D >= 3) {
        const int dz = slot / 5 - 2;
        const double zlo = g.origin[2] + (double)(c2 + dz) * s - slack, zhi = zlo + s + 2.0 * slack;
        const double gz = fmax(0.0, fmax(zlo - (double)hi[2], (double)lo[2] - zhi));
        rem -= gz * gz;
    }
    if (rem < 0.0) return false;
    const double rx = sqrt(rem) + slack;
    xl = (int)floor(((double)lo[0] - rx - g.origin[0]) * inv_s);
    xh = (int)floor(((double)hi[0] + rx - g.origin[0]) * inv_s);
    return true;
}

// Candidate cells of one stencil row of a tile, computed by lane `slot`
// (slot < rows_per_cell; other lanes return an empty window).  The tile is a
// run of cells [cf, cl] of one row; its bounding box is [lo, hi].  Rows out
// of reach of the box are empty; along dimension 0 only cells that can hold
// a point within the radius of the box are kept.  Returns the cell index
// range [a, e).
template <int D>
__device__ __forceinline__ int2 tile_window(int slot, int rows_per_cell, int cf, int cl, const float (&lo)[D],
                                            const float (&hi)[D], const InstGrid &g, int c1, int c2,
                                            const int2 *__restrict__ ranges, const int32_t *__restrict__ cell_c,
                                            int64_t nc) {
    if (slot >= rows_per_cell) return make_int2(0, 0);
    int a = ranges[(int64_t)cf * rows_per_cell + slot].x;
    int e = ranges[(int64_t)cl * rows_per_cell + slot].y;
    if (a >= e) return make_int2(0, 0);
    int xl, xh;
    if (!row_reach<D>(slot, lo, hi, g, c1, c2, xl, xh)) return make_int2(0, 0);
    while (a < e && cell_c[a] < xl) ++a;
    while (e > a && cell_c[e - 1] > xh) --e;
    return make_int2(a, e);
}

}  // namespace ndbx
