// Dense 2-D cells: cells that hold at least kDenseCell points.
//
// Their points are re-sorted into 2^L column strips (L from the cell's
// population, so that 64 consecutive points of a strip form a tile about as
// tall as the strip is wide) and by y at 2^kDenseYBits rows inside a strip.
// A table per dense cell gives, for every (strip, row) at L x R bits, the
// first point at or after it in that order; entry 2^(L+R) is the cell's end.
//
// The count kernel uses the tables for a tile against a dense cell: for each
// strip, the rows every point of the tile's box reaches within
// sqrt(e2 (1 - 2^-18)) are counted from the table without tests (the
// float32 predicate accepts every pair that close: its rounding error is
// below 2^-21 relative), rows that some point of the box may reach within
// sqrt(e2 (1 + 2^-18)) are tested with the predicate, and the rest hold no
// neighbour.  Strip and row bounds are widened by 2^-20 of a cell.
#pragma once

#include "ndbx_tile.h"

namespace ndbx {

constexpr int kDenseCell = 4096;  // points per cell from which a cell is dense
constexpr int kDenseYBits = 10;   // y resolution of the order inside a dense cell
constexpr int kDenseSubBits = 20; // key bits below the cell index: strip << 10 | y

struct DenseCell {
    int32_t cell;  // cell index
    int32_t L;     // strips: 2^L
    int32_t R;     // table rows: 2^R (R <= kDenseYBits)
    int32_t tbl;   // first table entry
};

struct DenseView {
    const int32_t *id;  // per cell: dense index, or -1
    const DenseCell *info;
    const int32_t *tbl;
};

constexpr int kDenseTile = 16;  // points per count tile of a dense cell

// Tests the kStage staged candidates [0, n) (slots past n hold +inf) against
// the lane's query.  NSUB = 1: every lane owns a query and tests them all
// (broadcast reads).  NSUB = 4: lane = (query lane & 15, slice lane >> 4) and
// slice s tests the candidates i with (i / 4) % 4 == s.
template <int D, int NSUB>
__device__ __forceinline__ int32_t stage_test(const float (&qx)[D], const float *stg, int n, float e2, int lane) {
    int32_t cnt = 0;
    if constexpr (NSUB == 1) {
        const int n8 = __builtin_amdgcn_readfirstlane((n + 7) & ~7);
        for (int i = 0; i < n8; i += 8) {
            float c[D][8];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float4 a = *reinterpret_cast<const float4 *>(stg + k * kStage + i);
                const float4 b = *reinterpret_cast<const float4 *>(stg + k * kStage + i + 4);
                c[k][0] = a.x, c[k][1] = a.y, c[k][2] = a.z, c[k][3] = a.w;
                c[k][4] = b.x, c[k][5] = b.y, c[k][6] = b.z, c[k][7] = b.w;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                float l2 = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float df = qx[k] - c[k][u];
                    l2 = __fadd_rn(l2, __fmul_rn(df, df));
                }
                cnt = add_if(cnt, l2 <= e2);
            }
        }
    } else {
        static_assert(NSUB == 4, "slices");
        const int sl = 4 * (lane >> 4);
        const int n32 = __builtin_amdgcn_readfirstlane((n + 31) & ~31);
        for (int i = 0; i < n32; i += 32) {
            float c[D][8];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float4 a = *reinterpret_cast<const float4 *>(stg + k * kStage + i + sl);
                const float4 b = *reinterpret_cast<const float4 *>(stg + k * kStage + i + 16 + sl);
                c[k][0] = a.x, c[k][1] = a.y, c[k][2] = a.z, c[k][3] = a.w;
                c[k][4] = b.x, c[k][5] = b.y, c[k][6] = b.z, c[k][7] = b.w;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                float l2 = 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float df = qx[k] - c[k][u];
                    l2 = __fadd_rn(l2, __fmul_rn(df, df));
                }
                cnt = add_if(cnt, l2 <= e2);
            }
        }
    }
    return cnt;
}

// Every lane's query against the candidates of two point ranges per lane,
// [s1, e1) and [s2, e2_), flattened over the wave (scan of their lengths) and
// staged through LDS kStage at a time with independent loads.  Returns the
// lane's count (per slice when NSUB = 4).
template <int D, int NSUB>
__device__ __forceinline__ int32_t staged_count2(int s1, int e1, int s2, int e2_, const float (&qx)[D],
                                                 const float *__restrict__ soa, int64_t npad, float e2, float *stg,
                                                 int lane, int64_t &issued) {
    const int len1 = e1 > s1 ? e1 - s1 : 0;
    const int len = len1 + (e2_ > s2 ? e2_ - s2 : 0);
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t2 = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t2;
    }
    const int pre = incl - len;
    const int total = __builtin_amdgcn_readlane(incl, 63);
    int32_t cnt = 0;
    for (int cb = 0; cb < total; cb += kStage) {
        const int n_here = total - cb < kStage ? total - cb : kStage;
#pragma unroll
        for (int u = 0; u < kStage / 64; ++u) {
            const int gi = cb + u * 64 + lane;
            const int gc = gi < total ? gi : total - 1;  // all lanes take part in the shuffles
            int lo_w = 0, hi_w = 63;                      // last lane whose ranges start at or before gc
#pragma unroll
            for (int it = 0; it < 6; ++it) {
                const int mid = (lo_w + hi_w + 1) >> 1;
                if (__shfl(pre, mid, 64) <= gc) lo_w = mid;
                else hi_w = mid - 1;
            }
            // every shuffle runs on all lanes (a lane outside a branch would
            // read as 0 to the lanes inside it)
            const int off = gc - __shfl(pre, lo_w, 64);
            const int l1 = __shfl(len1, lo_w, 64);
            const int b1 = __shfl(s1, lo_w, 64), b2 = __shfl(s2, lo_w, 64);
            const int j = off < l1 ? b1 + off : b2 + (off - l1);
            // slots past the end hold +inf, which no query accepts
#pragma unroll
            for (int k = 0; k < D; ++k)
                stg[k * kStage + u * 64 + lane] = gi < total ? soa[k * npad + j] : __builtin_inff();
        }
        __builtin_amdgcn_wave_barrier();
        cnt += stage_test<D, NSUB>(qx, stg, n_here, e2, lane);
        __builtin_amdgcn_wave_barrier();
    }
    issued += total;
    return cnt;
}

// Every lane's query against the contiguous candidates [ps, pe) (wave-uniform),
// staged kStage at a time with coalesced loads.
template <int D, int NSUB>
__device__ __forceinline__ int32_t staged_run(int ps, int pe, const float (&qx)[D], const float *__restrict__ soa,
                                              int64_t npad, float e2, float *stg, int lane, int64_t &issued) {
    int32_t cnt = 0;
    for (int cb = ps; cb < pe; cb += kStage) {
        const int n_here = pe - cb < kStage ? pe - cb : kStage;
#pragma unroll
        for (int u = 0; u < kStage / 64; ++u) {
            const int j = cb + u * 64 + lane;
#pragma unroll
            for (int k = 0; k < D; ++k) stg[k * kStage + u * 64 + lane] = j < pe ? soa[k * npad + j] : __builtin_inff();
        }
        __builtin_amdgcn_wave_barrier();
        cnt += stage_test<D, NSUB>(qx, stg, n_here, e2, lane);
        __builtin_amdgcn_wave_barrier();
    }
    issued += pe > ps ? pe - ps : 0;
    return cnt;
}

// Neighbours of the lanes' queries in dense cell (c0, c1) described by dc,
// for a tile whose box is [lo, hi].  Lanes take the cell's strips 64 at a
// time; all lanes must call it (wave-uniform arguments).  Points counted
// without tests (the same for every query of the tile) go to `bulk`; the
// return value is the lane's tested count (per slice when NSUB = 4).
template <int NSUB>
__device__ __forceinline__ int32_t dense_cell_count(const float (&qx)[2], const float (&lo)[2], const float (&hi)[2],
                                                    const InstGrid &g, int c0, int c1, const DenseCell dc,
                                                    const int32_t *__restrict__ tbl, const float *__restrict__ soa,
                                                    int64_t npad, float e2, float *stg, int lane, int32_t &bulk,
                                                    int64_t &issued) {
    const double s = 1.0 / g.inv_s, slack = s * 9.5367431640625e-07;  // 2^-20 cell
    const int ncol = 1 << dc.L, nrow = 1 << dc.R;
    const double hx = s / ncol, hy = s / nrow;
    const double r2o = search_r2(e2), r2i = (double)e2 * (1.0 - 3.814697265625e-06);
    const double X0 = g.origin[0] + (double)c0 * s, Y0 = g.origin[1] + (double)c1 * s;
    const double ro = sqrt(r2o);
    int cb = (int)floor(((double)lo[0] - ro - X0 - slack) / hx);
    int ce = (int)floor(((double)hi[0] + ro - X0 + slack) / hx);
    cb = cb < 0 ? 0 : cb;
    ce = ce > ncol - 1 ? ncol - 1 : ce;
    const int32_t *t = tbl + dc.tbl;
    int32_t cnt = 0;
    for (int col0 = cb; col0 <= ce; col0 += 64) {
        const int col = col0 + lane;
        int b1s = 0, b1e = 0, b2s = 0, b2e = 0, inner = 0;
        if (col <= ce) {
            const double cx0 = X0 + (double)col * hx - slack, cx1 = cx0 + hx + 2.0 * slack;
            const double dxmin = fmax(0.0, fmax(cx0 - (double)hi[0], (double)lo[0] - cx1));
            const double dxmax = fmax(cx1 - (double)lo[0], (double)hi[0] - cx0);
            if (dxmin * dxmin <= r2o) {
                const double dyo = sqrt(r2o - dxmin * dxmin);
                int rlo = (int)floor(((double)lo[1] - dyo - Y0 - slack) / hy);
                int rhi = (int)floor(((double)hi[1] + dyo - Y0 + slack) / hy);
                rlo = rlo < 0 ? 0 : rlo;
                rhi = rhi > nrow - 1 ? nrow - 1 : rhi;
                if (rlo <= rhi) {
                    const int base = col * nrow;
                    int ilo = 1, ihi = 0;  // no inner rows
                    if (dxmax * dxmax < r2i) {
                        const double dyi = sqrt(r2i - dxmax * dxmax);
                        const double ya = (double)hi[1] - dyi, yb = (double)lo[1] + dyi;
                        if (ya <= yb) {
                            int a = (int)ceil((ya - Y0 + slack) / hy), b = (int)floor((yb - Y0 - slack) / hy) - 1;
                            a = a < rlo ? rlo : a;
                            b = b > rhi ? rhi : b;
                            if (a <= b) {
                                ilo = a;
                                ihi = b;
                            }
                        }
                    }
                    const int p_lo = t[base + rlo], p_hi = t[base + rhi + 1];
                    if (ilo <= ihi) {
                        const int p_ilo = t[base + ilo], p_ihi = t[base + ihi + 1];
                        inner = p_ihi - p_ilo;
                        b1s = p_lo;
                        b1e = p_ilo;
                        b2s = p_ihi;
                        b2e = p_hi;
                    } else {
                        b1s = p_lo;
                        b1e = p_hi;
                    }
                }
            }
        }
        for (int o = 32; o >= 1; o >>= 1) inner += __shfl_xor(inner, o, 64);
        bulk += inner;
        cnt += staged_count2<2, NSUB>(b1s, b1e, b2s, b2e, qx, soa, npad, e2, stg, lane, issued);
    }
    return cnt;
}

// Host side: finds the dense cells of a 2-D grid, re-sorts their points
// (c.keys / c.vals hold the new order afterwards) and builds the tables.
// Returns the view (id == nullptr when no cell is dense).
DenseView dense_prepare(Ctx &c, const InstGrid *ig, int64_t n, int64_t nc, int32_t *d_total, hipStream_t s);

}  // namespace ndbx
