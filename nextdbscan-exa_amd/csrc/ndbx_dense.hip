// Dense 2-D cells: detection, strip re-sort and (strip, row) offset tables.
// No counterpart in the reference, whose counting is all-pairs
// (data_process.cpp:433-454); the tables only decide which pairs are tested
// and which are counted wholesale with a margin the float32 predicate cannot
// contradict.
#include <cmath>

#include "ndbx_dense.h"
#include "ndbx_internal.h"

namespace ndbx {

__global__ __launch_bounds__(256) void k_dense_flags(const int32_t *__restrict__ cell_start, int64_t nc,
                                                     int32_t *__restrict__ f) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) f[c] = cell_start[c + 1] - cell_start[c] >= kDenseCell;
}

// Strip bits L: 2^L close to sqrt(points / 8), so that in a cell of uniform
// density a tile of kDenseTile = 16 points of a strip is about twice as tall
// as the strip is wide (the tile's box height and the strip width plus the
// box width then add about equally to the band of tested rows).  Table rows:
// 2^(L+2), at most 2^kDenseYBits.
__global__ __launch_bounds__(256) void k_dense_ids(const int32_t *__restrict__ f, const int32_t *__restrict__ excl,
                                                   const int32_t *__restrict__ cell_start, int64_t nc,
                                                   int32_t *__restrict__ dense_id, DenseCell *__restrict__ info,
                                                   int32_t *__restrict__ tbl_size) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (!f[c]) {
        dense_id[c] = -1;
        return;
    }
    const int k = excl[c];
    dense_id[c] = k;
    const double cnt = (double)(cell_start[c + 1] - cell_start[c]);
    int L = (int)lrint(log2(sqrt(cnt / 8.0)));
    L = L < 3 ? 3 : (L > kDenseYBits ? kDenseYBits : L);
    const int R = L + 2 < kDenseYBits ? L + 2 : kDenseYBits;
    info[k] = DenseCell{(int32_t)c, L, R, 0};
    tbl_size[k] = (1 << (L + R)) + 1;
}

__global__ __launch_bounds__(256) void k_dense_tbl_off(const int32_t *__restrict__ off, int64_t nd,
                                                       DenseCell *__restrict__ info) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nd) info[k].tbl = off[k];
}

// New key of sorted position i: cell index above kDenseSubBits; for points of
// dense cells strip << kDenseYBits | y row below (same cell arithmetic as
// the grid keys, so the cell does not change).  Other points get 0 below and
// keep their order (the sort is stable).
__global__ __launch_bounds__(256) void k_dense_rekey(const float *__restrict__ x, const uint32_t *__restrict__ perm,
                                                     const int32_t *__restrict__ cell_of,
                                                     const uint64_t *__restrict__ cell_key,
                                                     const InstGrid *__restrict__ ig, int lin_bits,
                                                     const int32_t *__restrict__ dense_id,
                                                     const DenseCell *__restrict__ info, int64_t n,
                                                     uint64_t *__restrict__ key2, uint32_t *__restrict__ val2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t cell = cell_of[i];
    const uint32_t p = perm[i];
    const int32_t k = dense_id[cell];
    uint32_t sub = 0;
    if (k >= 0) {
        const InstGrid &g = ig[cell_key[cell] >> lin_bits];
        const int bits[2] = {info[k].L, kDenseYBits};
        uint32_t f2[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double t = ((double)x[(int64_t)p * 2 + d] - g.origin[d]) * g.inv_s;
            int64_t ck = (int64_t)floor(t);
            ck = ck < 0 ? 0 : (ck >= g.G[d] ? g.G[d] - 1 : ck);
            int fi = (int)floor((t - (double)ck) * (double)(1 << bits[d]));
            fi = fi < 0 ? 0 : (fi >= (1 << bits[d]) ? (1 << bits[d]) - 1 : fi);
            f2[d] = (uint32_t)fi;
        }
        sub = (f2[0] << kDenseYBits) | f2[1];
    }
    key2[i] = ((uint64_t)cell << kDenseSubBits) | sub;
    val2[i] = p;
}

// Table entry e of dense cell k, (strip, row) = (e >> R, e & (2^R - 1)): the
// first point of the cell whose (strip, y) is at or after (strip, row).
__global__ __launch_bounds__(256) void k_dense_table(const uint64_t *__restrict__ key2,
                                                     const int32_t *__restrict__ cell_start,
                                                     const DenseCell *__restrict__ info, int64_t nd, int64_t total,
                                                     int32_t *__restrict__ tbl) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int lo_k = 0, hi_k = (int)nd - 1;  // last dense cell whose table starts at or before e
    while (lo_k < hi_k) {
        const int mid = (lo_k + hi_k + 1) >> 1;
        if (info[mid].tbl <= e) lo_k = mid;
        else hi_k = mid - 1;
    }
    const DenseCell dc = info[lo_k];
    const int local = (int)(e - dc.tbl);
    const uint32_t col = (uint32_t)(local >> dc.R), row = (uint32_t)(local & ((1 << dc.R) - 1));
    const uint64_t target = ((uint64_t)col << kDenseYBits) | ((uint64_t)row << (kDenseYBits - dc.R));
    const uint64_t mask = (1ull << kDenseSubBits) - 1;
    int lo = cell_start[dc.cell], hi = cell_start[dc.cell + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((key2[mid] & mask) < target) lo = mid + 1;
        else hi = mid;
    }
    tbl[e] = lo;
}

DenseView dense_prepare(Ctx &c, const InstGrid *ig, int64_t n, int64_t nc, int32_t *d_total, hipStream_t s) {
    DenseView v{nullptr, nullptr, nullptr};
    const int ncb = (int)ceil_div(nc, 256);
    c.dense_tmp.ensure((size_t)(nc + 1) * 8);
    int32_t *flag = c.dense_tmp.as<int32_t>(), *excl = flag + (nc + 1);
    {
        Launch L(c.prof, s, "dense_prepare");
        k_dense_flags<<<ncb, 256, 0, s>>>(c.cell_start.as<int32_t>(), nc, flag);
    }
    scan_exclusive_i32(flag, excl, nc, d_total, c.scratch, s, c.prof);
    int32_t h = 0;
    NDBX_HIP(hipMemcpyAsync(&h, d_total, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    const int64_t nd = h;
    c.n_dense = nd;
    if (nd == 0) return v;
    c.dense_id.ensure((size_t)nc * 4);
    c.dense_info.ensure((size_t)nd * sizeof(DenseCell) + (size_t)(nd + 1) * 8);
    DenseCell *info = c.dense_info.as<DenseCell>();
    int32_t *size = reinterpret_cast<int32_t *>(info + nd), *off = size + (nd + 1);
    {
        Launch L(c.prof, s, "dense_prepare");
        k_dense_ids<<<ncb, 256, 0, s>>>(flag, excl, c.cell_start.as<int32_t>(), nc, c.dense_id.as<int32_t>(), info,
                                        size);
    }
    scan_exclusive_i32(size, off, nd, d_total, c.scratch, s, c.prof);
    NDBX_HIP(hipMemcpyAsync(&h, d_total, 4, hipMemcpyDeviceToHost, s));
    {
        Launch L(c.prof, s, "dense_prepare");
        k_dense_tbl_off<<<(unsigned)ceil_div(nd, 256), 256, 0, s>>>(off, nd, info);
        k_dense_rekey<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(
            c.coords, c.vals.as<uint32_t>(), c.cell_of.as<int32_t>(), c.cell_key.as<uint64_t>(), ig, c.lin_bits,
            c.dense_id.as<int32_t>(), info, n, c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>());
    }
    int cbits = 1;
    while ((1ll << cbits) < nc) ++cbits;
    bool alt = false;
    radix_sort_u64(c.keys2.as<uint64_t>(), c.vals2.as<uint32_t>(), c.keys.as<uint64_t>(), c.vals.as<uint32_t>(), n,
                   cbits + kDenseSubBits, c.scratch, s, c.prof, &alt);
    if (!alt) {
        std::swap(c.keys, c.keys2);
        std::swap(c.vals, c.vals2);
    }
    c.sync();
    const int64_t total = h;
    c.fine_off.ensure((size_t)total * 4);
    {
        Launch L(c.prof, s, "dense_prepare");
        k_dense_table<<<(unsigned)ceil_div(total, 256), 256, 0, s>>>(c.keys.as<uint64_t>(), c.cell_start.as<int32_t>(),
                                                                      info, nd, total, c.fine_off.as<int32_t>());
    }
    v.id = c.dense_id.as<int32_t>();
    v.info = info;
    v.tbl = c.fine_off.as<int32_t>();
    return v;
}

}  // namespace ndbx
