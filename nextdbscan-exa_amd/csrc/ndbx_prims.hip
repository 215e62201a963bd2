// Device primitives: fills, exclusive scan, stable LSD radix sort.
// All written for 64-lane wavefronts (gfx950); no library calls.
#include "ndbx_internal.h"

namespace ndbx {

// ----------------------------------------------------------- buffers --
thread_local bool g_no_alloc = false;
thread_local Arena *g_arena = nullptr;

void *DevBuf::ensure(size_t want) {
    if (want == 0) want = 256;
    if (want <= bytes) return p;
    if (g_arena) {
        // carve from the fixed region; a buffer that grows takes a new
        // piece (the old one is not reused within the call)
        const size_t at = (g_arena->used + 255) & ~(size_t)255;
        if (at > g_arena->size || want > g_arena->size - at)
            throw Error(NDBX_ENOMEM, "no-alloc scratch exhausted: " + std::to_string(at + want) + " bytes needed of " +
                                         std::to_string(g_arena->size));
        release();
        p = g_arena->base + at;
        bytes = want;
        borrowed = true;
        g_arena->used = at + want;
        return p;
    }
    if (g_no_alloc) throw Error(NDBX_EDEVICE, "device buffer would grow inside a graph capture");
    release();
    NDBX_HIP(hipMalloc(&p, want));
    bytes = want;
    return p;
}

void DevBuf::release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
}

// ---------------------------------------------------------- profiler --
int Profiler::slot_of(const char *name) {
    for (size_t i = 0; i < names.size(); ++i)
        if (names[i] == name) return (int)i;
    names.emplace_back(name);
    ms.push_back(0.0);
    calls.push_back(0);
    return (int)names.size() - 1;
}

hipEvent_t Profiler::take() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    NDBX_HIP(hipEventCreate(&e));
    return e;
}

void Profiler::resolve() {
    for (auto &pe : pending) {
        float t = 0.f;
        NDBX_HIP(hipEventElapsedTime(&t, pe.a, pe.b));
        ms[pe.slot] += t;
        calls[pe.slot] += 1;
        pool.push_back(pe.a);
        pool.push_back(pe.b);
    }
    pending.clear();
}

void Profiler::reset() {
    names.clear();
    ms.clear();
    calls.clear();
}

void Profiler::destroy() {
    for (auto &pe : pending) {
        pool.push_back(pe.a);
        pool.push_back(pe.b);
    }
    pending.clear();
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
}

Launch::Launch(Profiler &p, hipStream_t st, const char *name) : prof(p), s(st) {
    if (!prof.on || (!prof.only.empty() && prof.only != name)) return;
    slot = prof.slot_of(name);
    a = prof.take();
    b = prof.take();
    NDBX_HIP(hipEventRecord(a, s));
}

Launch::~Launch() noexcept(false) {
    hipError_t e = hipGetLastError();
    if (slot >= 0) {
        if (hipEventRecord(b, s) == hipSuccess)
            prof.pending.push_back({slot, a, b});
    }
    if (e != hipSuccess)
        throw Error(NDBX_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
}

// ------------------------------------------------------------- fills --
template <class T>
__global__ void k_fill(T *__restrict__ p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static int grid_for(int64_t n, int block, int cap = 8192) {
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s, Profiler &prof) {
    if (n <= 0) return;
    Launch L(prof, s, "fill");
    k_fill<int32_t><<<grid_for(n, 256), 256, 0, s>>>(p, v, n);
}

void fill_u64(uint64_t *p, uint64_t v, int64_t n, hipStream_t s, Profiler &prof) {
    if (n <= 0) return;
    Launch L(prof, s, "fill");
    k_fill<uint64_t><<<grid_for(n, 256), 256, 0, s>>>(p, v, n);
}

// Sum of int32 values into a 64-bit counter (totals that may exceed int32):
// one atomic per block.
__global__ __launch_bounds__(256) void k_sum_i32_u64(const int32_t *__restrict__ v, int64_t n,
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

void sum_i32_u64(const int32_t *v, int64_t n, unsigned long long *out, hipStream_t s, Profiler &prof) {
    fill_u64(reinterpret_cast<uint64_t *>(out), 0, 1, s, prof);
    if (n <= 0) return;
    Launch L(prof, s, "sum");
    k_sum_i32_u64<<<grid_for(n, 256, 1024), 256, 0, s>>>(v, n, out);
}

// -------------------------------------------------------------- scan --
// Three-phase exclusive scan: per-block sums, scan of the sums (recursive),
// per-block downsweep seeded with the block's prefix.  Tile = 256 x 8.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int32_t wave_incl_scan(int32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *sh_w, int32_t *block_total) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int32_t inc = wave_incl_scan(v);
    if (lane == 63) sh_w[w] = inc;
    __syncthreads();
    int32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        int32_t x = sh_w[i];
        if (i < w) base += x;
        tot += x;
    }
    *block_total = tot;
    return base + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t *__restrict__ in,
                                                              int32_t *__restrict__ part, int64_t n) {
    __shared__ int32_t sh_w[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    int32_t tot;
    (void)block_excl_scan(s, sh_w, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int32_t *__restrict__ in,
                                                            int32_t *__restrict__ out,
                                                            const int32_t *__restrict__ seed, int64_t n,
                                                            int32_t *__restrict__ total) {
    __shared__ int32_t sh[kScanTile + kScanTile / 32];
    __shared__ int32_t sh_w[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#define PAD(i) ((i) + ((i) >> 5))
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int li = k * kScanThreads + threadIdx.x;
        int64_t i = base + li;
        sh[PAD(li)] = (i < n) ? in[i] : 0;
    }
    __syncthreads();
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = sh[PAD(threadIdx.x * kScanItems + k)];
        s += v[k];
    }
    int32_t tot;
    int32_t run = block_excl_scan(s, sh_w, &tot) + (seed ? seed[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        sh[PAD(threadIdx.x * kScanItems + k)] = run;
        run += v[k];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total = run;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int li = k * kScanThreads + threadIdx.x;
        int64_t i = base + li;
        if (i < n) out[i] = sh[PAD(li)];
    }
#undef PAD
}

static void scan_rec(const int32_t *in, int32_t *out, int64_t n, int32_t *total, int32_t *tmp,
                     hipStream_t s, Profiler &prof) {
    int64_t nb = ceil_div(n, kScanTile);
    if (nb <= 1) {
        Launch L(prof, s, "scan");
        k_scan_down<<<1, kScanThreads, 0, s>>>(in, out, nullptr, n, total);
        return;
    }
    int32_t *part = tmp;
    int32_t *rest = tmp + nb;
    {
        Launch L(prof, s, "scan");
        k_scan_reduce<<<(unsigned)nb, kScanThreads, 0, s>>>(in, part, n);
    }
    scan_rec(part, part, nb, nullptr, rest, s, prof);
    Launch L(prof, s, "scan");
    k_scan_down<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, part, n, total);
}

void scan_exclusive_i32(const int32_t *in, int32_t *out, int64_t n, int32_t *total, DevBuf &scratch,
                        hipStream_t s, Profiler &prof) {
    if (n <= 0) {
        if (total) fill_i32(total, 0, 1, s, prof);
        return;
    }
    // scratch: sum over levels of ceil(n / tile^k)
    int64_t need = 0, m = n;
    while (m > 1) {
        m = ceil_div(m, kScanTile);
        need += m;
    }
    scratch.ensure((size_t)(need + 16) * sizeof(int32_t));
    scan_rec(in, out, n, total, scratch.as<int32_t>(), s, prof);
}

// -------------------------------------------------------- radix sort --
// LSD, 8 bits per pass.  Per pass: block histograms (digit-major), one
// exclusive scan, then a stable scatter in which every 256-item round is
// ranked with a wave-level multisplit (8 ballots) plus per-wave digit counts
// in LDS.
constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 8;  // 2048-key tiles: 29 KB of LDS in the scatter (10M 3-D: 0.075 ms a pass, 0.091 at 4096)
constexpr int kRadixTile = kRadixThreads * kRadixItems;

__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n,
                                                              int shift, int32_t *__restrict__ hist,
                                                              int64_t nblocks) {
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int k = 0; k < kRadixItems; ++k) {
        int64_t i = base + (int64_t)k * kRadixThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one block's tile by the digit at `shift`.  The tile's
// keys are first placed in LDS in digit order (a digit's keys keep their
// input order: rounds, then waves, then lanes), then written out with
// consecutive threads on consecutive positions, so each digit's run of the
// tile is one coalesced stretch of the output.  The tile's count of each
// digit is the difference of consecutive entries of the scanned histogram
// (digit-major, one entry per block).
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals, uint64_t *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, int64_t n, int shift, const int32_t *__restrict__ offs_g,
    int64_t nblocks) {
    __shared__ uint64_t sk[kRadixTile];
    __shared__ uint32_t sv[kRadixTile];
    __shared__ int32_t gofs[256], dstart[256], run[256];
    __shared__ int32_t wcnt[kRadixThreads / 64][256];
    __shared__ int32_t sh_w[kRadixThreads / 64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kRadixTile;
    const int tile_n = (int)(n - base < kRadixTile ? n - base : kRadixTile);
    // this block's global offset and tile count of digit threadIdx.x
    const int64_t idx = (int64_t)threadIdx.x * nblocks + blockIdx.x;
    const int32_t g0 = offs_g[idx];
    const int32_t g1 = idx + 1 < 256 * nblocks ? offs_g[idx + 1] : (int32_t)n;
    gofs[threadIdx.x] = g0;
    int32_t tot;
    dstart[threadIdx.x] = block_excl_scan(g1 - g0, sh_w, &tot);
    run[threadIdx.x] = 0;
    // the tile's keys in registers (all loads in flight together)
    uint64_t key[kRadixItems];
    uint32_t val[kRadixItems];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const int li = r * kRadixThreads + threadIdx.x;
        key[r] = li < tile_n ? keys[base + li] : 0;
        val[r] = li < tile_n ? vals[base + li] : 0;
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
#pragma unroll
        for (int q = 0; q < kRadixThreads / 64; ++q) wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        const bool valid = r * kRadixThreads + (int)threadIdx.x < tile_n;
        const unsigned dg = (unsigned)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const int rank_in_wave = __popcll(peers & lt_mask);
        const int cnt_in_wave = __popcll(peers);
        // the highest lane of each peer group publishes the group size
        if (valid && rank_in_wave == cnt_in_wave - 1) wcnt[w][dg] = cnt_in_wave;
        __syncthreads();
        if (valid) {
            int pos = dstart[dg] + run[dg] + rank_in_wave;
            for (int q = 0; q < w; ++q) pos += wcnt[q][dg];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int q = 0; q < kRadixThreads / 64; ++q) add += wcnt[q][threadIdx.x];
        run[threadIdx.x] += add;
        // every wave has read this round's wcnt (positions above) before the
        // next round clears it
        __syncthreads();
    }
    for (int p = threadIdx.x; p < tile_n; p += kRadixThreads) {
        const uint64_t k = sk[p];
        const unsigned dg = (unsigned)(k >> shift) & 255u;
        const int64_t gp = (int64_t)gofs[dg] + (p - dstart[dg]);
        keys_out[gp] = k;
        vals_out[gp] = sv[p];
    }
}

void radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt, int64_t n,
                    int nbits, DevBuf &scratch, hipStream_t s, Profiler &prof, bool *result_in_alt) {
    *result_in_alt = false;
    if (n <= 1 || nbits <= 0) return;
    const int64_t nblocks = ceil_div(n, kRadixTile);
    const int64_t hn = nblocks * 256;
    // scratch: hist + scan temp
    int64_t scan_need = 0, m = hn;
    while (m > 1) {
        m = ceil_div(m, kScanTile);
        scan_need += m;
    }
    scratch.ensure((size_t)(hn + scan_need + 32) * sizeof(int32_t));
    int32_t *hist = scratch.as<int32_t>();
    int32_t *tmp = hist + hn + 16;
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    bool alt = false;
    for (int shift = 0; shift < nbits; shift += 8) {
        {
            Launch L(prof, s, "radix_hist");
            k_radix_hist<<<(unsigned)nblocks, kRadixThreads, 0, s>>>(ka, n, shift, hist, nblocks);
        }
        scan_rec(hist, hist, hn, nullptr, tmp, s, prof);
        {
            Launch L(prof, s, "radix_scatter");
            k_radix_scatter<<<(unsigned)nblocks, kRadixThreads, 0, s>>>(ka, va, kb, vb, n, shift, hist,
                                                                         nblocks);
        }
        std::swap(ka, kb);
        std::swap(va, vb);
        alt = !alt;
    }
    *result_in_alt = alt;
}

}  // namespace ndbx
