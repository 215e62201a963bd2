// Internal declarations of the ndbx library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/ndbx.h"
#include "ndbx_limits.h"
#include "ndbx_slabs.h"

namespace ndbx {

// ---------------------------------------------------------------- errors --
struct Error : std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char *what() const noexcept override { return msg.c_str(); }
};

#define NDBX_HIP(call)                                                          \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            throw ::ndbx::Error(e_ == hipErrorOutOfMemory ? NDBX_ENOMEM         \
                                                          : NDBX_EDEVICE,       \
                                std::string(#call) + ": " +                     \
                                    hipGetErrorString(e_));                     \
    } while (0)

// ------------------------------------------------------- device buffers --
// Grow-only device allocation owned by a context.
// Set while a stream capture is open on this thread: a buffer that would
// have to grow then raises an error instead of allocating.
extern thread_local bool g_no_alloc;

// A fixed device region that buffers are carved from instead of allocated
// (bump allocation, 256-byte aligned): set on a thread for the duration of a
// call that must not allocate.  Running past its end raises NDBX_ENOMEM.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
};
extern thread_local Arena *g_arena;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;  // memory not owned (arena or caller's): never freed
    template <class T>
    T *as() const { return static_cast<T *>(p); }
    void *ensure(size_t want);
    void release();
    // Use caller-owned device memory of `bytes` bytes in place.
    void borrow(void *mem, size_t nbytes) {
        release();
        p = mem;
        bytes = nbytes;
        borrowed = true;
    }
};

// ------------------------------------------------------------- profiler --
struct Profiler {
    bool on = false;
    std::string only;  // non-empty: events only around launches of this name
    struct Pending {
        int slot;
        hipEvent_t a, b;
    };
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<int64_t> calls;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int slot_of(const char *name);
    hipEvent_t take();
    void resolve();  // after the stream is idle
    void reset();
    void destroy();
};

// Brackets one kernel launch with events when profiling is on.
struct Launch {
    Profiler &prof;
    hipStream_t s;
    int slot = -1;
    hipEvent_t a{}, b{};
    Launch(Profiler &p, hipStream_t st, const char *name);
    ~Launch() noexcept(false);
};

// --------------------------------------------------------- grid params --
// One DBSCAN instance's grid: cell side s, origin at the instance's minimum,
// G[k] cells along each grid dimension.  Cell keys are
//   key = (instance << lin_bits) | ((c2*G1 + c1)*G0 + c0).
struct InstGrid {
    double origin[3];
    double inv_s;
    int64_t G[3];
    int64_t point_begin;  // first point of the instance (input order)
    int64_t point_end;
    float e2;
    int32_t m;
    int32_t reach;  // stencil half-width in cells
    int32_t core_base;  // cores before this instance (input order), set on device
    int64_t row_base;   // first entry of the instance's rows in a direct row table (eager runs)
};

// Per-chunk bounds partials: min[3], max[3], special count.
struct BoundsPart {
    float lo[3];
    float hi[3];
    int32_t special;
    int32_t count;
};

constexpr int kTile = 64;         // points per counting tile (one wave)
constexpr int kMaxRowsPerCell = 25;  // (2*2+1)^(3-1) stencil rows for d<=3

// ----------------------------------------------------------- primitives --
// Exclusive scan of int32; `total` (device, may be null) receives the sum.
void scan_exclusive_i32(const int32_t *in, int32_t *out, int64_t n, int32_t *total,
                        DevBuf &scratch, hipStream_t s, Profiler &prof);
// Stable LSD radix sort of (key, val) on key bits [0, nbits).
void radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                    int64_t n, int nbits, DevBuf &scratch, hipStream_t s, Profiler &prof,
                    bool *result_in_alt);
void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s, Profiler &prof);
// *out (device, 64-bit) = sum of n int32 values.
void sum_i32_u64(const int32_t *v, int64_t n, unsigned long long *out, hipStream_t s, Profiler &prof);
void fill_u64(uint64_t *p, uint64_t v, int64_t n, hipStream_t s, Profiler &prof);

// ------------------------------------------------------------- context --
enum Path { PATH_GRID = 0, PATH_BRUTE = 1, PATH_MFMA = 2 };

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false: the caller's stream, not destroyed with the context
    int64_t n = 0;
    int d = 0;
    uint32_t flags = 0;
    int n_inst = 1;
    std::vector<int64_t> h_offsets;  // n_inst+1
    std::vector<float> h_eps;
    std::vector<int32_t> h_m;
    bool classified = false, labelled = false;
    int path = PATH_GRID;

    const float *coords = nullptr;  // device AoS n*d
    DevBuf own_coords;

    // results, input order
    DevBuf nn, labels, core;  // int32, int32, uint8
    DevBuf meta;              // int32[3*n_inst]
    std::vector<int32_t> h_meta;
    bool meta_valid = false;

    // grid path state
    DevBuf offsets_d, inst_d, bounds_part, keys, vals, keys2, vals2, flag, cid, cell_of,
        cell_start, cell_key, soa, row_keys, nranges, ranges, tile_off, tiles, nn_s,
        core_s, cell_core_min, cell_has_core, core_bits, border_tiles, parent, comp_min, cell_label, label_s, rank,
        misc, scratch, scratch2, soa_core, mfp_chi, mfp_cent, mfp_rad, mfp_trad, mfp_trho, mfp_tab, mfp_tcm, mfp_moff, mfp_mask, mfp_amask, mfp_tacc, mfp_trep, mfp_tband, mfp_mhi, mfp_mlo, mfp_mna, mfp_emask, mfp_dab, mfp_colpart, cell_c, gflag, ctile_off, ctiles, tile_tests, dense_id, dense_info, dense_tmp, fine_off, part_off, part_fine, big_id;
    int64_t n_cells = 0, n_tiles = 0, n_label_tiles = 0, pair_tests = 0, n_dense = 0;
    int64_t n_rows_direct = 0;  // rows of all instances' (c2, c1) grids (eager runs)
    bool row_direct = false;          // the last grid's row table: direct (else hashed, row_mask)
    uint64_t row_mask = 0;
    bool back_rows_pending = false;   // 3-D stencil planes below the centre not yet computed
    int lin_bits = 0, cell_bits = 0, sub_bits = 0, key_bits = 0, reach_max = 2, rows_per_cell = 1;
    int group_cells = 8;  // cells of one row per counting group
    int quad_mode = -1;   // small tiles four per wave: 1 yes, 0 no, -1 by the tiles' mean size
    bool tuned = false;   // group_cells chosen by the autotuner
    std::vector<InstGrid> h_inst;

    // captured runs (NDBX_GRAPH): an eager run sizes everything, then the
    // grid pipeline runs in capacity mode (device-side counts, launches over
    // the capacities) and is captured into a graph that later runs replay
    bool capacity_mode = false;
    int64_t cap_cells = 0, cap_tiles = 0, n_chunks = 0;
    int64_t cap_rows = 0;  // captured runs: direct row table entries (0: the row hash)
    DevBuf dcount;   // int32: cells, tiles, status (bit 1: a count over capacity, bit 2: plan outside layout)
    DevBuf plan_in;  // chunk table, first chunk per instance, eps, min_pts
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_replays = 0, graph_reruns = 0;

    // brute / mfma path state
    DevBuf b_parent, b_rootmin, pflag, bq, cgroups;
    int64_t bq_cap = 0, bq_last = 0;  // guard-band queue: capacity, last run's count
    unsigned long long bq_last64 = 0;  // (copied back after the count pass)
    int64_t bq_fixed = 0;              // > 0: capacity fixed (NDBX_BAND_QUEUE)
    int audit_stride = 0;              // > 0: the float64 audit takes every audit_stride-th group (NDBX_AUDIT_STRIDE)
    int audit_offset = 0;              // from this group (NDBX_AUDIT_OFFSET, below the stride)
    bool pruned = false;               // the last count pass skipped blocks by the distance bounds

    // audit (mfma)
    int64_t audit_band = 0, audit_fp64_flips = 0, audit_filter_flips = 0;
    double audit_max_err = 0.0;

    // chunked runs (ndbx_set_chunking): slabs of at most chunk_points points
    int64_t chunk_points = 0;
    int64_t n_slabs = 0;
    std::vector<Slab> h_slabs;
    DevBuf ch_perm, ch_grank, ch_parent, ch_gid, ch_coords, ch_ext, ch_inv;
    struct Ctx *sub = nullptr;      // the per-slab context (grid pipeline buffers for one slab)
    bool no_count = false;          // grid_classify: grid and tiles only, counts come from outside
    const int32_t *ext_label = nullptr;  // grid_label: cluster ids of the cores given per point

    Profiler prof;

    void sync();
    ~Ctx();
};

// Pipelines (each file owns its kernels).
void grid_plan(Ctx &c);            // bounds -> path choice and per-instance grids
void grid_classify(Ctx &c);        // keys, sort, cells, stencil, counts
void grid_label(Ctx &c);           // union-find over cells, border, labels
// 3-D: the stencil planes below the centre for the cells the border search
// reads (the first and last cell of each listed tile, the cell of each listed
// point); lists and their lengths on the device
void stencil_back(Ctx &c, const int32_t *blist, const int32_t *n_blist, const int32_t *plist,
                  const int32_t *n_plist);
void brute_classify(Ctx &c);
void brute_label(Ctx &c);
void mfma_classify(Ctx &c);        // d >= 4: partition, bf16 MFMA filter, counts
void mfma_label(Ctx &c);
bool mfma_supported(int d);
void chunked_classify(Ctx &c);     // slab by slab (chunk_points > 0)
void chunked_label(Ctx &c);
void cell_table(const uint64_t *keys, int64_t n, int sub_bits, int32_t *flag, int32_t *excl, int32_t *cell_of,
                int32_t *cell_start, uint64_t *cell_key, int32_t *total, DevBuf &scratch, hipStream_t s,
                Profiler &prof);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace ndbx
