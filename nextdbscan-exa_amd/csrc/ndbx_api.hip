// C ABI of the ndbx library: contexts, argument checks, status codes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <new>

#include "ndbx_internal.h"

namespace ndbx {

static thread_local std::string g_last_error;

void Ctx::sync() {
    NDBX_HIP(hipStreamSynchronize(stream));
    if (prof.on) prof.resolve();
}

Ctx::~Ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    prof.destroy();
    DevBuf *bufs[] = {&own_coords, &nn, &labels, &core, &meta, &offsets_d, &inst_d, &bounds_part, &keys, &vals,
                      &keys2, &vals2, &flag, &cid, &cell_of, &cell_start, &cell_key, &soa, &row_keys,
                      &nranges, &ranges, &tile_off, &tiles, &nn_s, &core_s, &cell_core_min, &cell_has_core,
                      &parent, &comp_min, &cell_label, &label_s, &rank, &misc, &scratch, &scratch2, &b_parent,
                      &b_rootmin, &soa_core, &cell_c, &gflag, &ctile_off, &ctiles, &tile_tests, &dense_id, &dense_info, &dense_tmp, &fine_off, &dcount, &plan_in, &part_off,
                      &pflag, &bq, &cgroups, &part_fine, &big_id, &core_bits, &border_tiles,
                      &ch_perm, &ch_grank, &ch_parent, &ch_gid, &ch_coords, &ch_ext, &ch_inv};
    delete sub;
    for (auto *b : bufs) b->release();
    if (stream && own_stream) (void)hipStreamDestroy(stream);
}

template <class F>
static int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return NDBX_OK;
    } catch (const Error &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return NDBX_ENOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return NDBX_EDEVICE;
    }
}

static void need(bool ok, const char *msg) {
    if (!ok) throw Error(NDBX_EINVAL, msg);
}

// Every flag bit include/ndbx.h defines.  NDBX_APPROX is accepted and runs
// the exact labelling: cores and noise are the same in both of the
// reference's modes (main.cpp:45), and the exact cluster count is the
// tightest of its upper-bound approximate counts (data_process.cpp:214-219).
constexpr uint32_t kKnownFlags = NDBX_APPROX | NDBX_FORCE_BRUTE | NDBX_DEVICE_INPUT | NDBX_PROFILE |
                                 NDBX_MFMA_FILTER | NDBX_FP64_AUDIT | NDBX_GRAPH | NDBX_AUTOTUNE;

static void open_common(Ctx &c, const float *coords, int64_t n_total, uint32_t flags) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(NDBX_EDEVICE, "no HIP device");
    NDBX_HIP(hipGetDevice(&c.device));
    NDBX_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.prof.on = (flags & NDBX_PROFILE) != 0;
    if (const char *gc = std::getenv("NDBX_GROUP_CELLS")) c.group_cells = std::max(1, std::atoi(gc));
    if (const char *qt = std::getenv("NDBX_QUAD_TILES")) c.quad_mode = std::atoi(qt) != 0;
    // fixed guard-band queue capacity of the MFMA count pass (tests of its overflow path)
    if (const char *bq = std::getenv("NDBX_BAND_QUEUE")) c.bq_fixed = std::max(1ll, std::atoll(bq));
    if (const char *as = std::getenv("NDBX_AUDIT_STRIDE")) c.audit_stride = std::max(1, std::atoi(as));
    if (const char *ao = std::getenv("NDBX_AUDIT_OFFSET")) c.audit_offset = std::max(0, std::atoi(ao));
    c.n = n_total;
    c.flags = flags;
    const size_t bytes = (size_t)n_total * (size_t)c.d * sizeof(float);
    if (flags & NDBX_DEVICE_INPUT) {
        c.coords = coords;
    } else if (n_total > 0) {
        c.own_coords.ensure(bytes);
        NDBX_HIP(hipMemcpyAsync(c.own_coords.p, coords, bytes, hipMemcpyHostToDevice, c.stream));
        c.coords = c.own_coords.as<float>();
    }
    c.nn.ensure(n_total * 4);
    c.labels.ensure(n_total * 4);
    c.core.ensure(n_total);
    c.meta.ensure((size_t)c.n_inst * 3 * 4);
    c.sync();
}

static void run_classify(Ctx &c) {
    c.labelled = false;
    c.meta_valid = false;
    if (c.chunk_points > 0 && c.n > 0) {
        chunked_classify(c);
        c.classified = true;
        return;
    }
    c.n_slabs = 0;
    grid_plan(c);
    if (c.n == 0) {
        c.classified = true;
        return;
    }
    if (c.path == PATH_GRID) grid_classify(c);
    else if (c.path == PATH_MFMA) mfma_classify(c);
    else brute_classify(c);
    c.classified = true;
}

static void run_label(Ctx &c) {
    if (!c.classified) throw Error(NDBX_ESTATE, "ndbx_label before ndbx_classify");
    // labels are a function of the last classify, whose working buffers the
    // label pass consumes (the MFMA path reuses its tile tables as scratch):
    // a second call keeps the labels it has
    if (c.labelled) return;
    if (c.n == 0) {
        fill_i32(c.meta.as<int32_t>(), 0, 3 * c.n_inst, c.stream, c.prof);
    } else if (c.chunk_points > 0) {
        chunked_label(c);
    } else if (c.path == PATH_GRID) {
        grid_label(c);
    } else if (c.path == PATH_MFMA) {
        mfma_label(c);
    } else {
        brute_label(c);
    }
    c.h_meta.assign((size_t)c.n_inst * 3, 0);
    NDBX_HIP(hipMemcpyAsync(c.h_meta.data(), c.meta.p, c.h_meta.size() * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    c.meta_valid = true;
    c.labelled = true;
}

// Tile-shape autotuner: cells per counting group (the run of cells of one
// row that share tiles) from {1, 2, 4, 8, 16}, and tiles per wave (one, or
// four tiles of at most 16 queries sharing a wave), chosen by the count
// kernel's time on the context's own input.  Profiling data recorded so far
// is kept.
static void autotune(Ctx &c) {
    c.sync();
    Profiler saved_state = c.prof;
    const bool was_on = c.prof.on;
    c.prof.pending.clear();
    c.prof.names.clear();
    c.prof.ms.clear();
    c.prof.calls.clear();
    c.prof.on = true;
    c.prof.only = "count";
    int best = c.group_cells, best_q = c.quad_mode;
    double best_ms = 1e300;
    for (int gc : {1, 2, 4, 8, 16})
        for (int q : {0, 1}) {
            c.group_cells = gc;
            c.quad_mode = q;
            run_classify(c);
            c.sync();
            const int slot = c.prof.slot_of("count");
            const double ms = c.prof.calls[slot] ? c.prof.ms[slot] : 1e300;
            if (ms < best_ms) {
                best_ms = ms;
                best = gc;
                best_q = q;
            }
            c.prof.ms.assign(c.prof.ms.size(), 0.0);
            c.prof.calls.assign(c.prof.calls.size(), 0);
        }
    c.group_cells = best;
    c.quad_mode = best_q;
    c.tuned = true;
    // keep the event pool, restore what the caller had recorded
    saved_state.pool = c.prof.pool;
    c.prof = saved_state;
    c.prof.on = was_on;
}

// Eager classify + label, then, where the run is capturable (grid path, no
// dense 2-D cells), a sizing run in capacity mode and the capture of that
// pipeline into a graph.  Later runs replay the graph; a replay whose counts
// exceed the capacities or whose geometry leaves the captured key layout is
// rerun eagerly and captured again.
static void destroy_graph(Ctx &c) {
    if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph = nullptr;
}

static void capacity_pipeline(Ctx &c) {
    fill_i32(c.dcount.as<int32_t>() + 2, 0, 1, c.stream, c.prof);
    grid_plan(c);
    grid_classify(c);
    grid_label(c);
}

static void read_meta(Ctx &c, int32_t *status) {
    c.h_meta.assign((size_t)c.n_inst * 3, 0);
    NDBX_HIP(hipMemcpyAsync(c.h_meta.data(), c.meta.p, c.h_meta.size() * 4, hipMemcpyDeviceToHost, c.stream));
    if (status) NDBX_HIP(hipMemcpyAsync(status, c.dcount.as<int32_t>() + 2, 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

static void batch_run_graph(Ctx &c) {
    if (c.graph_exec) {
        int32_t status = 0;
        NDBX_HIP(hipGraphLaunch(c.graph_exec, c.stream));
        read_meta(c, &status);
        if (status == 0) {
            c.classified = c.labelled = c.meta_valid = true;
            ++c.graph_replays;
            return;
        }
        ++c.graph_reruns;
        destroy_graph(c);
    }
    if ((c.flags & NDBX_AUTOTUNE) && !c.tuned) autotune(c);
    run_classify(c);
    run_label(c);
    if (c.path != PATH_GRID || c.d > 3 || c.n == 0 || c.n_dense > 0) return;
    c.dcount.ensure(64);
    c.cap_cells = std::min<int64_t>(c.n, c.n_cells + c.n_cells / 4 + 1024);
    c.cap_tiles = std::min<int64_t>(c.n + c.cap_cells, c.n_tiles + c.n_tiles / 4 + 1024);
    // a direct row table where the eager run took one
    c.cap_rows = c.row_direct ? c.n_rows_direct + c.n_rows_direct / 4 + 1024 : 0;
    const bool prof_on = c.prof.on;
    try {
        // sizing run: grows every buffer to its capacity size, same results
        c.capacity_mode = true;
        capacity_pipeline(c);
        int32_t status = 0;
        read_meta(c, &status);
        if (status != 0) {
            c.capacity_mode = false;
            return;
        }
        c.prof.on = false;
        NDBX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        g_no_alloc = true;
        try {
            capacity_pipeline(c);
        } catch (...) {
            g_no_alloc = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        g_no_alloc = false;
        NDBX_HIP(hipStreamEndCapture(c.stream, &c.graph));
        NDBX_HIP(hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0));
    } catch (...) {
        c.capacity_mode = false;
        c.prof.on = prof_on;
        destroy_graph(c);
        throw;
    }
    c.capacity_mode = false;
    c.prof.on = prof_on;
}

// ----------------------------------------------------- no-alloc calls --
// The fixed device array behind ndbx_dbscan_noalloc.  Host code never uses
// its name as an address (on the host it is a stand-in, not device memory):
// the device address comes from hipGetSymbolAddress, taken once per device
// with that device current, and only that address reaches kernels and copies.
static __device__ __align__(256) unsigned char scratch[NDBX_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::vector<char *> g_scratch_ptr;  // per device; null until set up

static char *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        NDBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(NDBX_EDEVICE, "no-alloc scratch: hipGetSymbolAddress gave a null address");
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    return g_scratch_ptr[device];
}

// Points DevBuf::ensure at an arena on this thread for one call.
struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a) : prev(g_arena) { g_arena = a; }
    ~ArenaScope() { g_arena = prev; }
};

}  // namespace ndbx

using namespace ndbx;

struct ndbx_ctx {
    Ctx c;
};

extern "C" {

int ndbx_version(void) { return 100; }

const char *ndbx_last_error(void) { return g_last_error.c_str(); }

int ndbx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ndbx_create(ndbx_ctx **out, const float *coords, int64_t n, int d, float eps, int min_pts, uint32_t flags) {
    return guarded([&] {
        need(out != nullptr, "ctx out-pointer is null");
        *out = nullptr;
        need(n >= 0, "n < 0");
        need(n < (int64_t)INT32_MAX - 64, "n must be below 2^31 - 64 (the reference's .bin header is int32)");
        need(d >= 1, "d < 1");
        need(n == 0 || coords != nullptr, "coords is null");
        need((flags & ~kKnownFlags) == 0, "unknown flag bits");
        auto *h = new ndbx_ctx();
        Ctx &c = h->c;
        c.d = d;
        c.n_inst = 1;
        c.h_offsets = {0, n};
        c.h_eps = {eps};
        c.h_m = {min_pts};
        try {
            open_common(c, coords, n, flags);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int ndbx_batch_create(ndbx_ctx **out, const float *coords, const int64_t *offsets, int n_inst, int d,
                      const float *eps, const int *min_pts, uint32_t flags) {
    return guarded([&] {
        need(out != nullptr, "ctx out-pointer is null");
        *out = nullptr;
        need(n_inst >= 1, "n_inst < 1");
        need(offsets && eps && min_pts, "null offsets/eps/min_pts");
        need(d >= 1, "d < 1");
        need(offsets[0] == 0, "offsets[0] must be 0");
        for (int i = 0; i < n_inst; ++i) need(offsets[i + 1] >= offsets[i], "offsets must be non-decreasing");
        const int64_t n = offsets[n_inst];
        need(n < (int64_t)INT32_MAX - 64, "total points must be below 2^31 - 64");
        need(n == 0 || coords != nullptr, "coords is null");
        need((flags & ~kKnownFlags) == 0, "unknown flag bits");
        auto *h = new ndbx_ctx();
        Ctx &c = h->c;
        c.d = d;
        c.n_inst = n_inst;
        c.h_offsets.assign(offsets, offsets + n_inst + 1);
        c.h_eps.assign(eps, eps + n_inst);
        c.h_m.assign(min_pts, min_pts + n_inst);
        try {
            open_common(c, coords, n, flags);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int ndbx_classify(ndbx_ctx *h) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        run_classify(h->c);
    });
}

int ndbx_label(ndbx_ctx *h) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        run_label(h->c);
    });
}

int ndbx_batch_run(ndbx_ctx *h) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        Ctx &c = h->c;
        if (c.flags & NDBX_GRAPH) {
            batch_run_graph(c);
            return;
        }
        if ((c.flags & NDBX_AUTOTUNE) && !c.tuned) autotune(c);
        run_classify(c);
        run_label(c);
    });
}

int ndbx_batch_set_coords(ndbx_ctx *h, const float *coords) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        Ctx &c = h->c;
        need(!(c.flags & NDBX_DEVICE_INPUT), "device input: write the caller's buffer instead");
        need(c.n == 0 || coords != nullptr, "coords is null");
        if (c.n == 0) return;
        NDBX_HIP(hipMemcpyAsync(c.own_coords.p, coords, (size_t)c.n * c.d * sizeof(float), hipMemcpyHostToDevice,
                                c.stream));
        c.sync();
        c.classified = c.labelled = c.meta_valid = false;
    });
}

int ndbx_batch_info(ndbx_ctx *h, int *group_cells, int64_t *graph_replays, int64_t *graph_reruns) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (group_cells) *group_cells = h->c.group_cells;
        if (graph_replays) *graph_replays = h->c.graph_replays;
        if (graph_reruns) *graph_reruns = h->c.graph_reruns;
    });
}

int ndbx_tile_shape(ndbx_ctx *h, int *group_cells, int *tiles_per_wave) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (group_cells) *group_cells = h->c.group_cells;
        if (tiles_per_wave) *tiles_per_wave = h->c.quad_mode < 0 ? 0 : h->c.quad_mode ? 4 : 1;
    });
}

int ndbx_result_meta(ndbx_ctx *h, int *cores, int *noise, int *clusters, int *n) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        Ctx &c = h->c;
        if (!c.meta_valid) throw Error(NDBX_ESTATE, "ndbx_result_meta before ndbx_label");
        int64_t a = 0, b = 0, k = 0;
        for (int i = 0; i < c.n_inst; ++i) {
            a += c.h_meta[3 * i];
            b += c.h_meta[3 * i + 1];
            k += c.h_meta[3 * i + 2];
        }
        if (cores) *cores = (int)a;
        if (noise) *noise = (int)b;
        if (clusters) *clusters = (int)k;
        if (n) *n = (int)c.n;
    });
}

int ndbx_batch_meta(ndbx_ctx *h, int32_t *meta3) {
    return guarded([&] {
        need(h != nullptr && meta3 != nullptr, "null argument");
        Ctx &c = h->c;
        if (!c.meta_valid) throw Error(NDBX_ESTATE, "ndbx_batch_meta before ndbx_batch_run");
        std::memcpy(meta3, c.h_meta.data(), c.h_meta.size() * 4);
    });
}

static void copy_out(Ctx &c, const DevBuf &b, void *out, size_t bytes, bool need_label) {
    if (!(need_label ? c.labelled : c.classified))
        throw Error(NDBX_ESTATE, need_label ? "labels requested before ndbx_label"
                                            : "results requested before ndbx_classify");
    if (bytes == 0) return;
    NDBX_HIP(hipMemcpyAsync(out, b.p, bytes, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

int ndbx_get_nn(ndbx_ctx *h, int32_t *out) {
    return guarded([&] {
        need(h != nullptr && (out != nullptr || h->c.n == 0), "null argument");
        copy_out(h->c, h->c.nn, out, (size_t)h->c.n * 4, false);
    });
}

int ndbx_get_labels(ndbx_ctx *h, int32_t *out) {
    return guarded([&] {
        need(h != nullptr && (out != nullptr || h->c.n == 0), "null argument");
        copy_out(h->c, h->c.labels, out, (size_t)h->c.n * 4, true);
    });
}

int ndbx_get_core(ndbx_ctx *h, uint8_t *out) {
    return guarded([&] {
        need(h != nullptr && (out != nullptr || h->c.n == 0), "null argument");
        copy_out(h->c, h->c.core, out, (size_t)h->c.n, false);
    });
}

int ndbx_device_results(ndbx_ctx *h, const int32_t **nn, const int32_t **labels) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (!h->c.labelled) throw Error(NDBX_ESTATE, "results requested before ndbx_label");
        if (nn) *nn = h->c.nn.as<int32_t>();
        if (labels) *labels = h->c.labels.as<int32_t>();
    });
}

int ndbx_path(ndbx_ctx *h, int *path) {
    return guarded([&] {
        need(h != nullptr && path != nullptr, "null argument");
        *path = h->c.path;
    });
}

int ndbx_sync(ndbx_ctx *h) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        h->c.sync();
    });
}

int ndbx_profile_count(ndbx_ctx *h, int *count) {
    return guarded([&] {
        need(h != nullptr && count != nullptr, "null argument");
        h->c.sync();
        *count = (int)h->c.prof.names.size();
    });
}

int ndbx_profile_entry(ndbx_ctx *h, int i, const char **name, double *total_ms, int64_t *launches) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        Profiler &p = h->c.prof;
        need(i >= 0 && i < (int)p.names.size(), "profile index out of range");
        if (name) *name = p.names[i].c_str();
        if (total_ms) *total_ms = p.ms[i];
        if (launches) *launches = p.calls[i];
    });
}

int ndbx_profile_reset(ndbx_ctx *h) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        h->c.sync();
        h->c.prof.reset();
    });
}

int ndbx_profile_select(ndbx_ctx *h, const char *name) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        h->c.prof.only = name ? name : "";
    });
}

int ndbx_stats(ndbx_ctx *h, int64_t *cells, int64_t *tiles, int64_t *pair_tests) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (cells) *cells = h->c.n_cells;
        if (tiles) *tiles = h->c.n_tiles;
        if (pair_tests) *pair_tests = h->c.pair_tests;
    });
}

int ndbx_audit(ndbx_ctx *h, int64_t *band_pairs, int64_t *fp64_flips, int64_t *filter_flips, double *max_abs_err) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (band_pairs) *band_pairs = h->c.audit_band;
        if (fp64_flips) *fp64_flips = h->c.audit_fp64_flips;
        if (filter_flips) *filter_flips = h->c.audit_filter_flips;
        if (max_abs_err) *max_abs_err = h->c.audit_max_err;
    });
}

int ndbx_destroy(ndbx_ctx *h) {
    return guarded([&] { delete h; });
}

int ndbx_set_chunking(ndbx_ctx *h, int64_t max_slab_points) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        need(max_slab_points >= 0, "max_slab_points < 0");
        Ctx &c = h->c;
        need(max_slab_points == 0 || c.n_inst == 1, "chunked runs take one instance (not a batch)");
        need(max_slab_points == 0 || !(c.flags & NDBX_FORCE_BRUTE), "chunked runs take the cell grid");
        c.chunk_points = max_slab_points;
        c.classified = c.labelled = c.meta_valid = false;
    });
}

int ndbx_chunk_info(ndbx_ctx *h, int64_t *slabs) {
    return guarded([&] {
        need(h != nullptr, "ctx is null");
        if (slabs) *slabs = h->c.n_slabs;
    });
}

int ndbx_dbscan_noalloc(const float *coords, int64_t n, int d, float eps, int min_pts, uint32_t flags,
                        void *stream, int32_t *nn, int32_t *labels, int32_t *meta3) {
    return guarded([&] {
        need(n >= 0, "n < 0");
        need(n <= NDBX_NOALLOC_MAX_POINTS, ("n = " + std::to_string(n) + " exceeds NDBX_NOALLOC_MAX_POINTS (" +
                                            std::to_string(NDBX_NOALLOC_MAX_POINTS) + ")").c_str());
        need(d >= 1 && d <= 3, "ndbx_dbscan_noalloc takes 1 <= d <= 3");
        need((flags & ~(uint32_t)(NDBX_APPROX | NDBX_FORCE_BRUTE)) == 0,
             "ndbx_dbscan_noalloc accepts only NDBX_APPROX and NDBX_FORCE_BRUTE");
        need(meta3 != nullptr, "meta3 is null");
        need(n == 0 || (coords && nn && labels), "null coords, nn or labels");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(NDBX_EDEVICE, "no HIP device");
        int dev = 0;
        NDBX_HIP(hipGetDevice(&dev));
        // the array's last MiB stays unused, so a vector load that reads a few
        // bytes past the last buffer still reads inside the array
        Arena arena{scratch_ptr(dev), (size_t)NDBX_NOALLOC_SCRATCH_BYTES - (1u << 20), 0};
        ArenaScope scope(&arena);
        Ctx c;
        c.device = dev;
        c.stream = static_cast<hipStream_t>(stream);
        c.own_stream = false;
        c.d = d;
        c.n = n;
        c.n_inst = 1;
        c.h_offsets = {0, n};
        c.h_eps = {eps};
        c.h_m = {min_pts};
        c.flags = flags | NDBX_DEVICE_INPUT;
        c.coords = coords;
        c.nn.borrow(nn, (size_t)n * 4);
        c.labels.borrow(labels, (size_t)n * 4);
        c.core.ensure((size_t)n);
        c.meta.ensure(3 * 4);
        run_classify(c);
        run_label(c);
        std::memcpy(meta3, c.h_meta.data(), 3 * sizeof(int32_t));
    });
}

}  // extern "C"
