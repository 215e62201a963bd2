// C ABI of the ndbx library: contexts, argument checks, status codes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
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
    prof.destroy();
    DevBuf *bufs[] = {&own_coords, &nn, &labels, &core, &meta, &offsets_d, &inst_d, &bounds_part, &keys, &vals,
                      &keys2, &vals2, &flag, &cid, &cell_of, &cell_start, &cell_key, &soa, &row_keys, &row_vals,
                      &nranges, &ranges, &tile_off, &tiles, &nn_s, &core_s, &cell_core_min, &cell_has_core,
                      &parent, &comp_min, &cell_label, &label_s, &rank, &misc, &scratch, &scratch2, &b_parent,
                      &b_rootmin, &soa_core, &cell_c, &gflag, &ctile_off, &ctiles, &tile_tests, &dense_id, &dense_info, &dense_tmp, &fine_off};
    for (auto *b : bufs) b->release();
    if (stream) (void)hipStreamDestroy(stream);
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

static void open_common(Ctx &c, const float *coords, int64_t n_total, uint32_t flags) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(NDBX_EDEVICE, "no HIP device");
    NDBX_HIP(hipGetDevice(&c.device));
    NDBX_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.prof.on = (flags & NDBX_PROFILE) != 0;
    if (const char *gc = std::getenv("NDBX_GROUP_CELLS")) c.group_cells = std::max(1, std::atoi(gc));
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
    if (c.n == 0) {
        fill_i32(c.meta.as<int32_t>(), 0, 3 * c.n_inst, c.stream, c.prof);
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
        run_classify(h->c);
        run_label(h->c);
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

}  // extern "C"
