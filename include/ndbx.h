/*
 * ndbx: MI355X (gfx950) DBSCAN hot path behind a C ABI.
 *
 * The reference (ernire/nextdbscan-exa) selects its compute backend at
 * compile time (data_process.cpp:24-30: OMP_ON -> magma_exa_omp.h, CUDA_ON ->
 * magma_exa_cu.h) and drives it through `class data_process`
 * (data_process.h:45-102) from exactly one caller, nextdbscan::start
 * (nextdbscan.cpp:46-59).  These entry points replace that class's CUDA
 * (Thrust) backend; the host `data_process` of this project calls them in the
 * same order nextdbscan::start calls the reference's methods.
 *
 * Conventions: C linkage; plain pointers and sizes; every call returns an
 * ndbx_status (0 = OK); ndbx_last_error() gives a thread-local message for the
 * last failing call on this thread.  A context owns its device buffers and one
 * HIP stream.  Contexts are independent and may be used from separate host
 * threads; a single context is not re-entrant.
 */
#ifndef NDBX_H
#define NDBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ndbx_ctx ndbx_ctx;

typedef enum ndbx_status {
    NDBX_OK = 0,
    NDBX_EINVAL = 1,      /* bad argument (null pointer, n < 0, d < 1, ...) */
    NDBX_ENOMEM = 2,      /* device or host allocation failed */
    NDBX_EDEVICE = 3,     /* HIP runtime / kernel launch error, or no device */
    NDBX_ESTATE = 4,      /* call out of order (e.g. label before classify) */
    NDBX_EUNSUPPORTED = 5 /* configuration outside what this build handles */
} ndbx_status;

/* Flags for ndbx_create / ndbx_batch_create. */
enum {
    /* -a 1 (main.cpp:45, data_process.cpp:214-219): the reference stops after
     * one label sweep, so its cluster count is an upper bound while cores and
     * noise are the same as in exact mode.  This build returns the exact
     * labelling in both modes: the tightest of those upper bounds.  Flag bits
     * not defined here are refused (NDBX_EINVAL). */
    NDBX_APPROX = 1u << 0,
    /* Force the all-pairs kernels even where the cell grid applies. */
    NDBX_FORCE_BRUTE = 1u << 1,
    /* `coords` is a device pointer (already resident in HBM); not copied H2D
     * and not owned: it must stay valid until ndbx_destroy. */
    NDBX_DEVICE_INPUT = 1u << 2,
    /* Record HIP events around every kernel launch; see ndbx_profile_*. */
    NDBX_PROFILE = 1u << 3,
    /* Accepted for compatibility; 4 <= d <= 128 always takes the BF16 MFMA
     * distance filter (FP32 recheck of guard-band pairs, so every decision
     * is the FP32 predicate's) unless NDBX_FORCE_BRUTE is set. */
    NDBX_MFMA_FILTER = 1u << 4,
    /* MFMA filter: recheck every filter decision against a float64 distance
     * and record the maximum error and flip counts (ndbx_audit). */
    NDBX_FP64_AUDIT = 1u << 5,
    /* ndbx_batch_run: after the first (eager) run, capture the grid pipeline
     * into a HIP graph and replay it on later runs.  A replay whose cell or
     * tile counts exceed the captured capacities, or whose cell geometry
     * leaves the captured key layout, is rerun eagerly and captured again. */
    NDBX_GRAPH = 1u << 6,
    /* ndbx_batch_run: choose the counting tile shape (cells of one row per
     * tile group) by timing the count kernel on the context's input, once. */
    NDBX_AUTOTUNE = 1u << 7
};

/* Cluster index of a point that belongs to no cluster (data_process.h:26). */
#define NDBX_NO_CLUSTER INT32_MAX

/* Library version (major*10000 + minor*100 + patch). */
int ndbx_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *ndbx_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int ndbx_device_count(void);

/*
 * Replaces data_process::data_process (data_process.h:61-76).
 * coords: n*d float32, row-major (AoS) as read_input leaves it
 * (magma_input.h:113-125).  Host memory is copied to the device; with
 * NDBX_DEVICE_INPUT it is a device pointer used in place.  e2 = eps*eps is
 * computed in float exactly as data_process.h:73 does.  n may be 0.
 */
int ndbx_create(ndbx_ctx **ctx, const float *coords, int64_t n, int d, float eps,
                int min_pts, uint32_t flags);

/* Replaces data_process::select_and_process (data_process.cpp:334-392) and
 * process_points (:395-458): per-point neighbour counts nn (self included)
 * and core flags nn >= min_pts. */
int ndbx_classify(ndbx_ctx *ctx);

/* Replaces data_process::select_cores_and_process (data_process.cpp:268-332)
 * and process_cores (:93-266): cluster index per point.  A core's index is
 * the rank (among cores in ascending point id) of the lowest-id core of its
 * connected component; a non-core point within eps of a core gets the
 * smallest index among its cores; others get NDBX_NO_CLUSTER. */
int ndbx_label(ndbx_ctx *ctx);

/* Replaces data_process::get_result_meta (data_process.cpp:50-91). Any
 * output pointer may be NULL. */
int ndbx_result_meta(ndbx_ctx *ctx, int *cores, int *noise, int *clusters, int *n);

/* Copies of the per-point results in input order (v_coord_nn,
 * v_coord_cluster_index, data_process.h:57-58) into host memory; core is
 * nn >= min_pts as 0/1 bytes. */
int ndbx_get_nn(ndbx_ctx *ctx, int32_t *out);
int ndbx_get_labels(ndbx_ctx *ctx, int32_t *out);
int ndbx_get_core(ndbx_ctx *ctx, uint8_t *out);

/* Device pointers to the same arrays (valid until the next classify/label
 * or destroy), for callers that keep results in HBM. */
int ndbx_device_results(ndbx_ctx *ctx, const int32_t **nn, const int32_t **labels);

/* Which kernels ran: 0 = cell grid, 1 = all-pairs, 2 = MFMA filter. */
int ndbx_path(ndbx_ctx *ctx, int *path);

/* Blocks until the context's stream is idle. */
int ndbx_sync(ndbx_ctx *ctx);

/* Per-kernel timing (NDBX_PROFILE): entries accumulate over calls until
 * ndbx_profile_reset.  `name` stays valid for the context's lifetime. */
int ndbx_profile_count(ndbx_ctx *ctx, int *count);
int ndbx_profile_entry(ndbx_ctx *ctx, int i, const char **name, double *total_ms,
                       int64_t *launches);
int ndbx_profile_reset(ndbx_ctx *ctx);
/* Restricts the events to launches named `name` (NULL or "" = every
 * launch), so a timed run can keep its one kernel of interest evented
 * without the per-launch overhead of the others. */
int ndbx_profile_select(ndbx_ctx *ctx, const char *name);

/* Grid statistics of the last classify: non-empty cells, tiles, pair tests
 * issued by the counting kernel (collected while NDBX_PROFILE events every
 * launch; otherwise the last collected value, 0 if none). */
int ndbx_stats(ndbx_ctx *ctx, int64_t *cells, int64_t *tiles, int64_t *pair_tests);

/*
 * Chunked runs (data_process.cpp:285-302, :335-362 bound the reference's
 * working copies to chunks of 100M points).  With max_slab_points > 0 the
 * next ndbx_classify / ndbx_label run slab by slab: points sorted by their
 * first coordinate, each slab a range of that order plus every point within
 * eps of it in that coordinate, at most max_slab_points points; the grid
 * pipeline's buffers are sized for one slab, and per-point state (counts,
 * core flags, labels, core ranks: 17 bytes per point, plus the coordinates)
 * stays resident.  Results are identical to the resident run.  Takes one
 * instance on the cell grid (d <= 3, finite coordinates; else NDBX_EUNSUPPORTED
 * from ndbx_classify), and a budget that holds the points within eps (in the
 * first coordinate) of every point.  0 restores resident runs.
 */
int ndbx_set_chunking(ndbx_ctx *ctx, int64_t max_slab_points);
/* Slabs of the last chunked classify (0 for resident runs). */
int ndbx_chunk_info(ndbx_ctx *ctx, int64_t *slabs);

/* MFMA filter audit (NDBX_FP64_AUDIT): guard-band pairs rechecked, decisions
 * that differ between the FP32 predicate and a float64 distance, decisions
 * of the BF16 estimate outside the band that the FP32 predicate contradicts
 * (must be 0), and the largest |bf16 estimate - float64 distance^2|. */
int ndbx_audit(ndbx_ctx *ctx, int64_t *band_pairs, int64_t *fp64_flips,
               int64_t *filter_flips, double *max_abs_err);

int ndbx_destroy(ndbx_ctx *ctx);

/*
 * Batched form: n_inst independent DBSCAN instances, one run.  Instance i is
 * points [offsets[i], offsets[i+1]) of `coords` (n_total*d float32 AoS, host
 * memory, or device memory with NDBX_DEVICE_INPUT), with its own eps[i] and
 * min_pts[i].  Cluster indices are per instance, as if each instance were run
 * alone.  No reference counterpart (the reference runs one dataset per
 * process); results are defined instance-wise by ndbx_classify/ndbx_label.
 */
int ndbx_batch_create(ndbx_ctx **ctx, const float *coords, const int64_t *offsets,
                      int n_inst, int d, const float *eps, const int *min_pts,
                      uint32_t flags);
/* Runs classify + label for every instance. */
int ndbx_batch_run(ndbx_ctx *ctx);
/* Per instance (cores, noise, clusters) triples, 3*n_inst ints. */
int ndbx_batch_meta(ndbx_ctx *ctx, int32_t *meta3);
/* New coordinates for the same instances (host memory, n_total*d floats);
 * not for NDBX_DEVICE_INPUT contexts, whose caller writes its own buffer. */
int ndbx_batch_set_coords(ndbx_ctx *ctx, const float *coords);
/* Tile shape in use (cells per counting group) and graph statistics:
 * replays that completed, and replays rerun eagerly. */
int ndbx_batch_info(ndbx_ctx *ctx, int *group_cells, int64_t *graph_replays,
                    int64_t *graph_reruns);
/* Tile shape of the count: cells per counting group, and tiles per wave
 * (1, or 4 when tiles of at most 16 queries share a wave; 0 while the
 * choice is left to the tiles' mean size).  NDBX_AUTOTUNE picks both by
 * timing the count on the context's input; the environment variables
 * NDBX_GROUP_CELLS and NDBX_QUAD_TILES (0 or 1) fix them at creation.
 * No counterpart in the reference (its count has no tiles). */
int ndbx_tile_shape(ndbx_ctx *ctx, int *group_cells, int *tiles_per_wave);

/* Largest input of ndbx_dbscan_noalloc, and the fixed device memory it
 * works in (held on each device the library runs on, until the process ends). */
#define NDBX_NOALLOC_MAX_POINTS 131072
#define NDBX_NOALLOC_SCRATCH_BYTES (64u << 20)

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): classify + label of one instance, as ndbx_create with
 * NDBX_DEVICE_INPUT + ndbx_classify + ndbx_label + ndbx_result_meta, with
 * every buffer beyond the caller's inputs and outputs taken from the
 * library's fixed device array of NDBX_NOALLOC_SCRATCH_BYTES; no hipMalloc,
 * hipFree or other allocating call is made.
 *   coords  n*d float32 AoS, device memory
 *   nn      n int32, device memory: neighbour counts (as ndbx_get_nn)
 *   labels  n int32, device memory: cluster index per point (as ndbx_get_labels)
 *   meta3   host: (cores, noise, clusters)
 *   stream  a hipStream_t to run on, or NULL for the null stream; the call
 *           returns after that stream is idle
 * Needs 1 <= d <= 3 and 0 <= n <= NDBX_NOALLOC_MAX_POINTS (else NDBX_EINVAL);
 * flags may hold NDBX_APPROX and NDBX_FORCE_BRUTE only.  Every call on one
 * device shares that device's array, which keeps the last call's contents:
 * calls that overlap on one device (two threads, or two streams) race.
 */
int ndbx_dbscan_noalloc(const float *coords, int64_t n, int d, float eps, int min_pts,
                        uint32_t flags, void *stream, int32_t *nn, int32_t *labels,
                        int32_t *meta3);

#ifdef __cplusplus
}
#endif

#endif /* NDBX_H */
