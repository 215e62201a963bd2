/*
 * CPU restatement of the reference DBSCAN path (ernire/nextdbscan-exa).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function restates the reference's arithmetic, not its parallel
 * schedule:
 *   - the neighbour predicate is data_process.cpp:441-449 (and :196-203):
 *     len = 0.0f; for d ascending: len += (p1[d]-p2[d])*(p1[d]-p2[d]);
 *     neighbour iff len <= e2, with e2 = e*e in float (data_process.h:73).
 *     It must be built without FMA contraction (-ffp-contract=off), as the
 *     reference's g++ build has none.
 *   - a candidate j whose coordinate 0 equals FLT_MAX is skipped as padding
 *     (data_process.cpp:440 for counting, :194 for the label sweep).
 *   - core <=> nn >= m, self included (data_process.cpp:282, :58).
 *   - cluster ids: each core's id is the rank, among cores in ascending point
 *     id, of the lowest-id core of its connected component (the fixed point of
 *     the reference's label propagation, data_process.cpp:149-260, with labels
 *     initialised to core rank at :151-170 and only ever lowered by atomic_min).
 *     Components are found with a textbook union-find instead of replaying the
 *     racy sweep (the sweep's termination test at :214-217 can miss a hook).
 *   - a non-core point within eps of at least one core gets the smallest
 *     cluster id among those cores (the last, event-free sweep applies
 *     atomic_min over all of them, data_process.cpp:204-209); otherwise it is
 *     noise, NO_CLUSTER = INT32_MAX (data_process.h:26).
 *   - meta: cores = #(nn>=m), noise = #(label==NO_CLUSTER), clusters =
 *     #components (data_process.cpp:50-91, label[k]==k roots).
 */
#include <float.h>
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_NO_CLUSTER INT32_MAX

/* data_process.cpp:443-447: summation in ascending d from 0.0f, '<=' test. */
static inline int oracle_is_neighbour(const float *p1, const float *p2, int d,
                                      float e2) {
    float length = 0.0f;
    for (int k = 0; k < d; ++k) {
        float t = (p1[k] - p2[k]) * (p1[k] - p2[k]);
        length = length + t;
    }
    return length <= e2;
}

/* data_process.cpp:440 / :194: a candidate with p[0]==FLT_MAX is padding. */
static inline int oracle_valid_source(const float *p) { return p[0] != FLT_MAX; }

/* data_process.h:73 */
float oracle_e2(float e) { return e * e; }

/* data_process.cpp:395-458: nn[i] = #{ j : valid_source(j) && pred(i, j) }.
 * The reference pushes from chunk point j with atomic adds; the sum is the
 * same integer, so this pulls per i. */
void oracle_classify(const float *x, int64_t n, int d, float e2, int32_t *nn) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float *p1 = x + i * d;
        int32_t hits = 0;
        for (int64_t j = 0; j < n; ++j) {
            const float *p2 = x + j * d;
            if (!oracle_valid_source(p2)) continue;
            if (oracle_is_neighbour(p1, p2, d, e2)) ++hits;
        }
        nn[i] = hits;
    }
}

/* oracle_classify restricted to the listed rows: nn[r] for point rows[r]
 * against every point (data_process.cpp:395-458), for checking a sample of a
 * run too large for the full O(n^2) count. */
void oracle_classify_rows(const float *x, int64_t n, int d, float e2,
                          const int64_t *rows, int64_t n_rows, int32_t *nn) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        const float *p1 = x + rows[r] * d;
        int32_t hits = 0;
        for (int64_t j = 0; j < n; ++j) {
            const float *p2 = x + j * d;
            if (!oracle_valid_source(p2)) continue;
            if (oracle_is_neighbour(p1, p2, d, e2)) ++hits;
        }
        nn[r] = hits;
    }
}

/* For each listed row: the smallest and largest labels[j] over core points j
 * (core[j] != 0, valid source) within eps of point rows[r]; INT32_MAX and
 * INT32_MIN when there is none.  Given a labelling of the cores, the
 * reference's rules say (data_process.cpp:177-260): a core's own label is
 * both (all cores within eps of a core share its cluster), and a non-core
 * point's label is the smallest (the last sweep's atomic_min over its cores,
 * :204-209), NO_CLUSTER when there is none. */
void oracle_core_label_range_rows(const float *x, int64_t n, int d, float e2,
                                  const uint8_t *core, const int32_t *labels,
                                  const int64_t *rows, int64_t n_rows,
                                  int32_t *lo, int32_t *hi) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        const float *p1 = x + rows[r] * d;
        int32_t a = INT32_MAX, b = INT32_MIN;
        for (int64_t j = 0; j < n; ++j) {
            if (!core[j]) continue;
            const float *p2 = x + j * d;
            if (!oracle_valid_source(p2)) continue;
            if (oracle_is_neighbour(p1, p2, d, e2)) {
                if (labels[j] < a) a = labels[j];
                if (labels[j] > b) b = labels[j];
            }
        }
        lo[r] = a;
        hi[r] = b;
    }
}

/* union-find over core ranks; the root of a set is always its smallest rank,
 * so the root is the canonical cluster id directly. */
static int32_t uf_find(int32_t *parent, int32_t a) {
    for (;;) {
        int32_t p = __atomic_load_n(&parent[a], __ATOMIC_RELAXED);
        if (p == a) return a;
        int32_t gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
        if (gp != p) /* benign halving: gp is an ancestor of a */
            __atomic_compare_exchange_n(&parent[a], &p, gp, 0, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED);
        a = gp;
    }
}

static void uf_unite(int32_t *parent, int32_t a, int32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) { int32_t t = a; a = b; b = t; }
        /* hang the larger root under the smaller one */
        int32_t expect = b;
        if (__atomic_compare_exchange_n(&parent[b], &expect, a, 0,
                                        __ATOMIC_ACQ_REL, __ATOMIC_RELAXED))
            return;
    }
}

/* data_process.cpp:268-332 and :93-266 (exact mode, converged): per-point
 * cluster id as described in the header.  Returns the number of clusters;
 * `labels` receives v_coord_cluster_index. */
int32_t oracle_labels(const float *x, int64_t n, int d, float e2, int32_t m,
                      const int32_t *nn, int32_t *labels) {
    /* data_process.cpp:276-283: ascending list of core ids (ordered copy_if) */
    int64_t n_core = 0;
    for (int64_t i = 0; i < n; ++i) n_core += (nn[i] >= m);
    int64_t *core_id = (int64_t *)malloc(sizeof(int64_t) * (n_core ? n_core : 1));
    int32_t *rank = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (n_core ? n_core : 1));
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (nn[i] >= m) { rank[i] = (int32_t)c; core_id[c++] = i; }
        else rank[i] = -1;
    }
    for (int64_t r = 0; r < n_core; ++r) parent[r] = (int32_t)r;

    /* core-core edges: point a (core) vs core b with valid_source(b), in either
     * role (data_process.cpp:194 skips padding cores only). */
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ra = 0; ra < n_core; ++ra) {
        const float *pa = x + core_id[ra] * d;
        for (int64_t rb = ra + 1; rb < n_core; ++rb) {
            const float *pb = x + core_id[rb] * d;
            if (!oracle_valid_source(pa) && !oracle_valid_source(pb)) continue;
            if (uf_find(parent, (int32_t)ra) == uf_find(parent, (int32_t)rb)) continue;
            if (oracle_is_neighbour(pa, pb, d, e2))
                uf_unite(parent, (int32_t)ra, (int32_t)rb);
        }
    }
    int32_t clusters = 0;
    for (int64_t r = 0; r < n_core; ++r) {
        int32_t root = uf_find(parent, (int32_t)r);
        parent[r] = root;
        clusters += (root == r);
    }

#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        if (rank[i] >= 0) { labels[i] = parent[rank[i]]; continue; }
        int32_t best = ORACLE_NO_CLUSTER;
        const float *p1 = x + i * d;
        for (int64_t r = 0; r < n_core; ++r) {
            if (parent[r] >= best) continue;
            const float *p2 = x + core_id[r] * d;
            if (!oracle_valid_source(p2)) continue;
            if (oracle_is_neighbour(p1, p2, d, e2)) best = parent[r];
        }
        labels[i] = best;
    }
    free(core_id);
    free(rank);
    free(parent);
    return clusters;
}

/* data_process.cpp:50-91 */
void oracle_meta(const int32_t *nn, const int32_t *labels, int64_t n, int32_t m,
                 int32_t clusters, int32_t *out_cores, int32_t *out_noise,
                 int32_t *out_clusters) {
    int32_t cores = 0, noise = 0;
    for (int64_t i = 0; i < n; ++i) {
        cores += (nn[i] >= m);
        noise += (labels[i] == ORACLE_NO_CLUSTER);
    }
    *out_cores = cores;
    *out_noise = noise;
    *out_clusters = clusters;
}

/* One call for the whole path: classify, labels, meta. */
void oracle_dbscan(const float *x, int64_t n, int d, float e, int32_t m,
                   int32_t *nn, int32_t *labels, int32_t *meta3) {
    float e2 = oracle_e2(e);
    oracle_classify(x, n, d, e2, nn);
    int32_t clusters = oracle_labels(x, n, d, e2, m, nn, labels);
    oracle_meta(nn, labels, n, m, clusters, &meta3[0], &meta3[1], &meta3[2]);
}
