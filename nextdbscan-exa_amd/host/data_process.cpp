#include "data_process.h"

#include <cstdio>
#include <cstdlib>

// The reference's methods are noexcept and have no error path; a failing
// device call here prints the library's message and exits, as main.cpp:83
// does for bad arguments.
void data_process::check(int rc, const char *what) const {
    if (rc == NDBX_OK) return;
    std::fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, ndbx_last_error());
    std::exit(EXIT_FAILURE);
}

data_process::data_process(h_vec<float> &v_coord, int const m, float const e, int const n_dim,
                           int const n_total_coord, bool const is_approximate)
    : m(m),
      n_dim(n_dim),
      n_coord(n_dim > 0 ? v_coord.size() / (std::size_t)n_dim : 0),
      n_total_coord((std::size_t)n_total_coord),
      e(e),
      e2(e * e),
      is_approximate(is_approximate) {
    uint32_t flags = is_approximate ? NDBX_APPROX : 0u;
    if (std::getenv("NDBX_FORCE_BRUTE")) flags |= NDBX_FORCE_BRUTE;
    check(ndbx_create(&ctx, v_coord.data(), (int64_t)n_coord, n_dim, e, m, flags), "ndbx_create");
    // slab-by-slab runs with at most this many points per slab
    if (const char *cp = std::getenv("NDBX_CHUNK_POINTS"))
        check(ndbx_set_chunking(ctx, std::atoll(cp)), "ndbx_set_chunking");
    this->v_coord = std::move(v_coord);
    allocated_bytes = (unsigned long long)n_coord * (unsigned long long)n_dim * sizeof(float);
}

data_process::~data_process() { ndbx_destroy(ctx); }

void data_process::select_and_process(magmaMPI) noexcept {
    check(ndbx_classify(ctx), "ndbx_classify");
    v_coord_id.resize(n_coord);
    for (std::size_t i = 0; i < n_coord; ++i) v_coord_id[i] = (int)i;
    v_coord_nn.resize(n_coord);
    check(ndbx_get_nn(ctx, v_coord_nn.data()), "ndbx_get_nn");
}

void data_process::select_cores_and_process(magmaMPI) noexcept {
    check(ndbx_label(ctx), "ndbx_label");
    v_coord_cluster_index.resize(n_coord);
    check(ndbx_get_labels(ctx, v_coord_cluster_index.data()), "ndbx_get_labels");
    // label table over core ranks: the cluster index of each core, in
    // ascending point id (every entry is already its root)
    v_cluster_label.clear();
    for (std::size_t i = 0; i < n_coord; ++i)
        if (v_coord_nn[i] >= m) v_cluster_label.push_back(v_coord_cluster_index[i]);
}

void data_process::get_result_meta(int &cores, int &noise, int &clusters, int &n, magmaMPI) noexcept {
    check(ndbx_result_meta(ctx, &cores, &noise, &clusters, &n), "ndbx_result_meta");
}

std::string data_process::path_name() const {
    int p = -1;
    ndbx_path(ctx, &p);
    return p == 0 ? "grid" : p == 1 ? "all-pairs" : p == 2 ? "mfma-filter" : "unknown";
}
