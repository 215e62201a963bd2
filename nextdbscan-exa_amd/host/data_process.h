// Host class with the reference's data_process interface (data_process.h:45-102),
// backed by the ndbx C ABI instead of the Thrust/OpenMP primitives.
//
// Same constructor arguments and method names; the per-point public vectors
// (v_coord_nn, v_coord_cluster_index) are filled from the device after the
// phase that produces them.  The MPI argument is kept for source
// compatibility; only a single rank is supported (the MPI layer is dropped).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "magma_mpi_stub.h"
#include "ndbx.h"

static const int NO_CLUSTER = NDBX_NO_CLUSTER;  // data_process.h:26

template <typename T>
using h_vec = std::vector<T>;

class data_process {
   private:
    int const m, n_dim;
    std::size_t const n_coord;
    std::size_t const n_total_coord;
    float const e, e2;
    bool const is_approximate;
    ndbx_ctx *ctx = nullptr;

    void check(int rc, const char *what) const;

   public:
    unsigned long long allocated_bytes = 0;
    // data_process.h:54-59.  v_coord holds the coordinates (moved in, as the
    // reference's OpenMP build does; the device copy is the context's).
    // v_coord_id is the point ids 0..n-1 after select_and_process (one rank).
    // v_cluster_label is the label table over core ranks after
    // select_cores_and_process: entry k is the cluster of the k-th core in
    // ascending point id, and a cluster's id is the core rank of its lowest
    // core, so the roots (entry k == k) are the clusters, as
    // get_result_meta counts them (data_process.cpp:79-88).
    h_vec<float> v_coord;
    h_vec<int> v_coord_id;
    h_vec<int> v_coord_nn;             // per point, input order
    h_vec<int> v_coord_cluster_index;  // per point, input order
    h_vec<int> v_cluster_label;        // per core rank

    // data_process.h:61-76: coordinates are copied to HBM here.
    data_process(h_vec<float> &v_coord, int m, float e, int n_dim, int n_total_coord, bool is_approximate);
    ~data_process();
    data_process(const data_process &) = delete;
    data_process &operator=(const data_process &) = delete;

    // data_process.cpp:334: neighbour counts and core flags.
    void select_and_process(magmaMPI mpi) noexcept;
    // data_process.cpp:268: cluster index per point.
    void select_cores_and_process(magmaMPI mpi) noexcept;
    // data_process.cpp:50
    void get_result_meta(int &cores, int &noise, int &clusters, int &n, magmaMPI mpi) noexcept;
    // Which kernels ran ("grid", "all-pairs", "mfma-filter").
    std::string path_name() const;
};
