// nextdbscan::start: the reference's phase order and timing prints
// (nextdbscan.cpp:28-68, magma_util.h:51-64), driving data_process over HIP.
#include "nextdbscan.h"

#include <chrono>
#include <iostream>
#include <utility>

#include "data_process.h"
#include "ndbx_input.h"

namespace {

using clock_type = std::chrono::high_resolution_clock;

// "<name><ms> milliseconds" as magma_util::measure_duration prints it.
template <class F>
long long timed(const char *name, bool verbose, F &&f) {
    if (verbose) std::cout << name << std::flush;
    auto t0 = clock_type::now();
    f();
    auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(clock_type::now() - t0).count();
    if (verbose) std::cout << ms << " milliseconds" << std::endl;
    return ms;
}

}  // namespace

nextdbscan::result nextdbscan::start(int const m, float const e, int const n_thread, std::string const &in_file,
                                     magmaMPI mpi, bool const is_approximate) noexcept {
    const bool root = mpi.rank == 0;
    if (root)
        std::cout << "Total of " << (n_thread * mpi.n_nodes) << " cores used on " << mpi.n_nodes << " node(s)."
                  << std::endl;
    auto t_start = clock_type::now();
    h_vec<float> v_coord;
    int n_total_coord = -1, n_dim = -1;
    timed("Read Input Data: ", root, [&] { magma_input::read_input(in_file, v_coord, n_total_coord, n_dim); });
    if (root)
        std::cout << "Read " << n_total_coord << " aggregated points with " << n_dim << " dimensions. " << std::endl;
    auto t_start_noio = clock_type::now();
    result res{};
    {
        data_process dp(v_coord, m, e, n_dim, n_total_coord, is_approximate);
        timed("Classify Points: ", root, [&] { dp.select_and_process(mpi); });
        timed("Assign Labels: ", root, [&] { dp.select_cores_and_process(mpi); });
        timed("Collect Results: ", root,
              [&] { dp.get_result_meta(res.core_count, res.noise, res.clusters, res.n, mpi); });
        res.point_clusters = std::move(dp.v_coord_cluster_index);
    }
    auto t_end = clock_type::now();
    auto total = std::chrono::duration_cast<std::chrono::milliseconds>(t_end - t_start).count();
    auto total_noio = std::chrono::duration_cast<std::chrono::milliseconds>(t_end - t_start_noio).count();
    if (root) {
        std::cout << "Total Execution Time: " << total << " milliseconds" << std::endl;
        std::cout << "Total Execution Time (without I/O): " << total_noio << " milliseconds" << std::endl;
    }
    return res;
}
