// nextdbscan::start and result with the reference's signature
// (nextdbscan.h:41-52), single rank.  `point_clusters` carries the per-point
// cluster index the reference's commented writer expects (main.cpp:108-123).
#pragma once

#include <string>
#include <vector>

#include "magma_mpi_stub.h"

namespace nextdbscan {

struct result {
    int clusters;
    int noise;
    int core_count;
    int n;
    std::vector<int> point_clusters;  // input order; NO_CLUSTER for noise
};

result start(int m, float e, int n_thread, const std::string &in_file, magmaMPI mpi, bool is_approximate) noexcept;

}  // namespace nextdbscan
