// Drives the host data_process (nextdbscan-exa_amd/host) the way
// nextdbscan::start drives the reference's (nextdbscan.cpp:46-59) and dumps
// its public members, so a test can compare them with the reference's own
// outputs for the same input.
//   members_check <in.bin> <m> <e-text> <out.bin>
//   out.bin: int32 n, d, v_coord.size(), cores, noise, clusters, n_labels,
//            roots (entries k with v_cluster_label[k] == k),
//            v_coord_id[n], v_coord_nn[n], v_coord_cluster_index[n],
//            v_cluster_label[n_labels], then v_coord (n*d float32)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "data_process.h"

int main(int argc, char **argv) {
    if (argc != 5) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int n = 0, d = 0;
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(&d, 4, 1, f) != 1) return 3;
    h_vec<float> v((size_t)n * d);
    if (std::fread(v.data(), 4, v.size(), f) != v.size()) return 3;
    std::fclose(f);
    const int m = std::atoi(argv[2]);
    const float e = std::strtof(argv[3], nullptr);
    auto mpi = magmaMPI::build();
    data_process dp(v, m, e, d, n, false);
    dp.select_and_process(mpi);
    dp.select_cores_and_process(mpi);
    int cores = 0, noise = 0, clusters = 0, nn = 0;
    dp.get_result_meta(cores, noise, clusters, nn, mpi);
    int roots = 0;
    for (size_t k = 0; k < dp.v_cluster_label.size(); ++k) roots += dp.v_cluster_label[k] == (int)k;
    FILE *o = std::fopen(argv[4], "wb");
    if (!o) return 4;
    int hdr[8] = {n, d, (int)dp.v_coord.size(), cores, noise, clusters, (int)dp.v_cluster_label.size(), roots};
    std::fwrite(hdr, 4, 8, o);
    std::fwrite(dp.v_coord_id.data(), 4, dp.v_coord_id.size(), o);
    std::fwrite(dp.v_coord_nn.data(), 4, dp.v_coord_nn.size(), o);
    std::fwrite(dp.v_coord_cluster_index.data(), 4, dp.v_coord_cluster_index.size(), o);
    std::fwrite(dp.v_cluster_label.data(), 4, dp.v_cluster_label.size(), o);
    std::fwrite(dp.v_coord.data(), 4, dp.v_coord.size(), o);
    std::fclose(o);
    return 0;
}
