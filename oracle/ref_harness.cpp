// Harness around the reference's own OpenMP data_process (test infrastructure).
//
// Built by oracle/Makefile from the reference sources where they lie
// (data_process.cpp, data_process.h, magma_*.h) with the reference's OpenMP
// flags; output only under oracle/_ref/.  It drives the public members of
// `class data_process` (data_process.h:45-102) in the order nextdbscan::start
// does (nextdbscan.cpp:46-59) and dumps the per-point state the reference CLI
// never writes: v_coord_nn, v_coord_cluster_index and v_cluster_label.
//
//   ref_harness full <in.bin> <m> <e-text> <threads> <approx 0|1> <out.bin>
//       out.bin: int32 n, cores, noise, clusters, n_labels,
//                int32 nn[n], int32 cci[n], int32 label[n_labels]
//   ref_harness classify_sample <in.bin> <m> <e-text> <threads> <rows> <seed>
//       times process_points (data_process.cpp:395-458) over a chunk holding
//       `rows` randomly chosen points plus the reference's padding, and prints
//       one JSON line: ms, pair tests and tests/s.
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "data_process.h"

static bool read_bin(const char *path, std::vector<float> &v, int &n, int &d) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(&d, 4, 1, f) != 1) {
        std::fclose(f);
        return false;
    }
    v.resize((size_t)n * (size_t)d);
    size_t got = std::fread(v.data(), 4, v.size(), f);
    std::fclose(f);
    return got == v.size();
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: see header\n");
        return 2;
    }
    std::string mode = argv[1];
    if (mode == "full" && argc == 8) {
        std::vector<float> v;
        int n, d;
        if (!read_bin(argv[2], v, n, d)) return 3;
        int m = std::atoi(argv[3]);
        float e = std::strtof(argv[4], nullptr);  // main.cpp:71
        omp_set_num_threads(std::atoi(argv[5]));
        bool approx = std::atoi(argv[6]) == 1;
        auto mpi = magmaMPI::build();
        data_process dp(v, m, e, d, n, approx);
        dp.select_and_process(mpi);
        dp.select_cores_and_process(mpi);
        int cores, noise, clusters, nn;
        dp.get_result_meta(cores, noise, clusters, nn, mpi);
        FILE *f = std::fopen(argv[7], "wb");
        if (!f) return 4;
        int hdr[5] = {n, cores, noise, clusters, (int)dp.v_cluster_label.size()};
        std::fwrite(hdr, 4, 5, f);
        std::fwrite(dp.v_coord_nn.data(), 4, dp.v_coord_nn.size(), f);
        std::fwrite(dp.v_coord_cluster_index.data(), 4, dp.v_coord_cluster_index.size(), f);
        std::fwrite(dp.v_cluster_label.data(), 4, dp.v_cluster_label.size(), f);
        std::fclose(f);
        return 0;
    }
    if (mode == "classify_sample" && argc == 8) {
        std::vector<float> v;
        int n, d;
        if (!read_bin(argv[2], v, n, d)) return 3;
        int m = std::atoi(argv[3]);
        float e = std::strtof(argv[4], nullptr);
        int threads = std::atoi(argv[5]);
        omp_set_num_threads(threads);
        long rows = std::atol(argv[6]);
        unsigned seed = (unsigned)std::atol(argv[7]);
        if (rows > n) rows = n;
        std::mt19937_64 rng(seed);
        std::vector<int> ids(rows);
        std::vector<float> data((size_t)(rows + 2) * d, std::numeric_limits<float>::max());
        std::vector<int> chunk_ids(rows + 2, -1);
        for (long r = 0; r < rows; ++r) {
            int i = (int)(rng() % (unsigned long long)n);
            chunk_ids[r] = i;
            for (int k = 0; k < d; ++k) data[(size_t)r * d + k] = v[(size_t)i * d + k];
        }
        data_process dp(v, m, e, d, n, false);
        dp.v_coord_nn.assign((size_t)n, 0);
        auto t0 = std::chrono::high_resolution_clock::now();
        dp.process_points(chunk_ids, data);
        auto t1 = std::chrono::high_resolution_clock::now();
        double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        double tests = (double)rows * (double)n;
        std::printf("{\"ms\": %.3f, \"rows\": %ld, \"n\": %d, \"d\": %d, \"threads\": %d, "
                    "\"pair_tests\": %.0f, \"tests_per_s\": %.6e}\n",
                    ms, rows, n, d, threads, tests, tests / (ms * 1e-3));
        return 0;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
