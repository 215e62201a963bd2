// Input loaders with the reference's formats and entry point
// (magma_input.h:144-167), single rank.
//
//   .bin  int32 n, int32 d, then n*d float32 row-major (magma_input.h:113-125)
//   .csv  one point per line, dimensions separated by spaces; d is the token
//         count of the first line; values parsed with atof then cast to float
//         (magma_input.h:52-68, :127-142)
//   .h5 / .hdf5  not supported by this build: prints the reference's message
//         and exits with -1 (magma_input.h:157-160)
#pragma once

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace magma_input {

inline bool ends_with(const std::string &s, const std::string &suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

inline bool read_input_bin(const std::string &path, std::vector<float> &v, int &n, int &d) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    int32_t hdr[2];
    bool ok = std::fread(hdr, sizeof(int32_t), 2, f) == 2 && hdr[0] >= 0 && hdr[1] >= 1;
    if (ok) {
        n = hdr[0];
        d = hdr[1];
        v.resize((size_t)n * (size_t)d);
        ok = std::fread(v.data(), sizeof(float), v.size(), f) == v.size();
    }
    std::fclose(f);
    return ok;
}

inline bool read_input_csv(const std::string &path, std::vector<float> &v, int &n, int &d) {
    std::ifstream is(path);
    if (!is) return false;
    std::string line, tok, last;
    n = 0;
    d = 0;
    v.clear();
    while (std::getline(is, line)) {
        std::istringstream ss(line);
        if (n == 0) {
            std::vector<std::string> toks;
            while (ss >> tok) toks.push_back(tok);
            d = (int)toks.size();
            for (auto &t : toks) v.push_back(static_cast<float>(std::atof(t.c_str())));
            if (!toks.empty()) last = toks.back();
        } else {
            for (int k = 0; k < d; ++k) {
                if (ss >> tok) last = tok;  // a short line repeats its last token
                v.push_back(static_cast<float>(std::atof(last.c_str())));
            }
        }
        ++n;
    }
    return d > 0 || n == 0;
}

// magma_input.h:144: dispatch on the file extension.
inline void read_input(const std::string &in_file, std::vector<float> &v_input, int &n, int &n_dim) {
    bool ok = false;
    if (ends_with(in_file, ".bin")) {
        ok = read_input_bin(in_file, v_input, n, n_dim);
    } else if (ends_with(in_file, ".h5") || ends_with(in_file, ".hdf5")) {
        std::cerr << "Error: HDF5 is not supported by this executable." << std::endl;
        std::exit(-1);
    } else if (ends_with(in_file, ".csv")) {
        std::cout << "WARNING: USING SLOW CSV I/O." << std::endl;
        ok = read_input_csv(in_file, v_input, n, n_dim);
    } else {
        std::cerr << "Error: unsupported input file type: " << in_file << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (!ok) {
        std::cerr << "Error: could not read " << in_file << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

}  // namespace magma_input
