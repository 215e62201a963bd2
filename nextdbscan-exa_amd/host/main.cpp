// Command line with the reference's flags and output (main.cpp:29-129):
//   -m minPoints -e epsilon [-t threads] [-a 0|1] [-o output] <input file>
// -e is parsed with strtof (main.cpp:71); arguments come in pairs and the
// input file is the last argument.  -o, documented but never parsed by the
// reference, writes one cluster index per line in input order (its commented
// writer, main.cpp:108-123); noise is NO_CLUSTER (2147483647).
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "nextdbscan.h"

static void usage() {
    std::cout << "NextDBSCAN compiled for HIP (MI355X, gfx950)" << std::endl << std::endl;
    std::cout << "Usage: [executable] -m minPoints -e epsilon -t threads [input file]" << std::endl;
    std::cout << "    -m minPoints : DBSCAN parameter, minimum number of points required to form a cluster, "
                 "postive integer, required"
              << std::endl;
    std::cout << "    -e epsilon   : DBSCAN parameter, maximum neighborhood search radius for cluster, positive "
                 "floating point, required"
              << std::endl;
    std::cout << "    -t threads   : Processing parameter, the number of threads to use, positive integer, "
                 "defaults to 1"
              << std::endl;
    std::cout << "    -a 0|1       : For faster results this option will produce an upper-bound approximation of "
                 "the final number of clusters. Noise and cores, however, will be identical to the "
                 "non-approximate case"
              << std::endl;
    std::cout << "    -o output    : Output file containing the cluster ids in the same order as the input"
              << std::endl;
    std::cout << "    -h help      : Show this help message" << std::endl << std::endl;
    std::cout << "Supported Input Types:" << std::endl;
    std::cout << ".csv: Text file with one sample/point per line and features/dimensions separated by a space "
                 "delimiter, i.e. ' '"
              << std::endl;
    std::cout << ".bin: Custom binary format for faster file reads. Use cvs2bin executable to transform csv into "
                 "bin files."
              << std::endl;
}

int main(int argc, char **argv) {
    int m = -1, t = 1, a = 0;
    float e = -1;
    std::string output_file;
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string flag(argv[i]);
        if (flag == "-m") m = std::stoi(argv[i + 1]);
        else if (flag == "-e") e = std::strtof(argv[i + 1], nullptr);
        else if (flag == "-t") t = std::stoi(argv[i + 1]);
        else if (flag == "-a") a = std::stoi(argv[i + 1]);
        else if (flag == "-o") output_file = argv[i + 1];
    }
    if (argc < 2 || m == -1 || e == -1) {
        std::cout << "Input Error: Please specify the m and e parameters" << std::endl << std::endl;
        usage();
        std::exit(EXIT_FAILURE);
    }
    const std::string input_file = argv[argc - 1];
    auto mpi = magmaMPI::build();
    std::cout << "Starting NextDBSCAN with file: " << input_file << " m: " << m << " e: " << e << " t: " << t
              << " a: " << a << std::endl;
    auto results = nextdbscan::start(m, e, t, input_file, mpi, a == 1);
    std::cout << std::endl;
    if (a == 0) std::cout << "Estimated Clusters: " << results.clusters << std::endl;
    else if (a == 1) std::cout << "Upper-bound Approximate Clusters: " << results.clusters << std::endl;
    std::cout << "Core Points: " << results.core_count << std::endl;
    std::cout << "Noise Points: " << results.noise << std::endl;
    if (!output_file.empty()) {
        std::cout << "Writing output to " << output_file << std::endl;
        std::ofstream os(output_file);
        for (int v : results.point_clusters) os << v << '\n';
        os.close();
        if (!os) {
            std::cerr << "Error: could not write " << output_file << std::endl;
            return EXIT_FAILURE;
        }
        std::cout << "Done!" << std::endl;
    }
    return 0;
}
