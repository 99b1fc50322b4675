// Runs the reference's own sequential scan (SkipSequentialScan with
// lower_bounds = false, tremor/algos/sss/SkipSequential.cpp) on one rank,
// compiled from the reference sources where they lie, and writes its window
// statistics and its kNN / threshold CSV files.  TEST INFRASTRUCTURE ONLY: the
// fixture generator under tests/golden/ runs it in the build container to pin
// the oracle; it never travels with the product and the product never calls it.
//
//   ref_sss <waveform.bin> <n> <m> <queries.bin> <count> <threads> <out_prefix> <merge> <k> <rho>...
//
// writes <out_prefix>.stats (float mean[W], then float std[W], W = n - m + 1),
// <out_prefix>.knn.csv (k > 0) and <out_prefix>.thr<i>_0.csv for each rho.
#include <mpi.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

// The statistics are private members of the scan object; the harness reads
// them after build().  Standard headers are included above, so only the
// class definition below sees this.
#define private public
#include "algos/sss/SkipSequential.hpp"
#undef private

int main(int argc, char **argv)
{
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s wave n m queries count threads out merge k rho...\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    const std::string wave = argv[1];
    const long long n = std::atoll(argv[2]);
    const int m = std::atoi(argv[3]);
    const std::string qfile = argv[4];
    const int count = std::atoi(argv[5]);
    const int threads = std::atoi(argv[6]);
    const std::string out = argv[7];
    const int merge = std::atoi(argv[8]);
    const int k = std::atoi(argv[9]);
    std::vector<float> queries(static_cast<size_t>(count) * m);
    FILE *f = std::fopen(qfile.c_str(), "rb");
    if (!f || std::fread(queries.data(), sizeof(float), queries.size(), f) != queries.size()) {
        std::fprintf(stderr, "cannot read %s\n", qfile.c_str());
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    std::fclose(f);
    // One scan object per search: the object's template counter
    // (next_query) is claimed to the end by a search and never reset, as in
    // the reference driver, which runs one search per process.
    for (int run = 0; run < argc - 9; run++) {
        // threads, paa_segments (unused without lower bounds), merge offset,
        // replication groups (0: one per rank), lower_bounds = false
        SkipSequentialScan scan(threads, 1, merge, 0, false);
        scan.build(wave, static_cast<idx_t>(n), m);
        if (run == 0) {
            FILE *s = std::fopen((out + ".stats").c_str(), "wb");
            std::fwrite(scan.means.data(), sizeof(float), scan.means.size(), s);
            std::fwrite(scan.stds.data(), sizeof(float), scan.stds.size(), s);
            std::fclose(s);
            if (k > 0)
                scan.search_knn(queries.data(), count, k, out + ".knn.csv");
        } else {
            const int i = 9 + run;
            scan.search_threshold(queries.data(), count, std::stof(argv[i]), out + ".thr" + std::to_string(run - 1));
        }
    }
    MPI_Finalize();
    return 0;
}
