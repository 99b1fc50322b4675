// Runs the reference's own sequential scan (SkipSequentialScan,
// tremor/algos/sss/SkipSequential.cpp, lower_bounds off or on) on one rank,
// compiled from the reference sources where they lie, and writes its window
// statistics and its kNN / threshold CSV files.  TEST INFRASTRUCTURE ONLY: the
// fixture generator under tests/golden/ runs it in the build container to pin
// the oracle; it never travels with the product and the product never calls it.
//
//   ref_sss [--lb <segments>] <waveform.bin> <n> <m> <queries.bin> <count> <threads> <out_prefix> <merge> <k> <rho>...
//
// writes <out_prefix>.stats (float mean[W], then float std[W], W = n - m + 1),
// <out_prefix>.knn.csv (k > 0) and <out_prefix>.thr<i>_0.csv for each rho.
// With --lb the scan runs with lower_bounds = true (the skip-sequential
// method) and also writes <out_prefix>.sax (uint8 word[W][segments]),
// <out_prefix>.paa (float paa[count][segments]), <out_prefix>.lb (float
// lower bound[count][W]) and <out_prefix>.bp (the 255 float breakpoints of
// cardinality 256 the bounds use).
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
#include "isax/SAXBreakpoints.hpp"

int main(int argc, char **argv)
{
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s wave n m queries count threads out merge k rho...\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    int segments = 1;
    bool lb = false;
    if (std::string(argv[1]) == "--lb") {
        lb = true;
        segments = std::atoi(argv[2]);
        argv += 2;
        argc -= 2;
    }
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
        // threads, paa_segments (used with lower bounds only), merge offset,
        // replication groups (0: one per rank), lower_bounds
        SkipSequentialScan scan(threads, segments, merge, 0, lb);
        scan.build(wave, static_cast<idx_t>(n), m);
        if (run == 0) {
            FILE *s = std::fopen((out + ".stats").c_str(), "wb");
            std::fwrite(scan.means.data(), sizeof(float), scan.means.size(), s);
            std::fwrite(scan.stds.data(), sizeof(float), scan.stds.size(), s);
            std::fclose(s);
            if (lb) {
                s = std::fopen((out + ".sax").c_str(), "wb");
                std::fwrite(scan.sax.data(), 1, scan.sax.size(), s);
                std::fclose(s);
                const int card = 256, off = ((card - 1) * (card - 2)) / 2;
                s = std::fopen((out + ".bp").c_str(), "wb");
                std::fwrite(&sax_breakpoints[off], sizeof(float), card - 1, s);
                std::fclose(s);
                FILE *fp = std::fopen((out + ".paa").c_str(), "wb");
                FILE *fl = std::fopen((out + ".lb").c_str(), "wb");
                for (int q = 0; q < count; q++) {
                    std::vector<ts_type> znorm, paa;
                    std::vector<int> order;
                    scan.prepare_query(queries.data() + static_cast<size_t>(q) * m, znorm, paa, order);
                    std::fwrite(paa.data(), sizeof(float), paa.size(), fp);
                    std::vector<float> bounds(static_cast<size_t>(scan.subsequences));
                    for (idx_t id = 0; id < scan.subsequences; id++)
                        bounds[id] = scan.lower_bound(paa.data(), id);
                    std::fwrite(bounds.data(), sizeof(float), bounds.size(), fl);
                }
                std::fclose(fp);
                std::fclose(fl);
            }
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
