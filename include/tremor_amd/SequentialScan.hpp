// C++ host API with the shape of the reference's sequential scan class
// (tremor/algos/sss/SkipSequential.hpp:32-78), backed by the MI355X C-ABI.
//
// Same constructor parameters, same build/search calls, same output files
// and formats, same exception types:
//   search_knn        writes <output>:          "tID, k, pos, corr"   (SkipSequential.cpp:378-392)
//   search_threshold  writes <output>_<rank>.csv: "Query, Counter, Position, Cross Correlation"
//                                                                    (SkipSequential.cpp:230-290)
// One process drives one GPU: the waveform forms a single partition (rank 0).
// threads is accepted for call-site compatibility.  lower_bounds (with
// paa_segments) puts the reference's iSAX lower bound on the
// early-abandoning scan (SkipSequential.cpp:200-211): templates on that scan
// skip the windows it rules out; the answers equal the sequential method's.
#ifndef TREMOR_AMD_SEQUENTIAL_SCAN_HPP
#define TREMOR_AMD_SEQUENTIAL_SCAN_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "tremor_amd/tmx.h"

namespace tremor_amd {

using idx_t = unsigned long long;

class SequentialScan {
public:
    // replication_groups: 0 or 1 (one GPU holds the whole waveform).
    SequentialScan(int threads, int paa_segments, int merge_offset, int replication_groups,
                   bool lower_bounds = true, int precision = TMX_PREC_BF16X3, int device = 0);
    ~SequentialScan();
    SequentialScan(const SequentialScan &) = delete;
    SequentialScan &operator=(const SequentialScan &) = delete;

    // Reads total_samples raw little-endian float32 samples from filename and
    // makes them resident on the device with their statistics.
    void build(const std::string &filename, idx_t total_samples, int window_length);
    void build(const float *samples, idx_t total_samples, int window_length);

    // Writes <output>_0.csv.
    void search_threshold(const float *queries, int query_count, float correlation, const std::string &output);
    // Writes <output>.
    void search_knn(const float *queries, int query_count, int k, const std::string &output);

    // In-memory forms: k slots per template (unfilled: 0, FLT_MAX) / hit rows.
    void knn(const float *queries, int query_count, int k, std::vector<idx_t> &positions,
             std::vector<float> &distances, tmx_stats *stats = nullptr);
    struct Hit {
        int query;
        long long counter;
        idx_t position;
        float distance;
    };
    void threshold(const float *queries, int query_count, float correlation, std::vector<Hit> &hits,
                   tmx_stats *stats = nullptr);

    // Candidate forms (answers unchanged): enum tmx_fft_mode / tmx_scan_mode.
    void set_forms(int fft_mode, int scan_mode);
    // The iSAX index (enum tmx_index_mode) with leaves of at most leaf_size
    // windows: Tremor's method (tremor_main.cpp --method tremor, --leaf-size).
    void set_index(int index_mode, int leaf_size);

    // Device bytes held beyond the raw samples (prefix statistics).
    double summary_memory_bytes() const { return summary_bytes_; }
    int window() const { return window_; }

private:
    void check(int rc) const;
    tmx_ctx *ctx_ = nullptr;
    int merge_offset_ = 0;
    int window_ = 0;
    idx_t samples_ = 0;
    double summary_bytes_ = 0.0;
};

}  // namespace tremor_amd

#endif
