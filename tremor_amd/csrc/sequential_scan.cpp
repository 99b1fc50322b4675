// tremor_amd::SequentialScan - the reference sequential scan's class shape
// (SkipSequential.hpp:32-78) on top of the MI355X C-ABI.
#include "tremor_amd/SequentialScan.hpp"

#include <cfloat>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace tremor_amd {

namespace {

// ed_to_cc (ZnormSearch.hpp:28-31), in FP32.
float corr_of(float ed, int m) { return 1.0f - (ed / (2.0f * static_cast<float>(m))); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

SequentialScan::SequentialScan(int threads, int paa_segments, int merge_offset, int replication_groups,
                               bool lower_bounds, int precision, int device)
    : merge_offset_(merge_offset)
{
    (void)threads;
    if (replication_groups < 0 || replication_groups > 1)
        throw std::invalid_argument("--replication-groups must be 0 or 1 on one GPU");
    if (merge_offset < 0)
        throw std::invalid_argument("--merge-offset must be non-negative");
    tmx_cfg cfg;
    tmx_cfg_default(&cfg);
    cfg.device = device;
    cfg.precision = precision;
    // lower_bounds: the iSAX bound on the early-abandoning scan (answers
    // unchanged; it only skips windows that cannot qualify).
    if (lower_bounds) {
        if (paa_segments < 1 || paa_segments > 64)
            throw std::invalid_argument("--paa-segments must be in 1..64 with lower bounds");
        cfg.lower_bounds = 1;
        cfg.paa_segments = paa_segments;
    }
    if (tmx_create(&ctx_, &cfg) != 0)
        throw std::runtime_error("no usable HIP device " + std::to_string(device));
}

SequentialScan::~SequentialScan() { tmx_destroy(ctx_); }

void SequentialScan::check(int rc) const
{
    if (rc != 0)
        throw std::runtime_error(tmx_last_error(ctx_));
}

void SequentialScan::set_forms(int fft_mode, int scan_mode)
{
    check(tmx_set_fft_mode(ctx_, fft_mode));
    check(tmx_set_scan_mode(ctx_, scan_mode));
}

void SequentialScan::set_index(int index_mode, int leaf_size) { check(tmx_set_index_mode(ctx_, index_mode, leaf_size)); }

void SequentialScan::build(const std::string &filename, idx_t total_samples, int window_length)
{
    std::vector<float> raw(static_cast<size_t>(total_samples));
    FILE *f = std::fopen(filename.c_str(), "rb");
    if (f == nullptr || std::fread(raw.data(), sizeof(float), raw.size(), f) != raw.size()) {
        if (f)
            std::fclose(f);
        throw std::runtime_error("cannot read " + filename);
    }
    std::fclose(f);
    build(raw.data(), total_samples, window_length);
}

void SequentialScan::build(const float *samples, idx_t total_samples, int window_length)
{
    if (window_length <= 0)
        throw std::invalid_argument("window length must be positive");
    if (total_samples < static_cast<idx_t>(window_length))
        throw std::invalid_argument("dataset has fewer samples than the window length");
    window_ = window_length;
    samples_ = total_samples;
    check(tmx_load_waveform(ctx_, samples, total_samples));
    summary_bytes_ = 16.0 * static_cast<double>(total_samples + 1);
}

void SequentialScan::knn(const float *queries, int query_count, int k, std::vector<idx_t> &positions,
                         std::vector<float> &distances, tmx_stats *stats)
{
    if (window_ == 0)
        throw std::runtime_error("build() must run before a search");
    std::vector<int32_t> lengths(static_cast<size_t>(query_count), window_);
    std::vector<uint64_t> offsets(static_cast<size_t>(query_count));
    for (int q = 0; q < query_count; q++)
        offsets[q] = static_cast<uint64_t>(q) * static_cast<uint64_t>(window_);
    positions.assign(static_cast<size_t>(query_count) * k, 0);
    distances.assign(static_cast<size_t>(query_count) * k, FLT_MAX);
    std::vector<uint64_t> pos(positions.size());
    check(tmx_knn(ctx_, queries, lengths.data(), offsets.data(), query_count, k, merge_offset_, pos.data(),
                  distances.data(), stats));
    for (size_t i = 0; i < pos.size(); i++)
        positions[i] = pos[i];
}

void SequentialScan::threshold(const float *queries, int query_count, float correlation, std::vector<Hit> &hits,
                               tmx_stats *stats)
{
    if (window_ == 0)
        throw std::runtime_error("build() must run before a search");
    std::vector<int32_t> lengths(static_cast<size_t>(query_count), window_);
    std::vector<uint64_t> offsets(static_cast<size_t>(query_count));
    for (int q = 0; q < query_count; q++)
        offsets[q] = static_cast<uint64_t>(q) * static_cast<uint64_t>(window_);
    tmx_hits *h = nullptr;
    check(tmx_threshold(ctx_, queries, lengths.data(), offsets.data(), query_count, correlation, merge_offset_, &h,
                        stats));
    hits.resize(static_cast<size_t>(h->count));
    for (int64_t r = 0; r < h->count; r++)
        hits[r] = Hit{h->tmpl[r], static_cast<long long>(h->counter[r]), h->pos[r], h->d[r]};
    tmx_hits_free(h);
}

void SequentialScan::search_knn(const float *queries, int query_count, int k, const std::string &output)
{
    const double start = now_s();
    std::vector<idx_t> pos;
    std::vector<float> dist;
    knn(queries, query_count, k, pos, dist);
    FILE *out = std::fopen(output.c_str(), "w");
    if (out == nullptr)
        throw std::runtime_error("cannot write kNN output " + output);
    std::fprintf(out, "tID, k, pos, corr\n");
    for (int q = 0; q < query_count; q++)
        for (int j = 0; j < k; j++) {
            const size_t o = static_cast<size_t>(q) * k + j;
            std::fprintf(out, "%d, %d, %llu, %f\n", q, j + 1, pos[o], corr_of(dist[o], window_));
        }
    std::fclose(out);
    std::printf("[Node 0] %d queries: %.6f s (gpu)\n", query_count, now_s() - start);
}

void SequentialScan::search_threshold(const float *queries, int query_count, float correlation,
                                      const std::string &output)
{
    const double start = now_s();
    FILE *out = std::fopen((output + "_0.csv").c_str(), "w");
    if (out == nullptr)
        throw std::runtime_error("cannot write threshold output " + output);
    std::vector<Hit> hits;
    threshold(queries, query_count, correlation, hits);
    std::fprintf(out, "Query, Counter, Position, Cross Correlation\n");
    for (const Hit &h : hits)
        std::fprintf(out, "%d, %lld, %llu, %f\n", h.query, h.counter, h.position, corr_of(h.distance, window_));
    std::fclose(out);
    std::printf("[Node 0] %d queries: %.6f s (gpu)\n", query_count, now_s() - start);
}

}  // namespace tremor_amd
