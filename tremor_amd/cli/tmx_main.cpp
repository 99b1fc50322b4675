// Command-line driver with the reference CLI's flags, inputs and outputs for
// the sequential method (tremor_main.cpp:246-470 flags and checks, :494-606
// inputs, :698-789 run and timings CSV), running on one MI355X.
//
//   tmx_main --dataset x.bin --queries q.bin --window-length 64 --mode knn --k 5 \
//            [--merge-offset 16] [--output out.csv] [--precision bf16x3|fp16|fp32]
//
// Inputs are raw little-endian float32 files (missing samples -100000).
// Outputs: kNN "tID, k, pos, corr" to --output; threshold rows to
// <output>_0.csv; timings "node_rank,index_time,query_time,index_gb" to
// --benchmark-output (default <output>.timings.csv).
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "tremor_amd/SequentialScan.hpp"

namespace {

using tremor_amd::idx_t;

struct Options {
    std::string dataset, queries, output, bench_output;
    bool dataset_size_set = false, query_count_set = false, no_bench = false;
    idx_t dataset_size = 0, query_count = 0, window = 0;
    bool knn = true;
    idx_t k = 10;
    float rho = 0.93f;
    int merge = 0, replication_groups = 0, paa = 16, workers = 4, index_threads = 4, leaf_size = 2000;
    std::string method = "sequential";
    bool adaptive = false;  // --refinement adaptive
    int precision = TMX_PREC_BF16X3;
    int fft_mode = TMX_FFT_AUTO;
    int scan_mode = TMX_SCAN_AUTO;
    bool scan_mode_given = false;
    int device = 0;
};

void usage(const char *prog)
{
    std::printf(
        "Usage: %s --dataset FILE --queries FILE --window-length M [options]\n"
        "  --dataset-size N           samples to use (default: whole file)\n"
        "  --query-count N            templates to use (default: whole file / M)\n"
        "  --mode knn|threshold       search mode (default knn)\n"
        "  --k K                      neighbours per template (default 10)\n"
        "  --threshold RHO            correlation threshold (default 0.93)\n"
        "  --merge-offset N           exclusion zone between answers (default 0)\n"
        "  --output PATH              kNN CSV, or threshold prefix (<PATH>_0.csv)\n"
        "  --benchmark-output PATH    timings CSV (default <output>.timings.csv)\n"
        "  --no-benchmark-report      skip the timings CSV\n"
        "  --method sequential|skip-sequential|tremor  (skip-sequential: every template on the\n"
        "                             early-abandoning scan with the iSAX lower bound; tremor: the iSAX\n"
        "                             index for window lengths divisible by 16 with 16 PAA segments;\n"
        "                             same answers)\n"
        "  --refinement leaf|adaptive tremor: every such template through the index (leaf, default),\n"
        "                             or where it leaves at most 1%% of the windows (adaptive)\n"
        "  --leaf-size N              iSAX leaf size (default 2000)\n"
        "  --precision bf16x3|fp16|fp32  candidate pass on the GPU (answers identical)\n"
        "  --fft-mode auto|off|always     FFT form of the candidate pass (default auto: cost model)\n"
        "  --scan-mode auto|off|always    early-abandoning scan form (default auto: cost model)\n"
        "  --device N                 HIP device\n"
        "  --search-workers/--index-threads/--paa-segments/--replication-groups accepted for compatibility\n",
        prog);
}

idx_t parse_u(const std::string &v, const std::string &flag)
{
    try {
        size_t used = 0;
        const unsigned long long x = std::stoull(v, &used);
        if (used != v.size() || (!v.empty() && v[0] == '-'))
            throw std::invalid_argument("");
        return x;
    } catch (...) {
        throw std::invalid_argument(flag + " must be a non-negative integer");
    }
}

int parse_i(const std::string &v, const std::string &flag)
{
    try {
        size_t used = 0;
        const int x = std::stoi(v, &used);
        if (used != v.size())
            throw std::invalid_argument("");
        return x;
    } catch (...) {
        throw std::invalid_argument(flag + " must be an integer");
    }
}

Options parse(int argc, char **argv)
{
    Options o;
    auto value = [&](int &i, const std::string &flag) -> std::string {
        if (i + 1 >= argc)
            throw std::invalid_argument(flag + " needs a value");
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") {
            usage(argv[0]);
            std::exit(0);
        } else if (a == "--dataset") {
            o.dataset = value(i, a);
        } else if (a == "--queries") {
            o.queries = value(i, a);
        } else if (a == "--output") {
            o.output = value(i, a);
        } else if (a == "--benchmark-output") {
            o.bench_output = value(i, a);
        } else if (a == "--no-benchmark-report") {
            o.no_bench = true;
        } else if (a == "--dataset-size") {
            o.dataset_size = parse_u(value(i, a), a);
            o.dataset_size_set = true;
        } else if (a == "--query-count") {
            o.query_count = parse_u(value(i, a), a);
            o.query_count_set = true;
        } else if (a == "--window-length") {
            o.window = parse_u(value(i, a), a);
        } else if (a == "--mode") {
            const std::string m = value(i, a);
            if (m == "knn")
                o.knn = true;
            else if (m == "threshold")
                o.knn = false;
            else
                throw std::invalid_argument("--mode must be knn or threshold");
        } else if (a == "--method") {
            o.method = value(i, a);
            if (o.method != "sequential" && o.method != "skip-sequential" && o.method != "tremor")
                throw std::invalid_argument("Unknown --method: " + o.method +
                                            " (expected tremor, skip-sequential, or sequential)");
        } else if (a == "--refinement") {
            const std::string r = value(i, a);
            if (r != "leaf" && r != "adaptive")
                throw std::invalid_argument("Unknown --refinement: " + r + " (expected leaf or adaptive)");
            o.adaptive = r == "adaptive";
        } else if (a == "--leaf-size") {
            o.leaf_size = parse_i(value(i, a), a);
        } else if (a == "--k") {
            o.k = parse_u(value(i, a), a);
        } else if (a == "--threshold") {
            const std::string v = value(i, a);
            try {
                o.rho = std::stof(v);
            } catch (...) {
                throw std::invalid_argument("--threshold must be a number");
            }
        } else if (a == "--merge-offset") {
            o.merge = parse_i(value(i, a), a);
        } else if (a == "--replication-groups") {
            o.replication_groups = parse_i(value(i, a), a);
        } else if (a == "--paa-segments") {
            o.paa = parse_i(value(i, a), a);
        } else if (a == "--search-workers") {
            o.workers = parse_i(value(i, a), a);
        } else if (a == "--index-threads") {
            o.index_threads = parse_i(value(i, a), a);
        } else if (a == "--fft-mode" || a == "--scan-mode") {
            const std::string v = value(i, a);
            const int mode = v == "auto" ? 0 : v == "off" ? 1 : v == "always" ? 2 : -1;
            if (mode < 0)
                throw std::invalid_argument(a + " must be auto, off or always");
            (a == "--fft-mode" ? o.fft_mode : o.scan_mode) = mode;
            if (a == "--scan-mode")
                o.scan_mode_given = true;
        } else if (a == "--precision") {
            const std::string p = value(i, a);
            if (p == "bf16x3")
                o.precision = TMX_PREC_BF16X3;
            else if (p == "fp16")
                o.precision = TMX_PREC_FP16;
            else if (p == "fp32")
                o.precision = TMX_PREC_FP32;
            else
                throw std::invalid_argument("--precision must be bf16x3, fp16 or fp32");
        } else if (a == "--device") {
            o.device = parse_i(value(i, a), a);
        } else {
            throw std::invalid_argument("unknown argument " + a);
        }
    }
    // Checks of the reference driver (tremor_main.cpp:379-455).
    if (o.dataset.empty())
        throw std::invalid_argument("--dataset is required");
    if (o.queries.empty())
        throw std::invalid_argument("--queries is required");
    if (o.window == 0)
        throw std::invalid_argument("--window-length must be greater than 0");
    if (o.window % 8 != 0)
        throw std::invalid_argument("--window-length must be divisible by 8");
    if (o.paa <= 0)
        throw std::invalid_argument("--paa-segments must be greater than 0");
    if (o.leaf_size <= 0)
        throw std::invalid_argument("--leaf-size must be greater than 0");
    if (o.window % static_cast<idx_t>(o.paa) != 0)
        throw std::invalid_argument("--window-length must be divisible by --paa-segments");
    if (o.dataset_size_set && o.dataset_size < o.window)
        throw std::invalid_argument("--dataset-size must be at least --window-length");
    if (o.query_count_set && o.query_count == 0)
        throw std::invalid_argument("--query-count must be greater than 0");
    if (o.k == 0)
        throw std::invalid_argument("--k must be greater than 0");
    if (o.rho < -1.0f || o.rho > 1.0f)
        throw std::invalid_argument("--threshold must be between -1.0 and 1.0");
    if (o.workers <= 0 || o.index_threads <= 0)
        throw std::invalid_argument("thread counts must be greater than 0");
    if (o.merge < 0)
        throw std::invalid_argument("--merge-offset must be non-negative");
    if (o.replication_groups < 0)
        throw std::invalid_argument("--replication-groups must be non-negative");
    if (o.window > 65536)
        throw std::invalid_argument("--window-length is too large for this build (at most 65536)");
    if (o.k > 4096)
        throw std::invalid_argument("--k is too large for this build (at most 4096)");
    if (o.output.empty())
        o.output = o.knn ? "tremor_knn.csv" : "tremor_threshold";
    if (o.bench_output.empty() && !o.no_bench)
        o.bench_output = o.output + ".timings.csv";
    return o;
}

idx_t float_count(const std::string &path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f)
        throw std::runtime_error("Cannot open " + path + ": " + std::strerror(errno));
    const std::streamoff bytes = f.tellg();
    if (bytes < 0)
        throw std::runtime_error("Cannot determine file size for " + path);
    if (bytes % static_cast<std::streamoff>(sizeof(float)) != 0)
        throw std::runtime_error(path + " size is not a multiple of sizeof(float)");
    return static_cast<idx_t>(bytes) / sizeof(float);
}

std::vector<float> read_floats(const std::string &path, idx_t count)
{
    std::vector<float> v(static_cast<size_t>(count));
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("Cannot open " + path + ": " + std::strerror(errno));
    f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(float)));
    if (!f)
        throw std::runtime_error("Failed to read " + path);
    return v;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int run(const Options &o)
{
    const double t_in = now_s();
    const idx_t file_n = float_count(o.dataset);
    const idx_t n = o.dataset_size_set ? o.dataset_size : file_n;
    if (n > file_n)
        throw std::runtime_error("--dataset-size exceeds float count in " + o.dataset);
    if (n < o.window)
        throw std::runtime_error("dataset has fewer samples than --window-length");
    const idx_t qfloats = float_count(o.queries);
    idx_t qcount = o.query_count;
    if (o.query_count_set) {
        if (qcount * o.window > qfloats)
            throw std::runtime_error("--query-count * --window-length exceeds float count in " + o.queries);
    } else {
        if (qfloats % o.window != 0)
            throw std::runtime_error("query file float count is not divisible by --window-length");
        qcount = qfloats / o.window;
        if (qcount == 0)
            throw std::runtime_error("query file contains no complete query vectors");
    }
    if (qcount > 65535)
        throw std::runtime_error("at most 65535 templates per run in this build");
    const std::vector<float> queries = read_floats(o.queries, qcount * o.window);
    const double input_s = now_s() - t_in;

    // Lower bounds for the methods other than sequential (tremor_main.cpp:302).
    tremor_amd::SequentialScan scan(o.workers, o.paa, o.merge, o.replication_groups, o.method != "sequential",
                                    o.precision, o.device);
    // The skip-sequential method is the bounded early-abandoning scan for
    // every template (unless --scan-mode says otherwise).
    const bool sss = o.method == "skip-sequential";
    scan.set_forms(o.fft_mode, sss && !o.scan_mode_given ? TMX_SCAN_ALWAYS : o.scan_mode);
    // The tremor method: the iSAX index (16 PAA segments), for every eligible
    // template with leaf refinement or by the 1% rule with adaptive.
    if (o.method == "tremor") {
        if (o.paa != 16)
            throw std::invalid_argument("this GPU driver's index uses 16 PAA segments");
        scan.set_index(o.adaptive ? TMX_INDEX_AUTO : TMX_INDEX_ALWAYS, o.leaf_size);
    }
    const double t_build = now_s();
    scan.build(o.dataset, n, static_cast<int>(o.window));
    const double build_s = now_s() - t_build;
    const double t_search = now_s();
    if (o.knn)
        scan.search_knn(queries.data(), static_cast<int>(qcount), static_cast<int>(o.k), o.output);
    else
        scan.search_threshold(queries.data(), static_cast<int>(qcount), o.rho, o.output);
    const double search_s = input_s + now_s() - t_search;
    if (!o.no_bench && !o.bench_output.empty()) {
        FILE *f = std::fopen(o.bench_output.c_str(), "w");
        if (f == nullptr) {
            std::fprintf(stderr, "Warning: cannot open benchmark report file %s: %s\n", o.bench_output.c_str(),
                         std::strerror(errno));
        } else {
            std::fprintf(f, "node_rank,index_time,query_time,index_gb\n");
            std::fprintf(f, "%d,%0.6f,%0.6f,%0.3f\n", 0, build_s, search_s, scan.summary_memory_bytes() / 1e9);
            std::fclose(f);
            std::printf("Benchmark timings CSV: %s\n", o.bench_output.c_str());
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(parse(argc, argv));
    } catch (const std::invalid_argument &e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        usage(argv[0]);
        return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
