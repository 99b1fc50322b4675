// Context internals shared by the synchronous C-ABI (tmx.cpp) and the
// serving loop (serve.cpp).
//
// A context holds the resident waveform and the structures derived from it
// (FP64 prefix sums, block-rebased FP32 prefix pairs, FFT piece spectra) and
// any number of lanes.  A lane is one HIP stream with its own per-batch work
// buffers; it runs one batch of templates at a time.  The synchronous calls
// use the context's direct lane, the serving loop one lane per worker thread,
// so batches of different lanes overlap on the device.
//
// Resident structures are guarded by a reader/writer lock: a batch holds it
// shared from start to end, a (re)build of a derived structure holds it
// exclusively, so nothing a running batch reads is ever reallocated.
#pragma once

#include <hipfft/hipfft.h>

#include <atomic>
#include <chrono>
#include <map>
#include <cstddef>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "tmx_internal.hpp"

namespace tmx {

struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t reserve(size_t want)
    {
        if (want <= bytes)
            return hipSuccess;
        release();
        const size_t b = want < 256 ? 256 : want;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess)
            bytes = b;
        else
            p = nullptr;
        return e;
    }
    template <class T>
    T *as() const
    {
        return static_cast<T *>(p);
    }
};

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One stream and its per-batch work buffers (grown on demand, reused).
struct Lane {
    int device = 0;
    int cus = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_pass_ms = 0.0f;
    std::string err;
    // templates of the batch
    DevMem raw, qz, order, tmpl, groups, afrag;
    // candidates, grouping, results
    DevMem cand, cand_val, cand_lo, cand_hi, cand_d, cand_count;
    DevMem cand2, val2, lo2, hi2, total;  // second candidate set (refinement)
    DevMem counts, offsets, cursor, keys_tmp, keys, tv_a, tv_b, sort_tmp;
    DevMem out_pos, out_d, errbits;
    uint64_t cand_cap = 0;
    // FFT form: template spectra, products, inverse transforms
    DevMem Y, pairs, qspec, qmax, qin, ids, slot;
    // fused FFT form: queue of chunks for the per-window pass
    DevMem chunk_rec, chunk_dt, chunk_n;
    // early-abandoning scan: its templates and the sample counters; the
    // templates' PAA and the count of windows the lower bound skipped
    DevMem scan_ids, scan_visit, lb_paa, lb_skip;
    uint64_t last_lb_skipped = 0;
    uint64_t last_ix_windows = 0;  // windows in the passing leaves of the last main pass
    uint64_t chunk_cap = 0;
    uint64_t chunk_need = 0;  // > chunk_cap after a main pass that overflowed the queue
    // matrix-core form: queue of level-1 survivors for the rigorous second level
    DevMem surv, surv_n;
    DevMem queue;  // the persistent matrix-core launch's per-XCD work counters
    DevMem ix_ids, ix_passed;  // templates on the iSAX index; windows of their passing leaves
    DevMem ix_dyn;             // running bounds (k = 1) of the scan and index searches, one per template
    uint64_t surv_cap = 0;
    uint64_t surv_need = 0;  // > surv_cap after a main pass that overflowed the queue
    uint64_t last_survivors = 0;  // level-1 survivors of the last main pass
    // template forward plans, one per FFT length (batches alternate between
    // the classes, and a plan takes milliseconds to make): one slot for each
    // length the FFT forms use, 2^14 .. 2^18, so none is ever evicted
    static constexpr int kPlanSlots = 5;
    hipfftHandle tfwd[kPlanSlots] = {};
    int tfwd_B[kPlanSlots] = {};
    // pinned host staging of the threshold hits (grown on demand)
    void *host_keys = nullptr;
    size_t host_keys_bytes = 0;

    Lane() = default;
    Lane(const Lane &) = delete;
    Lane &operator=(const Lane &) = delete;
    ~Lane() { destroy(); }
    hipError_t init(int dev)
    {
        device = dev;
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess)
            e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipEventCreate(&ev0);
        if (e == hipSuccess)
            e = hipEventCreate(&ev1);
        if (e == hipSuccess)
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        return e;
    }
    void destroy()
    {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        if (host_keys)
            (void)hipHostFree(host_keys);
        host_keys = nullptr;
        host_keys_bytes = 0;
        for (int i = 0; i < kPlanSlots; i++) {
            if (tfwd[i])
                (void)hipfftDestroy(tfwd[i]);
            tfwd[i] = 0;
            tfwd_B[i] = 0;
        }
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        ev0 = ev1 = nullptr;
        if (stream)
            (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    int fail(const std::string &msg, int code = -1)
    {
        err = msg;
        return code;
    }
    hipError_t reserve_host_keys(size_t want)
    {
        if (want <= host_keys_bytes)
            return hipSuccess;
        if (host_keys)
            (void)hipHostFree(host_keys);
        host_keys = nullptr;
        host_keys_bytes = 0;
        const size_t b = want < (1u << 20) ? (1u << 20) : want + want / 4;
        const hipError_t e = hipHostMalloc(&host_keys, b, hipHostMallocDefault);
        if (e == hipSuccess)
            host_keys_bytes = b;
        else
            host_keys = nullptr;
        return e;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return -2;
    }
};

// Overlap-save shape: pieces of B samples every B - Lmax + 1 windows, for
// templates up to Lmax samples.
struct FftShape {
    int B, Lmax;
};

// The iSAX index of one length (index.hip): window ids and keys in key order
// and the leaves.
struct IxIndex {
    DevMem keys;    // uint64 [W], sorted
    DevMem ids;     // uint32 [W], the windows in key order
    DevMem words;   // uint8 [W][16], their words in key order
    DevMem sums;    // WindowSums [W], their sums in key order (eS1 < 0: first sample missing)
    DevMem leaves;  // IxLeaf [nleaves], in key order
    int64_t nleaves = 0;
    int leaf_size = 0;
};

// iSAX words of every window of one length (skip-sequential lower bound and
// the index), and the index once built.
struct SaxWords {
    int m = 0, segments = 0;
    DevMem words;  // uint8 [n - m + 1][segments]
    IxIndex ix;
};

// Builds the index of the W windows of length m of w from their 16-segment
// words.
hipError_t build_isax_index(const Waveform &w, int m, const uint8_t *words, uint64_t W, int leaf_size, IxIndex &ix,
                            hipStream_t stream);

// Piece spectra of the resident waveform for one shape (B, Lmax).
struct FftClass {
    int B = 0, Lmax = 0, step = 0, bins = 0;
    uint64_t P = 0;
    bool fused = false;  // B = 2^14 or 2^15: fused LDS pass (twiddle table, no rebased prefix pairs)
    DevMem spectra, pinfo, pnorm, rb, norm2, work_real, table;
    hipfftHandle fwd = 0;
    FftClass() = default;
    FftClass(const FftClass &) = delete;
    FftClass &operator=(const FftClass &) = delete;
    ~FftClass()
    {
        if (fwd)
            (void)hipfftDestroy(fwd);
    }
};

// One template of a batch.
struct Query {
    const float *data;  // raw samples
    int32_t m;
    int32_t k;          // > 0: kNN with k answers; 0: threshold
    int32_t merge;      // merge offset
    float rho;          // threshold correlation (threshold templates)
};

// Its answer: kNN, k slots (unfilled (0, FLT_MAX)); threshold, the hits by
// position after the merge.
struct QueryOut {
    std::vector<uint64_t> pos;
    std::vector<float> d;
};

struct Server;
struct ServerDelete {
    void operator()(Server *s) const;  // serve.cpp
};

}  // namespace tmx

struct tmx_ctx {
    tmx_cfg cfg{};
    std::mutex err_mu;
    std::string err;
    std::mutex call_mu;            // one synchronous call at a time
    std::shared_mutex res_mu;      // resident structures (see above)
    std::mutex gate_mu;            // taken before res_mu: a waiting writer holds it, so readers cannot starve it
    tmx::Lane direct;              // lane of the synchronous calls
    // resident waveform
    tmx::DevMem x, p, pl, prefix_work;  // samples; double-double prefix sums (high, low parts)
    uint64_t n = 0;
    // block-rebased FP32 prefix pairs of the matrix-core pass
    tmx::DevMem rb, binfo;
    int rb_len = 0;
    uint64_t rb_n = 0;
    // FFT piece spectra, one class per block length
    std::vector<std::unique_ptr<tmx::FftClass>> fft;
    // iSAX words per (template length, segments), oldest first, and the
    // breakpoints of the lower bounds
    std::vector<std::unique_ptr<tmx::SaxWords>> sax;
    tmx::DevMem sax_bp;
    // serving loop (null when not running)
    std::unique_ptr<tmx::Server, tmx::ServerDelete> server;
    // lanes of stopped serving loops, kept with their grown buffers and FFT
    // plans for the next start (guarded by spare_mu)
    std::mutex spare_mu;
    std::vector<std::unique_ptr<tmx::Lane>> spare_lanes;
    // largest candidate buffer and chunk queue (entries) any lane has needed:
    // lanes reserve at least these, so one overflow and rerun teaches all
    std::atomic<uint64_t> cand_hint{0}, chunk_hint{0}, surv_hint{0};
    // autotuner: templates per matrix-core workgroup chosen per batch shape
    // (precision, template count, mean padded length), guarded by tune_mu
    std::mutex tune_mu;
    std::map<uint32_t, int> tuned_group;  // shape -> group size | launch form << 8

    int fail(const std::string &msg, int code = -1)
    {
        std::lock_guard<std::mutex> lk(err_mu);
        err = msg;
        return code;
    }
};

namespace tmx {

// Shared (reader) hold of the resident structures, through the gate.
inline std::shared_lock<std::shared_mutex> read_resident(tmx_ctx *ctx)
{
    std::lock_guard<std::mutex> gate(ctx->gate_mu);
    return std::shared_lock<std::shared_mutex>(ctx->res_mu);
}

// Exclusive (writer) hold: the gate stays held until the writer is done.
struct WriteResident {
    std::lock_guard<std::mutex> gate;
    std::unique_lock<std::shared_mutex> lock;
    explicit WriteResident(tmx_ctx *ctx) : gate(ctx->gate_mu), lock(ctx->res_mu) {}
};

// Return code of one batch pass whose candidates exceed what one pass may
// hold; run_batch then splits the batch's templates.
constexpr int kRcSplit = 3;

// The whole search for one batch of templates on lane L: preparation, kNN
// bound seeding, candidate pass, exact verification, selection.  out gets
// one entry per query, in query order.  Thread-safe across lanes.
int run_batch(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const Query *q, int count, std::vector<QueryOut> &out,
              tmx_stats *st);

// Stops the serving loop (if running): every submitted request is finished.
void server_shutdown(tmx_ctx *ctx);
bool server_running(const tmx_ctx *ctx);

}  // namespace tmx
