// C-ABI of the MI355X template-matching library (include/tremor_amd/tmx.h).
//
// A context owns one HIP stream on one device, the resident waveform with its
// FP64 prefix sums, and the per-call work buffers (grown on demand, reused).
// A search call runs on the device end to end:
//   prepare  - reference z-norm + block order of every template (prepare.hip)
//   seed     - kNN only: a strided sample of waveform blocks gives every
//              template a rigorous upper bound on its answer's k-th
//              (2k-1-th separated) distance (scan/mfma + select.hip)
//   main     - inner product of every template with every window and the
//              fused filter: a window becomes a candidate iff the lower
//              bound of its reference distance is within the bound
//   verify   - the reference's exact arithmetic on every candidate
//   select   - canonical kNN selection, or threshold hits by position
// Only the (small) results and the candidate count cross PCIe.
#include "tremor_amd/tmx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "filter.hpp"
#include "tmx_internal.hpp"

namespace tmx {
// select.hip
hipError_t launch_group_sort(const Cand *cand, const float *val, uint64_t count, const TmplDev *tmpl, int T, int pred,
                             int keykind, uint32_t *counts, uint32_t *offsets, uint32_t *cursor, uint64_t *keys_tmp,
                             uint64_t *keys, void *sort_tmp, size_t sort_tmp_bytes, uint64_t key_capacity,
                             hipStream_t stream);
size_t group_sort_tmp_bytes(uint64_t count, int T);
hipError_t launch_seed_bound(const uint64_t *keys, const uint32_t *offsets, TmplDev *tmpl, int T, int want, int merge,
                             hipStream_t stream);
hipError_t launch_knn_select(const uint64_t *keys, const uint32_t *offsets, int T, int k, int merge,
                             uint64_t *out_pos, float *out_d, hipStream_t stream);
// mfma.hip
hipError_t launch_mfma_pack(const TmplDev *tmpl, int count, const float *qz, int precision, void *a_frag,
                            hipStream_t stream);
hipError_t launch_scan_mfma(const Waveform &w, const TmplDev *tmpl, int count, const int32_t *group_begin, int groups,
                            int precision, const void *a_frag, const float2 *rb, const float4 *binfo, int rb_len,
                            int mode, int block_stride, Cand *cand, float *cand_val, unsigned long long *cand_count,
                            uint64_t cand_cap, hipStream_t stream);
int mfma_group_size();
uint64_t mfma_block_windows();
int64_t mfma_frag_count(int kpad);
int mfma_kpad(int m);
// misc kernels
hipError_t launch_zero_u64(unsigned long long *p, hipStream_t stream);
hipError_t launch_rho_err(const Cand *cand, const float *est, const float *exact, uint64_t count,
                          const TmplDev *tmpl, unsigned int *out_bits, hipStream_t stream);
}  // namespace tmx

using namespace tmx;

namespace {

enum GroupPredHost { kPredAllH = 0, kPredHitH = 1, kPredFiniteH = 2 };
enum GroupKeyHost { kKeyValuePosH = 0, kKeyPosValueH = 1 };

struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
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

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct tmx_ctx {
    tmx_cfg cfg{};
    hipStream_t stream = nullptr;
    std::string err;
    // resident waveform
    DevMem x, p, prefix_work;
    uint64_t n = 0;
    // block-rebased FP32 prefix pairs for the matrix-core pass (built on demand)
    DevMem rb, binfo;
    int rb_len = 0;
    uint64_t rb_n = 0;
    // per-call buffers
    DevMem raw, qz, order, tmpl, groups, afrag;
    DevMem cand, cand_val, cand_d, cand_count;
    DevMem counts, offsets, cursor, keys_tmp, keys, sort_tmp;
    DevMem out_pos, out_d, errbits;
    uint64_t cand_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_pass_ms = 0.0f;

    int fail(const std::string &msg, int code = -1)
    {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return -2;
    }
};

#define TMX_HIP(call, what)                                                                                          \
    do {                                                                                                             \
        hipError_t e__ = (call);                                                                                     \
        if (e__ != hipSuccess)                                                                                       \
            return ctx->hip_fail(e__, what);                                                                         \
    } while (0)

namespace {

// Host-side view of one call's templates, sorted by length (the device order).
struct CallPlan {
    int T = 0;
    std::vector<int> perm;          // device index -> caller index
    std::vector<TmplDev> tmpl;      // device order
    std::vector<int32_t> group_begin;
    int64_t raw_len = 0, order_len = 0;
    int64_t afrag_len = 0;          // 16-byte fragments
    int64_t comparisons = 0;
    double flops = 0.0;             // sum of 2 m (n - m + 1)
};

int build_plan(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets, int count,
               CallPlan &plan, std::vector<float> &raw_packed)
{
    plan.T = count;
    plan.perm.resize(count);
    std::iota(plan.perm.begin(), plan.perm.end(), 0);
    for (int i = 0; i < count; i++) {
        if (lengths[i] <= 0)
            return ctx->fail("template " + std::to_string(i) + " has length <= 0");
        if (lengths[i] > 65536)
            return ctx->fail("template " + std::to_string(i) + " is longer than 65536 samples");
    }
    std::stable_sort(plan.perm.begin(), plan.perm.end(), [&](int a, int b) { return lengths[a] < lengths[b]; });
    plan.tmpl.resize(count);
    raw_packed.clear();
    int64_t ooff = 0, aoff = 0;
    for (int d = 0; d < count; d++) {
        const int i = plan.perm[d];
        TmplDev t{};
        t.m = lengths[i];
        t.raw_off = static_cast<int64_t>(raw_packed.size());
        t.q_off = t.raw_off;
        t.order_off = ooff;
        t.kpad = mfma_kpad(t.m);
        t.a_off = aoff;
        t.bound = FLT_MAX;
        raw_packed.insert(raw_packed.end(), templates + offsets[i], templates + offsets[i] + lengths[i]);
        ooff += t.m / kBlock + 1;
        aoff += mfma_frag_count(t.kpad);
        if (static_cast<uint64_t>(t.m) <= ctx->n)
        {
            const uint64_t W = ctx->n - static_cast<uint64_t>(t.m) + 1;
            plan.comparisons += static_cast<int64_t>(W);
            plan.flops += 2.0 * static_cast<double>(t.m) * static_cast<double>(W);
        }
        plan.tmpl[d] = t;
    }
    plan.raw_len = static_cast<int64_t>(raw_packed.size());
    plan.order_len = ooff;
    plan.afrag_len = aoff;
    const int G = ctx->cfg.precision == TMX_PREC_FP32 ? 1 : mfma_group_size();
    plan.group_begin.clear();
    for (int d = 0; d < count; d += G)
        plan.group_begin.push_back(d);
    plan.group_begin.push_back(count);
    return 0;
}

// Uploads templates, prepares them on the device, packs MFMA operands.
int upload_and_prepare(tmx_ctx *ctx, CallPlan &plan, const std::vector<float> &raw_packed)
{
    const int T = plan.T;
    TMX_HIP(ctx->raw.reserve(sizeof(float) * (plan.raw_len + 64)), "hipMalloc templates");
    TMX_HIP(ctx->qz.reserve(sizeof(float) * (plan.raw_len + 64)), "hipMalloc templates");
    TMX_HIP(ctx->order.reserve(sizeof(int32_t) * (plan.order_len + 64)), "hipMalloc block orders");
    TMX_HIP(ctx->tmpl.reserve(sizeof(TmplDev) * T), "hipMalloc template table");
    TMX_HIP(ctx->groups.reserve(sizeof(int32_t) * plan.group_begin.size()), "hipMalloc groups");
    TMX_HIP(hipMemcpyAsync(ctx->raw.p, raw_packed.data(), sizeof(float) * plan.raw_len, hipMemcpyHostToDevice,
                           ctx->stream),
            "copy templates");
    TMX_HIP(hipMemcpyAsync(ctx->tmpl.p, plan.tmpl.data(), sizeof(TmplDev) * T, hipMemcpyHostToDevice, ctx->stream),
            "copy template table");
    TMX_HIP(hipMemcpyAsync(ctx->groups.p, plan.group_begin.data(), sizeof(int32_t) * plan.group_begin.size(),
                           hipMemcpyHostToDevice, ctx->stream),
            "copy groups");
    TMX_HIP(launch_prepare(ctx->raw.as<float>(), ctx->tmpl.as<TmplDev>(), T, ctx->qz.as<float>(),
                           ctx->order.as<int32_t>(), ctx->stream),
            "prepare kernel");
    if (ctx->cfg.precision != TMX_PREC_FP32) {
        TMX_HIP(ctx->afrag.reserve(16 * static_cast<size_t>(plan.afrag_len + 64)), "hipMalloc MFMA operands");
        TMX_HIP(launch_mfma_pack(ctx->tmpl.as<TmplDev>(), T, ctx->qz.as<float>(), ctx->cfg.precision, ctx->afrag.p,
                                 ctx->stream),
                "MFMA operand packing");
    }
    return 0;
}

int ensure_cand(tmx_ctx *ctx, uint64_t cap, int T)
{
    if (cap > ctx->cand_cap) {
        TMX_HIP(ctx->cand.reserve(sizeof(Cand) * cap), "hipMalloc candidates");
        TMX_HIP(ctx->cand_val.reserve(sizeof(float) * cap), "hipMalloc candidates");
        TMX_HIP(ctx->cand_d.reserve(sizeof(float) * cap), "hipMalloc candidates");
        TMX_HIP(ctx->keys_tmp.reserve(sizeof(uint64_t) * cap), "hipMalloc candidate keys");
        TMX_HIP(ctx->keys.reserve(sizeof(uint64_t) * cap), "hipMalloc candidate keys");
        ctx->cand_cap = cap;
    }
    TMX_HIP(ctx->cand_count.reserve(sizeof(unsigned long long)), "hipMalloc counter");
    TMX_HIP(ctx->counts.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc counts");
    TMX_HIP(ctx->offsets.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc offsets");
    TMX_HIP(ctx->cursor.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc cursors");
    TMX_HIP(ctx->errbits.reserve(sizeof(unsigned int)), "hipMalloc error word");
    return 0;
}

// Block-rebased prefix pairs long enough for the plan's longest template.
int ensure_block_prefix(tmx_ctx *ctx, const CallPlan &plan)
{
    int kmax = 0;
    for (const auto &t : plan.tmpl)
        kmax = std::max(kmax, t.kpad);
    const int nw = static_cast<int>(mfma_block_windows());
    const int need = ((nw + kmax + 1023) / 1024) * 1024;
    if (ctx->rb_n == ctx->n && ctx->rb_len >= need)
        return 0;
    const uint64_t blocks = (ctx->n + nw - 1) / nw;
    ctx->rb_n = 0;
    TMX_HIP(ctx->rb.reserve(sizeof(float2) * blocks * static_cast<uint64_t>(need)), "hipMalloc block prefix");
    TMX_HIP(ctx->binfo.reserve(sizeof(float4) * blocks), "hipMalloc block prefix");
    TMX_HIP(launch_block_prefix(ctx->p.as<double2>(), ctx->n, nw, need, ctx->binfo.as<float4>(), ctx->rb.as<float2>(),
                                ctx->stream),
            "block prefix");
    ctx->rb_len = need;
    ctx->rb_n = ctx->n;
    return 0;
}

// One inner-product pass (seed or main) over every group; returns the entry
// count through *count (may exceed the capacity).
int run_pass(tmx_ctx *ctx, const CallPlan &plan, int mode, int block_stride, uint64_t *count)
{
    if (ctx->cfg.precision != TMX_PREC_FP32) {
        const int rc = ensure_block_prefix(ctx, plan);
        if (rc)
            return rc;
    }
    Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->n};
    TMX_HIP(launch_zero_u64(ctx->cand_count.as<unsigned long long>(), ctx->stream), "reset counter");
    const int groups = static_cast<int>(plan.group_begin.size()) - 1;
    TMX_HIP(hipEventRecord(ctx->ev0, ctx->stream), "event");
    if (ctx->cfg.precision == TMX_PREC_FP32) {
        TMX_HIP(launch_scan_fp32(w, ctx->tmpl.as<TmplDev>(), ctx->qz.as<float>(), plan.T, ctx->groups.as<int32_t>(),
                                 groups, mode, block_stride, ctx->cand.as<Cand>(), ctx->cand_val.as<float>(),
                                 ctx->cand_count.as<unsigned long long>(), ctx->cand_cap, ctx->stream),
                "FP32 scan");
    } else {
        TMX_HIP(launch_scan_mfma(w, ctx->tmpl.as<TmplDev>(), plan.T, ctx->groups.as<int32_t>(), groups,
                                 ctx->cfg.precision, ctx->afrag.p, ctx->rb.as<float2>(), ctx->binfo.as<float4>(),
                                 ctx->rb_len, mode, block_stride, ctx->cand.as<Cand>(), ctx->cand_val.as<float>(),
                                 ctx->cand_count.as<unsigned long long>(), ctx->cand_cap, ctx->stream),
                "MFMA scan");
    }
    TMX_HIP(hipEventRecord(ctx->ev1, ctx->stream), "event");
    unsigned long long c = 0;
    TMX_HIP(hipMemcpyAsync(&c, ctx->cand_count.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "read counter");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "inner-product pass");
    TMX_HIP(hipEventElapsedTime(&ctx->last_pass_ms, ctx->ev0, ctx->ev1), "event time");
    *count = c;
    return 0;
}

// Main pass with buffer growth on overflow.
int run_main(tmx_ctx *ctx, const CallPlan &plan, uint64_t *count, int *retries)
{
    for (;;) {
        int rc = run_pass(ctx, plan, kModeMain, 1, count);
        if (rc)
            return rc;
        if (*count <= ctx->cand_cap)
            return 0;
        if (*count > (1ull << 31) - 1)
            return ctx->fail("more than 2^31 candidate windows; raise the bound or split the call", 1);
        (*retries)++;
        rc = ensure_cand(ctx, *count + (*count >> 3) + 1024, plan.T);
        if (rc)
            return rc;
    }
}

uint64_t waveform_blocks(tmx_ctx *ctx)
{
    const uint64_t wb = ctx->cfg.precision == TMX_PREC_FP32 ? static_cast<uint64_t>(kScanWindows)
                                                             : mfma_block_windows();
    return (ctx->n + wb - 1) / wb;
}

int group_sort(tmx_ctx *ctx, const float *val, uint64_t count, int T, int pred, int keykind)
{
    const size_t tmp = group_sort_tmp_bytes(count, T);
    TMX_HIP(ctx->sort_tmp.reserve(tmp + 256), "hipMalloc sort scratch");
    TMX_HIP(launch_group_sort(ctx->cand.as<Cand>(), val, count, ctx->tmpl.as<TmplDev>(), T, pred, keykind,
                              ctx->counts.as<uint32_t>(), ctx->offsets.as<uint32_t>(), ctx->cursor.as<uint32_t>(),
                              ctx->keys_tmp.as<uint64_t>(), ctx->keys.as<uint64_t>(), ctx->sort_tmp.p,
                              ctx->sort_tmp.bytes, ctx->cand_cap, ctx->stream),
            "group/sort");
    return 0;
}

int verify_and_err(tmx_ctx *ctx, const CallPlan &plan, uint64_t count, tmx_stats *st)
{
    TMX_HIP(launch_verify(ctx->x.as<float>(), ctx->n, ctx->tmpl.as<TmplDev>(), ctx->qz.as<float>(),
                          ctx->order.as<int32_t>(), ctx->cand.as<Cand>(), count, ctx->cand_d.as<float>(), ctx->stream),
            "verify kernel");
    TMX_HIP(launch_rho_err(ctx->cand.as<Cand>(), ctx->cand_val.as<float>(), ctx->cand_d.as<float>(), count,
                           ctx->tmpl.as<TmplDev>(), ctx->errbits.as<unsigned int>(), ctx->stream),
            "error reduction");
    unsigned int bits = 0;
    TMX_HIP(hipMemcpyAsync(&bits, ctx->errbits.p, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream), "read error");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "verify");
    float e;
    std::memcpy(&e, &bits, sizeof(e));
    if (st)
        st->max_abs_rho_err = e;
    (void)plan;
    return 0;
}

int check_common(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets, int count)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (count <= 0)
        return ctx->fail("template count must be positive");
    if (!templates || !lengths || !offsets)
        return ctx->fail("null template arguments");
    if (count > 65535)
        return ctx->fail("at most 65535 templates per call");
    return 0;
}

}  // namespace

extern "C" {

int tmx_cfg_default(tmx_cfg *cfg)
{
    if (!cfg)
        return -1;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->precision = TMX_PREC_BF16X3;
    cfg->cand_capacity = 0;
    cfg->seed_target = 0;
    cfg->verbose = 0;
    return 0;
}

int tmx_create(tmx_ctx **out, const tmx_cfg *cfg)
{
    if (!out)
        return -1;
    *out = nullptr;
    tmx_ctx *ctx = new tmx_ctx();
    if (cfg)
        ctx->cfg = *cfg;
    else
        tmx_cfg_default(&ctx->cfg);
    if (ctx->cfg.precision < TMX_PREC_FP32 || ctx->cfg.precision > TMX_PREC_FP16) {
        delete ctx;
        return -1;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || ctx->cfg.device < 0 || ctx->cfg.device >= ndev) {
        delete ctx;
        return -3;
    }
    if (hipSetDevice(ctx->cfg.device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return -3;
    }
    *out = ctx;
    return 0;
}

void tmx_destroy(tmx_ctx *ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->cfg.device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->ev0)
        (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1)
        (void)hipEventDestroy(ctx->ev1);
    delete ctx;
}

const char *tmx_last_error(const tmx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tmx_set_precision(tmx_ctx *ctx, int32_t precision)
{
    if (!ctx)
        return -1;
    if (precision < TMX_PREC_FP32 || precision > TMX_PREC_FP16)
        return ctx->fail("unknown precision");
    ctx->cfg.precision = precision;
    return 0;
}

int tmx_load_waveform(tmx_ctx *ctx, const float *x, uint64_t n)
{
    if (!ctx)
        return -1;
    if (!x || n == 0)
        return ctx->fail("empty waveform");
    if (n >= (1ull << 32))
        return ctx->fail("waveforms of 2^32 samples or more are not supported");
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    // Tail padding: inner-product tiles read past the last window.
    const size_t pad = 1 << 20;
    ctx->n = 0;
    ctx->rb_n = 0;
    TMX_HIP(ctx->x.reserve(sizeof(float) * (n + pad)), "hipMalloc waveform");
    TMX_HIP(ctx->p.reserve(sizeof(double2) * (n + 1)), "hipMalloc prefix sums");
    const size_t wl = prefix_work_len(n);
    TMX_HIP(ctx->prefix_work.reserve(sizeof(double) * wl), "hipMalloc prefix scratch");
    TMX_HIP(hipMemcpyAsync(ctx->x.p, x, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream), "copy waveform");
    TMX_HIP(hipMemsetAsync(ctx->x.as<float>() + n, 0, sizeof(float) * pad, ctx->stream), "clear padding");
    TMX_HIP(launch_prefix(ctx->x.as<float>(), n, ctx->p.as<double2>(), ctx->prefix_work.as<double>(), wl,
                          ctx->stream),
            "prefix sums");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "load waveform");
    ctx->n = n;
    return 0;
}

uint64_t tmx_waveform_length(const tmx_ctx *ctx) { return ctx ? ctx->n : 0; }

int tmx_knn(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets, int32_t count,
            int32_t k, int32_t merge_offset, uint64_t *out_pos, float *out_d, tmx_stats *stats)
{
    int rc = check_common(ctx, templates, lengths, offsets, count);
    if (rc)
        return rc;
    if (k <= 0 || k > 4096)
        return ctx->fail("k must be in [1, 4096]");
    if (merge_offset < 0)
        return ctx->fail("merge offset must be non-negative");
    if (!out_pos || !out_d)
        return ctx->fail("null outputs");
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    tmx_stats st{};
    const double t0 = now_s();
    CallPlan plan;
    std::vector<float> raw_packed;
    if ((rc = build_plan(ctx, templates, lengths, offsets, count, plan, raw_packed)))
        return rc;
    st.comparisons = plan.comparisons;
    if ((rc = upload_and_prepare(ctx, plan, raw_packed)))
        return rc;
    const uint64_t def_cap = ctx->cfg.cand_capacity > 0 ? static_cast<uint64_t>(ctx->cfg.cand_capacity)
                                                        : (1ull << 22);
    if ((rc = ensure_cand(ctx, def_cap, count)))
        return rc;

    // Seed: per-wavefront minima of the distance upper bound over a strided
    // sample of blocks, then the k-th (2k-1-th separated) smallest.
    const int want = merge_offset > 0 ? 2 * k - 1 : k;
    const uint64_t blocks = waveform_blocks(ctx);
    const uint64_t target = ctx->cfg.seed_target > 0 ? static_cast<uint64_t>(ctx->cfg.seed_target) : 1024;
    const uint64_t per_block = ctx->cfg.precision == TMX_PREC_FP32 ? 4 : 4;
    uint64_t stride = (blocks * per_block) / target;
    if (stride < 1)
        stride = 1;
    double t1 = now_s();
    uint64_t nseed = 0;
    if ((rc = run_pass(ctx, plan, kModeSeed, static_cast<int>(std::min<uint64_t>(stride, 1u << 30)), &nseed)))
        return rc;
    if (nseed > ctx->cand_cap)
        nseed = ctx->cand_cap;  // any subset of seed entries still gives a valid bound
    if ((rc = group_sort(ctx, ctx->cand_val.as<float>(), nseed, count, kPredFiniteH, kKeyValuePosH)))
        return rc;
    TMX_HIP(launch_seed_bound(ctx->keys.as<uint64_t>(), ctx->offsets.as<uint32_t>(), ctx->tmpl.as<TmplDev>(), count,
                              want, merge_offset, ctx->stream),
            "seed bound");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "seed");
    double t2 = now_s();
    st.seconds_seed = t2 - t1;

    uint64_t ncand = 0;
    if ((rc = run_main(ctx, plan, &ncand, &st.retries)))
        return rc;
    st.kernel_ms_main = ctx->last_pass_ms;
    st.flops_main = plan.flops;
    double t3 = now_s();
    st.seconds_main = t3 - t2;
    st.candidates = static_cast<int64_t>(ncand);
    if ((rc = verify_and_err(ctx, plan, ncand, &st)))
        return rc;
    double t4 = now_s();
    st.seconds_verify = t4 - t3;

    if ((rc = group_sort(ctx, ctx->cand_d.as<float>(), ncand, count, kPredAllH, kKeyValuePosH)))
        return rc;
    const size_t slots = static_cast<size_t>(count) * static_cast<size_t>(k);
    TMX_HIP(ctx->out_pos.reserve(sizeof(uint64_t) * slots), "hipMalloc results");
    TMX_HIP(ctx->out_d.reserve(sizeof(float) * slots), "hipMalloc results");
    TMX_HIP(launch_knn_select(ctx->keys.as<uint64_t>(), ctx->offsets.as<uint32_t>(), count, k, merge_offset,
                              ctx->out_pos.as<uint64_t>(), ctx->out_d.as<float>(), ctx->stream),
            "kNN selection");
    std::vector<uint64_t> hp(slots);
    std::vector<float> hd(slots);
    TMX_HIP(hipMemcpyAsync(hp.data(), ctx->out_pos.p, sizeof(uint64_t) * slots, hipMemcpyDeviceToHost, ctx->stream),
            "read results");
    TMX_HIP(hipMemcpyAsync(hd.data(), ctx->out_d.p, sizeof(float) * slots, hipMemcpyDeviceToHost, ctx->stream),
            "read results");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "kNN selection");
    for (int d = 0; d < count; d++) {
        const int i = plan.perm[d];
        std::memcpy(out_pos + static_cast<size_t>(i) * k, hp.data() + static_cast<size_t>(d) * k,
                    sizeof(uint64_t) * k);
        std::memcpy(out_d + static_cast<size_t>(i) * k, hd.data() + static_cast<size_t>(d) * k, sizeof(float) * k);
    }
    double t5 = now_s();
    st.seconds_select = t5 - t4;
    st.seconds_total = t5 - t0;
    if (ctx->cfg.verbose)
        std::fprintf(stderr, "tmx_knn: T=%d k=%d seed %.3f ms main %.3f ms (%llu cand) verify %.3f ms select %.3f ms\n",
                     count, k, 1e3 * st.seconds_seed, 1e3 * st.seconds_main,
                     static_cast<unsigned long long>(ncand), 1e3 * st.seconds_verify, 1e3 * st.seconds_select);
    if (stats)
        *stats = st;
    return 0;
}

int tmx_threshold(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets,
                  int32_t count, float rho, int32_t merge_offset, tmx_hits **out, tmx_stats *stats)
{
    int rc = check_common(ctx, templates, lengths, offsets, count);
    if (rc)
        return rc;
    if (!out)
        return ctx->fail("null output");
    if (merge_offset < 0)
        return ctx->fail("merge offset must be non-negative");
    if (!std::isfinite(rho))
        return ctx->fail("correlation threshold must be finite");
    *out = nullptr;
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    tmx_stats st{};
    const double t0 = now_s();
    CallPlan plan;
    std::vector<float> raw_packed;
    if ((rc = build_plan(ctx, templates, lengths, offsets, count, plan, raw_packed)))
        return rc;
    // Threshold distance of the reference: cc_to_ed (ZnormSearch.hpp:23-26).
    for (auto &t : plan.tmpl)
        t.bound = 2.0f * static_cast<float>(t.m) * (1.0f - rho);
    st.comparisons = plan.comparisons;
    if ((rc = upload_and_prepare(ctx, plan, raw_packed)))
        return rc;
    const uint64_t def_cap = ctx->cfg.cand_capacity > 0 ? static_cast<uint64_t>(ctx->cfg.cand_capacity)
                                                        : (1ull << 22);
    if ((rc = ensure_cand(ctx, def_cap, count)))
        return rc;
    const double t2 = now_s();
    uint64_t ncand = 0;
    if ((rc = run_main(ctx, plan, &ncand, &st.retries)))
        return rc;
    st.kernel_ms_main = ctx->last_pass_ms;
    st.flops_main = plan.flops;
    const double t3 = now_s();
    st.seconds_main = t3 - t2;
    st.candidates = static_cast<int64_t>(ncand);
    if ((rc = verify_and_err(ctx, plan, ncand, &st)))
        return rc;
    const double t4 = now_s();
    st.seconds_verify = t4 - t3;
    if ((rc = group_sort(ctx, ctx->cand_d.as<float>(), ncand, count, kPredHitH, kKeyPosValueH)))
        return rc;
    std::vector<uint32_t> offs(count + 1);
    TMX_HIP(hipMemcpyAsync(offs.data(), ctx->offsets.p, sizeof(uint32_t) * (count + 1), hipMemcpyDeviceToHost,
                           ctx->stream),
            "read hit offsets");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "hit offsets");
    const uint32_t nh = offs[count];
    std::vector<uint64_t> keys(nh);
    if (nh)
        TMX_HIP(hipMemcpyAsync(keys.data(), ctx->keys.p, sizeof(uint64_t) * nh, hipMemcpyDeviceToHost, ctx->stream),
                "read hits");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "read hits");
    // Merge by position (SkipSequential.cpp:279-290), templates in caller order.
    std::vector<int> inv(count);
    for (int d = 0; d < count; d++)
        inv[plan.perm[d]] = d;
    tmx_hits *h = static_cast<tmx_hits *>(std::calloc(1, sizeof(tmx_hits)));
    std::vector<int32_t> rt;
    std::vector<int64_t> rc_;
    std::vector<uint64_t> rp;
    std::vector<float> rd;
    for (int i = 0; i < count; i++) {
        const int d = inv[i];
        long long counter = 0;
        bool has_last = false;
        uint64_t last = 0;
        for (uint32_t e = offs[d]; e < offs[d + 1]; e++) {
            const uint64_t pos = keys[e] >> 32;
            float dist;
            const uint32_t bits = static_cast<uint32_t>(keys[e]);
            std::memcpy(&dist, &bits, sizeof(dist));
            if (has_last && merge_offset > 0 && pos < last + static_cast<uint64_t>(merge_offset + 1))
                continue;
            rt.push_back(i);
            rc_.push_back(counter++);
            rp.push_back(pos);
            rd.push_back(dist);
            last = pos;
            has_last = true;
        }
    }
    h->count = static_cast<int64_t>(rt.size());
    const size_t hn = rt.size() ? rt.size() : 1;
    h->tmpl = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * hn));
    h->counter = static_cast<int64_t *>(std::malloc(sizeof(int64_t) * hn));
    h->pos = static_cast<uint64_t *>(std::malloc(sizeof(uint64_t) * hn));
    h->d = static_cast<float *>(std::malloc(sizeof(float) * hn));
    if (!rt.empty()) {
        std::memcpy(h->tmpl, rt.data(), sizeof(int32_t) * rt.size());
        std::memcpy(h->counter, rc_.data(), sizeof(int64_t) * rt.size());
        std::memcpy(h->pos, rp.data(), sizeof(uint64_t) * rt.size());
        std::memcpy(h->d, rd.data(), sizeof(float) * rt.size());
    }
    *out = h;
    const double t5 = now_s();
    st.seconds_select = t5 - t4;
    st.seconds_total = t5 - t0;
    if (ctx->cfg.verbose)
        std::fprintf(stderr, "tmx_threshold: T=%d main %.3f ms (%llu cand) verify %.3f ms select %.3f ms\n", count,
                     1e3 * st.seconds_main, static_cast<unsigned long long>(ncand), 1e3 * st.seconds_verify,
                     1e3 * st.seconds_select);
    if (stats)
        *stats = st;
    return 0;
}

void tmx_hits_free(tmx_hits *hits)
{
    if (!hits)
        return;
    std::free(hits->tmpl);
    std::free(hits->counter);
    std::free(hits->pos);
    std::free(hits->d);
    std::free(hits);
}

int tmx_prepare_template(tmx_ctx *ctx, const float *tmpl, int32_t m, float *out_z, int32_t *out_order)
{
    if (!ctx)
        return -1;
    if (!tmpl || m <= 0 || m > 65536 || !out_z)
        return ctx->fail("bad template");
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    CallPlan plan;
    std::vector<float> raw_packed;
    const uint64_t off = 0;
    const uint64_t saved_n = ctx->n;
    int rc = build_plan(ctx, tmpl, &m, &off, 1, plan, raw_packed);
    if (rc)
        return rc;
    const int saved_prec = ctx->cfg.precision;
    ctx->cfg.precision = TMX_PREC_FP32;  // no MFMA packing needed
    rc = upload_and_prepare(ctx, plan, raw_packed);
    ctx->cfg.precision = saved_prec;
    (void)saved_n;
    if (rc)
        return rc;
    TMX_HIP(hipMemcpyAsync(out_z, ctx->qz.p, sizeof(float) * m, hipMemcpyDeviceToHost, ctx->stream), "read template");
    if (out_order && m / kBlock > 0)
        TMX_HIP(hipMemcpyAsync(out_order, ctx->order.p, sizeof(int32_t) * (m / kBlock), hipMemcpyDeviceToHost,
                               ctx->stream),
                "read block order");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "prepare template");
    return 0;
}

int tmx_distance_batch(tmx_ctx *ctx, const float *tmpl, int32_t m, const uint64_t *starts, int32_t count, float *out_d)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (!tmpl || m <= 0 || m > 65536 || !starts || !out_d || count < 0)
        return ctx->fail("bad arguments");
    if (count == 0)
        return 0;
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    CallPlan plan;
    std::vector<float> raw_packed;
    const uint64_t off = 0;
    int rc = build_plan(ctx, tmpl, &m, &off, 1, plan, raw_packed);
    if (rc)
        return rc;
    const int saved_prec = ctx->cfg.precision;
    ctx->cfg.precision = TMX_PREC_FP32;
    rc = upload_and_prepare(ctx, plan, raw_packed);
    ctx->cfg.precision = saved_prec;
    if (rc)
        return rc;
    DevMem dstarts, dout;
    TMX_HIP(dstarts.reserve(sizeof(uint64_t) * count), "hipMalloc starts");
    TMX_HIP(dout.reserve(sizeof(float) * count), "hipMalloc distances");
    TMX_HIP(hipMemcpyAsync(dstarts.p, starts, sizeof(uint64_t) * count, hipMemcpyHostToDevice, ctx->stream),
            "copy starts");
    TMX_HIP(launch_dist_batch(ctx->x.as<float>(), ctx->n, ctx->tmpl.as<TmplDev>(), ctx->qz.as<float>(),
                              ctx->order.as<int32_t>(), dstarts.as<uint64_t>(), count, dout.as<float>(), ctx->stream),
            "distance kernel");
    TMX_HIP(hipMemcpyAsync(out_d, dout.p, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream),
            "read distances");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "distance batch");
    return 0;
}

int tmx_window_stats(tmx_ctx *ctx, int32_t m, uint64_t first, uint64_t count, float *mean, float *std_out)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (m <= 0 || !mean || !std_out || static_cast<uint64_t>(m) > ctx->n ||
        first + count > ctx->n - static_cast<uint64_t>(m) + 1)
        return ctx->fail("window range out of bounds");
    if (count == 0)
        return 0;
    TMX_HIP(hipSetDevice(ctx->cfg.device), "hipSetDevice");
    DevMem a, b;
    TMX_HIP(a.reserve(sizeof(float) * count), "hipMalloc stats");
    TMX_HIP(b.reserve(sizeof(float) * count), "hipMalloc stats");
    TMX_HIP(launch_window_stats(ctx->x.as<float>(), m, first, count, a.as<float>(), b.as<float>(), ctx->stream),
            "stats kernel");
    TMX_HIP(hipMemcpyAsync(mean, a.p, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream), "read stats");
    TMX_HIP(hipMemcpyAsync(std_out, b.p, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream), "read stats");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "window stats");
    return 0;
}

int tmx_sync(tmx_ctx *ctx)
{
    if (!ctx)
        return -1;
    TMX_HIP(hipStreamSynchronize(ctx->stream), "sync");
    return 0;
}

}  // extern "C"
