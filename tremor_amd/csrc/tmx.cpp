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

#include <hipfft/hipfft.h>

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
// fft.hip
struct FftPair {
    uint32_t t;
    uint32_t p;
};
hipError_t launch_piece_input(const float *x, uint64_t n, int step, int B, uint64_t p0, int cnt, const float4 *pinfo,
                              float *out, double *norm2, hipStream_t stream);
hipError_t launch_piece_norm(const double *norm2, float2 *pnorm, uint64_t P, hipStream_t stream);
hipError_t launch_spec_max(const float2 *spec, int rows, int bins, float *out, int stride_out, hipStream_t stream);
hipError_t launch_tmpl_input(const TmplDev *tmpl, const int32_t *ids, int count, int B, const float *qz, float *out,
                             hipStream_t stream);
hipError_t launch_product(const FftPair *pairs, int npairs, const float2 *spectra, const float2 *qspec,
                          const int32_t *slot, int bins, float2 *Y, hipStream_t stream);
hipError_t launch_fft_filter(const Waveform &w, const TmplDev *tmpl, const FftPair *pairs, int npairs, const float *y,
                             int B, int step, const float2 *rb, int rb_len, const float4 *pinfo, const float2 *pnorm,
                             const float *qmax, const int32_t *slot, double kfft, int mode, Cand *cand,
                             float *cand_val, unsigned long long *cand_count, uint64_t cap, hipStream_t stream);
double fft_error_factor(int B);
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
    // FFT (overlap-save) form: one class (B, Lmax) of cached piece spectra
    struct Fft {
        int B = 0, Lmax = 0, step = 0, bins = 0;
        uint64_t P = 0, n = 0;
        DevMem spectra, pinfo, pnorm, rb, norm2, work_real, Y, y, pairs, qspec, qmax, qin, ids, slot;
        hipfftHandle fwd = 0, inv = 0;
        int fwd_batch = 0, inv_batch = 0;
        void release_plans()
        {
            if (fwd)
                (void)hipfftDestroy(fwd);
            if (inv)
                (void)hipfftDestroy(inv);
            fwd = inv = 0;
            fwd_batch = inv_batch = 0;
        }
        ~Fft() { release_plans(); }
    } fft;

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
    int Tg = 0;                     // templates [0, Tg) take the inner-product form, [Tg, T) the FFT form
    int fft_L = 0;                  // longest FFT template
};

// Cost model (per template, seconds per window), in the role of the
// reference's adaptive choice between scan and FFT (Tremor.cpp:2265-2286,
// fft_fallback.hpp:55-60): the matrix-core form costs in proportion to the
// padded template length plus a fixed epilogue, the FFT form in proportion
// to B log2 B per piece of (B - Lmax + 1) windows.  Constants measured on
// MI355X (matrix-core loop and epilogue of the split-BF16 / FP16 passes,
// batched hipFFT c2r with the product and filter kernels); FFT iff it is
// estimated below 0.9 x the inner-product form, as in the reference.
double gemm_seconds_per_window(int precision, int kpad)
{
    const double per_k = precision == TMX_PREC_FP16 ? 2.2e-15 : 4.6e-15;
    return per_k * kpad + 2.1e-12;
}

int fft_block_for(int L)
{
    int B = 1 << 15;
    while (B < 4 * L)
        B *= 2;
    return B;
}

double fft_seconds_per_window(int L)
{
    const int B = fft_block_for(L);
    double lg = 0.0;
    while ((1 << static_cast<int>(lg)) < B)
        lg += 1.0;
    const double per_piece = 0.45e-12 * B * lg + 1.6e-12 * B;
    return per_piece / static_cast<double>(B - L + 1);
}

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
    // Path choice: templates are sorted by length and the FFT form wins for
    // long templates, so the FFT set is a suffix of the device order.
    plan.Tg = count;
    if (ctx->cfg.precision != TMX_PREC_FP32 && ctx->cfg.fft_mode != TMX_FFT_OFF) {
        if (ctx->cfg.fft_mode == TMX_FFT_ALWAYS) {
            plan.Tg = 0;
        } else {
            while (plan.Tg > 0) {
                const TmplDev &t = plan.tmpl[plan.Tg - 1];
                if (!(fft_seconds_per_window(t.m) < 0.9 * gemm_seconds_per_window(ctx->cfg.precision, t.kpad)))
                    break;
                plan.Tg--;
            }
        }
    }
    for (int d = plan.Tg; d < count; d++)
        plan.fft_L = std::max(plan.fft_L, plan.tmpl[d].m);
    const int G = ctx->cfg.precision == TMX_PREC_FP32 ? 1 : mfma_group_size();
    plan.group_begin.clear();
    for (int d = 0; d < plan.Tg; d += G)
        plan.group_begin.push_back(d);
    plan.group_begin.push_back(plan.Tg);
    return 0;
}

int fft_prepare(tmx_ctx *ctx, const CallPlan &plan);

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
        TMX_HIP(launch_mfma_pack(ctx->tmpl.as<TmplDev>(), plan.Tg, ctx->qz.as<float>(), ctx->cfg.precision,
                                 ctx->afrag.p, ctx->stream),
                "MFMA operand packing");
    }
    return fft_prepare(ctx, plan);
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
    for (int d = 0; d < plan.Tg; d++)
        kmax = std::max(kmax, plan.tmpl[d].kpad);
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

// ---- FFT form ---------------------------------------------------------------

constexpr int kFftPieceBatch = 64;  // pieces / templates per forward transform launch

// Piece spectra of the current waveform for templates up to L samples:
// B = 2^15 doubled until >= 4 L (fft_fallback.cpp:124-126), Lmax = B / 4,
// step = B - Lmax + 1; per piece the centre, the rebased prefix pairs of the
// level-1 filter, ||x~||_2 and ||X||_inf.
int ensure_fft_class(tmx_ctx *ctx, int L)
{
    auto &f = ctx->fft;
    const int B = fft_block_for(L);
    if (f.n == ctx->n && f.B == B)
        return 0;
    f.n = 0;
    if (f.B != B)
        f.release_plans();
    f.B = B;
    f.Lmax = B / 4;
    f.step = B - f.Lmax + 1;
    f.bins = B / 2 + 1;
    f.P = (ctx->n + f.step - 1) / f.step;
    const uint64_t Palloc = ((f.P + kFftPieceBatch - 1) / kFftPieceBatch) * kFftPieceBatch;
    TMX_HIP(f.spectra.reserve(sizeof(float2) * Palloc * f.bins), "hipMalloc FFT spectra");
    TMX_HIP(f.pinfo.reserve(sizeof(float4) * f.P), "hipMalloc FFT pieces");
    TMX_HIP(f.pnorm.reserve(sizeof(float2) * f.P), "hipMalloc FFT pieces");
    TMX_HIP(f.norm2.reserve(sizeof(double) * f.P), "hipMalloc FFT pieces");
    TMX_HIP(f.rb.reserve(sizeof(float2) * f.P * static_cast<uint64_t>(B + 1)), "hipMalloc FFT prefix pairs");
    TMX_HIP(f.work_real.reserve(sizeof(float) * static_cast<size_t>(kFftPieceBatch) * B), "hipMalloc FFT input");
    if (!f.fwd) {
        int nn = B;
        if (hipfftPlanMany(&f.fwd, 1, &nn, nullptr, 1, B, nullptr, 1, f.bins, HIPFFT_R2C, kFftPieceBatch) !=
            HIPFFT_SUCCESS)
            return ctx->fail("hipFFT forward plan failed", -2);
        f.fwd_batch = kFftPieceBatch;
    }
    hipfftSetStream(f.fwd, ctx->stream);
    TMX_HIP(launch_block_prefix(ctx->p.as<double2>(), ctx->n, f.step, B + 1, f.pinfo.as<float4>(), f.rb.as<float2>(),
                                ctx->stream),
            "FFT prefix pairs");
    TMX_HIP(hipMemsetAsync(f.norm2.p, 0, sizeof(double) * f.P, ctx->stream), "clear norms");
    for (uint64_t p0 = 0; p0 < f.P; p0 += kFftPieceBatch) {
        const int cnt = static_cast<int>(std::min<uint64_t>(kFftPieceBatch, f.P - p0));
        TMX_HIP(launch_piece_input(ctx->x.as<float>(), ctx->n, f.step, B, p0, cnt, f.pinfo.as<float4>(),
                                   f.work_real.as<float>(), f.norm2.as<double>(), ctx->stream),
                "FFT piece input");
        if (hipfftExecR2C(f.fwd, f.work_real.as<hipfftReal>(),
                          reinterpret_cast<hipfftComplex *>(f.spectra.as<float2>() + p0 * f.bins)) != HIPFFT_SUCCESS)
            return ctx->fail("hipFFT r2c of the pieces failed", -2);
    }
    TMX_HIP(launch_piece_norm(f.norm2.as<double>(), f.pnorm.as<float2>(), f.P, ctx->stream), "FFT piece norms");
    TMX_HIP(launch_spec_max(f.spectra.as<float2>(), static_cast<int>(f.P), f.bins,
                            reinterpret_cast<float *>(f.pnorm.as<float2>()) + 1, 2, ctx->stream),
            "FFT spectral maxima");
    TMX_HIP(hipStreamSynchronize(ctx->stream), "FFT spectra");
    f.n = ctx->n;
    return 0;
}

// Spectra of the call's FFT templates [Tg, T) (reversed, zero padded).
int fft_prepare(tmx_ctx *ctx, const CallPlan &plan)
{
    const int F = plan.T - plan.Tg;
    if (F <= 0)
        return 0;
    int rc = ensure_fft_class(ctx, plan.fft_L);
    if (rc)
        return rc;
    auto &f = ctx->fft;
    const int Falloc = ((F + kFftPieceBatch - 1) / kFftPieceBatch) * kFftPieceBatch;
    TMX_HIP(f.qspec.reserve(sizeof(float2) * static_cast<size_t>(Falloc) * f.bins), "hipMalloc template spectra");
    TMX_HIP(f.qmax.reserve(sizeof(float) * F), "hipMalloc template spectra");
    TMX_HIP(f.qin.reserve(sizeof(float) * static_cast<size_t>(kFftPieceBatch) * f.B), "hipMalloc template input");
    std::vector<int32_t> ids(F), slot(plan.T, -1);
    for (int i = 0; i < F; i++) {
        ids[i] = plan.Tg + i;
        slot[plan.Tg + i] = i;
    }
    TMX_HIP(f.ids.reserve(sizeof(int32_t) * F), "hipMalloc FFT ids");
    TMX_HIP(f.slot.reserve(sizeof(int32_t) * plan.T), "hipMalloc FFT ids");
    TMX_HIP(hipMemcpyAsync(f.ids.p, ids.data(), sizeof(int32_t) * F, hipMemcpyHostToDevice, ctx->stream), "copy ids");
    TMX_HIP(hipMemcpyAsync(f.slot.p, slot.data(), sizeof(int32_t) * plan.T, hipMemcpyHostToDevice, ctx->stream),
            "copy slots");
    hipfftSetStream(f.fwd, ctx->stream);
    for (int i0 = 0; i0 < F; i0 += kFftPieceBatch) {
        const int cnt = std::min(kFftPieceBatch, F - i0);
        TMX_HIP(launch_tmpl_input(ctx->tmpl.as<TmplDev>(), f.ids.as<int32_t>() + i0, cnt, f.B, ctx->qz.as<float>(),
                                  f.qin.as<float>(), ctx->stream),
                "FFT template input");
        if (hipfftExecR2C(f.fwd, f.qin.as<hipfftReal>(),
                          reinterpret_cast<hipfftComplex *>(f.qspec.as<float2>() + static_cast<size_t>(i0) * f.bins)) !=
            HIPFFT_SUCCESS)
            return ctx->fail("hipFFT r2c of the templates failed", -2);
    }
    TMX_HIP(launch_spec_max(f.qspec.as<float2>(), F, f.bins, f.qmax.as<float>(), 1, ctx->stream),
            "template spectral maxima");
    return 0;
}

// The FFT form of a pass over every `stride`-th piece.
int run_fft_pass(tmx_ctx *ctx, const CallPlan &plan, int mode, int stride)
{
    const int F = plan.T - plan.Tg;
    if (F <= 0)
        return 0;
    auto &f = ctx->fft;
    std::vector<FftPair> pairs;
    for (uint64_t p = 0; p < f.P; p += static_cast<uint64_t>(stride))
        for (int i = 0; i < F; i++) {
            const int t = plan.Tg + i;
            const uint64_t m = static_cast<uint64_t>(plan.tmpl[t].m);
            if (m <= ctx->n && p * f.step < ctx->n - m + 1)
                pairs.push_back(FftPair{static_cast<uint32_t>(t), static_cast<uint32_t>(p)});
        }
    if (pairs.empty())
        return 0;
    const int NB = static_cast<int>(std::max<int64_t>(16, std::min<int64_t>(1024, (512ll << 20) / (12ll * f.B))));
    if (!f.inv || f.inv_batch != NB) {
        if (f.inv)
            (void)hipfftDestroy(f.inv);
        f.inv = 0;
        int nn = f.B;
        if (hipfftPlanMany(&f.inv, 1, &nn, nullptr, 1, f.bins, nullptr, 1, f.B, HIPFFT_C2R, NB) != HIPFFT_SUCCESS)
            return ctx->fail("hipFFT inverse plan failed", -2);
        f.inv_batch = NB;
    }
    hipfftSetStream(f.inv, ctx->stream);
    TMX_HIP(f.Y.reserve(sizeof(float2) * static_cast<size_t>(NB) * f.bins), "hipMalloc FFT products");
    TMX_HIP(f.y.reserve(sizeof(float) * static_cast<size_t>(NB) * f.B), "hipMalloc FFT outputs");
    TMX_HIP(f.pairs.reserve(sizeof(FftPair) * pairs.size()), "hipMalloc FFT pairs");
    TMX_HIP(hipMemcpyAsync(f.pairs.p, pairs.data(), sizeof(FftPair) * pairs.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            "copy FFT pairs");
    Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->n};
    const double kfft = fft_error_factor(f.B);
    for (size_t b0 = 0; b0 < pairs.size(); b0 += NB) {
        const int cnt = static_cast<int>(std::min<size_t>(NB, pairs.size() - b0));
        const FftPair *pp = f.pairs.as<FftPair>() + b0;
        TMX_HIP(launch_product(pp, cnt, f.spectra.as<float2>(), f.qspec.as<float2>(), f.slot.as<int32_t>(), f.bins,
                               f.Y.as<float2>(), ctx->stream),
                "FFT product");
        if (hipfftExecC2R(f.inv, reinterpret_cast<hipfftComplex *>(f.Y.as<float2>()), f.y.as<hipfftReal>()) !=
            HIPFFT_SUCCESS)
            return ctx->fail("hipFFT c2r failed", -2);
        TMX_HIP(launch_fft_filter(w, ctx->tmpl.as<TmplDev>(), pp, cnt, f.y.as<float>(), f.B, f.step,
                                  f.rb.as<float2>(), f.B + 1, f.pinfo.as<float4>(), f.pnorm.as<float2>(),
                                  f.qmax.as<float>(), f.slot.as<int32_t>(), kfft, mode, ctx->cand.as<Cand>(),
                                  ctx->cand_val.as<float>(), ctx->cand_count.as<unsigned long long>(), ctx->cand_cap,
                                  ctx->stream),
                "FFT filter");
    }
    return 0;
}

int fft_seed_stride(tmx_ctx *ctx, uint64_t target)
{
    const auto &f = ctx->fft;
    if (f.P == 0)
        return 1;
    const uint64_t chunks = (static_cast<uint64_t>(f.step) + 4095) / 4096;
    const uint64_t s = (f.P * chunks) / std::max<uint64_t>(1, target);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(s, 1u << 30)));
}

// One inner-product pass (seed or main) over every group; returns the entry
// count through *count (may exceed the capacity).
int run_pass(tmx_ctx *ctx, const CallPlan &plan, int mode, int block_stride, uint64_t *count, int fft_stride = 1)
{
    if (ctx->cfg.precision != TMX_PREC_FP32 && plan.Tg > 0) {
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
    {
        const int rc = run_fft_pass(ctx, plan, mode, fft_stride);
        if (rc)
            return rc;
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
    cfg->fft_mode = TMX_FFT_AUTO;
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
    if (ctx->cfg.precision < TMX_PREC_FP32 || ctx->cfg.precision > TMX_PREC_FP16 ||
        ctx->cfg.fft_mode < TMX_FFT_AUTO || ctx->cfg.fft_mode > TMX_FFT_ALWAYS) {
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

int tmx_set_fft_mode(tmx_ctx *ctx, int32_t mode)
{
    if (!ctx)
        return -1;
    if (mode < TMX_FFT_AUTO || mode > TMX_FFT_ALWAYS)
        return ctx->fail("unknown FFT mode");
    ctx->cfg.fft_mode = mode;
    return 0;
}

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
    ctx->fft.n = 0;
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
    if ((rc = run_pass(ctx, plan, kModeSeed, static_cast<int>(std::min<uint64_t>(stride, 1u << 30)), &nseed,
                       fft_seed_stride(ctx, target))))
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
    st.fft_templates = plan.T - plan.Tg;
    st.fft_block = plan.T > plan.Tg ? ctx->fft.B : 0;
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
    st.fft_templates = plan.T - plan.Tg;
    st.fft_block = plan.T > plan.Tg ? ctx->fft.B : 0;
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
