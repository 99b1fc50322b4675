// C-ABI of the MI355X template-matching library (include/tremor_amd/tmx.h):
// the batch pipeline and the synchronous entry points.
//
// A batch runs on one lane (stream) end to end:
//   prepare  - reference z-norm + block order of every template (prepare.hip)
//   seed     - kNN templates: a strided sample of waveform blocks gives every
//              one a rigorous upper bound on its answer's k-th (2k-1-th
//              separated) distance (scan/mfma/fft + select.hip)
//   main     - inner product of every template with every window and the
//              fused filter: a window becomes a candidate iff the lower
//              bound of its reference distance is within the bound
//   verify   - the reference's exact arithmetic on every candidate
//   select   - canonical kNN selection, or threshold hits by position
// Only the (small) results and the candidate count cross PCIe.  kNN and
// threshold templates with their own k, correlation and merge offset can
// share one batch.
#include "tremor_amd/tmx.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "isax.hpp"
#include "filter.hpp"

namespace tmx {
// select.hip
hipError_t launch_group_sort(const Cand *cand, const float *val, uint64_t count, const TmplDev *tmpl, int T,
                             int pred, int keykind, uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                             uint64_t *keys_tmp, uint64_t *keys, uint32_t *tv_a, uint32_t *tv_b,
                             unsigned long long *total, void *sort_tmp, size_t sort_tmp_bytes, hipStream_t stream);
hipError_t launch_refine(const Cand *cand, const float *est, const float *lo, const float *hi, uint64_t count,
                         const TmplDev *tmpl, Cand *cand2, float *est2, float *lo2, float *hi2,
                         unsigned long long *total, hipStream_t stream);
size_t group_sort_tmp_bytes(uint64_t count, int T);
hipError_t launch_seed_bound(const uint64_t *keys, const uint32_t *offsets, TmplDev *tmpl, int T, int max_want,
                             hipStream_t stream);
hipError_t launch_knn_select(const uint64_t *keys, const uint32_t *offsets, const TmplDev *tmpl, int T, int max_k,
                             uint64_t *out_pos, float *out_d, hipStream_t stream);
// mfma.hip
hipError_t launch_mfma_pack(const TmplDev *tmpl, int count, int max_kpad, const float *qz, int precision,
                            void *a_frag, hipStream_t stream);
// persistent > 0: that many workgroups claim the items from per-XCD queues
// (queue: 8 counters); else one workgroup per item.
hipError_t launch_scan_mfma(const Waveform &w, const TmplDev *tmpl, int count, const int32_t *group_begin, int groups,
                            int precision, const void *a_frag, const float2 *rb, const float4 *binfo, int rb_len,
                            int mode, int block_stride, const SurvOut &so, const CandOut &co, int persistent,
                            uint32_t *queue, hipStream_t stream);
hipError_t launch_mfma_consts(TmplDev *tmpl, int count, int precision, hipStream_t stream);
hipError_t launch_mfma_level2(const Waveform &w, const TmplDev *tmpl, int precision, const float4 *binfo,
                              const SurvOut &so, const CandOut &co, int cus, hipStream_t stream);
int mfma_group_size();
uint64_t mfma_block_windows();
int mfma_block_waves();
int64_t mfma_frag_count(int kpad);
int mfma_kpad(int m);
// fft.hip
hipError_t launch_piece_input(const float *x, uint64_t n, int step, int B, uint64_t p0, int cnt, float4 *pinfo,
                              float *out, double *norm2, hipStream_t stream);
// fft_fused.hip
bool fft_fused_supports(int B);
size_t fft_chunk_rec_bytes();
int fft_chunk_width();  // windows per chunk record of the fused pass
hipError_t launch_fft_fused(const Waveform &w, const TmplDev *tmpl, int t0, int F, const float2 *spectra, int bins,
                            const float2 *qspec, const float *qmax, const float4 *pinfo, const float2 *pnorm,
                            const float2 *table, uint64_t P, int step, int stride, double kfft, int mode, void *rec,
                            float *rec_dt, unsigned long long *nrec, uint64_t rec_cap, const CandOut &co, int grid,
                            hipStream_t stream);
hipError_t launch_piece_norm(const double *norm2, float2 *pnorm, uint64_t P, hipStream_t stream);
hipError_t launch_spec_max(const float2 *spec, int rows, int bins, float *out, int stride_out, hipStream_t stream);
hipError_t launch_tmpl_input(const TmplDev *tmpl, const int32_t *ids, int count, int B, const float *qz, float *out,
                             hipStream_t stream);
// fft_large.hip
bool fft_large_supports(int B);
hipError_t launch_fft_large(const Waveform &w, const TmplDev *tmpl, const FftPair *pairs, int npairs,
                            const float2 *spectra, const float2 *qspec, const int32_t *slot, int B, const float2 *E,
                            float2 *U, int step, const float2 *rb, int rb_len, const float4 *pinfo,
                            const float2 *pnorm, const float *qmax, double kfft, int mode, const CandOut &co,
                            hipStream_t stream);
hipError_t launch_fft_large_probe(int B, const float2 *E, const float2 *zin, int count, float2 *U, float2 *zout,
                                  hipStream_t stream);
hipError_t launch_zero_f64(double *p, uint64_t count, hipStream_t stream);
double fft_error_factor(int B);
// noalloc.hip
int noalloc_max_template();
hipError_t noalloc_setup(int device);
hipError_t noalloc_distance_batch(int device, const float *x, uint64_t n, const float *d_tmpl, int m,
                                  const uint64_t *d_starts, int count, float *d_out, hipStream_t stream);
// misc.hip
hipError_t launch_zero_u64(unsigned long long *p, hipStream_t stream);
hipError_t launch_scan_ea(const Waveform &w, const TmplDev *tmpl, int T, const float *qz, const int32_t *order,
                          const int32_t *ids, int count, uint64_t max_windows, int sample, uint64_t sample_n,
                          const CandOut &co, unsigned long long *visited, const float *bp, const float *paa,
                          unsigned long long *skipped, unsigned int *dyn, hipStream_t stream);
hipError_t launch_mark_ea(TmplDev *tmpl, const int32_t *ids, int count, int form, hipStream_t stream);
hipError_t launch_rho_err(const Cand *cand, const float *est, const float *exact, uint64_t count,
                          const TmplDev *tmpl, unsigned int *out_bits, hipStream_t stream);
}  // namespace tmx

using namespace tmx;

#define LANE_HIP(call, what)                                                                                         \
    do {                                                                                                             \
        hipError_t e__ = (call);                                                                                     \
        if (e__ != hipSuccess)                                                                                       \
            return L.hip_fail(e__, what);                                                                            \
    } while (0)

namespace {

enum GroupPredHost { kPredAllH = 0, kPredHitH = 1, kPredFiniteH = 2, kPredQueryH = 3 };
enum GroupKeyHost { kKeyValuePosH = 0, kKeyPosValueH = 1, kKeyQueryH = 2 };

constexpr int kMaxTemplate = 65536;
constexpr int kMaxK = 4096;

// Host-side view of one batch, templates sorted by length (the device order).
struct CallPlan {
    int T = 0;
    std::vector<int> perm;          // device index -> query index
    std::vector<TmplDev> tmpl;      // device order
    std::vector<int32_t> group_begin;
    int launch = TMX_LAUNCH_GRID;   // matrix-core launch form (tmx_launch)
    int64_t raw_len = 0, order_len = 0;
    int64_t afrag_len = 0;          // 16-byte fragments
    int64_t comparisons = 0;
    double flops = 0.0;             // sum of 2 m (n - m + 1)
    int Tg = 0;                     // templates [0, Tg) take the inner-product form, [Tg, T) the FFT form
    int fft_L = 0;                  // longest FFT template
    // The FFT templates by overlap-save shape: [t0, t1) consecutive (sorted
    // by length), their spectra from qspec_off (float2 units) in the lane's
    // template-spectrum buffer.
    struct FftGroup {
        int t0, t1;
        FftShape shape;
        size_t qspec_off;
    };
    std::vector<FftGroup> fft;
    int max_want = 0;               // largest kNN bound set (k or 2k - 1)
    int max_k = 0;                  // largest k
    int64_t out_slots = 0;          // sum of k over the kNN templates
    bool has_threshold = false;
    std::vector<int32_t> scan_ids;  // templates on the early-abandoning scan (set after seeding)
    std::vector<int32_t> ix_ids;    // templates on the iSAX index (set after seeding)
    int64_t ix_max_leaves = 0;      // most leaves among their indexes
    uint64_t scan_windows = 0;      // largest window count among them
    double scan_blocks = 0.0;       // their mean sampled blocks per window
};

// Cost model (per template, seconds per window), in the role of the
// reference's adaptive choice between scan and FFT (Tremor.cpp:2265-2286,
// fft_fallback.hpp:55-60): the matrix-core form costs in proportion to the
// padded template length, plus the staging and epilogue of each waveform
// block shared by the templates of a group (at most 8, fewer in a small
// batch); the fused FFT form (B = 2^14 for templates up to 4096 samples,
// 2^15 up to 8192) a constant per window of the class; the FFT form with hipFFT transforms (longer
// templates) in proportion to B log2 B per piece of (B - Lmax + 1) windows.
// Constants measured on MI355X; FFT iff it is estimated below 0.9 x the
// inner-product form, as in the reference.
double gemm_seconds_per_window(int precision, int kpad, int group)
{
    const double per_k = precision == TMX_PREC_FP16 ? 1.64e-15 : 4.0e-15;
    return per_k * kpad + 5.7e-12 / std::max(1, std::min(group, 8));
}

// Overlap-save shape of the FFT form for a batch whose longest FFT template
// has L samples: pieces of B samples every B - Lmax + 1 windows.  The fused
// lengths take the smallest Lmax that holds L (a piece's transform costs the
// same whatever Lmax, so a shorter Lmax gives more windows per piece); longer
// templates double B from 2^16 until B >= 4 L, with Lmax = B / 4 (the
// reference's rule, fft_fallback.cpp:124-126, which starts at B = 2^15).
FftShape fft_shape_for(int L)
{
    if (L <= 1024)
        return FftShape{1 << 14, 1024};
    if (L <= 2048)
        return FftShape{1 << 14, 2048};
    if (L <= 4096)
        return FftShape{1 << 15, 4096};
    if (L <= 8192)
        return FftShape{1 << 15, 8192};
    int B = 1 << 16;
    while (B < 4 * L)
        B *= 2;
    return FftShape{B, B / 4};
}

// Fused lengths: the per-window cost measured at Lmax = B / 4 (3.05e-12 s for
// B = 2^14, 3.4e-12 for 2^15), times the ratio of windows per piece.
double fft_seconds_per_window(int L)
{
    const FftShape sh = fft_shape_for(L);
    const int B = sh.B;
    const double step = static_cast<double>(B - sh.Lmax + 1);
    const double step_quarter = static_cast<double>(B - B / 4 + 1);
    if (B == (1 << 14))
        return 3.05e-12 * step_quarter / step;
    if (B == (1 << 15))
        return 3.4e-12 * step_quarter / step;
    double lg = 0.0;
    while ((1 << static_cast<int>(lg)) < B)
        lg += 1.0;
    const double per_piece = 0.45e-12 * B * lg + 1.6e-12 * B;
    return per_piece / static_cast<double>(B - L + 1);
}

constexpr int kFusedChunk = 16;    // templates per fused FFT launch
constexpr int kFftPieceBatch = 64;  // pieces / templates per forward transform launch
constexpr uint64_t kSurvQueueMin = 1u << 22;   // level-1 survivors a lane's matrix-core pass starts with (16 B each)
constexpr uint64_t kChunkQueueMin = 1u << 20;  // chunk records a lane's fused pass starts with (280 B each)
constexpr int kFftMinLength = 160;  // shortest template the cost model sends to the FFT form

// The early-abandoning scan (a third form, chosen once the bounds are known,
// as the reference decides after its sampled early-abandon test,
// Tremor.cpp:2265-2286): a fixed cost per window (statistics) plus one per
// 32-sample block visited, the blocks measured on a sample of windows.
constexpr double kScanWindow = 9.0e-12;
constexpr double kScanBlock = 3.6e-12;
constexpr uint64_t kScanSample = 4096;

double scan_seconds_per_window(double blocks) { return kScanWindow + kScanBlock * blocks; }
// The iSAX lower bound's own cost per window: the table form (16 segments,
// measured 2.6e-12 on 1.6e9 windows of 256 samples), the direct form for
// other counts (estimated at four times that).
constexpr double kScanBound16 = 3.0e-12, kScanBound = 1.2e-11;

// The planned form's estimate for template d of the plan.
double planned_seconds_per_window(const tmx_cfg &cfg, const CallPlan &plan, int d)
{
    const TmplDev &t = plan.tmpl[d];
    if (d >= plan.Tg)
        return fft_seconds_per_window(t.m);
    if (cfg.precision == TMX_PREC_FP32)
        return 8.0e-15 * t.m + 2.0e-12;
    return gemm_seconds_per_window(cfg.precision, t.kpad, plan.Tg);
}

// Consecutive templates (length-sorted) of the matrix-core form in groups of
// G: one workgroup per (block, group).
void set_groups(CallPlan &plan, int G)
{
    plan.group_begin.clear();
    for (int d = 0; d < plan.Tg; d += G)
        plan.group_begin.push_back(d);
    plan.group_begin.push_back(plan.Tg);
}

// A template length the skip-sequential lower bound applies to: a whole
// number of PAA segments (the reference requires it of every query,
// tremor_main.cpp:391-402) and at most the waveform.
bool lb_eligible(const tmx_cfg &cfg, int m, uint64_t n)
{
    return cfg.paa_segments >= 1 && cfg.paa_segments <= kMaxSegments && m % cfg.paa_segments == 0 &&
           static_cast<uint64_t>(m) <= n;
}

bool plan_has_lb(const CallPlan &plan)
{
    for (const auto &t : plan.tmpl)
        if (t.lb_seg > 0)
            return true;
    return false;
}

// A template length the iSAX index applies to: 16 PAA segments (the index's
// keys; also the lower bound's segments when those are on), a whole number
// of them, and fewer than 2^31 windows.
// Device bytes per window while the index of one length is built (its words
// 16, keys 8, ids 4, words 16 and sums 32 in key order, and the sort's
// unsorted keys and ids with their scratch), an upper estimate.
constexpr uint64_t kIxBuildBytes = 100;

SaxWords *find_sax(tmx_ctx *ctx, int m, int segments);

bool ix_eligible(const tmx_cfg &cfg, int m, uint64_t n)
{
    return cfg.index_mode != TMX_INDEX_OFF && (!cfg.lower_bounds || cfg.paa_segments == 16) && m % 16 == 0 &&
           static_cast<uint64_t>(m) <= n && n - static_cast<uint64_t>(m) + 1 < (1ull << 31);
}

bool plan_has_ix(const CallPlan &plan)
{
    for (const auto &t : plan.tmpl)
        if (t.ix > 0)
            return true;
    return false;
}

// PAA segments of the words a template needs (0: none).
int words_segments(const TmplDev &t) { return t.lb_seg > 0 ? t.lb_seg : (t.ix > 0 ? 16 : 0); }

int build_plan(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const Query *q, int count, CallPlan &plan,
               std::vector<float> &raw_packed)
{
    plan.T = count;
    plan.perm.resize(count);
    std::iota(plan.perm.begin(), plan.perm.end(), 0);
    for (int i = 0; i < count; i++) {
        if (q[i].m <= 0 || !q[i].data)
            return L.fail("template " + std::to_string(i) + " is empty");
        if (q[i].m > kMaxTemplate)
            return L.fail("template " + std::to_string(i) + " is longer than 65536 samples");
        if (q[i].k < 0 || q[i].k > kMaxK)
            return L.fail("k must be in [1, 4096]");
        if (q[i].merge < 0)
            return L.fail("merge offset must be non-negative");
        if (q[i].k == 0 && !std::isfinite(q[i].rho))
            return L.fail("correlation threshold must be finite");
    }
    std::stable_sort(plan.perm.begin(), plan.perm.end(), [&](int a, int b) { return q[a].m < q[b].m; });
    plan.tmpl.resize(count);
    raw_packed.clear();
    int64_t ooff = 0, aoff = 0, out_off = 0;
    for (int d = 0; d < count; d++) {
        const Query &qq = q[plan.perm[d]];
        TmplDev t{};
        t.m = qq.m;
        t.raw_off = static_cast<int64_t>(raw_packed.size());
        t.q_off = t.raw_off;
        t.order_off = ooff;
        t.kpad = mfma_kpad(t.m);
        t.a_off = aoff;
        t.k = qq.k;
        t.merge = qq.merge;
        // kNN: the seed pass sets the bound; threshold: cc_to_ed of the
        // reference in FP32 (ZnormSearch.hpp:23-26).
        t.bound = qq.k > 0 ? FLT_MAX : 2.0f * static_cast<float>(t.m) * (1.0f - qq.rho);
        t.out_off = out_off;
        if (qq.k > 0) {
            out_off += qq.k;
            plan.max_k = std::max(plan.max_k, qq.k);
            plan.max_want = std::max(plan.max_want, qq.merge > 0 ? 2 * qq.k - 1 : qq.k);
        } else {
            plan.has_threshold = true;
        }
        raw_packed.insert(raw_packed.end(), qq.data, qq.data + qq.m);
        ooff += t.m / kBlock + 1;
        aoff += mfma_frag_count(t.kpad);
        if (static_cast<uint64_t>(t.m) <= ctx->n) {
            const uint64_t W = ctx->n - static_cast<uint64_t>(t.m) + 1;
            plan.comparisons += static_cast<int64_t>(W);
            plan.flops += 2.0 * static_cast<double>(t.m) * static_cast<double>(W);
        }
        plan.tmpl[d] = t;
    }
    plan.out_slots = out_off;
    plan.raw_len = static_cast<int64_t>(raw_packed.size());
    plan.order_len = ooff;
    plan.afrag_len = aoff;
    // Skip-sequential lower bounds: templates whose length is a whole number
    // of segments; their PAA at d * kMaxSegments in the lane's array, their
    // words bound once the resident words exist (bind_sax).  The index:
    // eligible lengths, in auto mode only those whose index is built or fits
    // (a length without one needs about kIxBuildBytes per window while it is
    // built: its words, keys, ids, words and sums in key order, and the
    // sort's scratch) within 80% of the device's free memory, so that a call
    // at a size the index cannot take runs on the dense forms instead of
    // failing.  The same for the skip-sequential words when the scan is only
    // a candidate form (scan auto): lengths whose words are resident or fit;
    // none when the scan is off, since nothing would read them.
    std::vector<std::pair<int, bool>> ix_fit, lb_fit;  // length -> whether it is taken
    size_t free_b = 0;
    bool have_free = false;
    uint64_t planned = 0;  // device bytes the new structures of this call need
    const auto fits_memory = [&](uint64_t need) {
        size_t total_b = 0;
        if (!have_free)
            have_free = hipMemGetInfo(&free_b, &total_b) == hipSuccess;
        const bool ok = have_free && static_cast<double>(planned + need) <= 0.8 * static_cast<double>(free_b);
        if (ok)
            planned += need;
        return ok;
    };
    const auto ix_takes = [&](int m) {
        if (cfg.index_mode != TMX_INDEX_AUTO)
            return true;
        for (const auto &f : ix_fit)
            if (f.first == m)
                return f.second;
        const SaxWords *sw = find_sax(ctx, m, 16);
        const bool fits = (sw && sw->ix.nleaves > 0 && sw->ix.leaf_size == cfg.leaf_size) ||
                          fits_memory(kIxBuildBytes * (ctx->n - static_cast<uint64_t>(m) + 1));
        ix_fit.emplace_back(m, fits);
        return fits;
    };
    const auto lb_takes = [&](int m) {
        if (cfg.scan_mode == TMX_SCAN_OFF)
            return false;
        if (cfg.scan_mode == TMX_SCAN_ALWAYS)
            return true;
        for (const auto &f : lb_fit)
            if (f.first == m)
                return f.second;
        const uint64_t seg = static_cast<uint64_t>(cfg.paa_segments);
        const bool fits = find_sax(ctx, m, cfg.paa_segments) ||
                          fits_memory(seg * (ctx->n - static_cast<uint64_t>(m) + 1));
        lb_fit.emplace_back(m, fits);
        return fits;
    };
    for (int d = 0; d < count; d++) {
        TmplDev &t = plan.tmpl[d];
        if (cfg.lower_bounds && lb_eligible(cfg, t.m, ctx->n) && lb_takes(t.m)) {
            t.lb_seg = cfg.paa_segments;
            t.lb_paa = static_cast<int64_t>(d) * kMaxSegments;
        }
        if (ix_eligible(cfg, t.m, ctx->n) && ix_takes(t.m)) {
            t.ix = 1;
            t.lb_paa = static_cast<int64_t>(d) * kMaxSegments;
        }
    }
    // Path choice: templates are sorted by length and the FFT form wins for
    // long templates, so the FFT set is a suffix of the device order.
    plan.Tg = count;
    if (cfg.precision != TMX_PREC_FP32 && cfg.fft_mode != TMX_FFT_OFF) {
        if (cfg.fft_mode == TMX_FFT_ALWAYS) {
            plan.Tg = 0;
        } else {
            // Each FFT template runs in the overlap-save shape of its own
            // length (fft_shape_for), so pays that shape's cost.
            // Templates shorter than kFftMinLength stay on the matrix cores:
            // the FFT form's 64-window chunk filter needs a common core of
            // about a hundred samples to prune (single templates at 2^31
            // samples: m = 160 takes 8.7-11.8 ms on the FFT form against
            // 13.5-16.5 on the matrix cores; m = 128 is mixed).
            while (plan.Tg > 0) {
                const TmplDev &t = plan.tmpl[plan.Tg - 1];
                if (t.m < kFftMinLength ||
                    !(fft_seconds_per_window(t.m) < 0.9 * gemm_seconds_per_window(cfg.precision, t.kpad, plan.Tg)))
                    break;
                plan.Tg--;
            }
        }
    }
    for (int d = plan.Tg; d < count; d++)
        plan.fft_L = std::max(plan.fft_L, plan.tmpl[d].m);
    plan.fft.clear();
    size_t qoff = 0;
    for (int d = plan.Tg; d < count;) {
        const FftShape sh = fft_shape_for(plan.tmpl[d].m);
        int e = d + 1;
        while (e < count && fft_shape_for(plan.tmpl[e].m).B == sh.B && fft_shape_for(plan.tmpl[e].m).Lmax == sh.Lmax)
            e++;
        plan.fft.push_back(CallPlan::FftGroup{d, e, sh, qoff});
        // r2c transforms write whole batches of kFftPieceBatch templates
        qoff += static_cast<size_t>((e - d + kFftPieceBatch - 1) / kFftPieceBatch) * kFftPieceBatch *
                static_cast<size_t>(sh.B / 2 + 1);
        d = e;
    }
    set_groups(plan, cfg.precision == TMX_PREC_FP32 ? 1 : mfma_group_size());
    return 0;
}

// ---- resident structures (built under the exclusive lock) ------------------

int block_prefix_len(const CallPlan &plan)
{
    int kmax = 0;
    for (int d = 0; d < plan.Tg; d++)
        kmax = std::max(kmax, plan.tmpl[d].kpad);
    const int nw = static_cast<int>(mfma_block_windows());
    return ((nw + kmax + 1023) / 1024) * 1024;
}

FftClass *find_fft_class(tmx_ctx *ctx, FftShape sh)
{
    for (auto &c : ctx->fft)
        if (c->B == sh.B && c->Lmax == sh.Lmax)
            return c.get();
    return nullptr;
}

SaxWords *find_sax(tmx_ctx *ctx, int m, int segments)
{
    for (auto &w : ctx->sax)
        if (w->m == m && w->segments == segments)
            return w.get();
    return nullptr;
}

bool sax_ready(tmx_ctx *ctx, const tmx_cfg &cfg, const CallPlan &plan)
{
    for (const auto &t : plan.tmpl) {
        const int seg = words_segments(t);
        if (seg > 0 && !find_sax(ctx, t.m, seg))
            return false;
        if (t.ix > 0) {
            const SaxWords *sw = find_sax(ctx, t.m, 16);
            if (!sw || sw->ix.nleaves == 0 || sw->ix.leaf_size != cfg.leaf_size)
                return false;
        }
    }
    return !(plan_has_lb(plan) || plan_has_ix(plan)) || ctx->sax_bp.p;
}

bool resident_ready(tmx_ctx *ctx, const tmx_cfg &cfg, const CallPlan &plan)
{
    if (!sax_ready(ctx, cfg, plan))
        return false;
    if (cfg.precision != TMX_PREC_FP32 && plan.Tg > 0 &&
        !(ctx->rb_n == ctx->n && ctx->rb_len >= block_prefix_len(plan)))
        return false;
    for (const auto &g : plan.fft)
        if (!find_fft_class(ctx, g.shape))
            return false;
    return true;
}

// Block-rebased prefix pairs long enough for the plan's longest template.
int build_block_prefix(tmx_ctx *ctx, Lane &L, int need)
{
    if (ctx->rb_n == ctx->n && ctx->rb_len >= need)
        return 0;
    const int nw = static_cast<int>(mfma_block_windows());
    const uint64_t blocks = (ctx->n + nw - 1) / nw;
    ctx->rb_n = 0;
    LANE_HIP(ctx->rb.reserve(sizeof(float2) * blocks * static_cast<uint64_t>(need)), "hipMalloc block prefix");
    LANE_HIP(ctx->binfo.reserve(sizeof(float4) * blocks), "hipMalloc block prefix");
    LANE_HIP(launch_block_prefix(ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n, nw, need, ctx->binfo.as<float4>(), ctx->rb.as<float2>(),
                                 L.stream),
             "block prefix");
    LANE_HIP(launch_block_absmax(ctx->x.as<float>(), ctx->n, nw, need, ctx->binfo.as<float4>(), L.stream),
             "block maxima");
    LANE_HIP(hipStreamSynchronize(L.stream), "block prefix");
    ctx->rb_len = need;
    ctx->rb_n = ctx->n;
    return 0;
}


// Piece spectra of the current waveform for templates up to Lt samples, in
// the shape fft_shape_for(Lt) (B, Lmax; step = B - Lmax + 1); per piece the
// centre, the rebased prefix pairs of the level-1 filter (hipFFT lengths),
// ||x~||_2 and ||X||_inf.
// e^{2 pi i num / den}, rounded from FP64.
float2 unit_root(int64_t num, int64_t den)
{
    const double a = 2.0 * M_PI * static_cast<double>(num) / static_cast<double>(den);
    return make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
}

// The FFT forms' table for length B (M = B / 2): E[t] = e^{i pi t / M},
// t < B; for B >= 2^16 followed by the long-template form's twiddles
// (fft_large.hip): w_256^t (t < 256), w_M^t (t < 512), w_M^{512 t}
// (t < M / 512), w_{N2}^t (t < N2 = M / 256).
std::vector<float2> fft_table(int B)
{
    const int64_t M = B / 2;
    std::vector<float2> tab;
    tab.reserve(static_cast<size_t>(B) + 256 + 512 + M / 512 + M / 256);
    for (int64_t t = 0; t < B; t++)
        tab.push_back(unit_root(t, 2 * M));
    if (B >= (1 << 16)) {
        for (int64_t t = 0; t < 256; t++)
            tab.push_back(unit_root(t, 256));
        for (int64_t t = 0; t < 512; t++)
            tab.push_back(unit_root(t, M));
        for (int64_t t = 0; t < M / 512; t++)
            tab.push_back(unit_root(512 * t, M));
        for (int64_t t = 0; t < M / 256; t++)
            tab.push_back(unit_root(t, M / 256));
    }
    return tab;
}

// The forward transform the FFT forms run (hipFFT), made in one place so that
// the diagnostic probe (tmx_fft_probe) measures exactly this plan: the r2c
// forward of kFftPieceBatch real pieces (or templates) of B samples, out of
// place.
hipfftResult make_r2c_plan(hipfftHandle *h, int B)
{
    int nn = B;
    return hipfftPlanMany(h, 1, &nn, nullptr, 1, B, nullptr, 1, B / 2 + 1, HIPFFT_R2C, kFftPieceBatch);
}

// (piece, template) pairs per launch pair of the long-template form: their
// intermediates (B / 2 complex values each) within 128 MB, so that they stay
// in the Infinity Cache between the two kernels.
int inverse_batch(int B)
{
    return static_cast<int>(std::max<int64_t>(16, std::min<int64_t>(1024, (128ll << 20) / (4ll * B))));
}

int build_fft_class(tmx_ctx *ctx, Lane &L, int Lt)
{
    const FftShape sh = fft_shape_for(Lt);
    if (find_fft_class(ctx, sh))
        return 0;
    const int B = sh.B;
    auto cls = std::make_unique<FftClass>();
    FftClass &f = *cls;
    f.B = B;
    f.Lmax = sh.Lmax;
    f.step = B - f.Lmax + 1;
    f.bins = B / 2 + 1;
    f.P = (ctx->n + f.step - 1) / f.step;
    f.fused = fft_fused_supports(B);
    const uint64_t Palloc = ((f.P + kFftPieceBatch - 1) / kFftPieceBatch) * kFftPieceBatch;
    LANE_HIP(f.spectra.reserve(sizeof(float2) * Palloc * f.bins), "hipMalloc FFT spectra");
    LANE_HIP(f.pinfo.reserve(sizeof(float4) * f.P), "hipMalloc FFT pieces");
    LANE_HIP(f.pnorm.reserve(sizeof(float2) * f.P), "hipMalloc FFT pieces");
    LANE_HIP(f.norm2.reserve(sizeof(double) * f.P), "hipMalloc FFT pieces");
    if (!f.fused)
        LANE_HIP(f.rb.reserve(sizeof(float2) * f.P * static_cast<uint64_t>(B + 1)), "hipMalloc FFT prefix pairs");
    LANE_HIP(f.work_real.reserve(sizeof(float) * static_cast<size_t>(kFftPieceBatch) * B), "hipMalloc FFT input");
    {
        // E[t] = e^{i pi t / M}, M = B / 2, rounded from FP64 (the fused pass's
        // twiddles and both forms' real-inverse factors), then the
        // long-template form's compact twiddle tables.
        std::vector<float2> tab = fft_table(B);
        const int len = static_cast<int>(tab.size());
        LANE_HIP(f.table.reserve(sizeof(float2) * len), "hipMalloc FFT table");
        LANE_HIP(hipMemcpyAsync(f.table.p, tab.data(), sizeof(float2) * len, hipMemcpyHostToDevice, L.stream),
                 "copy FFT table");
        LANE_HIP(hipStreamSynchronize(L.stream), "copy FFT table");
    }
    if (make_r2c_plan(&f.fwd, B) != HIPFFT_SUCCESS)
        return L.fail("hipFFT forward plan failed", -2);
    hipfftSetStream(f.fwd, L.stream);
    LANE_HIP(launch_block_prefix(ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n, f.step, f.fused ? 0 : B + 1, f.pinfo.as<float4>(),
                                 f.rb.as<float2>(), L.stream),
             "FFT prefix pairs");
    LANE_HIP(launch_zero_f64(f.norm2.as<double>(), f.P, L.stream), "clear norms");
    for (uint64_t p0 = 0; p0 < f.P; p0 += kFftPieceBatch) {
        const int cnt = static_cast<int>(std::min<uint64_t>(kFftPieceBatch, f.P - p0));
        LANE_HIP(launch_piece_input(ctx->x.as<float>(), ctx->n, f.step, B, p0, cnt, f.pinfo.as<float4>(),
                                    f.work_real.as<float>(), f.norm2.as<double>(), L.stream),
                 "FFT piece input");
        if (hipfftExecR2C(f.fwd, f.work_real.as<hipfftReal>(),
                          reinterpret_cast<hipfftComplex *>(f.spectra.as<float2>() + p0 * f.bins)) != HIPFFT_SUCCESS)
            return L.fail("hipFFT r2c of the pieces failed", -2);
    }
    LANE_HIP(launch_piece_norm(f.norm2.as<double>(), f.pnorm.as<float2>(), f.P, L.stream), "FFT piece norms");
    LANE_HIP(launch_spec_max(f.spectra.as<float2>(), static_cast<int>(f.P), f.bins,
                             reinterpret_cast<float *>(f.pnorm.as<float2>()) + 1, 2, L.stream),
             "FFT spectral maxima");
    LANE_HIP(hipStreamSynchronize(L.stream), "FFT spectra");
    f.work_real.release();
    ctx->fft.push_back(std::move(cls));
    return 0;
}

// iSAX words for every lower-bound template length of the plan that has
// none yet (SkipSequential.cpp:155-190, on the device), keeping at most
// kSaxKept lengths: the oldest that the plan does not use are dropped first
// (no batch runs meanwhile), and a plan that uses more lengths keeps them
// all.
constexpr size_t kSaxKept = 8;

int build_sax(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const CallPlan &plan)
{
    if (!plan_has_lb(plan) && !plan_has_ix(plan))
        return 0;
    if (!ctx->sax_bp.p) {
        float bp[255];
        sax_breakpoints(bp);
        LANE_HIP(ctx->sax_bp.reserve(sizeof(bp)), "hipMalloc breakpoints");
        LANE_HIP(hipMemcpy(ctx->sax_bp.p, bp, sizeof(bp), hipMemcpyHostToDevice), "copy breakpoints");
    }
    for (const auto &t : plan.tmpl) {
        const int seg = words_segments(t);
        if (seg <= 0 || find_sax(ctx, t.m, seg))
            continue;
        while (ctx->sax.size() >= kSaxKept) {
            const auto unused = std::find_if(ctx->sax.begin(), ctx->sax.end(), [&](const auto &sw) {
                return std::none_of(plan.tmpl.begin(), plan.tmpl.end(), [&](const TmplDev &u) {
                    return u.m == sw->m && words_segments(u) == sw->segments;
                });
            });
            if (unused == ctx->sax.end())
                break;
            ctx->sax.erase(unused);
        }
        auto w = std::make_unique<SaxWords>();
        w->m = t.m;
        w->segments = seg;
        const uint64_t W = ctx->n - static_cast<uint64_t>(t.m) + 1;
        LANE_HIP(w->words.reserve(W * static_cast<uint64_t>(seg) + 16), "hipMalloc iSAX words");
        LANE_HIP(launch_sax_words(ctx->x.as<float>(), ctx->n, t.m, seg, ctx->sax_bp.as<float>(),
                                  w->words.as<uint8_t>(), L.stream),
                 "iSAX words");
        LANE_HIP(hipStreamSynchronize(L.stream), "iSAX words");
        ctx->sax.push_back(std::move(w));
    }
    // The index of every index length (built once per length and leaf size).
    for (const auto &t : plan.tmpl) {
        if (t.ix <= 0)
            continue;
        SaxWords *sw = find_sax(ctx, t.m, 16);
        if (sw->ix.nleaves > 0 && sw->ix.leaf_size == cfg.leaf_size)
            continue;
        const uint64_t W = ctx->n - static_cast<uint64_t>(t.m) + 1;
        sw->ix.nleaves = 0;
        const Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
        LANE_HIP(build_isax_index(w, t.m, sw->words.as<uint8_t>(), W, cfg.leaf_size, sw->ix, L.stream), "iSAX index");
    }
    return 0;
}

int build_resident(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const CallPlan &plan)
{
    if (const int rc = build_sax(ctx, L, cfg, plan))
        return rc;
    if (cfg.precision != TMX_PREC_FP32 && plan.Tg > 0) {
        const int rc = build_block_prefix(ctx, L, block_prefix_len(plan));
        if (rc)
            return rc;
    }
    for (const auto &g : plan.fft) {
        const int rc = build_fft_class(ctx, L, plan.tmpl[g.t1 - 1].m);
        if (rc)
            return rc;
    }
    return 0;
}

// ---- per-batch work on lane L (under the shared lock) ----------------------

// A plan cache of Lane::kPlanSlots slots keyed by FFT length: plans are made
// on first use and kept; slots fill in order and, when all are taken, the
// oldest plan is replaced.  make(&handle) creates the plan for length B.
template <typename Make>
int cached_plan(Lane &L, hipfftHandle *plans, int *keys, int B, Make make, const char *what, hipfftHandle *out)
{
    constexpr int S = Lane::kPlanSlots;
    int slot = -1;
    for (int i = 0; i < S && slot < 0; i++)
        if (plans[i] && keys[i] == B)
            slot = i;
    if (slot < 0) {
        for (int i = 0; i < S && slot < 0; i++)
            if (!plans[i])
                slot = i;
        if (slot < 0) {
            // Work enqueued with the oldest plan may still be running.
            LANE_HIP(hipStreamSynchronize(L.stream), "FFT plan eviction");
            (void)hipfftDestroy(plans[0]);
            for (int i = 0; i + 1 < S; i++) {
                plans[i] = plans[i + 1];
                keys[i] = keys[i + 1];
            }
            slot = S - 1;
            plans[slot] = 0;
        }
        keys[slot] = 0;
        if (make(&plans[slot]) != HIPFFT_SUCCESS) {
            plans[slot] = 0;
            return L.fail(what, -2);
        }
        keys[slot] = B;
    }
    hipfftSetStream(plans[slot], L.stream);
    *out = plans[slot];
    return 0;
}

// The lane's forward plan of kFftPieceBatch templates at length f.B.
int template_plan(Lane &L, const FftClass &f, hipfftHandle *out)
{
    return cached_plan(
        L, L.tfwd, L.tfwd_B, f.B,
        [&](hipfftHandle *h) { return make_r2c_plan(h, f.B); }, "hipFFT template plan failed", out);
}

// Spectra of the batch's FFT templates [Tg, T) (reversed, zero padded to the
// B of each one's shape group); slot[t] = t's index within its group, qmax
// indexed by t - Tg.
int fft_prepare(Lane &L, const CallPlan &plan, const std::vector<const FftClass *> &fcls)
{
    const int F = plan.T - plan.Tg;
    if (F <= 0)
        return 0;
    const CallPlan::FftGroup &last = plan.fft.back();
    const size_t qlen = last.qspec_off + static_cast<size_t>((last.t1 - last.t0 + kFftPieceBatch - 1) /
                                                               kFftPieceBatch) *
                                             kFftPieceBatch * static_cast<size_t>(last.shape.B / 2 + 1);
    int Bmax = 0;
    for (const auto &g : plan.fft)
        Bmax = std::max(Bmax, g.shape.B);
    LANE_HIP(L.qspec.reserve(sizeof(float2) * qlen), "hipMalloc template spectra");
    LANE_HIP(L.qmax.reserve(sizeof(float) * F), "hipMalloc template spectra");
    LANE_HIP(L.qin.reserve(sizeof(float) * static_cast<size_t>(kFftPieceBatch) * Bmax), "hipMalloc template input");
    std::vector<int32_t> ids(F), slot(plan.T, -1);
    for (const auto &g : plan.fft)
        for (int t = g.t0; t < g.t1; t++) {
            ids[t - plan.Tg] = t;
            slot[t] = t - g.t0;
        }
    LANE_HIP(L.ids.reserve(sizeof(int32_t) * F), "hipMalloc FFT ids");
    LANE_HIP(L.slot.reserve(sizeof(int32_t) * plan.T), "hipMalloc FFT ids");
    LANE_HIP(hipMemcpyAsync(L.ids.p, ids.data(), sizeof(int32_t) * F, hipMemcpyHostToDevice, L.stream), "copy ids");
    LANE_HIP(hipMemcpyAsync(L.slot.p, slot.data(), sizeof(int32_t) * plan.T, hipMemcpyHostToDevice, L.stream),
             "copy slots");
    for (size_t gi = 0; gi < plan.fft.size(); gi++) {
        const CallPlan::FftGroup &g = plan.fft[gi];
        const FftClass &f = *fcls[gi];
        hipfftHandle plan_r2c = 0;
        if (int rc = template_plan(L, f, &plan_r2c))
            return rc;
        const int Fg = g.t1 - g.t0;
        float2 *qs = L.qspec.as<float2>() + g.qspec_off;
        for (int i0 = 0; i0 < Fg; i0 += kFftPieceBatch) {
            const int cnt = std::min(kFftPieceBatch, Fg - i0);
            LANE_HIP(launch_tmpl_input(L.tmpl.as<TmplDev>(), L.ids.as<int32_t>() + (g.t0 - plan.Tg) + i0, cnt, f.B,
                                       L.qz.as<float>(), L.qin.as<float>(), L.stream),
                     "FFT template input");
            if (hipfftExecR2C(plan_r2c, L.qin.as<hipfftReal>(),
                              reinterpret_cast<hipfftComplex *>(qs + static_cast<size_t>(i0) * f.bins)) !=
                HIPFFT_SUCCESS)
                return L.fail("hipFFT r2c of the templates failed", -2);
        }
        LANE_HIP(launch_spec_max(qs, Fg, f.bins, L.qmax.as<float>() + (g.t0 - plan.Tg), 1, L.stream),
                 "template spectral maxima");
    }
    return 0;
}

// Uploads the batch's templates, prepares them on the device, packs the
// matrix-core operands and the FFT template spectra.
int upload_and_prepare(Lane &L, const tmx_cfg &cfg, const CallPlan &plan, const std::vector<float> &raw_packed,
                       const std::vector<const FftClass *> &fcls)
{
    const int T = plan.T;
    LANE_HIP(L.raw.reserve(sizeof(float) * (plan.raw_len + 64)), "hipMalloc templates");
    LANE_HIP(L.qz.reserve(sizeof(float) * (plan.raw_len + 64)), "hipMalloc templates");
    LANE_HIP(L.order.reserve(sizeof(int32_t) * (plan.order_len + 64)), "hipMalloc block orders");
    LANE_HIP(L.tmpl.reserve(sizeof(TmplDev) * T), "hipMalloc template table");
    LANE_HIP(L.groups.reserve(sizeof(int32_t) * plan.group_begin.size()), "hipMalloc groups");
    LANE_HIP(hipMemcpyAsync(L.raw.p, raw_packed.data(), sizeof(float) * plan.raw_len, hipMemcpyHostToDevice, L.stream),
             "copy templates");
    LANE_HIP(hipMemcpyAsync(L.tmpl.p, plan.tmpl.data(), sizeof(TmplDev) * T, hipMemcpyHostToDevice, L.stream),
             "copy template table");
    LANE_HIP(hipMemcpyAsync(L.groups.p, plan.group_begin.data(), sizeof(int32_t) * plan.group_begin.size(),
                            hipMemcpyHostToDevice, L.stream),
             "copy groups");
    LANE_HIP(launch_prepare(L.raw.as<float>(), L.tmpl.as<TmplDev>(), T, L.qz.as<float>(), L.order.as<int32_t>(),
                            L.stream),
             "prepare kernel");
    if (plan_has_lb(plan) || plan_has_ix(plan)) {
        LANE_HIP(L.lb_paa.reserve(sizeof(float) * kMaxSegments * static_cast<size_t>(T)), "hipMalloc template PAA");
        LANE_HIP(launch_tmpl_paa(L.tmpl.as<TmplDev>(), T, L.qz.as<float>(), L.lb_paa.as<float>(), L.stream),
                 "template PAA");
    }
    if (cfg.precision != TMX_PREC_FP32 && plan.Tg > 0) {
        LANE_HIP(L.afrag.reserve(16 * static_cast<size_t>(plan.afrag_len + 64)), "hipMalloc MFMA operands");
        int max_kpad = 0;
        for (int d = 0; d < plan.Tg; d++)
            max_kpad = std::max(max_kpad, plan.tmpl[d].kpad);
        LANE_HIP(launch_mfma_pack(L.tmpl.as<TmplDev>(), plan.Tg, max_kpad, L.qz.as<float>(), cfg.precision, L.afrag.p,
                                  L.stream),
                 "MFMA operand packing");
    }
    return fft_prepare(L, plan, fcls);
}

// fn(i) for i in [0, n) on up to 16 host threads when the work (rows to
// move) is large enough to pay for them; items must be independent.
template <class Fn>
void parallel_items(int n, size_t work, const Fn &fn)
{
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = work < (1u << 20) ? 1 : static_cast<int>(std::min<unsigned>(hw, static_cast<unsigned>(n)));
    if (nt <= 1) {
        for (int i = 0; i < n; i++)
            fn(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
        th.emplace_back([&] {
            for (int i = next++; i < n; i = next++)
                fn(i);
        });
    for (auto &t : th)
        t.join();
}

void raise_hint(std::atomic<uint64_t> &hint, uint64_t v)
{
    uint64_t cur = hint.load();
    while (cur < v && !hint.compare_exchange_weak(cur, v)) {
    }
}

int ensure_cand(Lane &L, uint64_t cap, int T)
{
    if (cap > L.cand_cap) {
        L.cand_cap = 0;
        for (DevMem *d : {&L.cand, &L.cand2})
            LANE_HIP(d->reserve(sizeof(Cand) * cap), "hipMalloc candidates");
        for (DevMem *d : {&L.cand_val, &L.cand_lo, &L.cand_hi, &L.cand_d, &L.val2, &L.lo2, &L.hi2})
            LANE_HIP(d->reserve(sizeof(float) * cap), "hipMalloc candidates");
        for (DevMem *d : {&L.keys_tmp, &L.keys})
            LANE_HIP(d->reserve(sizeof(uint64_t) * cap), "hipMalloc candidate keys");
        for (DevMem *d : {&L.tv_a, &L.tv_b})
            LANE_HIP(d->reserve(sizeof(uint32_t) * cap), "hipMalloc candidate keys");
        L.cand_cap = cap;
    }
    LANE_HIP(L.cand_count.reserve(sizeof(unsigned long long)), "hipMalloc counter");
    LANE_HIP(L.total.reserve(sizeof(unsigned long long)), "hipMalloc counter");
    LANE_HIP(L.counts.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc counts");
    LANE_HIP(L.offsets.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc offsets");
    LANE_HIP(L.cursor.reserve(sizeof(uint32_t) * (T + 1)), "hipMalloc cursors");
    LANE_HIP(L.errbits.reserve(sizeof(unsigned int)), "hipMalloc error word");
    return 0;
}

// The FFT form of a pass over every `stride`-th piece.
CandOut cand_out(Lane &L)
{
    return CandOut{L.cand.as<Cand>(), L.cand_val.as<float>(), L.cand_lo.as<float>(), L.cand_hi.as<float>(),
                   L.cand_count.as<unsigned long long>(), L.cand_cap};
}

SurvOut surv_out(Lane &L)
{
    return SurvOut{L.surv.as<SurvRec>(), L.surv_n.as<unsigned long long>(), L.surv_cap};
}

// The matrix-core pass's queue of level-1 survivors, its counter cleared.
int reserve_surv_queue(tmx_ctx *ctx, Lane &L)
{
    const uint64_t want = std::max<uint64_t>({L.surv_need, ctx->surv_hint.load(), kSurvQueueMin});
    if (want > L.surv_cap) {
        L.surv_cap = 0;
        LANE_HIP(L.surv.reserve(sizeof(SurvRec) * want), "hipMalloc survivor queue");
        L.surv_cap = want;
        if (L.surv_need > 0)
            raise_hint(ctx->surv_hint, want);
    }
    LANE_HIP(L.surv_n.reserve(sizeof(unsigned long long)), "hipMalloc survivor queue");
    LANE_HIP(launch_zero_u64(L.surv_n.as<unsigned long long>(), L.stream), "reset survivor counter");
    return 0;
}

// The chunk queue of the fused passes of one candidate pass: room for every
// launch's records (launches of at most kFusedChunk templates, each followed
// by the per-window pass that empties the queue), and the overflow counter
// nrec[1] cleared once for the whole pass.
int reserve_chunk_queue(tmx_ctx *ctx, Lane &L, const CallPlan &plan, const std::vector<const FftClass *> &fcls,
                        int mode, const std::vector<int> &strides)
{
    uint64_t want = std::max<uint64_t>({L.chunk_need, ctx->chunk_hint.load(), uint64_t{1024}});
    for (size_t gi = 0; gi < plan.fft.size(); gi++) {
        const FftClass &f = *fcls[gi];
        if (!f.fused)
            continue;
        const int Fc = std::min(plan.fft[gi].t1 - plan.fft[gi].t0, kFusedChunk);
        const uint64_t stride = static_cast<uint64_t>(strides[gi]);
        const uint64_t nwork = (f.P + stride - 1) / stride;
        // Seed passes queue at most one record per wave (16 per piece and
        // template) and must keep them all: a template without seed entries
        // would keep an infinite bound.  Main passes: room for kChunkQueueMin
        // records (or every chunk a launch can produce, if fewer), more once
        // a pass has overflowed.
        const uint64_t cw = static_cast<uint64_t>(fft_chunk_width());
        const uint64_t most = static_cast<uint64_t>(Fc) * nwork * ((f.step + cw - 1) / cw + 1);
        want = std::max(want, mode == kModeSeed ? nwork * static_cast<uint64_t>(Fc) * 16 : std::min(most, kChunkQueueMin));
    }
    if (want > L.chunk_cap) {
        L.chunk_cap = 0;
        LANE_HIP(L.chunk_rec.reserve(fft_chunk_rec_bytes() * want), "hipMalloc chunk queue");
        LANE_HIP(L.chunk_dt.reserve(sizeof(float) * fft_chunk_width() * want), "hipMalloc chunk queue");
        L.chunk_cap = want;
        if (L.chunk_need > 0)
            raise_hint(ctx->chunk_hint, want);
    }
    LANE_HIP(L.chunk_n.reserve(2 * sizeof(unsigned long long)), "hipMalloc chunk queue");
    LANE_HIP(launch_zero_u64(L.chunk_n.as<unsigned long long>() + 1, L.stream), "reset chunk counter");
    return 0;
}

// The FFT form of a pass for one shape group, over every `stride`-th piece
// (the chunk queue of the fused lengths reserved beforehand).
int run_fft_pass(tmx_ctx *ctx, Lane &L, const CallPlan &plan, const CallPlan::FftGroup &g, const FftClass &f,
                 int mode, int stride)
{
    const int F = g.t1 - g.t0;
    bool any = false;
    for (int t = g.t0; t < g.t1 && !any; t++)
        any = static_cast<uint64_t>(plan.tmpl[t].m) <= ctx->n && !plan.tmpl[t].ea;
    if (!any)
        return 0;
    Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
    const double kfft = fft_error_factor(f.B);
    const float2 *qs = L.qspec.as<float2>() + g.qspec_off;
    float *qmax = L.qmax.as<float>() + (g.t0 - plan.Tg);
    if (f.fused) {
        // Launches of at most kFusedChunk templates: their spectra (8 (B/2 + 1)
        // bytes each) stay in L2 while the pieces stream past.
        for (int c0 = 0; c0 < F; c0 += kFusedChunk) {
            const int Fc = std::min(kFusedChunk, F - c0);
            LANE_HIP(launch_fft_fused(w, L.tmpl.as<TmplDev>(), g.t0 + c0, Fc, f.spectra.as<float2>(), f.bins,
                                      qs + static_cast<size_t>(c0) * f.bins, qmax + c0, f.pinfo.as<float4>(),
                                      f.pnorm.as<float2>(), f.table.as<float2>(), f.P, f.step, stride, kfft, mode,
                                      L.chunk_rec.p, L.chunk_dt.as<float>(), L.chunk_n.as<unsigned long long>(),
                                      L.chunk_cap, cand_out(L), L.cus > 0 ? L.cus : 256, L.stream),
                     "fused FFT pass");
        }
        return 0;
    }
    std::vector<FftPair> pairs;
    for (uint64_t p = 0; p < f.P; p += static_cast<uint64_t>(stride))
        for (int i = 0; i < F; i++) {
            const int t = g.t0 + i;
            const uint64_t m = static_cast<uint64_t>(plan.tmpl[t].m);
            if (m <= ctx->n && !plan.tmpl[t].ea && p * f.step < ctx->n - m + 1)
                pairs.push_back(FftPair{static_cast<uint32_t>(t), static_cast<uint32_t>(p)});
        }
    if (pairs.empty())
        return 0;
    if (!fft_large_supports(f.B))
        return L.fail("no FFT form for length " + std::to_string(f.B), -2);
    // Pairs piece-major (the templates of one piece consecutive: its
    // spectrum and prefix pairs are shared in L2); the hand-written inverse
    // and filter (fft_large.hip) in launches of NB pairs, NB from B alone.
    const int NB = inverse_batch(f.B);
    LANE_HIP(L.Y.reserve(sizeof(float2) * static_cast<size_t>(NB) * (f.B / 2)), "hipMalloc FFT intermediates");
    LANE_HIP(L.pairs.reserve(sizeof(FftPair) * pairs.size()), "hipMalloc FFT pairs");
    LANE_HIP(hipMemcpyAsync(L.pairs.p, pairs.data(), sizeof(FftPair) * pairs.size(), hipMemcpyHostToDevice, L.stream),
             "copy FFT pairs");
    for (size_t b0 = 0; b0 < pairs.size(); b0 += NB) {
        const int cnt = static_cast<int>(std::min<size_t>(NB, pairs.size() - b0));
        const FftPair *pp = L.pairs.as<FftPair>() + b0;
        LANE_HIP(launch_fft_large(w, L.tmpl.as<TmplDev>(), pp, cnt, f.spectra.as<float2>(), qs,
                                  L.slot.as<int32_t>(), f.B, f.table.as<float2>(), L.Y.as<float2>(), f.step,
                                  f.rb.as<float2>(), f.B + 1, f.pinfo.as<float4>(), f.pnorm.as<float2>(), qmax, kfft,
                                  mode, cand_out(L), L.stream),
                 "long-template FFT pass");
    }
    return 0;
}

// ---- autotuner: templates per matrix-core workgroup and launch form -------
//
// The group size trades workgroups (a larger grid) against the per-workgroup
// work shared by a group (the block's largest |x~|, its statistics reads and
// staging); the launch form is one workgroup per (block, group) item, or one
// persistent workgroup per CU claiming items from per-XCD queues (with
// stealing across XCDs at the end).  The first batch of a shape - precision,
// matrix-core template count and mean padded length, each in powers of two -
// times every (size, form) candidate on a strided sample of blocks of the
// main pass, with the batch's real bounds (after seeding), and the context
// keeps the fastest for that shape.  Answers depend on neither.

uint64_t waveform_blocks(tmx_ctx *ctx, const tmx_cfg &cfg);

constexpr int kTuneGroups[] = {4, 8, 16};
constexpr int kTuneLaunch[] = {TMX_LAUNCH_GRID, TMX_LAUNCH_QUEUE};
constexpr uint64_t kTuneBlocks = 512;  // blocks per timed sample (every S-th block)

// Workgroups of the persistent form for this plan (0: one per item).
int persistent_wgs(const Lane &L, const CallPlan &plan)
{
    return plan.launch == TMX_LAUNCH_QUEUE ? (L.cus > 0 ? L.cus : 256) : 0;
}

uint32_t tune_key(const tmx_cfg &cfg, const CallPlan &plan)
{
    double ksum = 0.0;
    for (int d = 0; d < plan.Tg; d++)
        ksum += static_cast<double>(plan.tmpl[d].kpad);
    const int kmean = static_cast<int>(ksum / std::max(plan.Tg, 1));
    uint32_t tb = 0, kb = 0;
    while ((1 << tb) < plan.Tg)
        tb++;
    while ((1 << kb) < kmean)
        kb++;
    return static_cast<uint32_t>(cfg.precision) << 16 | tb << 8 | kb;
}

int upload_groups(Lane &L, const CallPlan &plan)
{
    LANE_HIP(L.groups.reserve(sizeof(int32_t) * plan.group_begin.size()), "hipMalloc groups");
    LANE_HIP(hipMemcpyAsync(L.groups.p, plan.group_begin.data(), sizeof(int32_t) * plan.group_begin.size(),
                            hipMemcpyHostToDevice, L.stream),
             "copy groups");
    return 0;
}

int autotune_groups(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, CallPlan &plan, tmx_stats &st)
{
    LANE_HIP(L.queue.reserve(8 * sizeof(uint32_t)), "hipMalloc work queues");
    if (cfg.precision == TMX_PREC_FP32) {
        st.tuned_group = 1;
        st.tuned_launch = TMX_LAUNCH_GRID;
        return 0;
    }
    // A launch form asked for, else the grid unless tuned below.
    plan.launch = cfg.mfma_launch == TMX_LAUNCH_QUEUE ? TMX_LAUNCH_QUEUE : TMX_LAUNCH_GRID;
    st.tuned_group = mfma_group_size();
    st.tuned_launch = plan.launch;
    int inner = 0;
    for (int d = 0; d < plan.Tg; d++)
        inner += plan.tmpl[d].ea ? 0 : 1;
    // A fixed size asked for (measurements of one configuration).
    for (const int g : kTuneGroups)
        if (cfg.autotune == g) {
            st.tuned_group = g;
            set_groups(plan, g);
            return upload_groups(L, plan);
        }
    // Up to the smallest candidate every size gives one group: only the
    // launch form could differ, and one group leaves little to balance.
    if (!cfg.autotune || inner <= kTuneGroups[0])
        return 0;
    const uint32_t key = tune_key(cfg, plan);
    int G = 0, F = 0;
    {
        std::lock_guard<std::mutex> lk(ctx->tune_mu);
        const auto it = ctx->tuned_group.find(key);
        if (it != ctx->tuned_group.end()) {
            G = it->second & 0xff;
            F = it->second >> 8;
        }
    }
    if (G == 0) {
        const double t0 = now_s();
        Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
        const uint64_t blocks = waveform_blocks(ctx, cfg);
        const int S = static_cast<int>(std::max<uint64_t>(1, blocks / kTuneBlocks));
        float best = 0.0f;
        for (const int g : kTuneGroups) {
            set_groups(plan, g);
            if (int rc = upload_groups(L, plan))
                return rc;
            const int groups = static_cast<int>(plan.group_begin.size()) - 1;
            for (const int f : kTuneLaunch) {
                if (cfg.mfma_launch != TMX_LAUNCH_AUTO && f != cfg.mfma_launch)
                    continue;
                plan.launch = f;
                LANE_HIP(launch_zero_u64(L.cand_count.as<unsigned long long>(), L.stream), "reset counter");
                if (int rc = reserve_surv_queue(ctx, L))
                    return rc;
                LANE_HIP(launch_mfma_consts(L.tmpl.as<TmplDev>(), plan.Tg, cfg.precision, L.stream),
                         "level-1 constants");
                LANE_HIP(hipEventRecord(L.ev0, L.stream), "event");
                LANE_HIP(launch_scan_mfma(w, L.tmpl.as<TmplDev>(), plan.T, L.groups.as<int32_t>(), groups,
                                          cfg.precision, L.afrag.p, ctx->rb.as<float2>(), ctx->binfo.as<float4>(),
                                          ctx->rb_len, kModeMain, S, surv_out(L), cand_out(L),
                                          persistent_wgs(L, plan), L.queue.as<uint32_t>(), L.stream),
                         "MFMA scan (tuning)");
                LANE_HIP(hipEventRecord(L.ev1, L.stream), "event");
                LANE_HIP(hipEventSynchronize(L.ev1), "tuning");
                float ms = 0.0f;
                LANE_HIP(hipEventElapsedTime(&ms, L.ev0, L.ev1), "event time");
                if (G == 0 || ms < best) {
                    best = ms;
                    G = g;
                    F = f;
                }
            }
        }
        std::lock_guard<std::mutex> lk(ctx->tune_mu);
        ctx->tuned_group[key] = G | F << 8;
        st.seconds_tune = now_s() - t0;
    }
    st.tuned_group = G;
    st.tuned_launch = F;
    plan.launch = F;
    set_groups(plan, G);
    return upload_groups(L, plan);
}

// Piece stride of the FFT form's seed pass.  The default target (1024
// entries) is raised for large k on long waveforms, where it samples too few
// pieces for a tight k-th bound: up to one piece in 150 and 48 entries per
// answer the bound needs (at most 16384).  At 2^31 samples and k = 100 this
// cut the main pass's candidates from 1.9e5 to 4e4 and the request's time by
// ~1 ms for 0.1 ms more seeding.
int fft_seed_stride(const FftClass *f, uint64_t target, int max_want, bool default_target)
{
    if (!f || f->P == 0)
        return 1;
    const uint64_t chunks = (static_cast<uint64_t>(f->step) + 4095) / 4096;
    if (default_target) {
        const uint64_t dense = (f->P * chunks) / 150;
        const uint64_t by_k = 48ull * static_cast<uint64_t>(std::max(max_want, 0));
        target = std::max<uint64_t>(target, std::min<uint64_t>(16384, std::min(dense, by_k)));
    }
    const uint64_t s = (f->P * chunks) / std::max<uint64_t>(1, target);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(s, 1u << 30)));
}

// One candidate pass (seed or main) over every template; returns the entry
// count through *count (may exceed the capacity).
int run_pass(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const CallPlan &plan, const std::vector<const FftClass *> &fcls,
             int mode, int block_stride, uint64_t *count, const std::vector<int> *fft_strides = nullptr)
{
    Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
    LANE_HIP(launch_zero_u64(L.cand_count.as<unsigned long long>(), L.stream), "reset counter");
    const int groups = static_cast<int>(plan.group_begin.size()) - 1;
    bool inner = false, fft_any = false, mfma = false;
    for (int d = 0; d < plan.T; d++)
        (d < plan.Tg ? inner : fft_any) |= !plan.tmpl[d].ea;
    LANE_HIP(hipEventRecord(L.ev0, L.stream), "event");
    if (groups > 0 && inner) {
        if (cfg.precision == TMX_PREC_FP32) {
            LANE_HIP(launch_scan_fp32(w, L.tmpl.as<TmplDev>(), L.qz.as<float>(), plan.T, L.groups.as<int32_t>(),
                                      groups, mode, block_stride, cand_out(L), L.stream),
                     "FP32 scan");
        } else {
            if (int rc = reserve_surv_queue(ctx, L))
                return rc;
            LANE_HIP(launch_mfma_consts(L.tmpl.as<TmplDev>(), plan.Tg, cfg.precision, L.stream), "level-1 constants");
            LANE_HIP(launch_scan_mfma(w, L.tmpl.as<TmplDev>(), plan.T, L.groups.as<int32_t>(), groups, cfg.precision,
                                      L.afrag.p, ctx->rb.as<float2>(), ctx->binfo.as<float4>(), ctx->rb_len, mode,
                                      block_stride, surv_out(L), cand_out(L), persistent_wgs(L, plan),
                                      L.queue.as<uint32_t>(), L.stream),
                     "MFMA scan");
            if (mode == kModeMain)
                LANE_HIP(launch_mfma_level2(w, L.tmpl.as<TmplDev>(), cfg.precision, ctx->binfo.as<float4>(),
                                            surv_out(L), cand_out(L), L.cus, L.stream),
                         "MFMA level 2");
            mfma = true;
        }
    }
    bool fused = false;
    if (fft_any) {
        const std::vector<int> ones(plan.fft.size(), 1);
        const std::vector<int> &strides = fft_strides ? *fft_strides : ones;
        for (const FftClass *f : fcls)
            fused |= f->fused;
        if (fused) {
            const int rc = reserve_chunk_queue(ctx, L, plan, fcls, mode, strides);
            if (rc)
                return rc;
        }
        for (size_t gi = 0; gi < plan.fft.size(); gi++) {
            const int rc = run_fft_pass(ctx, L, plan, plan.fft[gi], *fcls[gi], mode, strides[gi]);
            if (rc)
                return rc;
        }
    }
    const bool lb_count = mode == kModeMain && !plan.scan_ids.empty() && plan_has_lb(plan);
    if (lb_count) {
        LANE_HIP(L.lb_skip.reserve(sizeof(unsigned long long)), "hipMalloc skip counter");
        LANE_HIP(launch_zero_u64(L.lb_skip.as<unsigned long long>(), L.stream), "reset skip counter");
    }
    if (mode == kModeMain && (!plan.scan_ids.empty() || !plan.ix_ids.empty()))
        LANE_HIP(L.ix_dyn.reserve(sizeof(unsigned int) * static_cast<size_t>(plan.T)), "hipMalloc running bounds");
    const bool scan_k1 = std::any_of(plan.scan_ids.begin(), plan.scan_ids.end(),
                                     [&](int32_t d) { return plan.tmpl[d].k == 1; });
    if (mode == kModeMain && !plan.scan_ids.empty())
        LANE_HIP(launch_scan_ea(w, L.tmpl.as<TmplDev>(), plan.T, L.qz.as<float>(), L.order.as<int32_t>(),
                                L.scan_ids.as<int32_t>(), static_cast<int>(plan.scan_ids.size()), plan.scan_windows,
                                0, 0, cand_out(L), nullptr, ctx->sax_bp.as<float>(), L.lb_paa.as<float>(),
                                lb_count ? L.lb_skip.as<unsigned long long>() : nullptr,
                                scan_k1 ? L.ix_dyn.as<unsigned int>() : nullptr, L.stream),
                 "early-abandoning scan");
    const bool ix = mode == kModeMain && !plan.ix_ids.empty();
    if (ix) {
        LANE_HIP(launch_zero_f64(L.ix_passed.as<double>(), static_cast<uint64_t>(plan.T), L.stream),
                 "clear index counts");
        LANE_HIP(launch_ix_search(w, L.tmpl.as<TmplDev>(), L.ix_ids.as<int32_t>(), static_cast<int>(plan.ix_ids.size()),
                                  plan.ix_max_leaves, L.qz.as<float>(), L.order.as<int32_t>(),
                                  ctx->sax_bp.as<float>(), L.lb_paa.as<float>(),
                                  L.ix_passed.as<unsigned long long>(),
                                  std::any_of(plan.ix_ids.begin(), plan.ix_ids.end(),
                                              [&](int32_t d) { return plan.tmpl[d].k == 1; })
                                      ? L.ix_dyn.as<unsigned int>()
                                      : nullptr,
                                  cand_out(L), L.cus, L.stream),
                 "iSAX index search");
    }
    LANE_HIP(hipEventRecord(L.ev1, L.stream), "event");
    unsigned long long c = 0, nchunk = 0, nsurv = 0, nskip = 0;
    LANE_HIP(hipMemcpyAsync(&c, L.cand_count.p, sizeof(c), hipMemcpyDeviceToHost, L.stream), "read counter");
    if (lb_count)
        LANE_HIP(hipMemcpyAsync(&nskip, L.lb_skip.p, sizeof(nskip), hipMemcpyDeviceToHost, L.stream),
                 "read skip counter");
    std::vector<unsigned long long> ixp(ix ? plan.T : 0);
    if (ix)
        LANE_HIP(hipMemcpyAsync(ixp.data(), L.ix_passed.p, sizeof(unsigned long long) * plan.T, hipMemcpyDeviceToHost,
                                L.stream),
                 "read index counts");
    if (mfma && mode == kModeMain)
        LANE_HIP(hipMemcpyAsync(&nsurv, L.surv_n.p, sizeof(nsurv), hipMemcpyDeviceToHost, L.stream),
                 "read survivor counter");
    if (fused)
        LANE_HIP(hipMemcpyAsync(&nchunk, L.chunk_n.as<unsigned long long>() + 1, sizeof(nchunk),
                                hipMemcpyDeviceToHost, L.stream),
                 "read chunk counter");
    LANE_HIP(hipStreamSynchronize(L.stream), "candidate pass");
    // A main pass whose chunk queue overflowed lost candidates: rerun.
    L.chunk_need = (fused && mode == kModeMain && nchunk > L.chunk_cap) ? nchunk + (nchunk >> 3) + 1024 : 0;
    // Likewise a matrix-core pass whose survivor queue overflowed.
    L.surv_need = nsurv > L.surv_cap ? nsurv + (nsurv >> 3) + 1024 : 0;
    L.last_survivors = nsurv;
    if (mode == kModeMain) {
        L.last_lb_skipped = nskip;
        L.last_ix_windows = 0;
        for (unsigned long long v : ixp)
            L.last_ix_windows += v;
    }
    LANE_HIP(hipEventElapsedTime(&L.last_pass_ms, L.ev0, L.ev1), "event time");
    *count = c;
    return 0;
}

// Most candidates one pass may hold: the configured limit, at most 2^31 - 1
// (the grouping's 32-bit offsets).
uint64_t cand_limit(const tmx_cfg &cfg)
{
    constexpr uint64_t hard = (1ull << 31) - 1;
    return cfg.max_candidates > 0 ? std::min<uint64_t>(static_cast<uint64_t>(cfg.max_candidates), hard) : hard;
}

// Device bytes per candidate slot across the lane's candidate buffers
// (two candidate sets, their estimates and bounds, the exact distances, the
// sort keys and values), with room for the sort scratch.
constexpr uint64_t kCandSlotBytes = 96;

// Main pass with buffer growth on overflow.  More candidates than one pass
// may hold (the limit, or what free device memory allows) returns
// kRcSplit: the caller splits the batch's templates.
int run_main(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const CallPlan &plan,
             const std::vector<const FftClass *> &fcls, uint64_t *count, int *retries)
{
    const uint64_t limit = cand_limit(cfg);
    for (;;) {
        int rc = run_pass(ctx, L, cfg, plan, fcls, kModeMain, 1, count);
        if (rc)
            return rc;
        if (L.chunk_need > 0 || L.surv_need > 0) {
            (*retries)++;
            continue;  // the next pass reserves the larger queues
        }
        if (*count > limit)
            return L.fail("more than " + std::to_string(limit) + " candidate windows in one pass", kRcSplit);
        if (*count <= L.cand_cap)
            return 0;
        (*retries)++;
        const uint64_t cap = std::min(*count + (*count >> 3) + 1024, limit);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
            (cap - L.cand_cap) * kCandSlotBytes > static_cast<uint64_t>(free_b) / 10 * 9)
            return L.fail("candidate buffers for " + std::to_string(*count) + " windows exceed free device memory",
                          kRcSplit);
        rc = ensure_cand(L, cap, plan.T);
        if (rc)
            return rc;
        raise_hint(ctx->cand_hint, cap);
    }
}

void swap_mem(DevMem &a, DevMem &b)
{
    std::swap(a.p, b.p);
    std::swap(a.bytes, b.bytes);
}

int group_sort(Lane &L, const float *val, uint64_t count, int T, int pred, int keykind)
{
    const size_t tmp = group_sort_tmp_bytes(count, T);
    LANE_HIP(L.sort_tmp.reserve(tmp + 256), "hipMalloc sort scratch");
    LANE_HIP(launch_group_sort(L.cand.as<Cand>(), val, count, L.tmpl.as<TmplDev>(), T, pred, keykind,
                               L.counts.as<uint32_t>(), L.offsets.as<uint32_t>(), L.cursor.as<uint32_t>(),
                               L.keys_tmp.as<uint64_t>(), L.keys.as<uint64_t>(), L.tv_a.as<uint32_t>(),
                               L.tv_b.as<uint32_t>(), L.total.as<unsigned long long>(), L.sort_tmp.p,
                               L.sort_tmp.bytes, L.stream),
             "group/sort");
    return 0;
}

int verify_and_err(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, uint64_t count, tmx_stats *st)
{
    LANE_HIP(launch_verify(ctx->x.as<float>(), ctx->n, L.tmpl.as<TmplDev>(), L.qz.as<float>(), L.order.as<int32_t>(),
                           L.cand.as<Cand>(), count, L.cand_d.as<float>(), cfg.verify_seq != 0, L.stream),
             "verify kernel");
    LANE_HIP(launch_rho_err(L.cand.as<Cand>(), L.cand_val.as<float>(), L.cand_d.as<float>(), count,
                            L.tmpl.as<TmplDev>(), L.errbits.as<unsigned int>(), L.stream),
             "error reduction");
    unsigned int bits = 0;
    LANE_HIP(hipMemcpyAsync(&bits, L.errbits.p, sizeof(bits), hipMemcpyDeviceToHost, L.stream), "read error");
    LANE_HIP(hipStreamSynchronize(L.stream), "verify");
    float e;
    std::memcpy(&e, &bits, sizeof(e));
    st->max_abs_rho_err = e;
    return 0;
}

uint64_t waveform_blocks(tmx_ctx *ctx, const tmx_cfg &cfg)
{
    const uint64_t wb = cfg.precision == TMX_PREC_FP32 ? static_cast<uint64_t>(kScanWindows) : mfma_block_windows();
    return (ctx->n + wb - 1) / wb;
}

// Which templates take the iSAX index (TREMOR's adaptive refinement,
// Tremor.cpp:2236-2262): with the bounds known, every eligible template for
// which at most the reference's fraction f = 0.01 of a strided sample of 2^16
// windows have word bounds within the template's bound (sample_candidates,
// Tremor.cpp:2045-2059; more than that and the reference scans instead), or
// every eligible one (TMX_INDEX_ALWAYS).  Marks them on the device (ea = 2)
// and in the plan.
constexpr double kIndexFraction = 0.01;
constexpr int kIndexSample = 1 << 16;
constexpr uint64_t kIndexHome = 4096;  // windows of the approximate search per kNN template
// The index's own cost on this GPU: seconds per window of a passing leaf (its
// word bound), per window whose word bound passes too (the early-abandoning
// window kernel), and per leaf (one thread's region bound, ~100
// instructions).  Measured on the exact search of 16 templates of 256 samples
// at n = 10^8: threshold rho >= 0.95, 9.89e8 windows in passing leaves and
// 6.0e7 passing their words, 8.2 ms (8.3e-12 per window all told); kNN k = 1
// with its running bound, 1.74e8 windows in passing leaves, 6.6 ms (then
// 3.8e-11 per word that passes, counting the 1.35e8 that pass the seeded
// bound; 7.6e-11 before the running bound, the figure kept for other
// templates, whose bounds stay as seeded).  Auto
// mode also asks that the index beat 0.9 x the planned form, as the scan must
// (the reference's rule is set for its CPU costs: on the GPU the dense forms
// cost picoseconds per window).
constexpr double kIndexPerWindow = 8.5e-12, kIndexPerLeaf = 1e-12;
constexpr double kIndexPerSurvivorK1 = 4.0e-11, kIndexPerSurvivor = 7.6e-11;

int choose_index(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, CallPlan &plan, tmx_stats &st)
{
    plan.ix_ids.clear();
    plan.ix_max_leaves = 0;
    if (cfg.index_mode == TMX_INDEX_OFF)
        return 0;
    std::vector<int32_t> ids;
    for (int d = 0; d < plan.T; d++)
        if (plan.tmpl[d].ix > 0)
            ids.push_back(d);
    if (ids.empty())
        return 0;
    LANE_HIP(L.ix_ids.reserve(sizeof(int32_t) * ids.size()), "hipMalloc index ids");
    LANE_HIP(hipMemcpyAsync(L.ix_ids.p, ids.data(), sizeof(int32_t) * ids.size(), hipMemcpyHostToDevice, L.stream),
             "copy index ids");
    const size_t pb = sizeof(unsigned long long) * static_cast<size_t>(plan.T);
    LANE_HIP(L.ix_passed.reserve(pb), "hipMalloc index counts");
    LANE_HIP(launch_zero_f64(L.ix_passed.as<double>(), static_cast<uint64_t>(plan.T), L.stream), "clear index counts");
    std::vector<unsigned long long> passed(plan.T, 0ull);
    if (cfg.index_mode != TMX_INDEX_ALWAYS) {
        LANE_HIP(launch_ix_sample(L.tmpl.as<TmplDev>(), L.ix_ids.as<int32_t>(), static_cast<int>(ids.size()),
                                  ctx->sax_bp.as<float>(), L.lb_paa.as<float>(), ctx->n, kIndexSample,
                                  L.ix_passed.as<unsigned long long>(), L.stream),
                 "index sample");
        LANE_HIP(hipMemcpyAsync(passed.data(), L.ix_passed.p, pb, hipMemcpyDeviceToHost, L.stream),
                 "read index sample");
        LANE_HIP(hipStreamSynchronize(L.stream), "index sample");
    }
    std::vector<int32_t> chosen;
    std::vector<double> sampled(plan.T, 0.0);  // fraction of the sample whose word bounds pass
    for (int d : ids) {
        const uint64_t W = ctx->n - static_cast<uint64_t>(plan.tmpl[d].m) + 1;
        const double samples = static_cast<double>(std::min<uint64_t>(kIndexSample, W));
        sampled[d] = static_cast<double>(passed[d]) / samples;
        if (cfg.index_mode == TMX_INDEX_ALWAYS || static_cast<double>(passed[d]) <= kIndexFraction * samples)
            chosen.push_back(d);
    }
    if (cfg.index_mode == TMX_INDEX_AUTO && !chosen.empty()) {
        // The windows in the passing leaves of the templates the rule kept.
        int64_t max_leaves = 0;
        for (int d : chosen)
            max_leaves = std::max<int64_t>(max_leaves, plan.tmpl[d].ix_nleaves);
        LANE_HIP(hipMemcpyAsync(L.ix_ids.p, chosen.data(), sizeof(int32_t) * chosen.size(), hipMemcpyHostToDevice,
                                L.stream),
                 "copy index ids");
        LANE_HIP(launch_zero_f64(L.ix_passed.as<double>(), static_cast<uint64_t>(plan.T), L.stream),
                 "clear index counts");
        LANE_HIP(launch_ix_count(L.tmpl.as<TmplDev>(), L.ix_ids.as<int32_t>(), static_cast<int>(chosen.size()),
                                 max_leaves, ctx->sax_bp.as<float>(), L.lb_paa.as<float>(),
                                 L.ix_passed.as<unsigned long long>(), L.stream),
                 "index leaf count");
        LANE_HIP(hipMemcpyAsync(passed.data(), L.ix_passed.p, pb, hipMemcpyDeviceToHost, L.stream),
                 "read index counts");
        LANE_HIP(hipStreamSynchronize(L.stream), "index leaf count");
        std::vector<int32_t> kept;
        for (int d : chosen) {
            const double W = static_cast<double>(ctx->n - static_cast<uint64_t>(plan.tmpl[d].m) + 1);
            const double per_survivor = plan.tmpl[d].k == 1 ? kIndexPerSurvivorK1 : kIndexPerSurvivor;
            const double ix_s = static_cast<double>(passed[d]) * kIndexPerWindow + sampled[d] * W * per_survivor +
                                static_cast<double>(plan.tmpl[d].ix_nleaves) * kIndexPerLeaf;
            if (ix_s < 0.9 * planned_seconds_per_window(cfg, plan, d) * W)
                kept.push_back(d);
        }
        chosen.swap(kept);
    }
    if (chosen.empty())
        return 0;
    LANE_HIP(hipMemcpyAsync(L.ix_ids.p, chosen.data(), sizeof(int32_t) * chosen.size(), hipMemcpyHostToDevice,
                            L.stream),
             "copy index ids");
    LANE_HIP(launch_mark_ea(L.tmpl.as<TmplDev>(), L.ix_ids.as<int32_t>(), static_cast<int>(chosen.size()), 2,
                            L.stream),
             "mark index templates");
    for (int d : chosen) {
        plan.tmpl[d].ea = 2;
        plan.ix_max_leaves = std::max<int64_t>(plan.ix_max_leaves, plan.tmpl[d].ix_nleaves);
    }
    plan.ix_ids = chosen;
    st.index_templates = static_cast<int32_t>(chosen.size());
    // The approximate search first (the reference visits the query's own
    // leaf before the others): the kNN templates' nearest windows in key
    // order, exactly verified; the k-th (2k-1-th separated) smallest of their
    // distances can only tighten each bound, and the exact search after it
    // then visits fewer leaves and windows.
    std::vector<int32_t> knn;
    for (int d : chosen)
        if (plan.tmpl[d].k > 0)
            knn.push_back(d);
    if (!knn.empty()) {
        const int per = static_cast<int>(std::min<uint64_t>(kIndexHome, L.cand_cap / knn.size()));
        LANE_HIP(L.scan_ids.reserve(sizeof(int32_t) * knn.size()), "hipMalloc index ids");
        LANE_HIP(hipMemcpyAsync(L.scan_ids.p, knn.data(), sizeof(int32_t) * knn.size(), hipMemcpyHostToDevice,
                                L.stream),
                 "copy index ids");
        LANE_HIP(launch_zero_u64(L.cand_count.as<unsigned long long>(), L.stream), "reset counter");
        Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
        LANE_HIP(launch_ix_home(w, L.tmpl.as<TmplDev>(), L.scan_ids.as<int32_t>(), static_cast<int>(knn.size()),
                                ctx->sax_bp.as<float>(), L.lb_paa.as<float>(), per, cand_out(L), L.stream),
                 "index approximate search");
        unsigned long long nh = 0;
        LANE_HIP(hipMemcpyAsync(&nh, L.cand_count.p, sizeof(nh), hipMemcpyDeviceToHost, L.stream), "read counter");
        LANE_HIP(hipStreamSynchronize(L.stream), "index approximate search");
        nh = std::min<unsigned long long>(nh, L.cand_cap);
        if (nh > 0) {
            LANE_HIP(launch_verify(ctx->x.as<float>(), ctx->n, L.tmpl.as<TmplDev>(), L.qz.as<float>(),
                                   L.order.as<int32_t>(), L.cand.as<Cand>(), nh, L.cand_d.as<float>(),
                                   cfg.verify_seq != 0, L.stream),
                     "verify kernel");
            if (int rc = group_sort(L, L.cand_d.as<float>(), nh, plan.T, kPredFiniteH, kKeyValuePosH))
                return rc;
            LANE_HIP(launch_seed_bound(L.keys.as<uint64_t>(), L.offsets.as<uint32_t>(), L.tmpl.as<TmplDev>(), plan.T,
                                       plan.max_want, L.stream),
                     "index bound");
            LANE_HIP(hipStreamSynchronize(L.stream), "index bound");
        }
    }
    return 0;
}

// Which templates take the early-abandoning scan: every one (TMX_SCAN_ALWAYS)
// or, with the bounds known, those whose sampled scan estimate is not beaten
// by 10% by the planned form (the reference's rule, Tremor.cpp:2286: the FFT
// only when fft < 0.9 scan).  Marks them on the device and in the plan.
int choose_scan(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, CallPlan &plan)
{
    plan.scan_ids.clear();
    plan.scan_windows = 0;
    plan.scan_blocks = 0.0;
    if (cfg.scan_mode == TMX_SCAN_OFF)
        return 0;
    std::vector<int32_t> ids;
    for (int d = 0; d < plan.T; d++) {
        if (static_cast<uint64_t>(plan.tmpl[d].m) > ctx->n || plan.tmpl[d].ea)
            continue;
        // The scan visits at least one block per window, unless a lower
        // bound skips windows before any block.
        if (cfg.scan_mode == TMX_SCAN_ALWAYS || plan.tmpl[d].lb_seg > 0 ||
            planned_seconds_per_window(cfg, plan, d) >= 0.9 * scan_seconds_per_window(1.0))
            ids.push_back(d);
    }
    if (ids.empty())
        return 0;
    LANE_HIP(L.scan_ids.reserve(sizeof(int32_t) * ids.size()), "hipMalloc scan ids");
    LANE_HIP(hipMemcpyAsync(L.scan_ids.p, ids.data(), sizeof(int32_t) * ids.size(), hipMemcpyHostToDevice, L.stream),
             "copy scan ids");
    std::vector<int32_t> chosen;
    double blocks_sum = 0.0;
    {
        const size_t vb = sizeof(unsigned long long) * 2 * static_cast<size_t>(plan.T);
        LANE_HIP(L.scan_visit.reserve(vb), "hipMalloc scan sample");
        LANE_HIP(launch_zero_f64(L.scan_visit.as<double>(), 2 * static_cast<uint64_t>(plan.T), L.stream),
                 "clear scan sample");
        Waveform w{ctx->x.as<float>(), ctx->p.as<double2>(), ctx->pl.as<double2>(), ctx->n};
        LANE_HIP(launch_scan_ea(w, L.tmpl.as<TmplDev>(), plan.T, L.qz.as<float>(), L.order.as<int32_t>(),
                                L.scan_ids.as<int32_t>(), static_cast<int>(ids.size()), 0, 1, kScanSample,
                                cand_out(L), L.scan_visit.as<unsigned long long>(), ctx->sax_bp.as<float>(),
                                L.lb_paa.as<float>(), nullptr, nullptr, L.stream),
                 "scan sample");
        std::vector<unsigned long long> visit(2 * static_cast<size_t>(plan.T));
        LANE_HIP(hipMemcpyAsync(visit.data(), L.scan_visit.p, vb, hipMemcpyDeviceToHost, L.stream), "read scan sample");
        LANE_HIP(hipStreamSynchronize(L.stream), "scan sample");
        for (int d : ids) {
            const double nw = static_cast<double>(std::max<unsigned long long>(1, visit[plan.T + d]));
            const double blocks = static_cast<double>(visit[d]) / nw;
            const int seg = plan.tmpl[d].lb_seg;
            const double bound_cost = seg == 0 ? 0.0 : seg == 16 ? kScanBound16 : kScanBound;
            if (cfg.scan_mode == TMX_SCAN_ALWAYS ||
                !(planned_seconds_per_window(cfg, plan, d) < 0.9 * (scan_seconds_per_window(blocks) + bound_cost))) {
                chosen.push_back(d);
                blocks_sum += blocks;
            }
        }
        if (chosen.empty())
            return 0;
        LANE_HIP(hipMemcpyAsync(L.scan_ids.p, chosen.data(), sizeof(int32_t) * chosen.size(), hipMemcpyHostToDevice,
                                L.stream),
                 "copy scan ids");
    }
    plan.scan_blocks = blocks_sum / static_cast<double>(chosen.size());
    LANE_HIP(launch_mark_ea(L.tmpl.as<TmplDev>(), L.scan_ids.as<int32_t>(), static_cast<int>(chosen.size()), 1,
                            L.stream),
             "mark scan templates");
    for (int d : chosen) {
        plan.tmpl[d].ea = 1;
        plan.scan_windows = std::max<uint64_t>(plan.scan_windows, ctx->n - static_cast<uint64_t>(plan.tmpl[d].m) + 1);
    }
    plan.scan_ids = chosen;
    return 0;
}

// The batch under the shared lock (resident structures ready).
int batch_locked(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, CallPlan &plan,
                 const std::vector<float> &raw_packed, std::vector<QueryOut> &out, tmx_stats &st)
{
    const int T = plan.T;
    std::vector<const FftClass *> fcls;
    for (const auto &g : plan.fft)
        fcls.push_back(find_fft_class(ctx, g.shape));
    for (auto &t : plan.tmpl) {
        if (words_segments(t) > 0)
            t.lb_words = find_sax(ctx, t.m, words_segments(t))->words.as<uint8_t>();
        if (t.ix > 0) {
            const IxIndex &ix = find_sax(ctx, t.m, 16)->ix;
            t.ix_leaves = ix.leaves.as<IxLeaf>();
            t.ix_ids = ix.ids.as<uint32_t>();
            t.ix_keys = ix.keys.as<uint64_t>();
            t.ix_words = ix.words.as<uint8_t>();
            t.ix_sums = ix.sums.as<WindowSums>();
            t.ix_nleaves = ix.nleaves;
        }
    }
    int rc = upload_and_prepare(L, cfg, plan, raw_packed, fcls);
    if (rc)
        return rc;
    const uint64_t def_cap = cfg.cand_capacity > 0 ? static_cast<uint64_t>(cfg.cand_capacity)
                                                   : std::max<uint64_t>(1ull << 22, ctx->cand_hint.load());
    if ((rc = ensure_cand(L, def_cap, T)))
        return rc;

    // Seed (kNN templates): per-wavefront minima of the distance upper bound
    // over a strided sample of blocks, then the k-th (2k-1-th separated)
    // smallest per template.
    const double t1 = now_s();
    if (plan.max_want > 0) {
        // Seed entries per template: 512 by default for the block passes
        // (c1: 5.4 ms of seeding instead of 10 for 1024, the looser bound
        // costing 1.6 ms in the main pass and verification), 1024 as the FFT
        // form's base (its seed pass is cheap per entry).
        const uint64_t blocks = waveform_blocks(ctx, cfg);
        const uint64_t target = cfg.seed_target > 0 ? static_cast<uint64_t>(cfg.seed_target) : 1024;
        const uint64_t block_target = cfg.seed_target > 0 ? target : 512;
        // One entry per wavefront of a sampled block.
        const uint64_t per_block = cfg.precision == TMX_PREC_FP32 ? 4 : static_cast<uint64_t>(mfma_block_waves());
        uint64_t stride = (blocks * per_block) / block_target;
        if (stride < 1)
            stride = 1;
        std::vector<int> fft_strides;
        for (const FftClass *f : fcls)
            fft_strides.push_back(fft_seed_stride(f, target, plan.max_want, cfg.seed_target <= 0));
        uint64_t nseed = 0;
        if ((rc = run_pass(ctx, L, cfg, plan, fcls, kModeSeed, static_cast<int>(std::min<uint64_t>(stride, 1u << 30)),
                           &nseed, &fft_strides)))
            return rc;
        if (nseed > L.cand_cap)
            nseed = L.cand_cap;  // any subset of seed entries still gives a valid bound
        if ((rc = group_sort(L, L.cand_val.as<float>(), nseed, T, kPredFiniteH, kKeyValuePosH)))
            return rc;
        LANE_HIP(launch_seed_bound(L.keys.as<uint64_t>(), L.offsets.as<uint32_t>(), L.tmpl.as<TmplDev>(), T,
                                   plan.max_want, L.stream),
                 "seed bound");
        LANE_HIP(hipStreamSynchronize(L.stream), "seed");
    }
    if ((rc = choose_index(ctx, L, cfg, plan, st)))
        return rc;
    if ((rc = choose_scan(ctx, L, cfg, plan)))
        return rc;
    st.scan_templates = static_cast<int32_t>(plan.scan_ids.size());
    st.scan_blocks = plan.scan_blocks;
    st.seconds_seed = now_s() - t1;
    if ((rc = autotune_groups(ctx, L, cfg, plan, st)))
        return rc;
    const double t2 = now_s();

    uint64_t ncand = 0;
    if ((rc = run_main(ctx, L, cfg, plan, fcls, &ncand, &st.retries)))
        return rc;
    st.lb_skipped = static_cast<int64_t>(L.last_lb_skipped);
    st.index_windows = static_cast<int64_t>(L.last_ix_windows);
    st.kernel_ms_main = L.last_pass_ms;
    st.flops_main = plan.flops;
    st.fft_templates = 0;
    for (int d = plan.Tg; d < plan.T; d++)
        st.fft_templates += plan.tmpl[d].ea ? 0 : 1;
    st.fft_block = fcls.empty() ? 0 : fcls.back()->B;
    st.fft_lmax = fcls.empty() ? 0 : fcls.back()->Lmax;
    const double t3 = now_s();
    st.seconds_main = t3 - t2;
    st.candidates = static_cast<int64_t>(ncand);
    // Refinement (kNN templates): the candidates' own upper bounds give a
    // bound at least as tight as the seed's (the k-th, or 2k-1-th separated,
    // smallest); candidates whose lower bound exceeds it cannot be answers
    // and skip the exact recomputation.
    if (plan.max_want > 0 && ncand > 2 * static_cast<uint64_t>(plan.out_slots) + 1024) {
        if ((rc = group_sort(L, L.cand_hi.as<float>(), ncand, T, kPredFiniteH, kKeyValuePosH)))
            return rc;
        LANE_HIP(launch_seed_bound(L.keys.as<uint64_t>(), L.offsets.as<uint32_t>(), L.tmpl.as<TmplDev>(), T,
                                   plan.max_want, L.stream),
                 "refined bound");
        LANE_HIP(launch_refine(L.cand.as<Cand>(), L.cand_val.as<float>(), L.cand_lo.as<float>(),
                               L.cand_hi.as<float>(), ncand, L.tmpl.as<TmplDev>(), L.cand2.as<Cand>(),
                               L.val2.as<float>(), L.lo2.as<float>(), L.hi2.as<float>(),
                               L.total.as<unsigned long long>(), L.stream),
                 "refinement");
        unsigned long long kept = 0;
        LANE_HIP(hipMemcpyAsync(&kept, L.total.p, sizeof(kept), hipMemcpyDeviceToHost, L.stream), "read count");
        LANE_HIP(hipStreamSynchronize(L.stream), "refinement");
        swap_mem(L.cand, L.cand2);
        swap_mem(L.cand_val, L.val2);
        swap_mem(L.cand_lo, L.lo2);
        swap_mem(L.cand_hi, L.hi2);
        ncand = kept;
    }
    st.verified = static_cast<int64_t>(ncand);
    if ((rc = verify_and_err(ctx, L, cfg, ncand, &st)))
        return rc;
    const double t4 = now_s();
    st.seconds_verify = t4 - t3;

    if ((rc = group_sort(L, L.cand_d.as<float>(), ncand, T, kPredQueryH, kKeyQueryH)))
        return rc;
    std::vector<uint64_t> hp;
    std::vector<float> hd;
    if (plan.out_slots > 0) {
        const size_t slots = static_cast<size_t>(plan.out_slots);
        LANE_HIP(L.out_pos.reserve(sizeof(uint64_t) * slots), "hipMalloc results");
        LANE_HIP(L.out_d.reserve(sizeof(float) * slots), "hipMalloc results");
        LANE_HIP(launch_knn_select(L.keys.as<uint64_t>(), L.offsets.as<uint32_t>(), L.tmpl.as<TmplDev>(), T,
                                   plan.max_k, L.out_pos.as<uint64_t>(), L.out_d.as<float>(), L.stream),
                 "kNN selection");
        hp.resize(slots);
        hd.resize(slots);
        LANE_HIP(hipMemcpyAsync(hp.data(), L.out_pos.p, sizeof(uint64_t) * slots, hipMemcpyDeviceToHost, L.stream),
                 "read results");
        LANE_HIP(hipMemcpyAsync(hd.data(), L.out_d.p, sizeof(float) * slots, hipMemcpyDeviceToHost, L.stream),
                 "read results");
    }
    std::vector<uint32_t> offs;
    const uint64_t *keys = nullptr;
    if (plan.has_threshold) {
        offs.resize(T + 1);
        LANE_HIP(hipMemcpyAsync(offs.data(), L.offsets.p, sizeof(uint32_t) * (T + 1), hipMemcpyDeviceToHost, L.stream),
                 "read hit offsets");
        LANE_HIP(hipStreamSynchronize(L.stream), "hit offsets");
        if (offs[T]) {
            LANE_HIP(L.reserve_host_keys(sizeof(uint64_t) * offs[T]), "hipHostMalloc hit staging");
            LANE_HIP(hipMemcpyAsync(L.host_keys, L.keys.p, sizeof(uint64_t) * offs[T], hipMemcpyDeviceToHost,
                                    L.stream),
                     "read hits");
            keys = static_cast<const uint64_t *>(L.host_keys);
        }
    }
    LANE_HIP(hipStreamSynchronize(L.stream), "results");
    out.assign(T, QueryOut{});
    parallel_items(T, offs.empty() ? 0 : offs[T], [&](int d) {
        const TmplDev &t = plan.tmpl[d];
        QueryOut &o = out[plan.perm[d]];
        if (t.k > 0) {
            o.pos.assign(hp.begin() + t.out_off, hp.begin() + t.out_off + t.k);
            o.d.assign(hd.begin() + t.out_off, hd.begin() + t.out_off + t.k);
            return;
        }
        // Threshold hits by position; greedy merge of the reference
        // (SkipSequential.cpp:279-290): drop a hit closer than merge + 1 to
        // the last kept one.
        bool has_last = false;
        uint64_t last = 0;
        o.pos.reserve(offs[d + 1] - offs[d]);
        o.d.reserve(offs[d + 1] - offs[d]);
        for (uint32_t e = offs[d]; e < offs[d + 1]; e++) {
            const uint64_t pos = keys[e] >> 32;
            const uint32_t bits = static_cast<uint32_t>(keys[e]);
            float dist;
            std::memcpy(&dist, &bits, sizeof(dist));
            if (has_last && t.merge > 0 && pos < last + static_cast<uint64_t>(t.merge) + 1)
                continue;
            o.pos.push_back(pos);
            o.d.push_back(dist);
            last = pos;
            has_last = true;
        }
    });
    st.seconds_select = now_s() - t4;
    return 0;
}

}  // namespace

namespace tmx {

namespace {

// One batch as one candidate pass per phase.
int run_batch_once(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const Query *q, int count,
                   std::vector<QueryOut> &out, tmx_stats *stats)
{
    if (count <= 0)
        return L.fail("template count must be positive");
    if (count > 65535)
        return L.fail("at most 65535 templates per batch");
    LANE_HIP(hipSetDevice(L.device), "hipSetDevice");
    tmx_stats st{};
    const double t0 = now_s();
    std::shared_lock<std::shared_mutex> rl = read_resident(ctx);
    if (ctx->n == 0)
        return L.fail("no waveform loaded");
    CallPlan plan;
    std::vector<float> raw_packed;
    int rc = build_plan(ctx, L, cfg, q, count, plan, raw_packed);
    if (rc)
        return rc;
    // The batch itself, with the resident structures held (shared, or
    // exclusively after repeated builds).
    const auto run_held = [&]() {
        st.comparisons = plan.comparisons;
        const int brc = batch_locked(ctx, L, cfg, plan, raw_packed, out, st);
        if (brc)
            return brc;
        st.seconds_total = now_s() - t0;
        if (cfg.verbose)
            std::fprintf(stderr,
                         "tmx batch: T=%d seed %.3f ms main %.3f ms (%lld cand) verify %.3f ms select %.3f ms\n",
                         count, 1e3 * st.seconds_seed, 1e3 * st.seconds_main, static_cast<long long>(st.candidates),
                         1e3 * st.seconds_verify, 1e3 * st.seconds_select);
        if (stats)
            *stats = st;
        return 0;
    };
    int builds = 0;
    while (!resident_ready(ctx, cfg, plan)) {
        rl.unlock();
        {
            WriteResident wl(ctx);
            rc = build_resident(ctx, L, cfg, plan);
            if (rc)
                return rc;
            // Lanes whose batches use more template lengths between them
            // than the words cache keeps can drop each other's words between
            // a build and the next shared hold.  After a second build the
            // batch runs under this exclusive hold, so every batch finishes.
            if (++builds >= 2 && resident_ready(ctx, cfg, plan))
                return run_held();
        }
        rl = read_resident(ctx);
    }
    return run_held();
}

void add_stats(tmx_stats &a, const tmx_stats &b)
{
    a.seconds_total += b.seconds_total;
    a.seconds_seed += b.seconds_seed;
    a.seconds_main += b.seconds_main;
    a.seconds_verify += b.seconds_verify;
    a.seconds_select += b.seconds_select;
    a.comparisons += b.comparisons;
    a.candidates += b.candidates;
    a.retries += b.retries;
    a.max_abs_rho_err = std::max(a.max_abs_rho_err, b.max_abs_rho_err);
    a.kernel_ms_main += b.kernel_ms_main;
    a.flops_main += b.flops_main;
    a.fft_templates += b.fft_templates;
    a.fft_block = std::max(a.fft_block, b.fft_block);
    a.verified += b.verified;
    a.scan_blocks = (a.scan_blocks * a.scan_templates + b.scan_blocks * b.scan_templates) /
                    std::max(1, a.scan_templates + b.scan_templates);
    a.scan_templates += b.scan_templates;
    a.tuned_group = std::max(a.tuned_group, b.tuned_group);
    a.tuned_launch = std::max(a.tuned_launch, b.tuned_launch);
    a.seconds_tune += b.seconds_tune;
    a.fft_lmax = std::max(a.fft_lmax, b.fft_lmax);
    a.splits += b.splits;
    a.lb_skipped += b.lb_skipped;
    a.index_templates += b.index_templates;
    a.index_windows += b.index_windows;
}

}  // namespace

// A batch whose candidates exceed what one pass may hold is split into two
// halves of its templates (caller order), each run as a batch of its own,
// recursively: every template's answer is the same in any batch.
int run_batch(tmx_ctx *ctx, Lane &L, const tmx_cfg &cfg, const Query *q, int count, std::vector<QueryOut> &out,
              tmx_stats *stats)
{
    int rc = run_batch_once(ctx, L, cfg, q, count, out, stats);
    if (rc != kRcSplit)
        return rc;
    if (count == 1)
        return L.fail("one template has more candidate windows than one pass may hold (" +
                          std::to_string(cand_limit(cfg)) + "; raise max_candidates or the threshold)",
                      -4);
    const int half = count / 2;
    std::vector<QueryOut> a, b;
    tmx_stats sa{}, sb{};
    if ((rc = run_batch(ctx, L, cfg, q, half, a, &sa)))
        return rc;
    if ((rc = run_batch(ctx, L, cfg, q + half, count - half, b, &sb)))
        return rc;
    out = std::move(a);
    out.insert(out.end(), std::make_move_iterator(b.begin()), std::make_move_iterator(b.end()));
    add_stats(sa, sb);
    sa.splits += 1;
    if (stats)
        *stats = sa;
    return 0;
}

}  // namespace tmx

namespace {

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
    return 0;
}

// Most templates one batch takes (the per-template launches index templates
// by the grid's second dimension).
constexpr int kBatchTemplates = 65535;

// Runs a synchronous call on the direct lane as batches of at most
// kBatchTemplates templates in caller order (every template's answer is the
// same in any batch); errors land in ctx->err.
int direct_batch(tmx_ctx *ctx, const std::vector<Query> &q, std::vector<QueryOut> &out, tmx_stats *stats)
{
    Lane &L = ctx->direct;
    const size_t count = q.size();
    out.clear();
    out.reserve(count);
    tmx_stats total{};
    for (size_t c0 = 0; c0 < count; c0 += kBatchTemplates) {
        const int cnt = static_cast<int>(std::min<size_t>(kBatchTemplates, count - c0));
        std::vector<QueryOut> part;
        tmx_stats st{};
        const int rc = run_batch(ctx, L, ctx->cfg, q.data() + c0, cnt, part, &st);
        if (rc) {
            ctx->fail(L.err, rc);
            return rc;
        }
        out.insert(out.end(), std::make_move_iterator(part.begin()), std::make_move_iterator(part.end()));
        if (c0 == 0)
            total = st;
        else
            add_stats(total, st);
    }
    if (stats)
        *stats = total;
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
    cfg->scan_mode = TMX_SCAN_AUTO;
    cfg->verify_seq = 0;
    cfg->autotune = 1;
    cfg->mfma_launch = TMX_LAUNCH_AUTO;
    cfg->index_mode = TMX_INDEX_OFF;
    cfg->leaf_size = 2000;
    cfg->max_candidates = 0;
    cfg->lower_bounds = 0;
    cfg->paa_segments = 16;
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
        ctx->cfg.fft_mode < TMX_FFT_AUTO || ctx->cfg.fft_mode > TMX_FFT_ALWAYS ||
        ctx->cfg.mfma_launch < TMX_LAUNCH_AUTO || ctx->cfg.mfma_launch > TMX_LAUNCH_QUEUE ||
        ctx->cfg.index_mode < TMX_INDEX_OFF || ctx->cfg.index_mode > TMX_INDEX_ALWAYS ||
        (ctx->cfg.index_mode != TMX_INDEX_OFF && ctx->cfg.leaf_size < 1) ||
        (ctx->cfg.lower_bounds && (ctx->cfg.paa_segments < 1 || ctx->cfg.paa_segments > kMaxSegments))) {
        delete ctx;
        return -1;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || ctx->cfg.device < 0 || ctx->cfg.device >= ndev) {
        delete ctx;
        return -3;
    }
    if (ctx->direct.init(ctx->cfg.device) != hipSuccess) {
        delete ctx;
        return -3;
    }
    // The device address of the allocation-free entry point's scratch, once
    // per device, with the device current (Lane::init made it so).
    if (noalloc_setup(ctx->cfg.device) != hipSuccess) {
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
    server_shutdown(ctx);
    (void)hipSetDevice(ctx->cfg.device);
    ctx->direct.destroy();
    delete ctx;
}

const char *tmx_last_error(const tmx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tmx_set_fft_mode(tmx_ctx *ctx, int32_t mode)
{
    if (!ctx)
        return -1;
    if (mode < TMX_FFT_AUTO || mode > TMX_FFT_ALWAYS)
        return ctx->fail("unknown FFT mode");
    std::lock_guard<std::mutex> lk(ctx->call_mu);
    ctx->cfg.fft_mode = mode;
    return 0;
}

int tmx_set_scan_mode(tmx_ctx *ctx, int32_t mode)
{
    if (!ctx)
        return -1;
    if (mode < TMX_SCAN_AUTO || mode > TMX_SCAN_ALWAYS)
        return ctx->fail("unknown scan mode");
    std::lock_guard<std::mutex> lk(ctx->call_mu);
    ctx->cfg.scan_mode = mode;
    return 0;
}

int tmx_set_index_mode(tmx_ctx *ctx, int32_t mode, int32_t leaf_size)
{
    if (!ctx)
        return -1;
    if (mode < TMX_INDEX_OFF || mode > TMX_INDEX_ALWAYS)
        return ctx->fail("unknown index mode");
    if (leaf_size < 1)
        return ctx->fail("the leaf size must be positive");
    std::lock_guard<std::mutex> lk(ctx->call_mu);
    ctx->cfg.index_mode = mode;
    ctx->cfg.leaf_size = leaf_size;
    return 0;
}

int tmx_set_precision(tmx_ctx *ctx, int32_t precision)
{
    if (!ctx)
        return -1;
    if (precision < TMX_PREC_FP32 || precision > TMX_PREC_FP16)
        return ctx->fail("unknown precision");
    std::lock_guard<std::mutex> lk(ctx->call_mu);
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
    if (server_running(ctx))
        return ctx->fail("stop the serving loop before loading another waveform");
    std::lock_guard<std::mutex> call(ctx->call_mu);
    WriteResident wl(ctx);
    Lane &L = ctx->direct;
    auto hip_err = [&](hipError_t e, const char *what) { return ctx->fail(std::string(what) + ": " + hipGetErrorString(e), -2); };
    hipError_t e;
    if ((e = hipSetDevice(ctx->cfg.device)) != hipSuccess)
        return hip_err(e, "hipSetDevice");
    // Tail padding: inner-product tiles read past the last window.
    const size_t pad = 1 << 20;
    ctx->n = 0;
    ctx->rb_n = 0;
    ctx->fft.clear();
    ctx->sax.clear();
    const size_t wl_len = prefix_work_len(n);
    if ((e = ctx->x.reserve(sizeof(float) * (n + pad))) != hipSuccess)
        return hip_err(e, "hipMalloc waveform");
    if ((e = ctx->p.reserve(sizeof(double2) * (n + 1))) != hipSuccess)
        return hip_err(e, "hipMalloc prefix sums");
    if ((e = ctx->pl.reserve(sizeof(double2) * (n + 1))) != hipSuccess)
        return hip_err(e, "hipMalloc prefix sums");
    if ((e = ctx->prefix_work.reserve(sizeof(double) * wl_len)) != hipSuccess)
        return hip_err(e, "hipMalloc prefix scratch");
    if ((e = hipMemcpyAsync(ctx->x.p, x, sizeof(float) * n, hipMemcpyHostToDevice, L.stream)) != hipSuccess)
        return hip_err(e, "copy waveform");
    if ((e = hipMemsetAsync(ctx->x.as<float>() + n, 0, sizeof(float) * pad, L.stream)) != hipSuccess)
        return hip_err(e, "clear padding");
    if ((e = launch_prefix(ctx->x.as<float>(), n, ctx->p.as<double2>(), ctx->pl.as<double2>(),
                           ctx->prefix_work.as<double>(), wl_len,
                           L.stream)) != hipSuccess)
        return hip_err(e, "prefix sums");
    double2 total;
    if ((e = hipMemcpyAsync(&total, ctx->p.as<double2>() + n, sizeof(double2), hipMemcpyDeviceToHost, L.stream)) !=
        hipSuccess)
        return hip_err(e, "read prefix total");
    if ((e = hipStreamSynchronize(L.stream)) != hipSuccess)
        return hip_err(e, "load waveform");
    // A non-finite sample makes the FP64 sum of squares non-finite from there
    // on (finite float32 squares cannot overflow it), so the last prefix pair
    // tells whether the waveform holds one.
    if (!std::isfinite(total.x) || !std::isfinite(total.y)) {
        uint64_t i = 0;
        while (i < n && std::isfinite(x[i]))
            i++;
        return ctx->fail("waveform sample " + std::to_string(i) +
                         " is not finite (NaN or infinity): only finite samples and the missing-sample sentinel "
                         "-100000 are accepted");
    }
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
    if (k <= 0 || k > kMaxK)
        return ctx->fail("k must be in [1, 4096]");
    if (merge_offset < 0)
        return ctx->fail("merge offset must be non-negative");
    if (!out_pos || !out_d)
        return ctx->fail("null outputs");
    std::lock_guard<std::mutex> call(ctx->call_mu);
    std::vector<Query> q(count);
    for (int i = 0; i < count; i++)
        q[i] = Query{templates + offsets[i], lengths[i], k, merge_offset, 0.0f};
    std::vector<QueryOut> out;
    if ((rc = direct_batch(ctx, q, out, stats)))
        return rc;
    for (int i = 0; i < count; i++) {
        std::memcpy(out_pos + static_cast<size_t>(i) * k, out[i].pos.data(), sizeof(uint64_t) * k);
        std::memcpy(out_d + static_cast<size_t>(i) * k, out[i].d.data(), sizeof(float) * k);
    }
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
    std::lock_guard<std::mutex> call(ctx->call_mu);
    std::vector<Query> q(count);
    for (int i = 0; i < count; i++)
        q[i] = Query{templates + offsets[i], lengths[i], 0, merge_offset, rho};
    std::vector<QueryOut> res;
    if ((rc = direct_batch(ctx, q, res, stats)))
        return rc;
    std::vector<size_t> first(count + 1, 0);
    for (int i = 0; i < count; i++)
        first[i + 1] = first[i] + res[i].pos.size();
    const size_t rows = first[count];
    tmx_hits *h = static_cast<tmx_hits *>(std::calloc(1, sizeof(tmx_hits)));
    const size_t hn = rows ? rows : 1;
    h->tmpl = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * hn));
    h->counter = static_cast<int64_t *>(std::malloc(sizeof(int64_t) * hn));
    h->pos = static_cast<uint64_t *>(std::malloc(sizeof(uint64_t) * hn));
    h->d = static_cast<float *>(std::malloc(sizeof(float) * hn));
    parallel_items(count, rows, [&](int i) {
        const size_t r0 = first[i];
        const size_t ni = res[i].pos.size();
        for (size_t e = 0; e < ni; e++) {
            h->tmpl[r0 + e] = i;
            h->counter[r0 + e] = static_cast<int64_t>(e);
        }
        std::memcpy(h->pos + r0, res[i].pos.data(), sizeof(uint64_t) * ni);
        std::memcpy(h->d + r0, res[i].d.data(), sizeof(float) * ni);
    });
    h->count = static_cast<int64_t>(rows);
    *out = h;
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

namespace {

// Uploads and prepares one template on the direct lane (no candidate pass).
int prepare_one(tmx_ctx *ctx, const float *tmpl, int32_t m)
{
    Lane &L = ctx->direct;
    tmx_cfg cfg = ctx->cfg;
    cfg.precision = TMX_PREC_FP32;  // no matrix-core operands needed
    cfg.fft_mode = TMX_FFT_OFF;
    const Query q{tmpl, m, 1, 0, 0.0f};
    CallPlan plan;
    std::vector<float> raw_packed;
    int rc = build_plan(ctx, L, cfg, &q, 1, plan, raw_packed);
    if (!rc)
        rc = upload_and_prepare(L, cfg, plan, raw_packed, std::vector<const FftClass *>{});
    if (rc)
        ctx->fail(L.err, rc);
    return rc;
}

}  // namespace

int tmx_prepare_template(tmx_ctx *ctx, const float *tmpl, int32_t m, float *out_z, int32_t *out_order)
{
    if (!ctx)
        return -1;
    if (!tmpl || m <= 0 || m > kMaxTemplate || !out_z)
        return ctx->fail("bad template");
    std::lock_guard<std::mutex> call(ctx->call_mu);
    Lane &L = ctx->direct;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    int rc = prepare_one(ctx, tmpl, m);
    if (rc)
        return rc;
    if (hipMemcpyAsync(out_z, L.qz.p, sizeof(float) * m, hipMemcpyDeviceToHost, L.stream) != hipSuccess ||
        (out_order && m / kBlock > 0 &&
         hipMemcpyAsync(out_order, L.order.p, sizeof(int32_t) * (m / kBlock), hipMemcpyDeviceToHost, L.stream) !=
             hipSuccess) ||
        hipStreamSynchronize(L.stream) != hipSuccess)
        return ctx->fail("prepare template: copy failed", -2);
    return 0;
}

int tmx_distance_batch(tmx_ctx *ctx, const float *tmpl, int32_t m, const uint64_t *starts, int32_t count, float *out_d)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (!tmpl || m <= 0 || m > kMaxTemplate || !starts || !out_d || count < 0)
        return ctx->fail("bad arguments");
    if (count == 0)
        return 0;
    std::lock_guard<std::mutex> call(ctx->call_mu);
    std::shared_lock<std::shared_mutex> rl = read_resident(ctx);
    Lane &L = ctx->direct;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    int rc = prepare_one(ctx, tmpl, m);
    if (rc)
        return rc;
    DevMem dstarts, dout;
    hipError_t e = dstarts.reserve(sizeof(uint64_t) * count);
    if (e == hipSuccess)
        e = dout.reserve(sizeof(float) * count);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dstarts.p, starts, sizeof(uint64_t) * count, hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess)
        e = launch_dist_batch(ctx->x.as<float>(), ctx->n, L.tmpl.as<TmplDev>(), L.qz.as<float>(),
                              L.order.as<int32_t>(), dstarts.as<uint64_t>(), count, dout.as<float>(), L.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out_d, dout.p, sizeof(float) * count, hipMemcpyDeviceToHost, L.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess)
        return ctx->fail(std::string("distance batch: ") + hipGetErrorString(e), -2);
    return 0;
}

int tmx_distance_batch_noalloc(tmx_ctx *ctx, const float *d_tmpl, int32_t m, const uint64_t *d_starts,
                               int32_t count, float *d_out, void *stream)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (m <= 0 || m > noalloc_max_template())
        return ctx->fail("template length must be in [1, " + std::to_string(noalloc_max_template()) +
                         "] for the allocation-free distance batch");
    if (!d_tmpl || !d_starts || !d_out || count < 0)
        return ctx->fail("bad arguments");
    if (count == 0)
        return 0;
    std::lock_guard<std::mutex> call(ctx->call_mu);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->direct.stream;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    const hipError_t e = noalloc_distance_batch(ctx->cfg.device, ctx->x.as<float>(), ctx->n, d_tmpl, m, d_starts,
                                                count, d_out, s);
    if (e != hipSuccess)
        return ctx->fail(std::string("allocation-free distance batch: ") + hipGetErrorString(e), -2);
    return 0;
}

int tmx_window_stats(tmx_ctx *ctx, int32_t m, uint64_t first, uint64_t count, float *mean, float *std_out)
{
    if (!ctx)
        return -1;
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    if (m <= 0 || !mean || !std_out || static_cast<uint64_t>(m) > ctx->n)
        return ctx->fail("window range out of bounds");
    // Written so that no sum can wrap.
    const uint64_t windows = ctx->n - static_cast<uint64_t>(m) + 1;
    if (first > windows || count > windows - first)
        return ctx->fail("window range out of bounds");
    if (count == 0)
        return 0;
    std::lock_guard<std::mutex> call(ctx->call_mu);
    std::shared_lock<std::shared_mutex> rl = read_resident(ctx);
    Lane &L = ctx->direct;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    DevMem a, b;
    hipError_t e = a.reserve(sizeof(float) * count);
    if (e == hipSuccess)
        e = b.reserve(sizeof(float) * count);
    if (e == hipSuccess)
        e = launch_window_stats(ctx->x.as<float>(), m, first, count, a.as<float>(), b.as<float>(), L.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(mean, a.p, sizeof(float) * count, hipMemcpyDeviceToHost, L.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(std_out, b.p, sizeof(float) * count, hipMemcpyDeviceToHost, L.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess)
        return ctx->fail(std::string("window stats: ") + hipGetErrorString(e), -2);
    return 0;
}

int tmx_index_probe(tmx_ctx *ctx, int32_t m, int32_t leaf_size, uint32_t *ids, uint64_t ids_cap,
                    uint32_t *leaf_start, uint32_t *leaf_count, int32_t *leaf_depth, int64_t leaf_cap,
                    int64_t *nleaves, uint8_t *words)
{
    if (!ctx)
        return -1;
    if (!nleaves || (leaf_cap > 0 && (!leaf_start || !leaf_count || !leaf_depth)) || leaf_size < 1)
        return ctx->fail("index probe: bad arguments");
    std::lock_guard<std::mutex> call(ctx->call_mu);
    if (ctx->n == 0)
        return ctx->fail("no waveform loaded");
    tmx_cfg cfg = ctx->cfg;
    cfg.index_mode = TMX_INDEX_ALWAYS;
    cfg.leaf_size = leaf_size;
    if (m <= 0 || !ix_eligible(cfg, m, ctx->n))
        return ctx->fail("index probe: the length must be a multiple of 16, at most n, with fewer than 2^31 windows");
    const uint64_t W = ctx->n - static_cast<uint64_t>(m) + 1;
    if ((ids || words) && ids_cap < W)
        return ctx->fail("index probe: ids / words hold fewer than n - m + 1 entries");
    Lane &L = ctx->direct;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    CallPlan plan;
    plan.T = 1;
    plan.tmpl.assign(1, TmplDev{});
    plan.tmpl[0].m = m;
    plan.tmpl[0].ix = 1;
    WriteResident wl(ctx);
    if (const int rc = build_sax(ctx, L, cfg, plan))
        return ctx->fail(L.err, rc);
    const SaxWords *sw = find_sax(ctx, m, 16);
    const IxIndex &ix = sw->ix;
    *nleaves = ix.nleaves;
    hipError_t e = hipSuccess;
    if (ids)
        e = hipMemcpy(ids, ix.ids.p, sizeof(uint32_t) * W, hipMemcpyDeviceToHost);
    if (e == hipSuccess && words)
        e = hipMemcpy(words, sw->words.p, 16 * W, hipMemcpyDeviceToHost);
    const int64_t nl = std::min<int64_t>(leaf_cap, ix.nleaves);
    if (e == hipSuccess && nl > 0) {
        std::vector<IxLeaf> lv(static_cast<size_t>(nl));
        e = hipMemcpy(lv.data(), ix.leaves.p, sizeof(IxLeaf) * lv.size(), hipMemcpyDeviceToHost);
        for (int64_t i = 0; e == hipSuccess && i < nl; i++) {
            leaf_start[i] = lv[i].start;
            leaf_count[i] = lv[i].count;
            leaf_depth[i] = lv[i].depth;
        }
    }
    if (e != hipSuccess)
        return ctx->fail(std::string("index probe: ") + hipGetErrorString(e), -2);
    return 0;
}

int tmx_fft_probe(tmx_ctx *ctx, int32_t B, int32_t kind, const float *in, int32_t count, float *out)
{
    if (!ctx)
        return -1;
    if (B < (1 << 14) || B > (1 << 18) || (B & (B - 1)) != 0)
        return ctx->fail("FFT probe: B must be a power of two in [2^14, 2^18]");
    if (kind != TMX_FFT_PROBE_R2C && kind != TMX_FFT_PROBE_INVERSE)
        return ctx->fail("FFT probe: unknown transform");
    if (kind == TMX_FFT_PROBE_INVERSE && !fft_large_supports(B))
        return ctx->fail("FFT probe: the long-template inverse has B = 2^16, 2^17 or 2^18");
    if (!in || !out || count <= 0 || count > 64)
        return ctx->fail("FFT probe: bad arguments");
    std::lock_guard<std::mutex> call(ctx->call_mu);
    Lane &L = ctx->direct;
    if (hipSetDevice(ctx->cfg.device) != hipSuccess)
        return ctx->fail("hipSetDevice failed", -2);
    // Input and output element counts (floats) per transform.
    const size_t in_len = static_cast<size_t>(B);  // B real samples, or B / 2 complex values
    const size_t out_len = kind == TMX_FFT_PROBE_R2C ? static_cast<size_t>(B / 2 + 1) * 2 : static_cast<size_t>(B);
    const int batch = kind == TMX_FFT_PROBE_R2C ? kFftPieceBatch : count;
    DevMem din, dout, dmid, dtab;
    hipError_t e = din.reserve(sizeof(float) * in_len * batch);
    if (e == hipSuccess)
        e = dout.reserve(sizeof(float) * out_len * batch);
    if (e == hipSuccess)
        e = hipMemsetAsync(din.p, 0, sizeof(float) * in_len * batch, L.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(din.p, in, sizeof(float) * in_len * count, hipMemcpyHostToDevice, L.stream);
    if (e != hipSuccess)
        return ctx->fail(std::string("FFT probe: ") + hipGetErrorString(e), -2);
    if (kind == TMX_FFT_PROBE_R2C) {
        hipfftHandle h = 0;
        if (make_r2c_plan(&h, B) != HIPFFT_SUCCESS)
            return ctx->fail("FFT probe: plan failed", -2);
        hipfftSetStream(h, L.stream);
        const hipfftResult xr = hipfftExecR2C(h, din.as<hipfftReal>(), reinterpret_cast<hipfftComplex *>(dout.p));
        e = xr == HIPFFT_SUCCESS ? hipStreamSynchronize(L.stream) : hipErrorLaunchFailure;
        (void)hipfftDestroy(h);
    } else {
        // The table the long-template form uses.
        const std::vector<float2> tab = fft_table(B);
        e = dtab.reserve(sizeof(float2) * tab.size());
        if (e == hipSuccess)
            e = dmid.reserve(sizeof(float) * in_len * count);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dtab.p, tab.data(), sizeof(float2) * tab.size(), hipMemcpyHostToDevice, L.stream);
        if (e == hipSuccess)
            e = launch_fft_large_probe(B, dtab.as<float2>(), din.as<float2>(), count, dmid.as<float2>(),
                                       dout.as<float2>(), L.stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(L.stream);
    }
    if (e == hipSuccess)
        e = hipMemcpy(out, dout.p, sizeof(float) * out_len * count, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return ctx->fail(std::string("FFT probe: ") + hipGetErrorString(e), -2);
    return 0;
}

int tmx_sync(tmx_ctx *ctx)
{
    if (!ctx)
        return -1;
    if (hipStreamSynchronize(ctx->direct.stream) != hipSuccess)
        return ctx->fail("sync failed", -2);
    return 0;
}

}  // extern "C"
