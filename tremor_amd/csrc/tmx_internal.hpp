// Internal declarations shared by the C-ABI layer (tmx.cpp) and the device
// code (stats.hip, prepare.hip, scan.hip, mfma.hip, verify.hip, select.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "tremor_amd/tmx.h"

#include <cfloat>
#include <cstdint>

namespace tmx {

// Missing-sample sentinel (ZnormSearch.hpp:21).
constexpr float kMissing = -100000.0f;
// Early-abandon block of the reference kernel (ZnormSearch.hpp:65).
constexpr int kBlock = 32;

// FP32 constants of the matrix-core form's level-1 test for one template,
// computed on the device once its bound is final (k_mfma_consts).
struct L1Consts {
    float gmin;    // (Q2 + m - bound - E_ref cap) / 2, shrunk: the dt / sigma-scaled margin
    float Kc;      // gmin / sqrt(m) less the relative dot-product error terms
    float eta_sq;  // eta * sqrt(Q2): relative error of the dot product
    float eabs_q;  // FP16: template values that may flush, per sqrt(sum x~^2)
    float sqrt_m, rsqrt_m, inv_mf;
    float absS;    // |sum q|, grown
    float exq;     // FP16: staged samples that may flush, per unit of the block's scale
    float pad;
};

// Per-template record on the device.  The z-normalised samples live in one
// packed float array (q_off), the block order in one packed int array
// (order_off).  Q2 = sum q^2 and S = sum q in FP64 feed the candidate filter.
struct TmplDev {
    int64_t q_off;      // offset into the packed z-normalised samples
    int64_t order_off;  // offset into the packed block orders
    int64_t raw_off;    // offset into the packed raw samples of this call
    int32_t m;          // length
    int32_t kpad;       // length padded for the MFMA operand layout (0 = none)
    int64_t a_off;      // offset (in 16-byte fragments) of the MFMA operand blocks
    double Q2;          // sum of q^2
    double S;           // sum of q
    float bound;        // candidate cut: threshold distance, or the kNN seed bound
    int32_t k;          // kNN answers wanted (0: threshold template)
    int32_t merge;      // exclusion zone of the answers (merge offset)
    int32_t ea;         // 1: the early-abandoning scan takes this template, 2: the iSAX index does (the
                        // other forms skip it either way)
    int64_t out_off;    // first kNN output slot of this template
    L1Consts l1;        // matrix-core level-1 constants (device-computed per pass)
    int32_t lb_seg;     // skip-sequential lower bound: PAA segments (0 = none)
    int32_t ix;         // 1: the iSAX index may take this template (16 segments; its PAA and words kept)
    int64_t lb_paa;     // offset of the template's PAA in the lane's PAA array
    const uint8_t *lb_words;  // resident iSAX words of this length (window j at j * lb_seg, or * 16)
    const struct IxLeaf *ix_leaves;  // the index of this length (ix = 1): leaves in key order,
    const uint32_t *ix_ids;          // the window ids in key order, their sorted keys,
    const uint64_t *ix_keys;
    const uint8_t *ix_words;         // their words in key order,
    const struct WindowSums *ix_sums;  // their sums in key order (eS1 < 0: first sample missing),
    int64_t ix_nleaves;              // and the leaf count
};

// One (template, FFT piece) pair of the FFT form.
struct FftPair {
    uint32_t t;  // template (call index)
    uint32_t p;  // piece
};

// One candidate window of one template.
struct Cand {
    uint32_t t;  // template index within the call
    uint32_t j;  // window start
};

// Output of a candidate pass.  Main mode: the window, the fast estimate of
// its reference distance and bounds [lo, hi] on it (floats rounded outward;
// -inf / +inf when unknown).  Seed mode: est holds the entry's upper bound.
struct CandOut {
    Cand *cand;
    float *est;
    float *lo;
    float *hi;
    unsigned long long *count;
    uint64_t cap;
};

// Level-1 survivors of the matrix-core pass, for its rigorous second level:
// the window, its dot product (real units) and the block's FP16 scale (1 for
// split BF16).
struct SurvRec {
    uint32_t t;  // template index within the call
    uint32_t j;  // window start
    float dt;    // dot product of the centred window with the template
    float scale; // FP16 scale of the block (bounds the flushed-operand error)
};

struct SurvOut {
    SurvRec *rec;
    unsigned long long *count;  // may exceed cap: the pass must be rerun with more room
    uint64_t cap;
};

__device__ inline float round_down_f(double v)
{
    float f = static_cast<float>(v);
    if (static_cast<double>(f) > v)
        f = nextafterf(f, -INFINITY);
    return f;
}

__device__ inline float round_up_f(double v)
{
    float f = static_cast<float>(v);
    if (static_cast<double>(f) < v)
        f = nextafterf(f, INFINITY);
    return f;
}

// One slot of *counter for each calling lane, with one atomic per wavefront
// (the lanes active at the call are the claimants).
__device__ inline unsigned long long wave_claim(unsigned long long *counter)
{
    const unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0ull;
    if (lane == leader)
        base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    return base + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Adds 1 to counts[key] for each calling lane, with one atomic per distinct
// key in the wavefront; returns the value counts[key] had for this lane (the
// lanes of one key get consecutive values, in lane order).
__device__ inline uint32_t wave_add_by_key(uint32_t *counts, uint32_t key)
{
    unsigned long long pending = __ballot(1);
    const int lane = __lane_id();
    uint32_t mine = 0u;
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const uint32_t kl = __shfl(key, leader);
        const unsigned long long same = __ballot(key == kl) & pending;
        uint32_t base = 0u;
        if (lane == leader)
            base = atomicAdd(&counts[kl], static_cast<uint32_t>(__popcll(same)));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull)
            mine = base + static_cast<uint32_t>(__popcll(same & ((1ull << lane) - 1ull)));
        pending &= ~same;
    }
    return mine;
}

// Stores one entry in a slot already claimed from o.count.
__device__ inline void cand_put(const CandOut &o, unsigned long long slot, uint32_t t, uint32_t j, float est,
                                double lo, double hi)
{
    if (slot < o.cap) {
        o.cand[slot].t = t;
        o.cand[slot].j = j;
        o.est[slot] = est;
        o.lo[slot] = round_down_f(lo);
        o.hi[slot] = round_up_f(hi);
    }
}

// Filter mode of the inner-product passes.
enum FilterMode : int {
    kModeMain = 0,  // emit every window whose distance lower bound <= bound
    kModeSeed = 1,  // emit per-thread minima of the distance upper bound
};

// ---- double-double arithmetic (the resident prefix sums) ------------------
// A value is hi + lo with |lo| <= ulp(hi) / 2.  Every function needs
// unfused, unreordered FP64 (-ffp-contract=off, no fast-math).

// s + e = a + b exactly (Knuth's two-sum).
__host__ __device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// s + e = a + b exactly, for |a| >= |b| or a = 0.
__host__ __device__ inline void quick_two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    e = b - (s - a);
}

// (ah + al) + (bh + bl) as a normalised pair, relative error <= 2^-104 of
// |a| + |b| (the accurate double-double addition: both parts two-summed).
__host__ __device__ inline double2 dd_add(double2 a, double2 b)
{
    double s1, s2, t1, t2;
    two_sum(a.x, b.x, s1, s2);
    two_sum(a.y, b.y, t1, t2);
    s2 += t1;
    quick_two_sum(s1, s2, s1, s2);
    s2 += t2;
    quick_two_sum(s1, s2, s1, s2);
    return make_double2(s1, s2);
}

// Relative error bound of the double-double prefix sums: every prefix value
// (hi + lo) is within kPrefixGamma * (sum of |terms|) of the exact sum.  Any
// prefix value is built by at most ~600 double-double additions along its
// sequential and tree paths, each within 2^-104 of its operands (whose
// magnitudes are at most the sum of |terms|): 600 * 2^-104 < 2^-94.
constexpr double kPrefixGamma = 0x1p-94;

struct Waveform {
    const float *x;     // raw samples, n + tail padding (missing kept)
    const double2 *p;   // prefix sums (x0, x0^2), high parts, missing as 0, n + 1 pairs
    const double2 *pl;  // their low parts: P[i] = p[i] + pl[i] (double-double), same layout
    uint64_t n;
};

// Sums of x0 and x0^2 over the samples [a, b) from the double-double prefix
// sums, with bounds on their absolute errors: the prefix values' own error
// (kPrefixGamma times the sum of |terms| before each end, sum_{k<i} |x_k| <=
// sqrt(i P2[i])) plus the roundings of the three FP64 operations per sum.
struct WindowSums {
    double S1, S2;    // sum x0, sum x0^2
    double eS1, eS2;  // |S - exact| bounds
};

__device__ inline WindowSums window_sums(double2 pa, double2 pla, double2 pb, double2 plb, uint64_t a, uint64_t b)
{
    constexpr double u64 = 1.1102230246251565e-16;
    WindowSums r;
    const double d1 = pb.x - pa.x, l1 = plb.x - pla.x;
    const double d2 = pb.y - pa.y, l2 = plb.y - pla.y;
    r.S1 = d1 + l1;
    r.S2 = d2 + l2;
    const double P2a = fabs(pa.y) * (1.0 + 4.0 * u64), P2b = fabs(pb.y) * (1.0 + 4.0 * u64);
    r.eS1 = kPrefixGamma * (sqrt(static_cast<double>(b) * P2b) + sqrt(static_cast<double>(a) * P2a)) * 1.01 +
            2.0 * u64 * (fabs(d1) + fabs(l1) + fabs(r.S1));
    r.eS2 = kPrefixGamma * (P2b + P2a) + 2.0 * u64 * (fabs(d2) + fabs(l2) + fabs(r.S2));
    return r;
}

__device__ inline WindowSums window_sums(const Waveform &w, uint64_t a, uint64_t b)
{
    return window_sums(w.p[a], w.pl[a], w.p[b], w.pl[b], a, b);
}

// The same from the high parts alone (one load per end, for register-tight
// kernels): each high part is within u64 |P| + kPrefixGamma * (sum of |terms|)
// of the exact prefix, |P| <= sum of |terms|.  Looser after loud stretches;
// a caller whose bound fails there falls back to window_sums.
__device__ inline WindowSums window_sums_hi(double2 pa, double2 pb, uint64_t a, uint64_t b)
{
    constexpr double u64 = 1.1102230246251565e-16;
    constexpr double g = u64 + kPrefixGamma;
    WindowSums r;
    r.S1 = pb.x - pa.x;
    r.S2 = pb.y - pa.y;
    const double P2a = fabs(pa.y) * (1.0 + 4.0 * u64), P2b = fabs(pb.y) * (1.0 + 4.0 * u64);
    r.eS1 = g * (sqrt(static_cast<double>(b) * P2b) + sqrt(static_cast<double>(a) * P2a)) * 1.01 +
            2.0 * u64 * fabs(r.S1);
    r.eS2 = g * (P2b + P2a) * 1.01 + 2.0 * u64 * fabs(r.S2);
    return r;
}

// The window's sums from the high parts where they bound its variance to
// 1e-3 (almost every window: one load per end), else from the full
// double-double sums (a quiet stretch after a loud one).
__device__ inline WindowSums window_sums_fast(const Waveform &w, uint64_t a, uint64_t b)
{
    const WindowSums r = window_sums_hi(w.p[a], w.p[b], a, b);
    const double L = static_cast<double>(b - a);
    const double V = r.S2 - r.S1 * r.S1 / L;
    const double eV = r.eS2 + (2.0 * fabs(r.S1) * r.eS1 + r.eS1 * r.eS1) / L;
    if (eV < 1e-3 * V)
        return r;
    return window_sums(w, a, b);
}

// ---- launchers (all asynchronous on `stream`) ----------------------------

// Double-double inclusive prefix sums (P1, P2) of x0 and x0^2 (P[0] = 0):
// high parts into p, low parts into pl, n + 1 pairs each.
hipError_t launch_prefix(const float *x, uint64_t n, double2 *p, double2 *pl, double *work, size_t work_len,
                         hipStream_t stream);
size_t prefix_work_len(uint64_t n);

// Block-rebased FP32 prefix pairs (and per-block centre and error bounds) for
// the level-1 filter of the matrix-core pass; `block` windows per block,
// `len` entries per block.
hipError_t launch_block_prefix(const double2 *p, const double2 *pl, uint64_t n, int block, int len, float4 *binfo,
                               float2 *rb, hipStream_t stream);

// Largest |x0 - c_b| over each block's len samples into binfo[b].w (after
// launch_block_prefix on the same stream).
hipError_t launch_block_absmax(const float *x, uint64_t n, int block, int len, float4 *binfo, hipStream_t stream);

// Reference template preparation: z-norm (ZnormSearch.hpp:33-62), block
// order (:72-90), plus Q2 and S.  One workgroup per template.
hipError_t launch_prepare(const float *raw, TmplDev *tmpl, int count, float *qz, int32_t *order,
                          hipStream_t stream);

// Plain FP32 sliding-dot pass with the fused candidate filter.
hipError_t launch_scan_fp32(const Waveform &w, const TmplDev *tmpl, const float *qz, int count,
                            const int32_t *group_begin, int groups, int mode, int block_stride, const CandOut &co,
                            hipStream_t stream);
constexpr int kScanWindows = 4096;  // windows per workgroup of the FP32 scan

// Exact recomputation of candidates in the reference's order.
hipError_t launch_verify(const float *x, uint64_t n, const TmplDev *tmpl, const float *qz, const int32_t *order,
                         const Cand *cand, uint64_t count, float *d_out, bool force_seq, hipStream_t stream);

// Exact reference stats of windows [first, first + count) for length m.
hipError_t launch_window_stats(const float *x, int m, uint64_t first, uint64_t count, float *mean, float *sd,
                               hipStream_t stream);

// Exact distances of one prepared template against the given windows.
hipError_t launch_dist_batch(const float *x, uint64_t n, const TmplDev *tmpl, const float *qz,
                             const int32_t *order, const uint64_t *starts, int count, float *out,
                             hipStream_t stream);

}  // namespace tmx
