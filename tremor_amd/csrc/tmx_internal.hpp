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
    int32_t ea;         // 1: the early-abandoning scan takes this template (the other forms skip it)
    int64_t out_off;    // first kNN output slot of this template
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

// Relative error bound of the FP64 prefix sums: every prefix value is within
// kPrefixGamma * (sum of |terms|) of the exact sum (a few thousand sequential
// and tree additions at most, each 2^-53).
constexpr double kPrefixGamma = 8192.0 * 1.1102230246251565e-16;

struct Waveform {
    const float *x;     // raw samples, n + tail padding (missing kept)
    const double2 *p;   // prefix sums (x0, x0^2) interleaved, missing as 0, n + 1 pairs
    uint64_t n;
};

// ---- launchers (all asynchronous on `stream`) ----------------------------

// FP64 inclusive prefix sums (P1, P2) of x0 and x0^2, interleaved (P[0] = 0).
hipError_t launch_prefix(const float *x, uint64_t n, double2 *p, double *work, size_t work_len, hipStream_t stream);
size_t prefix_work_len(uint64_t n);

// Block-rebased FP32 prefix pairs (and per-block centre and error bounds) for
// the level-1 filter of the matrix-core pass; `block` windows per block,
// `len` entries per block.
hipError_t launch_block_prefix(const double2 *p, uint64_t n, int block, int len, float4 *binfo, float2 *rb,
                               hipStream_t stream);

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
