// iSAX summaries and the skip-sequential lower bound (isax.hip): device
// helpers shared with the early-abandoning scan, and the launchers.
#pragma once

#include "tmx_internal.hpp"

namespace tmx {

// Most PAA segments a template's lower bound may use.
constexpr int kMaxSegments = 64;

// Symbol of a PAA value at cardinality 256: the number of the 255 sorted
// breakpoints strictly below it (SAX.cpp:46-66, the bsearch with the compare
// of SAX.cpp:7-22; 0 for NaN), by a branchless binary search.
__device__ inline int sax_symbol(const float *bp, float v)
{
    int pos = 0;
#pragma unroll
    for (int step = 128; step > 0; step >>= 1) {
        const int c = pos + step;
        if (c <= 255 && bp[c - 1] < v)
            pos = c;
    }
    return pos;
}

// One segment's distance of the template's PAA value p from the symbol
// region [bp[v - 1], bp[v]] of word symbol v (MINVAL / MAXVAL = -/+2e6 at
// the ends, iSAXTypes.hpp:17-18), in FP32 as the reference subtracts.
__device__ inline float sax_gap(const float *bp, float p, int v)
{
    const float lo = v == 0 ? -2000000.0f : bp[v - 1];
    const float hi = v == 255 ? 2000000.0f : bp[v];
    return lo > p ? lo - p : (hi < p ? hi - p : 0.0f);
}

// SkipSequentialScan::lower_bound (SkipSequential.cpp:200-211).  16 segments
// (minidist_paa_to_isax_rawa_SIMD): FP32 squares summed as the two DPPS of
// each 8-lane half, ((s0 + s1) + (s2 + s3)) per four, the halves added
// lane-wise and then lanes 0 and 4; other counts (minidist_paa_to_isax):
// each FP32 gap widened to FP64 and accumulated by one fused multiply-add
// into the FP32 sum.  Then times m / segments in FP32.
__device__ inline float sax_lower_bound(const float *bp, const float *paa, const uint8_t *word, int m, int segments)
{
    const float ratio = static_cast<float>(m) / static_cast<float>(segments);
    if (segments == 16) {
        const uint4 wv = *reinterpret_cast<const uint4 *>(word);
        const uint32_t wd[4] = {wv.x, wv.y, wv.z, wv.w};
        float h[4];
#pragma unroll
        for (int g = 0; g < 4; g++) {
            float q[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float d = sax_gap(bp, paa[4 * g + e], static_cast<int>((wd[g] >> (8 * e)) & 0xffu));
                q[e] = d * d;
            }
            h[g] = (q[0] + q[1]) + (q[2] + q[3]);
        }
        return ((h[0] + h[2]) + (h[1] + h[3])) * ratio;
    }
    float dist = 0.0f;
    for (int s = 0; s < segments; s++) {
        const double e = static_cast<double>(sax_gap(bp, paa[s], word[s]));
        dist = static_cast<float>(fma(e, e, static_cast<double>(dist)));
    }
    return ratio * dist;
}

// The 16-segment bound from a table of the template's gaps, gap[256 s + v]
// = sax_gap(bp, paa[s], v): the same FP32 squares and the same sums.
__device__ inline float sax_lower_bound16_table(const float *gap, const uint8_t *word, int m)
{
    const float ratio = static_cast<float>(m) / 16.0f;
    const uint4 wv = *reinterpret_cast<const uint4 *>(word);
    const uint32_t wd[4] = {wv.x, wv.y, wv.z, wv.w};
    float h[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        float q[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float d = gap[256 * (4 * g + e) + static_cast<int>((wd[g] >> (8 * e)) & 0xffu)];
            q[e] = d * d;
        }
        h[g] = (q[0] + q[1]) + (q[2] + q[3]);
    }
    return ((h[0] + h[2]) + (h[1] + h[3])) * ratio;
}

// ---- the iSAX index (index.hip) ------------------------------------------

constexpr int kIxPlanes = 4;       // symbol bits per segment in the 64-bit key (16 segments)
constexpr int kIxRootDepth = 16;   // the root level: one bit per segment

// A leaf: the sorted windows [start, start + count) share the top `depth`
// bits of their keys; key is its first window's.
struct IxLeaf {
    uint64_t key;
    uint32_t start, count;
    int32_t depth;
    int32_t pad;
};

// The leaf's region bound (minidist_paa_to_isax, SAX.cpp:112-170, with the
// segment cardinalities of the node): segment s known to b = ceil((depth - s)
// / 16) bits (at most 4) is the symbol range [v 2^(8-b), (v + 1) 2^(8-b) - 1]
// of its top bits v, whose breakpoints (MINVAL / MAXVAL at the ends) bound the
// PAA value's distance; FP32 squares summed in segment order, times m / 16.
__device__ inline float ix_leaf_bound(const float *bp, const float *paa, uint64_t key, int depth, int m)
{
    float dist = 0.0f;
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int b = depth > s ? min(kIxPlanes, (depth - s + 15) / 16) : 0;
        uint32_t v = 0;
        for (int p = 0; p < b; p++)
            v = v << 1 | static_cast<uint32_t>((key >> (63 - (16 * p + s))) & 1ull);
        const int lo_sym = static_cast<int>(v << (8 - b));
        const int hi_sym = lo_sym + (1 << (8 - b)) - 1;
        const float lo = lo_sym == 0 ? -2000000.0f : bp[lo_sym - 1];
        const float hi = hi_sym == 255 ? 2000000.0f : bp[hi_sym];
        const float d = lo > paa[s] ? lo - paa[s] : (hi < paa[s] ? hi - paa[s] : 0.0f);
        dist += d * d;
    }
    return static_cast<float>(m) / 16.0f * dist;
}

// Whether a lower bound lets a window (or leaf) through against the
// template's bound: pruned only when it exceeds the bound by more than 1e-5
// relative (a margin over the FP32 roundings of the bound and of the exact
// distance it bounds); NaN passes.
__device__ inline bool ix_pass(float lb, float bound) { return !(lb * 0.99999f > bound); }

// The 255 breakpoints of cardinality 256 (SAXBreakpoints.hpp, its
// sub-table at (c - 1)(c - 2) / 2): the standard normal quantiles
// Phi^-1(i / 256) rounded to float (host).
void sax_breakpoints(float *bp);

// Words of the n - m + 1 windows of x (uint8 [W][segments]).
hipError_t launch_sax_words(const float *x, uint64_t n, int m, int segments, const float *bp, uint8_t *out,
                            hipStream_t stream);

// PAA of the templates [0, count) that carry a lower bound (lb_seg > 0)
// into paa[tmpl.lb_paa ..].
hipError_t launch_tmpl_paa(const TmplDev *tmpl, int count, const float *qz, float *paa, hipStream_t stream);

// Templates ids[0, count) of the index: per leaf its bound against the
// template's, the windows of the passing leaves added to passed[template].
hipError_t launch_ix_count(const TmplDev *tmpl, const int32_t *ids, int count, int64_t max_leaves, const float *bp,
                           const float *paa, unsigned long long *passed, hipStream_t stream);

// The exact index search of templates ids[0, count) (cus: compute units of
// the device): the windows of the passing leaves added to passed[template]
// as in launch_ix_count, and each of those windows whose own bound passes
// goes through the early-abandoning window kernel (a candidate unless its
// distance is proved above the bound; windows starting with a missing sample
// are never searched).  dyn: one word per template of the batch (ids index
// it), the running bound of k = 1 templates; null when none is among ids.
hipError_t launch_ix_search(const Waveform &w, const TmplDev *tmpl, const int32_t *ids, int count, int64_t max_leaves,
                            const float *qz, const int32_t *order, const float *bp, const float *paa,
                            unsigned long long *passed, unsigned int *dyn, const CandOut &co, int cus,
                            hipStream_t stream);

// The approximate search before the exact one: for each template ids[i],
// the `per` windows nearest its own key in key order (its leaf and the
// neighbouring ones) as candidates.
hipError_t launch_ix_home(const Waveform &w, const TmplDev *tmpl, const int32_t *ids, int count, const float *bp,
                          const float *paa, int per, const CandOut &co, hipStream_t stream);

// The routing sample: windows j * (W / samples), j < samples, whose word
// bound does not exceed the template's bound, added to passed[template].
hipError_t launch_ix_sample(const TmplDev *tmpl, const int32_t *ids, int count, const float *bp, const float *paa,
                            uint64_t n, int samples, unsigned long long *passed, hipStream_t stream);

}  // namespace tmx
