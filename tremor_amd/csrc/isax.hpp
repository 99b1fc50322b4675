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

}  // namespace tmx
