// iSAX summaries of the windows and the skip-sequential lower bound
// (SkipSequential.cpp:155-211 with lower_bounds = true, SAX.cpp:24-85,
// 112-180, 362-490): per window of length m, the PAA of its z-normalised
// samples over `segments` equal segments, each value replaced by its symbol
// at cardinality 256 (the number of the 255 normal-quantile breakpoints
// strictly below it); per template, the PAA of its z-normalised samples.  The
// bound of a window is m / segments times the squared distance of the
// template's PAA from the word's symbol regions, in the reference's
// arithmetic, so a window the reference would skip is skipped here.
//
// The words are resident per (m, segments) once built (n - m + 1 windows of
// `segments` bytes); the bound is evaluated inside the early-abandoning scan
// (ea.hip) before any sample of the window is read.
#include "isax.hpp"

#include <cmath>

namespace tmx {
namespace {

constexpr int kThreads = 256;

// One thread per window: the reference's two-pass FP64 statistics (MISSING
// as 0, sigma <= 1e-12 -> divide by 1), each sample's (v - mean) / sigma
// rounded to float, FP32 segment sums in index order divided by the segment
// length, then the symbols.  Windows starting with MISSING keep zero words
// (never searched).
__global__ __launch_bounds__(kThreads) void k_sax_words(const float *__restrict__ x, uint64_t W, int m, int segments,
                                                        const float *__restrict__ bp_g, uint8_t *__restrict__ out)
{
    __shared__ float bp[256];
    for (int i = threadIdx.x; i < 255; i += kThreads)
        bp[i] = bp_g[i];
    __syncthreads();
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (j >= W)
        return;
    const float *w = x + j;
    uint8_t *o = out + j * static_cast<uint64_t>(segments);
    if (w[0] == kMissing) {
        for (int s = 0; s < segments; s++)
            o[s] = 0;
        return;
    }
    double sum = 0.0;
    for (int d = 0; d < m; d++)
        sum += w[d] == kMissing ? 0.0 : static_cast<double>(w[d]);
    const double mean = sum / m;
    double squares = 0.0;
    for (int d = 0; d < m; d++) {
        const double v = (w[d] == kMissing ? 0.0 : static_cast<double>(w[d])) - mean;
        squares = fma(v, v, squares);
    }
    const double sd = sqrt(squares / m);
    const double div = sd <= 1e-12 ? 1.0 : sd;
    const int len = m / segments;
    for (int s = 0; s < segments; s++) {
        float acc = 0.0f;
        for (int i = 0; i < len; i++) {
            const float v = w[s * len + i];
            acc += static_cast<float>(((v == kMissing ? 0.0 : static_cast<double>(v)) - mean) / div);
        }
        o[s] = static_cast<uint8_t>(sax_symbol(bp, acc / static_cast<float>(len)));
    }
}

// PAA of each template's z-normalised samples (paa_from_ts), one thread per
// (template, segment): the lower bound's segments, or 16 for the index.
__global__ void k_tmpl_paa(const TmplDev *__restrict__ tmpl, int count, const float *__restrict__ qz,
                           float *__restrict__ paa)
{
    const int i = blockIdx.x;
    if (i >= count)
        return;
    const TmplDev &t = tmpl[i];
    const int s = threadIdx.x;
    const int seg = t.lb_seg > 0 ? t.lb_seg : (t.ix > 0 ? 16 : 0);  // the index's are 16
    if (seg <= 0 || s >= seg)
        return;
    const int len = t.m / seg;
    const float *q = qz + t.q_off + static_cast<int64_t>(s) * len;
    float acc = 0.0f;
    for (int k = 0; k < len; k++)
        acc += q[k];
    paa[t.lb_paa + s] = acc / static_cast<float>(len);
}

}  // namespace

void sax_breakpoints(float *bp)
{
    // Phi^-1(i / 256) for i < 128 by bisection on Phi(x) = erfc(-x / sqrt 2) / 2
    // down to adjacent doubles, the median 0 and the upper half mirrored.
    bp[127] = 0.0f;
    for (int i = 1; i < 128; i++) {
        const double p = static_cast<double>(i) / 256.0;
        double lo = -10.0, hi = 10.0;
        for (;;) {
            const double mid = 0.5 * (lo + hi);
            if (!(mid > lo && mid < hi))
                break;
            if (0.5 * std::erfc(-mid * 0.70710678118654752440) < p)
                lo = mid;
            else
                hi = mid;
        }
        bp[i - 1] = static_cast<float>(0.5 * (lo + hi));
        bp[255 - i] = -bp[i - 1];
    }
}

hipError_t launch_sax_words(const float *x, uint64_t n, int m, int segments, const float *bp, uint8_t *out,
                            hipStream_t stream)
{
    if (static_cast<uint64_t>(m) > n || segments <= 0 || segments > kMaxSegments || m % segments != 0)
        return hipErrorInvalidValue;
    const uint64_t W = n - static_cast<uint64_t>(m) + 1;
    const uint64_t blocks = (W + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sax_words, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, stream, x, W, m,
                       segments, bp, out);
    return hipGetLastError();
}

hipError_t launch_tmpl_paa(const TmplDev *tmpl, int count, const float *qz, float *paa, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_tmpl_paa, dim3(static_cast<unsigned>(count)), dim3(kMaxSegments), 0, stream, tmpl, count, qz,
                       paa);
    return hipGetLastError();
}

}  // namespace tmx
