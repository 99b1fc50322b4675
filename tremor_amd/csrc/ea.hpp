// One window of the early-abandoning scan, shared by the scan pass (ea.hip)
// and the iSAX index search (index.hip).
//
// Early-abandoning scan as a candidate pass: the reference's skip-sequential
// scan without lower bounds (SkipSequential.cpp:223-294 threshold, :296-393
// kNN, per window the kernel of ZnormSearch.hpp:103-165 with its abandon
// test after every 32-sample block, blocks in the template's energy order of
// ZnormSearch.hpp:72-90).  One thread per window, so consecutive lanes read
// consecutive samples and the template's values are wavefront-uniform
// (scalar loads).
//
// The window statistics come from the FP64 prefix sums (as in filter.hpp)
// instead of the reference's two-pass sums, so the running sum differs from
// the reference's by a bounded amount.  A window is abandoned only once its
// running sum proves that the reference's distance exceeds the bound; a
// window that runs to the end becomes a candidate with an interval [lo, hi]
// of its reference distance when lo <= bound, and the exact recomputation
// decides it.  The candidate set therefore contains every answer, and the
// answers are those of the reference.
//
// Error bound (u = 2^-24).  With the exact window statistics mu, sigma and
// w_i = (x_i - mu) / sigma (so sum over the window of w_i^2 = m):
//   this pass:  dz_i = fma(-(x_i - mean_f), inv_f, q_i) with mean_f, inv_f
//               from the prefix sums; |dz_i - (q_i - w_i)| <= |w_i| (e_sig + u)
//               (1 + u) + |mean_f - mu| inv_f (1 + u) + u |dz_i|, e_sig the
//               relative error of inv_f; block sums of 32 fused products
//               added in sequence: relative error gamma_{nb + 34};
//   reference:  the same terms with its float statistics (relative error of
//               1 / std_f <= 2u + 1e-10, |mean_f - mu| <= u |mu| + 1e-11
//               (|mu| + sigma)); 16 FMA chains, lane sum and tail: relative
//               error gamma_{m/16 + 40}.
// Over the blocks done, the 2-norm of the difference of the two term vectors
// is at most T = (e_sig + 5u + 1e-9) sqrt(m) + sqrt(m) (dmu + dmu_ref) /
// sigma_lo (1.001) (dmu, dmu_ref: the two float means' distances from mu),
// hence with s this pass's running sum
//   d_ref >= (1 - gamma_ref) ((sqrt(s) (1 - gamma_acc - 1.01u) - T) / (1 + u))^2,
// and the window is abandoned once s exceeds
//   thr = ((sqrt(bound / (1 - gamma_ref)) (1 + u) + T) / (1 - gamma_acc - 1.01u))^2.
// The reference's partial sums only grow (fused products of squares onto
// non-negative accumulators), so its final distance is at least this bound
// whichever block the test fires at.
#pragma once

#include "isax.hpp"

namespace tmx {

constexpr int kModeSample = 2;  // count the blocks each sampled window visits

__device__ inline double gamma32(double k)
{
    constexpr double u = 5.9604644775390625e-08;
    return k * u / (1.0 - k * u);
}

__device__ inline double sq(double v) { return v * v; }

// The window kernel proper for window j (present, not skipped by a lower
// bound), from its sums and their error bounds (filter.hpp's, with products
// by the rounded 1 / m, whose extra roundings the 8 u64 terms cover), against
// bound_f (the template's bound, or a tighter valid one).  Returns an upper
// bound of the reference's distance, rounded up to float, for a window that
// became a candidate with a finite interval, else +inf.
__device__ __forceinline__ float scan_window_sums(const Waveform &w, const TmplDev *__restrict__ tmpl, int T, int t,
                                                  int m, const float *__restrict__ qz,
                                                  const int32_t *__restrict__ order, int mode, const CandOut &co,
                                                  unsigned long long *__restrict__ visited, uint64_t j,
                                                  const WindowSums &ws, float bound_f)
{
    constexpr double u = 5.9604644775390625e-08;
    constexpr double u64 = 1.1102230246251565e-16;
    const int nb = m / kBlock;
    const double md = static_cast<double>(m);
    const double invm = 1.0 / md;
    const double S1 = ws.S1, S2 = ws.S2, eS1 = ws.eS1, eS2 = ws.eS2;
    const double S1sq = S1 * S1 * invm;
    const double V = S2 - S1sq;
    const double eV = eS2 + (2.0 * fabs(S1) * eS1 + eS1 * eS1) * invm * (1.0 + 4.0 * u64) + 8.0 * u64 * (fabs(S2) + S1sq);
    if (!(V - eV > md * 4e-24) || !(eV < 0.25 * V)) {
        // Undecidable here (near-constant window or swamped prefix
        // differences): the exact recomputation decides.
        if (mode == kModeSample) {
            atomicAdd(&visited[t], static_cast<unsigned long long>(nb + 1));
            atomicAdd(&visited[T + t], 1ull);
        } else {
            cand_put(co, wave_claim(co.count), static_cast<uint32_t>(t), static_cast<uint32_t>(j), NAN, -INFINITY,
                     INFINITY);
        }
        return INFINITY;
    }
    // sigma = sqrt(V_exact / m) with |V_exact - V| <= eV, rV = eV / V <= 1/4:
    // |sigma / sig - 1| <= rV (sqrt(1 + x) - 1 <= x / 2, 1 - sqrt(1 - x) <= x).
    const double sig = sqrt(V * invm);
    const double rV = eV / V + 4.0 * u64;
    const double mu = S1 * invm;
    const double emu = eS1 * invm * (1.0 + 4.0 * u64) + 3.0 * u64 * fabs(mu);
    const float mean_f = static_cast<float>(mu);
    const float inv_f = 1.0f / static_cast<float>(sig);  // 1 / sigma within (rV + 3u)
    const double e_sig = rV * (1.0 + 4.0 * u) + 4.0 * u;
    const double dmu = u * fabs(mu) + emu;
    const double dmu_ref = u * (fabs(mu) + emu) + 1e-11 * (fabs(mu) + emu + sig * (1.0 + rV));
    const double rm = sqrt(md);
    // 1 / sig_lo <= inv_f (1 + 3u) (1 + 2 rV) for rV <= 1/2.
    const double inv_lo = static_cast<double>(inv_f) * (1.0 + 3.0 * u) * (1.0 + 2.0 * rV);
    const double Tm = (e_sig + 5.0 * u + 1e-9) * rm + rm * (dmu + dmu_ref) * 1.001 * inv_lo;
    const double g_acc = gamma32(static_cast<double>(nb + m % kBlock + 34));
    const double g_ref = gamma32(floor(md / 16.0) + 40.0);
    const double bound = static_cast<double>(bound_f);
    const double shrink = 1.0 - g_acc - 1.01 * u;
    float thr = INFINITY;
    if (bound < 3.0e38) {
        const double th = sq((sqrt(fmax(bound, 0.0) / (1.0 - g_ref)) * (1.0 + u) + Tm) / shrink) * (1.0 + 1e-5);
        thr = th < 3.0e38 ? static_cast<float>(th) * (1.0f + 1e-6f) : INFINITY;
    }

    const float *q = qz + tmpl[t].q_off;
    const int32_t *ord = order + tmpl[t].order_off;
    const float *r = w.x + j;
    float s = 0.0f;
    int b = 0;
    bool abandoned = false;
    for (; b < nb; b++) {
        const int off = ord[b];
        float bs = 0.0f;
#pragma unroll
        for (int h = 0; h < kBlock; h++) {
            float rv = r[off + h];
            rv = rv == kMissing ? 0.0f : rv;
            const float dz = __builtin_fmaf(-(rv - mean_f), inv_f, q[off + h]);
            bs = __builtin_fmaf(dz, dz, bs);
        }
        s += bs;
        if (s > thr) {
            abandoned = true;
            b++;
            break;
        }
    }
    if (mode == kModeSample) {
        atomicAdd(&visited[t], static_cast<unsigned long long>(abandoned ? b : nb + 1));
        atomicAdd(&visited[T + t], 1ull);
        return INFINITY;
    }
    if (abandoned)
        return INFINITY;
    float ts = 0.0f;
    for (int k = nb * kBlock; k < m; k++) {
        float rv = r[k];
        rv = rv == kMissing ? 0.0f : rv;
        const float dz = __builtin_fmaf(-(rv - mean_f), inv_f, q[k]);
        ts = __builtin_fmaf(dz, dz, ts);
    }
    s += ts;
    const double rs = sqrt(static_cast<double>(s));
    const double lo = (1.0 - g_ref) * sq(fmax(0.0, rs * shrink - Tm) / (1.0 + u));
    const double hi = (1.0 + g_ref) * sq((rs * (1.0 + g_acc + 1.01 * u) + Tm) / (1.0 - u));
    if (lo > bound)
        return INFINITY;
    cand_put(co, wave_claim(co.count), static_cast<uint32_t>(t), static_cast<uint32_t>(j), s, lo, hi);
    // hi >= 0: the next float up is the next bit pattern.
    float hf = static_cast<float>(hi);
    if (static_cast<double>(hf) < hi)
        hf = __uint_as_float(__float_as_uint(hf) + 1u);
    return hf;
}


// One window of the early-abandoning scan (index i of the template's
// windows, or of its sample) against bound_f (the template's bound or a
// tighter valid one); true when the lower bound skipped it.  *hi_out gets
// the candidate's upper bound as scan_window_sums returns it (+inf if none).
__device__ __forceinline__ bool scan_window(const Waveform &w, const TmplDev *__restrict__ tmpl, int T, int t, int m,
                                            const float *__restrict__ qz, const int32_t *__restrict__ order, int mode,
                                            uint64_t sample_n, const CandOut &co,
                                            unsigned long long *__restrict__ visited, const float *bp,
                                            const float *paa, const float *gap, int lb_seg, uint64_t i,
                                            float bound_f, float *hi_out)
{
    *hi_out = INFINITY;
    const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
    uint64_t j;
    if (mode == kModeSample) {
        if (i >= sample_n)
            return false;
        j = (i * W) / sample_n;
    } else {
        if (i >= W)
            return false;
        j = i;
    }
    if (w.x[j] == kMissing)
        return false;  // never searched (SkipSequential.cpp:256)
    if (lb_seg > 0) {
        const uint8_t *word = tmpl[t].lb_words + j * static_cast<uint64_t>(lb_seg);
        const float lb = lb_seg == 16 ? sax_lower_bound16_table(gap, word, m)
                                      : sax_lower_bound(bp, paa, word, m, lb_seg);
        if (lb > bound_f) {
            if (mode == kModeSample)
                atomicAdd(&visited[T + t], 1ull);  // sampled, no block visited
            return true;
        }
    }
    // Statistics from the prefix sums with their error bounds.
    *hi_out = scan_window_sums(w, tmpl, T, t, m, qz, order, mode, co, visited, j,
                               window_sums_fast(w, j, j + static_cast<uint64_t>(m)), bound_f);
    return false;
}

}  // namespace tmx
