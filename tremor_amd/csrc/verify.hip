// Exact recomputation of candidate windows in the reference's arithmetic.
//
// Statistics: SkipSequential.cpp:157-179 - index-order FP64 sum with missing
// samples as 0, mean = sum / m, second pass with the square fused into the
// accumulation, std = sqrt(squares / m), stored as float, std <= 1e-12 -> 0.
// Distance: ZnormSearch.hpp:103-165 with no early abandoning - inv = 1/std in
// FP32, delta = q - (r - mean) * inv with the product fused into the
// subtraction, 16 accumulators (two AVX2 registers of 8 lanes) fed with
// fma(delta, delta, acc) block by block in the template's block order, lane
// sum ((v0+v4)+(v2+v6))+((v1+v5)+(v3+v7)) of acc0+acc1, then the tail in
// index order, fused.  The reference's GCC build contracts exactly these
// products (pinned by the oracle tests against the reference header).
// Short templates: one thread per candidate.  Long templates (kWaveMin
// samples and more): one wavefront per candidate.  The 16 FMA chains of the
// distance then run on 16 lanes in the reference's order (bit-identical), and
// the statistics come from a 64-lane FP64 sum with a rigorous bound on its
// distance from the reference's sequential sum (first the worst case
// gamma_{m-1} sum |v|, then, where that cannot decide, a bound from the
// partial sums): when the whole interval of possible reference values rounds
// to one float, that float is the reference's; otherwise the wavefront runs
// the sequential sums themselves (wave_exact_stats).
#include "tmx_internal.hpp"

namespace tmx {
namespace {

constexpr int kThreads = 256;

#pragma clang fp contract(off)
__device__ inline void exact_stats(const float *w, int m, float &mean_f, float &sd_f)
{
    if (w[0] == kMissing) {
        mean_f = 0.0f;
        sd_f = 0.0f;
        return;
    }
    double sum = 0.0;
    for (int d = 0; d < m; d++) {
        const float v = w[d];
        sum += v == kMissing ? 0.0 : static_cast<double>(v);
    }
    const double mean = sum / static_cast<double>(m);
    double squares = 0.0;
    for (int d = 0; d < m; d++) {
        const float x = w[d];
        const double v = (x == kMissing ? 0.0 : static_cast<double>(x)) - mean;
        squares = __builtin_fma(v, v, squares);
    }
    const double sd = __builtin_sqrt(squares / static_cast<double>(m));
    mean_f = static_cast<float>(mean);
    sd_f = static_cast<float>(sd <= 1e-12 ? 0.0 : sd);
}

__device__ inline float exact_dist(const float *q, const float *r, int m, float mean, float sd, const int32_t *order)
{
    if (sd <= 1e-12f)
        return FLT_MAX;
    const float inv = 1.0f / sd;
    float a0[8], a1[8];
#pragma unroll
    for (int l = 0; l < 8; l++) {
        a0[l] = 0.0f;
        a1[l] = 0.0f;
    }
    const int blocks = m / kBlock;
    for (int b = 0; b < blocks; b++) {
        const int off = order[b];
#pragma unroll
        for (int h = 0; h < kBlock; h += 16) {
#pragma unroll
            for (int l = 0; l < 8; l++) {
                float r0 = r[off + h + l];
                float r1 = r[off + h + 8 + l];
                r0 = r0 == kMissing ? 0.0f : r0;
                r1 = r1 == kMissing ? 0.0f : r1;
                const float d0 = __builtin_fmaf(-(r0 - mean), inv, q[off + h + l]);
                const float d1 = __builtin_fmaf(-(r1 - mean), inv, q[off + h + 8 + l]);
                a0[l] = __builtin_fmaf(d0, d0, a0[l]);
                a1[l] = __builtin_fmaf(d1, d1, a1[l]);
            }
        }
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; l++)
        v[l] = a0[l] + a1[l];
    const float s0 = v[0] + v[4], s1 = v[1] + v[5], s2 = v[2] + v[6], s3 = v[3] + v[7];
    float dist = (s0 + s2) + (s1 + s3);
    for (int i = blocks * kBlock; i < m; i++) {
        float rv = r[i];
        rv = rv == kMissing ? 0.0f : rv;
        const float diff = __builtin_fmaf(-(rv - mean), inv, q[i]);
        dist = __builtin_fmaf(diff, diff, dist);
    }
    return dist;
}

constexpr int kWaveMin = 512;

// exact_stats by a wavefront, bit-identical (the same operations in the same
// order on every lane): coalesced loads of 8 x 64 samples, then the
// sequential FP64 sums over them, each sample read from its lane.  No lane
// waits on a global load per sample.
__device__ void wave_exact_stats(const float *w, int m, float &mean_f, float &sd_f)
{
    constexpr int kU = 8;
    const int lane = threadIdx.x & 63;
    const double md = static_cast<double>(m);
    double sum = 0.0;
    for (int c0 = 0; c0 < m; c0 += 64 * kU) {
        float v[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int d = c0 + 64 * k + lane;
            v[k] = d < m ? w[d] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int cnt = m - (c0 + 64 * k) < 64 ? m - (c0 + 64 * k) : 64;
            for (int l = 0; l < cnt; l++) {
                const float x = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v[k]), l));
                sum += x == kMissing ? 0.0 : static_cast<double>(x);
            }
        }
    }
    const double mean = sum / md;
    double squares = 0.0;
    for (int c0 = 0; c0 < m; c0 += 64 * kU) {
        float v[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int d = c0 + 64 * k + lane;
            v[k] = d < m ? w[d] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int cnt = m - (c0 + 64 * k) < 64 ? m - (c0 + 64 * k) : 64;
            for (int l = 0; l < cnt; l++) {
                const float x = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v[k]), l));
                const double t = (x == kMissing ? 0.0 : static_cast<double>(x)) - mean;
                squares = __builtin_fma(t, t, squares);
            }
        }
    }
    const double sd = __builtin_sqrt(squares / md);
    mean_f = static_cast<float>(mean);
    sd_f = static_cast<float>(sd <= 1e-12 ? 0.0 : sd);
}

__device__ inline double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

// gamma_k of the FP64 error analysis (k u / (1 - k u), u = 2^-53), rounded up.
__device__ inline double gamma64(double k)
{
    constexpr double u = 1.1102230246251565e-16;
    return k * u / (1.0 - k * u) * (1.0 + 1e-10);
}

// Bound on |sum_seq - S| for the reference's index-order FP64 sum sum_seq of
// the window's samples (S the exact sum, A >= sum |v|), from its partial
// sums: each addition's rounding is at most u times the partial sum it
// produces, and |s_k| <= |S_k| + gamma_k A, so |sum_seq - S| <= u / (1 - u)
// (sum_k |S_k| + gamma_m m A).  The exact prefix sums S_k are bounded by this
// wavefront's own (rows of 64 scanned in FP64 on top of the sum of the rows
// before: within gamma_{R+7} A of exact, R = ceil(m / 64)).  For sums that
// wander like a random walk this is about sqrt(m) times below gamma_{m-1} A.
__device__ double seq_sum_error(const float *w, int m, double A)
{
    constexpr double u = 1.1102230246251565e-16;
    const int lane = threadIdx.x & 63;
    const int R = (m + 63) / 64;
    double run = 0.0, sabs = 0.0;
    for (int r = 0; r < R; r++) {
        const int d = 64 * r + lane;
        double y = 0.0;
        if (d < m) {
            const float v = w[d];
            y = v == kMissing ? 0.0 : static_cast<double>(v);
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double t = __shfl_up(y, off);
            if (lane >= off)
                y += t;
        }
        if (d < m)
            sabs += fabs(run + y);
        run += __shfl(y, 63);
    }
    sabs = __shfl(wave_sum_f64(sabs), 0);
    const double md = static_cast<double>(m);
    const double sum_abs = sabs * (1.0 + 2.0 * gamma64(static_cast<double>(R + 6))) +
                           md * gamma64(static_cast<double>(R + 7)) * A;
    return u / (1.0 - u) * (sum_abs + gamma64(md) * md * A) * (1.0 + 1e-10);
}

// The reference statistics (exact_stats) of one window by a wavefront; w[0]
// is not missing.  Returns false when the fast path cannot decide a value.
__device__ bool wave_stats(const float *w, int m, float &mean_f, float &sd_f)
{
    constexpr double u = 1.1102230246251565e-16;
    const int lane = threadIdx.x & 63;
    const double md = static_cast<double>(m);
    // Lane sums in index order; eight reads of a round issued before the
    // first is added (the order of the additions is unchanged).
    constexpr int kU = 8;
    double s = 0.0, a = 0.0;
    for (int c0 = lane; c0 < m; c0 += 64 * kU) {
        float v[kU];
#pragma unroll
        for (int k = 0; k < kU; k++)
            v[k] = c0 + 64 * k < m ? w[c0 + 64 * k] : 0.0f;
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const double dv = v[k] == kMissing ? 0.0 : static_cast<double>(v[k]);
            s += dv;
            a += fabs(dv);
        }
    }
    // Every lane takes lane 0's sums, so the decisions below are uniform.
    s = __shfl(wave_sum_f64(s), 0);
    a = __shfl(wave_sum_f64(a), 0);
    // Parallel sum: ceil(m / 64) sequential terms then 6 tree levels.
    const double gp = gamma64(static_cast<double>((m + 63) / 64 + 6));
    const double A = a * (1.0 + 2.0 * gp);                      // >= sum |v|
    double Es = (gamma64(static_cast<double>(m - 1)) + gp) * A;  // |sum_seq - s|
    // mean_ref = fl(sum_seq / m) lies in [ml, mh].
    double ml = (s - Es) / md, mh = (s + Es) / md;
    ml -= fabs(ml) * 4.0 * u;
    mh += fabs(mh) * 4.0 * u;
    mean_f = static_cast<float>(ml);
    if (static_cast<float>(mh) != mean_f) {
        // Undecided (common for means near 0, where the float spacing is
        // fine): the partial-sum bound instead.
        Es = seq_sum_error(w, m, A) + gp * A;
        ml = (s - Es) / md;
        mh = (s + Es) / md;
        ml -= fabs(ml) * 4.0 * u;
        mh += fabs(mh) * 4.0 * u;
        mean_f = static_cast<float>(ml);
        if (static_cast<float>(mh) != mean_f)
            return false;
    }
    // Squares about the midpoint, then the reference's sum about mean_ref.
    const double mu = 0.5 * (ml + mh);
    const double W = 0.5 * (mh - ml) * (1.0 + 4.0 * u);
    double q = 0.0;
    for (int c0 = lane; c0 < m; c0 += 64 * kU) {
        float v[kU];
#pragma unroll
        for (int k = 0; k < kU; k++)
            v[k] = c0 + 64 * k < m ? w[c0 + 64 * k] : 0.0f;
#pragma unroll
        for (int k = 0; k < kU; k++) {
            if (c0 + 64 * k < m) {
                const double dv = (v[k] == kMissing ? 0.0 : static_cast<double>(v[k])) - mu;
                q = __builtin_fma(dv, dv, q);
            }
        }
    }
    q = __shfl(wave_sum_f64(q), 0);
    const double e2 = 2.01 * u + gp;
    const double Qlo0 = q / (1.0 + e2), Qhi0 = q / (1.0 - e2);       // Q(mu)
    const double T = (fabs(s - md * mu) + Es) * (1.0 + 8.0 * u) + 4.0 * u * fabs(md * mu);  // >= |S - m mu|
    const double Qlo = Qlo0 - 2.0 * W * T, Qhi = Qhi0 + md * W * W + 2.0 * W * T;  // Q(mean_ref)
    const double e1 = 2.01 * u + gamma64(md);
    const double sqlo = Qlo * (1.0 - e1), sqhi = Qhi * (1.0 + e1);  // the reference's squares
    if (!(sqlo >= 0.0))
        return false;
    double sdl = sqrt(sqlo / md), sdh = sqrt(sqhi / md);
    sdl *= 1.0 - 4.0 * u;
    sdh *= 1.0 + 4.0 * u;
    if (sdh <= 1e-12) {
        sd_f = 0.0f;
        return true;
    }
    if (!(sdl > 1e-12))
        return false;
    sd_f = static_cast<float>(sdl);
    return static_cast<float>(sdh) == sd_f;
}

// The reference distance of one window by lanes 0..15 of a wavefront (the
// 16 accumulators), combined and finished with the tail on lane 0; the
// result is valid on lane 0.
__device__ float wave_dist(const float *q, const float *r, int m, float mean, float sd, const int32_t *order)
{
    const int lane = threadIdx.x & 63;
    if (sd <= 1e-12f)
        return FLT_MAX;
    const float inv = 1.0f / sd;
    float acc = 0.0f;
    const int blocks = m / kBlock;
    if (lane < 16) {
        // Unrolled so that several blocks' loads are in flight at once.
#pragma unroll 8
        for (int b = 0; b < blocks; b++) {
            const int off = order[b];
#pragma unroll
            for (int h = 0; h < kBlock; h += 16) {
                float rv = r[off + h + lane];
                rv = rv == kMissing ? 0.0f : rv;
                const float dd = __builtin_fmaf(-(rv - mean), inv, q[off + h + lane]);
                acc = __builtin_fmaf(dd, dd, acc);
            }
        }
    }
    // v[l] = a0[l] + a1[l]: accumulators l and l + 8.
    float v = acc + __shfl(acc, (lane + 8) & 63);
    float vl[8];
#pragma unroll
    for (int l = 0; l < 8; l++)
        vl[l] = __shfl(v, l);
    float dist = 0.0f;
    if (lane == 0) {
        const float s0 = vl[0] + vl[4], s1 = vl[1] + vl[5], s2 = vl[2] + vl[6], s3 = vl[3] + vl[7];
        dist = (s0 + s2) + (s1 + s3);
        for (int i = blocks * kBlock; i < m; i++) {
            float rv = r[i];
            rv = rv == kMissing ? 0.0f : rv;
            const float diff = __builtin_fmaf(-(rv - mean), inv, q[i]);
            dist = __builtin_fmaf(diff, diff, dist);
        }
    }
    return dist;
}

// Long-template candidates, one wavefront each (others skip).
__global__ __launch_bounds__(kThreads) void k_verify_wave(const float *__restrict__ x, uint64_t n,
                                                          const TmplDev *__restrict__ tmpl,
                                                          const float *__restrict__ qz,
                                                          const int32_t *__restrict__ order,
                                                          const Cand *__restrict__ cand, uint64_t count,
                                                          float *__restrict__ d_out, int force_seq)
{
    const uint64_t i = (static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= count)
        return;
    const Cand c = cand[i];
    const int m = tmpl[c.t].m;
    if (m < kWaveMin)
        return;
    float out = FLT_MAX;
    if (static_cast<uint64_t>(c.j) + static_cast<uint64_t>(m) <= n && x[c.j] != kMissing) {
        const float *w = x + c.j;
        float mean, sd;
        if (force_seq || !wave_stats(w, m, mean, sd))
            wave_exact_stats(w, m, mean, sd);
        out = wave_dist(qz + tmpl[c.t].q_off, w, m, mean, sd, order + tmpl[c.t].order_off);
    }
    if (lane == 0)
        d_out[i] = out;
}

__global__ __launch_bounds__(kThreads) void k_verify(const float *__restrict__ x, uint64_t n,
                                                     const TmplDev *__restrict__ tmpl, const float *__restrict__ qz,
                                                     const int32_t *__restrict__ order, const Cand *__restrict__ cand,
                                                     uint64_t count, float *__restrict__ d_out)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= count)
        return;
    const Cand c = cand[i];
    const int m = tmpl[c.t].m;
    if (m >= kWaveMin)
        return;  // k_verify_wave
    if (static_cast<uint64_t>(c.j) + static_cast<uint64_t>(m) > n) {
        d_out[i] = FLT_MAX;
        return;
    }
    const float *w = x + c.j;
    if (w[0] == kMissing) {
        d_out[i] = FLT_MAX;
        return;
    }
    float mean, sd;
    exact_stats(w, m, mean, sd);
    d_out[i] = exact_dist(qz + tmpl[c.t].q_off, w, m, mean, sd, order + tmpl[c.t].order_off);
}

__global__ __launch_bounds__(kThreads) void k_window_stats(const float *__restrict__ x, int m, uint64_t first,
                                                           uint64_t count, float *__restrict__ mean,
                                                           float *__restrict__ sd)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= count)
        return;
    float a, b;
    exact_stats(x + first + i, m, a, b);
    mean[i] = a;
    sd[i] = b;
}

__global__ __launch_bounds__(kThreads) void k_dist_batch(const float *__restrict__ x, uint64_t n,
                                                         const TmplDev *__restrict__ tmpl,
                                                         const float *__restrict__ qz,
                                                         const int32_t *__restrict__ order,
                                                         const uint64_t *__restrict__ starts, int count,
                                                         float *__restrict__ out)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count)
        return;
    const int m = tmpl[0].m;
    const uint64_t j = starts[i];
    // Written so that no sum can wrap: a start near 2^64 is out of range too.
    if (static_cast<uint64_t>(m) > n || j > n - static_cast<uint64_t>(m) || x[j] == kMissing) {
        out[i] = FLT_MAX;
        return;
    }
    float mean, sd;
    exact_stats(x + j, m, mean, sd);
    out[i] = exact_dist(qz + tmpl[0].q_off, x + j, m, mean, sd, order + tmpl[0].order_off);
}

}  // namespace

hipError_t launch_verify(const float *x, uint64_t n, const TmplDev *tmpl, const float *qz, const int32_t *order,
                         const Cand *cand, uint64_t count, float *d_out, bool force_seq, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const uint64_t grid = (count + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_verify, dim3(static_cast<unsigned>(grid)), dim3(kThreads), 0, stream, x, n, tmpl, qz, order,
                       cand, count, d_out);
    const uint64_t wgrid = (count + kThreads / 64 - 1) / (kThreads / 64);
    for (uint64_t g0 = 0; g0 < wgrid; g0 += (1u << 31) - 1) {
        const uint64_t g = wgrid - g0 < (1u << 31) - 1 ? wgrid - g0 : (1u << 31) - 1;
        const uint64_t first = g0 * (kThreads / 64);
        hipLaunchKernelGGL(k_verify_wave, dim3(static_cast<unsigned>(g)), dim3(kThreads), 0, stream, x, n, tmpl, qz,
                           order, cand + first, count - first, d_out + first, force_seq ? 1 : 0);
    }
    return hipGetLastError();
}

hipError_t launch_window_stats(const float *x, int m, uint64_t first, uint64_t count, float *mean, float *sd,
                               hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const uint64_t grid = (count + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_window_stats, dim3(static_cast<unsigned>(grid)), dim3(kThreads), 0, stream, x, m, first,
                       count, mean, sd);
    return hipGetLastError();
}

hipError_t launch_dist_batch(const float *x, uint64_t n, const TmplDev *tmpl, const float *qz,
                             const int32_t *order, const uint64_t *starts, int count, float *out, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    const int grid = (count + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_dist_batch, dim3(grid), dim3(kThreads), 0, stream, x, n, tmpl, qz, order, starts, count,
                       out);
    return hipGetLastError();
}

}  // namespace tmx
