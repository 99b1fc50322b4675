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
// One thread per candidate: the FP64 sums are sequential by definition.
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
    if (j + static_cast<uint64_t>(m) > n || x[j] == kMissing) {
        out[i] = FLT_MAX;
        return;
    }
    float mean, sd;
    exact_stats(x + j, m, mean, sd);
    out[i] = exact_dist(qz + tmpl[0].q_off, x + j, m, mean, sd, order + tmpl[0].order_off);
}

}  // namespace

hipError_t launch_verify(const float *x, uint64_t n, const TmplDev *tmpl, const float *qz, const int32_t *order,
                         const Cand *cand, uint64_t count, float *d_out, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const uint64_t grid = (count + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_verify, dim3(static_cast<unsigned>(grid)), dim3(kThreads), 0, stream, x, n, tmpl, qz, order,
                       cand, count, d_out);
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
