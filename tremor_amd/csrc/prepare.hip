// Template preparation on the device, bit-identical to the reference:
//   znormalize_inplace   ZnormSearch.hpp:33-62  (single-pass FP64 sum and
//                        fused sum of squares in index order; variance with
//                        mean*mean fused; sigma <= 1e-12 -> zeros)
//   query_block_order    ZnormSearch.hpp:72-90  (FP32 block energies with
//                        fused squares; stable sort by descending energy)
// plus Q2 = sum q^2 and S = sum q in FP64 for the candidate filter.
// One workgroup per template.  The index-order FP64 accumulation is done by
// one thread: its rounding sequence is what the reference's values depend on.
#include "tmx_internal.hpp"

namespace tmx {
namespace {

constexpr int kPrepThreads = 256;
constexpr int kMaxBlocks = 2048;  // templates up to 65536 samples
constexpr int kStage = 4096;      // samples staged per chunk of the sequential sums

#pragma clang fp contract(off)
__global__ __launch_bounds__(kPrepThreads) void k_prepare(const float *__restrict__ raw_all, TmplDev *tmpl,
                                                          float *__restrict__ qz_all, int32_t *__restrict__ order_all)
{
    __shared__ double s_mean, s_sd;
    __shared__ float energy[kMaxBlocks];
    __shared__ float stage[kStage];
    __shared__ double red_a[kPrepThreads], red_b[kPrepThreads];
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int m = tmpl[t].m;
    const float *raw = raw_all + tmpl[t].raw_off;
    float *q = qz_all + tmpl[t].q_off;
    int32_t *order = order_all + tmpl[t].order_off;

    // Sequential FP64 sums (their order fixes the rounding), over chunks of
    // the samples staged in LDS by the whole workgroup, so the one summing
    // thread never waits on a global load per sample.
    double sum = 0.0, sum_sq = 0.0;
    for (int c0 = 0; c0 < m; c0 += kStage) {
        const int len = m - c0 < kStage ? m - c0 : kStage;
        __syncthreads();  // the previous chunk has been summed
        for (int i = tid; i < len; i += kPrepThreads)
            stage[i] = raw[c0 + i];
        __syncthreads();
        if (tid == 0) {
#pragma unroll 8
            for (int i = 0; i < len; i++) {
                const double v = static_cast<double>(stage[i]);
                sum += v;
                sum_sq = __builtin_fma(v, v, sum_sq);
            }
        }
    }
    if (tid == 0) {
        const double mean = sum / static_cast<double>(m);
        double var = __builtin_fma(-mean, mean, sum_sq / static_cast<double>(m));
        if (var < 0.0)
            var = 0.0;
        s_mean = mean;
        s_sd = __builtin_sqrt(var);
    }
    __syncthreads();
    const double mean = s_mean, sd = s_sd;
    double qa = 0.0, qb = 0.0;
    for (int i = tid; i < m; i += kPrepThreads) {
        const float v = sd <= 1e-12 ? 0.0f : static_cast<float>((static_cast<double>(raw[i]) - mean) / sd);
        q[i] = v;
        qa += static_cast<double>(v);
        qb += static_cast<double>(v) * static_cast<double>(v);
    }
    red_a[tid] = qa;
    red_b[tid] = qb;
    __syncthreads();
    for (int s = kPrepThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red_a[tid] += red_a[tid + s];
            red_b[tid] += red_b[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        tmpl[t].S = red_a[0];
        tmpl[t].Q2 = red_b[0];
    }
    // Global writes of this workgroup are visible to its threads after the barrier.
    __threadfence_block();
    __syncthreads();
    const int blocks = m / kBlock;
    for (int b = tid; b < blocks; b += kPrepThreads) {
        float e = 0.0f;
        for (int i = b * kBlock; i < (b + 1) * kBlock; i++)
            e = __builtin_fmaf(q[i], q[i], e);
        energy[b] = e;
    }
    __syncthreads();
    // Stable descending order: rank = #(strictly larger) + #(equal, earlier).
    for (int b = tid; b < blocks; b += kPrepThreads) {
        const float e = energy[b];
        int rank = 0;
        for (int c = 0; c < blocks; c++) {
            const float f = energy[c];
            rank += (f > e) || (f == e && c < b);
        }
        order[rank] = b * kBlock;
    }
}

}  // namespace

hipError_t launch_prepare(const float *raw, TmplDev *tmpl, int count, float *qz, int32_t *order, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_prepare, dim3(count), dim3(kPrepThreads), 0, stream, raw, tmpl, qz, order);
    return hipGetLastError();
}

}  // namespace tmx
