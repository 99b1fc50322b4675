// Plain FP32 sliding dot product with the fused candidate filter: the
// non-abandoning GPU restatement of the reference scan loop
// (SkipSequential.cpp:253-265, 317-333).  Every (template, window) pair costs
// m FP32 FMAs; it is the baseline the MFMA Hankel GEMM is measured against,
// and an exact path by itself once its candidates are verified.
//
// Workgroup: 256 threads, 4096 consecutive windows (16 per thread), one
// group of templates (sorted by length) looped inside the workgroup.  The
// centred waveform tile x0 - c is staged in LDS in chunks of 1024 template
// samples; each thread keeps a sliding register window of 32 samples and
// multiplies it by 16 template samples per step (256 FMAs per 4 LDS reads).
// Template samples are the same for all threads (LDS broadcast).
#include "filter.hpp"

namespace tmx {
namespace {

constexpr int kThreads = 256;
constexpr int kR = 16;                   // windows per thread
constexpr int kWB = kThreads * kR;       // windows per workgroup
constexpr int kKC = 1024;                // template samples per staged chunk
static_assert(kWB == kScanWindows, "window block size");

// LDS layout of the staged tile: 4 floats of padding after every 16, so the
// 16-byte reads of consecutive threads (80 bytes apart) hit distinct banks.
__device__ inline int xs_at(int e) { return e + ((e >> 4) << 2); }
constexpr int kXsLen = (kWB + kKC) + ((kWB + kKC) >> 2);

__device__ inline void emit(const CandOut &co, uint32_t t, uint32_t j, float est, double lo, double hi)
{
    cand_put(co, wave_claim(co.count), t, j, est, lo, hi);
}

__global__ __launch_bounds__(kThreads) void k_scan_fp32(Waveform w, const TmplDev *__restrict__ tmpl,
                                                        const float *__restrict__ qz,
                                                        const int32_t *__restrict__ group_begin, int mode,
                                                        int block_stride, CandOut co)
{
    __shared__ float xs[kXsLen];
    __shared__ float qs[kKC];
    const int tid = threadIdx.x;
    const int g = blockIdx.y;
    const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * static_cast<uint64_t>(block_stride);
    const uint64_t j0 = blk * kWB;
    if (j0 >= w.n)
        return;
    const uint64_t jend = j0 + kWB < w.n ? j0 + kWB : w.n;
    const float c = static_cast<float>((w.p[jend].x - w.p[j0].x) / static_cast<double>(jend - j0));
    const int tb = group_begin[g], te = group_begin[g + 1];
    for (int t = tb; t < te; t++) {
        const int m = tmpl[t].m;
        if (static_cast<uint64_t>(m) > w.n || tmpl[t].ea)
            continue;
        const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
        if (j0 >= W)
            continue;
        const int64_t qoff = tmpl[t].q_off;
        float acc[kR];
#pragma unroll
        for (int r = 0; r < kR; r++)
            acc[r] = 0.0f;
        for (int i0 = 0; i0 < m; i0 += kKC) {
            const int kc = m - i0 < kKC ? m - i0 : kKC;
            const int kcp = (kc + 15) & ~15;
            __syncthreads();
            for (int e = tid; e < kWB + kcp; e += kThreads) {
                const uint64_t p = j0 + static_cast<uint64_t>(i0) + static_cast<uint64_t>(e);
                float v = p < w.n ? w.x[p] : 0.0f;
                v = v == kMissing ? 0.0f : v;
                xs[xs_at(e)] = v - c;
            }
            for (int e = tid; e < kcp; e += kThreads)
                qs[e] = e < kc ? qz[qoff + i0 + e] : 0.0f;
            __syncthreads();
            const int base = tid * kR;
            float xv[2 * kR];
#pragma unroll
            for (int r = 0; r < kR; r++)
                xv[r] = xs[xs_at(base + r)];
            for (int ii = 0; ii < kcp; ii += kR) {
#pragma unroll
                for (int r = 0; r < kR; r++)
                    xv[kR + r] = xs[xs_at(base + ii + kR + r)];
#pragma unroll
                for (int s = 0; s < kR; s++) {
                    const float qv = qs[ii + s];
#pragma unroll
                    for (int r = 0; r < kR; r++)
                        acc[r] = __builtin_fmaf(qv, xv[r + s], acc[r]);
                }
#pragma unroll
                for (int r = 0; r < kR; r++)
                    xv[r] = xv[kR + r];
            }
        }
        // Fused filter.
        const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
        const double eta = eta_fp32(m);
        const double bound = static_cast<double>(tmpl[t].bound);
        float best = INFINITY;
        uint32_t best_j = 0xffffffffu;
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const uint64_t j = j0 + static_cast<uint64_t>(tid * kR + r);
            if (j >= W || w.x[j] == kMissing)
                continue;
            const DistInterval iv = dist_interval(w, j, m, Q2, S, c, static_cast<double>(acc[r]), eta);
            if (mode == kModeMain) {
                if (iv.lo <= bound)
                    emit(co, static_cast<uint32_t>(t), static_cast<uint32_t>(j), static_cast<float>(iv.est), iv.lo,
                         iv.hi);
            } else {
                const float hi = iv.hi < 3.0e38 ? static_cast<float>(iv.hi) : INFINITY;
                if (hi < best) {
                    best = hi;
                    best_j = static_cast<uint32_t>(j);
                }
            }
        }
        if (mode == kModeSeed) {
            // One entry per wavefront: the minimum upper bound of its 1024 windows.
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(best, off);
                const uint32_t oj = __shfl_xor(best_j, off);
                if (ov < best || (ov == best && oj < best_j)) {
                    best = ov;
                    best_j = oj;
                }
            }
            if ((tid & 63) == 0 && best < INFINITY)
                emit(co, static_cast<uint32_t>(t), best_j, best, best, best);
        }
    }
}

}  // namespace

hipError_t launch_scan_fp32(const Waveform &w, const TmplDev *tmpl, const float *qz, int count,
                            const int32_t *group_begin, int groups, int mode, int block_stride, const CandOut &co,
                            hipStream_t stream)
{
    (void)count;
    if (groups <= 0 || w.n == 0)
        return hipSuccess;
    const uint64_t blocks = (w.n + kWB - 1) / kWB;
    const uint64_t launched = (blocks + block_stride - 1) / block_stride;
    hipLaunchKernelGGL(k_scan_fp32, dim3(static_cast<unsigned>(launched), static_cast<unsigned>(groups)),
                       dim3(kThreads), 0, stream, w, tmpl, qz, group_begin, mode, block_stride, co);
    return hipGetLastError();
}

}  // namespace tmx
