// FP64 prefix sums of the waveform (x0 = x with missing samples as 0) and of
// x0^2, interleaved as pairs.  These are the resident statistics from which
// the inner-product passes derive every window's mean and standard deviation
// for any template length (with a rigorous error bound), replacing the
// reference's per-length two-pass statistics loop (SkipSequential.cpp:157-190,
// indexing.cpp:128-153) on the filtering side.  The exact two-pass statistics
// are recomputed only for candidate windows (verify.hip).
//
// Three launches, HBM-bound (4 bytes read, 16 bytes written per sample): chunk
// sums, a scan of the chunk sums in one workgroup, and the chunk-local scans
// with the chunk offsets added.  Every output is a sum of its terms through
// at most ~600 sequential or tree additions, well inside kPrefixGamma.  The
// waveform buffer carries zero padding past n, so chunk reads need no guard.
#include "tmx_internal.hpp"

namespace tmx {
namespace {

constexpr int kPThreads = 256;
constexpr int kPItems = 16;
constexpr int kPChunk = kPThreads * kPItems;  // samples per chunk
constexpr int kTopThreads = 1024;

// The thread's 16 consecutive samples (missing as 0), 64-byte aligned loads.
__device__ inline void load16(const float *x, uint64_t base, double *v)
{
    const float4 *src = reinterpret_cast<const float4 *>(x + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 f = src[q];
        const float e[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
        for (int u = 0; u < 4; u++)
            v[4 * q + u] = e[u] == kMissing ? 0.0 : static_cast<double>(e[u]);
    }
}

// Inclusive Hillis-Steele scan of two doubles per thread across the workgroup.
template <int N>
__device__ inline void block_scan2(double &a, double &b, double *sa, double *sb)
{
    const int tid = threadIdx.x;
    sa[tid] = a;
    sb[tid] = b;
    __syncthreads();
    for (int off = 1; off < N; off <<= 1) {
        double ta = 0.0, tb = 0.0;
        if (tid >= off) {
            ta = sa[tid - off];
            tb = sb[tid - off];
        }
        __syncthreads();
        if (tid >= off) {
            sa[tid] += ta;
            sb[tid] += tb;
        }
        __syncthreads();
    }
    a = sa[tid];
    b = sb[tid];
}

__global__ __launch_bounds__(kPThreads) void k_chunk_sums(const float *__restrict__ x, double *__restrict__ s1,
                                                          double *__restrict__ s2)
{
    __shared__ double sa[kPThreads], sb[kPThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kPChunk + static_cast<uint64_t>(threadIdx.x) * kPItems;
    double v[kPItems];
    load16(x, base, v);
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        a += v[e];
        b = fma(v[e], v[e], b);
    }
    block_scan2<kPThreads>(a, b, sa, sb);
    if (threadIdx.x == kPThreads - 1) {
        s1[blockIdx.x] = a;
        s2[blockIdx.x] = b;
    }
}

// Exclusive scan of the chunk sums (in place), one workgroup.
__global__ __launch_bounds__(kTopThreads) void k_chunk_scan(double *__restrict__ s1, double *__restrict__ s2,
                                                            uint64_t chunks)
{
    __shared__ double sa[kTopThreads], sb[kTopThreads];
    const uint64_t per = (chunks + kTopThreads - 1) / kTopThreads;
    const uint64_t lo = per * threadIdx.x;
    const uint64_t hi = lo + per < chunks ? lo + per : chunks;
    double a = 0.0, b = 0.0;
    for (uint64_t c = lo; c < hi; c++) {
        a += s1[c];
        b += s2[c];
    }
    block_scan2<kTopThreads>(a, b, sa, sb);
    double ra = threadIdx.x > 0 ? sa[threadIdx.x - 1] : 0.0;
    double rb = threadIdx.x > 0 ? sb[threadIdx.x - 1] : 0.0;
    for (uint64_t c = lo; c < hi; c++) {
        const double va = s1[c], vb = s2[c];
        s1[c] = ra;
        s2[c] = rb;
        ra += va;
        rb += vb;
    }
}

constexpr int kRound = 64 * kPItems;  // pairs staged per round (64 threads' worth)

__global__ __launch_bounds__(kPThreads) void k_chunk_final(const float *__restrict__ x, uint64_t n,
                                                           const double *__restrict__ off1,
                                                           const double *__restrict__ off2, double2 *__restrict__ p)
{
    __shared__ double sa[kPThreads], sb[kPThreads];
    __shared__ double2 stage[kRound];
    const int tid = threadIdx.x;
    const uint64_t cbase = static_cast<uint64_t>(blockIdx.x) * kPChunk;
    const uint64_t base = cbase + static_cast<uint64_t>(tid) * kPItems;
    double v[kPItems];
    load16(x, base, v);
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        a += v[e];
        b = fma(v[e], v[e], b);
    }
    block_scan2<kPThreads>(a, b, sa, sb);
    double ra = off1[blockIdx.x] + (tid > 0 ? sa[tid - 1] : 0.0);
    double rb = off2[blockIdx.x] + (tid > 0 ? sb[tid - 1] : 0.0);
    if (blockIdx.x == 0 && tid == 0)
        p[0] = make_double2(0.0, 0.0);
    double2 out[kPItems];
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        ra += v[e];
        rb = fma(v[e], v[e], rb);
        out[e] = make_double2(ra, rb);
    }
    // Transpose through LDS in rounds of 64 threads so that the stores of
    // consecutive threads are consecutive pairs.
    for (int q = 0; q < kPThreads / 64; q++) {
        __syncthreads();
        if ((tid >> 6) == q) {
#pragma unroll
            for (int e = 0; e < kPItems; e++)
                stage[(tid & 63) * kPItems + e] = out[e];
        }
        __syncthreads();
        for (int k = tid; k < kRound; k += kPThreads) {
            const uint64_t i = cbase + static_cast<uint64_t>(q) * kRound + static_cast<uint64_t>(k);
            if (i < n)
                p[i + 1] = stage[k];
        }
    }
}

// Block-rebased FP32 prefix pairs for the level-1 filter of the matrix-core
// pass: for waveform block b (first window J0 = b * block), centre c_b = mean
// of x0 over the block's windows, and for i in [0, len)
//   rb[b * len + i] = ( sum_{J0 <= k < J0 + i} (x0_k - c_b),
//                       sum_{J0 <= k < J0 + i} (x0_k - c_b)^2 )
// from the FP64 prefix sums, rounded to float.  binfo[b] = (c_b, E1, E2, 0):
// E1, E2 bound the absolute error of every entry's FP64 value (prefix-sum
// error, kPrefixGamma, plus the rounding of the rebasing), so that the FP32
// consumer can bound a window's statistics: entry error <= E + 2^-24 |entry|.
// len = 0: the centres and bounds only (rb untouched).
__global__ void k_block_prefix(const double2 *__restrict__ p, uint64_t n, uint64_t koff, int block, int len,
                               float4 *__restrict__ binfo, float2 *__restrict__ rb)
{
    constexpr double u64 = 1.1102230246251565e-16;
    const uint64_t b = blockIdx.y;
    const uint64_t J0 = b * static_cast<uint64_t>(block);
    const uint64_t jend = J0 + block < n ? J0 + block : n;
    const double2 p0 = p[J0];
    const float c = static_cast<float>((p[jend].x - p0.x) / static_cast<double>(jend - J0));
    const double cd = static_cast<double>(c);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t klast = J0 + static_cast<uint64_t>(len) < n ? J0 + static_cast<uint64_t>(len) : n;
        const double pmax = p[klast].y;  // P2 is non-decreasing: bounds every P2 of the block
        const double L = static_cast<double>(klast - J0);
        const double eP2 = kPrefixGamma * pmax;
        const double eP1 = kPrefixGamma * sqrt(static_cast<double>(koff + klast) * pmax);  // absolute index
        const double s1max = sqrt(L * pmax);
        const double E1 = 2.0 * eP1 + 4.0 * u64 * (s1max + fabs(cd) * L);
        const double E2 = 2.0 * eP2 + 4.0 * fabs(cd) * eP1 + 4.0 * u64 * (pmax + 2.0 * fabs(cd) * s1max + cd * cd * L);
        binfo[b] = make_float4(c, static_cast<float>(E1 * 1.01 + 1e-30), static_cast<float>(E2 * 1.01 + 1e-30), 0.0f);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        const uint64_t k = J0 + static_cast<uint64_t>(i) < n ? J0 + static_cast<uint64_t>(i) : n;
        const double2 pk = p[k];
        const double cnt = static_cast<double>(k - J0);
        const double s1 = pk.x - p0.x;
        const double s2 = pk.y - p0.y;
        const double t1 = s1 - cd * cnt;
        const double t2 = s2 - 2.0 * cd * s1 + cd * cd * cnt;
        rb[b * static_cast<uint64_t>(len) + static_cast<uint64_t>(i)] =
            make_float2(static_cast<float>(t1), static_cast<float>(t2));
    }
}

}  // namespace

hipError_t launch_block_prefix(const double2 *p, uint64_t n, int block, int len, float4 *binfo, float2 *rb,
                               hipStream_t stream)
{
    const uint64_t blocks = (n + block - 1) / block;
    if (blocks == 0)
        return hipErrorInvalidValue;
    unsigned gx = static_cast<unsigned>((len + 255) / 256 < 64 ? (len + 255) / 256 : 64);
    if (gx == 0)
        gx = 1;
    // grid.y is limited to 65535: larger waveforms take several launches.
    for (uint64_t b0 = 0; b0 < blocks; b0 += 65535) {
        const uint64_t nb = blocks - b0 < 65535 ? blocks - b0 : 65535;
        hipLaunchKernelGGL(k_block_prefix, dim3(gx, static_cast<unsigned>(nb)), dim3(256), 0, stream,
                           p + b0 * block, n - b0 * block, b0 * block, block, len, binfo + b0,
                           rb + b0 * static_cast<uint64_t>(len));
    }
    return hipGetLastError();
}

size_t prefix_work_len(uint64_t n)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    return static_cast<size_t>(2 * chunks + 2);
}

hipError_t launch_prefix(const float *x, uint64_t n, double2 *p, double *work, size_t work_len, hipStream_t stream)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    if (chunks == 0 || work_len < 2 * chunks)
        return hipErrorInvalidValue;
    double *s1 = work, *s2 = work + chunks;
    hipLaunchKernelGGL(k_chunk_sums, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, s1, s2);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(kTopThreads), 0, stream, s1, s2, chunks);
    hipLaunchKernelGGL(k_chunk_final, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, n, s1, s2,
                       p);
    return hipGetLastError();
}

}  // namespace tmx
