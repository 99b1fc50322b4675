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

}  // namespace

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
