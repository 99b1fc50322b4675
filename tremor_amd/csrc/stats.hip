// FP64 prefix sums of the waveform (x0 = x with missing samples as 0) and of
// x0^2.  These are the resident statistics from which the inner-product
// passes derive every window's mean and standard deviation for any template
// length (with a rigorous error bound), replacing the reference's per-length
// two-pass statistics loop (SkipSequential.cpp:157-190, indexing.cpp:128-153)
// on the filtering side.  The exact two-pass statistics are recomputed only
// for candidate windows (verify.hip).
//
// Three launches, HBM-bound: chunk sums, a scan of the chunk sums in one
// workgroup, and the chunk-local scans with the chunk offsets added.  Every
// output is a sum of its terms through at most ~600 sequential or tree
// additions, well inside kPrefixGamma.
#include "tmx_internal.hpp"

namespace tmx {
namespace {

constexpr int kPThreads = 256;
constexpr int kPItems = 16;
constexpr int kPChunk = kPThreads * kPItems;  // samples per chunk
constexpr int kTopThreads = 1024;

__device__ inline double x0_at(const float *x, uint64_t i, uint64_t n)
{
    if (i >= n)
        return 0.0;
    const float v = x[i];
    return v == kMissing ? 0.0 : static_cast<double>(v);
}

// Inclusive Hillis-Steele scan of one double per thread across the workgroup.
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

__global__ __launch_bounds__(kPThreads) void k_chunk_sums(const float *__restrict__ x, uint64_t n,
                                                          double *__restrict__ s1, double *__restrict__ s2)
{
    __shared__ double sa[kPThreads], sb[kPThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kPChunk + static_cast<uint64_t>(threadIdx.x) * kPItems;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        const double v = x0_at(x, base + e, n);
        a += v;
        b = fma(v, v, b);
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
    double ia = a, ib = b;
    block_scan2<kTopThreads>(ia, ib, sa, sb);
    double ra = ia - a, rb = ib - b;  // exclusive prefix of this thread's range
    if (threadIdx.x > 0) {
        ra = sa[threadIdx.x - 1];
        rb = sb[threadIdx.x - 1];
    } else {
        ra = 0.0;
        rb = 0.0;
    }
    for (uint64_t c = lo; c < hi; c++) {
        const double va = s1[c], vb = s2[c];
        s1[c] = ra;
        s2[c] = rb;
        ra += va;
        rb += vb;
    }
}

__global__ __launch_bounds__(kPThreads) void k_chunk_final(const float *__restrict__ x, uint64_t n,
                                                           const double *__restrict__ off1,
                                                           const double *__restrict__ off2, double *__restrict__ p1,
                                                           double *__restrict__ p2)
{
    __shared__ double sa[kPThreads], sb[kPThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kPChunk + static_cast<uint64_t>(threadIdx.x) * kPItems;
    double v[kPItems];
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        v[e] = x0_at(x, base + e, n);
        a += v[e];
        b = fma(v[e], v[e], b);
    }
    double ia = a, ib = b;
    block_scan2<kPThreads>(ia, ib, sa, sb);
    double ra = off1[blockIdx.x] + (ia - a);
    double rb = off2[blockIdx.x] + (ib - b);
    if (threadIdx.x > 0) {
        ra = off1[blockIdx.x] + sa[threadIdx.x - 1];
        rb = off2[blockIdx.x] + sb[threadIdx.x - 1];
    } else {
        ra = off1[blockIdx.x];
        rb = off2[blockIdx.x];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p1[0] = 0.0;
        p2[0] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < kPItems; e++) {
        ra += v[e];
        rb = fma(v[e], v[e], rb);
        const uint64_t i = base + e;
        if (i < n) {
            p1[i + 1] = ra;
            p2[i + 1] = rb;
        }
    }
}

}  // namespace

size_t prefix_work_len(uint64_t n)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    return static_cast<size_t>(2 * chunks + 2);
}

hipError_t launch_prefix(const float *x, uint64_t n, double *p1, double *p2, double *work, size_t work_len,
                         hipStream_t stream)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    if (chunks == 0)
        return hipMemsetAsync(p1, 0, sizeof(double), stream) == hipSuccess
                   ? hipMemsetAsync(p2, 0, sizeof(double), stream)
                   : hipErrorInvalidValue;
    if (work_len < 2 * chunks)
        return hipErrorInvalidValue;
    double *s1 = work, *s2 = work + chunks;
    hipLaunchKernelGGL(k_chunk_sums, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, n, s1, s2);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(kTopThreads), 0, stream, s1, s2, chunks);
    hipLaunchKernelGGL(k_chunk_final, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, n, s1, s2,
                       p1, p2);
    return hipGetLastError();
}

}  // namespace tmx
