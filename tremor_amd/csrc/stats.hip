// Double-double prefix sums of the waveform (x0 = x with missing samples as
// 0) and of x0^2: high parts interleaved as pairs in p, low parts in pl.
// These are the resident statistics from which the inner-product passes
// derive every window's mean and standard deviation for any template length
// (with a rigorous error bound), replacing the reference's per-length
// two-pass statistics loop (SkipSequential.cpp:157-190, indexing.cpp:128-153)
// on the filtering side.  The exact two-pass statistics are recomputed only
// for candidate windows (verify.hip).
//
// Double-double rather than FP64: a window's sums are differences of two
// prefix values, and after a loud stretch (a x1e6 burst, say) an FP64 prefix
// carries an absolute error far above the sums of the quiet windows that
// follow, so none of them could be judged and every one would become a
// candidate.  With about 2^-94 relative accuracy the quiet windows resolve.
//
// Three launches, HBM-bound (4 bytes read, 32 bytes written per sample): chunk
// sums, a scan of the chunk sums in one workgroup, and the chunk-local scans
// with the chunk offsets added.  Every output is a sum of its terms through
// at most ~600 sequential or tree double-double additions (kPrefixGamma).
// x0 and x0^2 are exact in FP64.  The waveform buffer carries zero padding
// past n, so chunk reads need no guard.
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

// One (sum x0, sum x0^2) pair of double-double values.
struct Pair {
    double2 a, b;  // (hi, lo) each
};

__device__ inline Pair pair_add(const Pair &p, const Pair &q)
{
    return Pair{dd_add(p.a, q.a), dd_add(p.b, q.b)};
}

__device__ inline Pair pair_of(double v)
{
    return Pair{make_double2(v, 0.0), make_double2(v * v, 0.0)};  // v * v is exact (v is a float)
}

// Inclusive Hillis-Steele scan of one pair per thread across the workgroup.
template <int N>
__device__ inline Pair block_scan(Pair p, Pair *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = p;
    __syncthreads();
    for (int off = 1; off < N; off <<= 1) {
        Pair t{};
        if (tid >= off)
            t = sh[tid - off];
        __syncthreads();
        if (tid >= off) {
            p = pair_add(t, p);
            sh[tid] = p;
        }
        __syncthreads();
    }
    return sh[tid];
}

__device__ inline Pair thread_sum(const double *v)
{
    Pair s{};
#pragma unroll
    for (int e = 0; e < kPItems; e++)
        s = pair_add(s, pair_of(v[e]));
    return s;
}

// Chunk totals, as double-double: work[4 c .. 4 c + 3] = (a.hi, a.lo, b.hi, b.lo).
__global__ __launch_bounds__(kPThreads) void k_chunk_sums(const float *__restrict__ x, double *__restrict__ work)
{
    __shared__ Pair sh[kPThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kPChunk + static_cast<uint64_t>(threadIdx.x) * kPItems;
    double v[kPItems];
    load16(x, base, v);
    const Pair s = block_scan<kPThreads>(thread_sum(v), sh);
    if (threadIdx.x == kPThreads - 1) {
        double *o = work + 4 * static_cast<uint64_t>(blockIdx.x);
        o[0] = s.a.x;
        o[1] = s.a.y;
        o[2] = s.b.x;
        o[3] = s.b.y;
    }
}

// Exclusive scan of the chunk totals (in place), one workgroup.
__global__ __launch_bounds__(kTopThreads) void k_chunk_scan(double *__restrict__ work, uint64_t chunks)
{
    __shared__ Pair sh[kTopThreads];
    const uint64_t per = (chunks + kTopThreads - 1) / kTopThreads;
    const uint64_t lo = per * threadIdx.x;
    const uint64_t hi = lo + per < chunks ? lo + per : chunks;
    auto at = [&](uint64_t c) {
        const double *o = work + 4 * c;
        return Pair{make_double2(o[0], o[1]), make_double2(o[2], o[3])};
    };
    Pair s{};
    for (uint64_t c = lo; c < hi; c++)
        s = pair_add(s, at(c));
    const Pair inc = block_scan<kTopThreads>(s, sh);
    // Exclusive offset of this thread's first chunk.
    Pair r = threadIdx.x > 0 ? sh[threadIdx.x - 1] : Pair{};
    for (uint64_t c = lo; c < hi; c++) {
        const Pair v = at(c);
        double *o = work + 4 * c;
        o[0] = r.a.x;
        o[1] = r.a.y;
        o[2] = r.b.x;
        o[3] = r.b.y;
        r = pair_add(r, v);
    }
    (void)inc;
}

constexpr int kRound = 64 * kPItems;  // entries staged per round (64 threads' worth)

__global__ __launch_bounds__(kPThreads) void k_chunk_final(const float *__restrict__ x, uint64_t n,
                                                           const double *__restrict__ work,
                                                           double2 *__restrict__ p, double2 *__restrict__ pl)
{
    __shared__ Pair sh[kPThreads];
    __shared__ double2 stage_h[kRound];
    __shared__ double2 stage_l[kRound];
    const int tid = threadIdx.x;
    const uint64_t cbase = static_cast<uint64_t>(blockIdx.x) * kPChunk;
    const uint64_t base = cbase + static_cast<uint64_t>(tid) * kPItems;
    double v[kPItems];
    load16(x, base, v);
    block_scan<kPThreads>(thread_sum(v), sh);
    const double *o = work + 4 * static_cast<uint64_t>(blockIdx.x);
    Pair r{make_double2(o[0], o[1]), make_double2(o[2], o[3])};
    if (tid > 0)
        r = pair_add(r, sh[tid - 1]);
    if (blockIdx.x == 0 && tid == 0) {
        p[0] = make_double2(0.0, 0.0);
        pl[0] = make_double2(0.0, 0.0);
    }
    // Transpose through LDS in rounds of 64 threads so that the stores of
    // consecutive threads are consecutive pairs; each thread's running sums
    // advance one sample at a time as its round comes.
    for (int q = 0; q < kPThreads / 64; q++) {
        __syncthreads();
        if ((tid >> 6) == q) {
#pragma unroll
            for (int e = 0; e < kPItems; e++) {
                r = pair_add(r, pair_of(v[e]));
                stage_h[(tid & 63) * kPItems + e] = make_double2(r.a.x, r.b.x);
                stage_l[(tid & 63) * kPItems + e] = make_double2(r.a.y, r.b.y);
            }
        }
        __syncthreads();
        for (int k = tid; k < kRound; k += kPThreads) {
            const uint64_t i = cbase + static_cast<uint64_t>(q) * kRound + static_cast<uint64_t>(k);
            if (i < n) {
                p[i + 1] = stage_h[k];
                pl[i + 1] = stage_l[k];
            }
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
__global__ void k_block_prefix(const double2 *__restrict__ p, const double2 *__restrict__ pl, uint64_t n,
                               uint64_t koff, int block, int len, float4 *__restrict__ binfo,
                               float2 *__restrict__ rb)
{
    constexpr double u64 = 1.1102230246251565e-16;
    const uint64_t b = blockIdx.y;
    const uint64_t J0 = b * static_cast<uint64_t>(block);
    const uint64_t jend = J0 + block < n ? J0 + block : n;
    const double2 p0 = p[J0], pl0 = pl[J0];
    const double2 pe = p[jend], ple = pl[jend];
    const float c = static_cast<float>(((pe.x - p0.x) + (ple.x - pl0.x)) / static_cast<double>(jend - J0));
    const double cd = static_cast<double>(c);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t klast = J0 + static_cast<uint64_t>(len) < n ? J0 + static_cast<uint64_t>(len) : n;
        const double pmax = p[klast].y * (1.0 + 4.0 * u64);  // P2 is non-decreasing: bounds every P2 up to klast
        // The block's own sum of x0^2 (non-decreasing in the entry index):
        // bounds every entry's s2 and, with L entries, every |s1|.
        const double loc2 = ((p[klast].y - p0.y) + (pl[klast].y - pl0.y)) * (1.0 + 8.0 * u64) + pmax * 1e-30;
        const double L = static_cast<double>(klast - J0);
        // Prefix values' error (absolute index), then each entry's three FP64
        // roundings in the differences and the rebasing.
        const double eP2 = kPrefixGamma * pmax;
        const double eP1 = kPrefixGamma * sqrt(static_cast<double>(koff + klast) * pmax);
        const double s1max = sqrt(L * loc2);
        const double E1 = 2.0 * eP1 + 6.0 * u64 * (s1max + fabs(cd) * L);
        const double E2 = 2.0 * eP2 + 4.0 * fabs(cd) * eP1 + 6.0 * u64 * (loc2 + 2.0 * fabs(cd) * s1max + cd * cd * L);
        binfo[b] = make_float4(c, static_cast<float>(E1 * 1.01 + 1e-30), static_cast<float>(E2 * 1.01 + 1e-30), 0.0f);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        const uint64_t k = J0 + static_cast<uint64_t>(i) < n ? J0 + static_cast<uint64_t>(i) : n;
        const double2 pk = p[k], plk = pl[k];
        const double cnt = static_cast<double>(k - J0);
        const double s1 = (pk.x - p0.x) + (plk.x - pl0.x);
        const double s2 = (pk.y - p0.y) + (plk.y - pl0.y);
        const double t1 = s1 - cd * cnt;
        const double t2 = s2 - 2.0 * cd * s1 + cd * cd * cnt;
        rb[b * static_cast<uint64_t>(len) + static_cast<uint64_t>(i)] =
            make_float2(static_cast<float>(t1), static_cast<float>(t2));
    }
}

// Largest |x0 - c_b| over the len samples of each block b (x0 = x with
// missing samples as 0, the zero padding past n included), into binfo[b].w
// (float bits, non-negative: an integer max), after k_block_prefix set c_b.
__global__ void k_block_absmax(const float *__restrict__ x, int block, int len, float4 *__restrict__ binfo)
{
    const uint64_t b = blockIdx.y;
    const uint64_t J0 = b * static_cast<uint64_t>(block);
    const float c = binfo[b].x;
    float mx = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        float v = x[J0 + static_cast<uint64_t>(i)];
        v = v == kMissing ? 0.0f : v;
        mx = fmaxf(mx, fabsf(v - c));
    }
    for (int off = 32; off > 0; off >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned int *>(&binfo[b].w), __float_as_uint(mx));
}

}  // namespace

hipError_t launch_block_absmax(const float *x, uint64_t n, int block, int len, float4 *binfo, hipStream_t stream)
{
    const uint64_t blocks = (n + block - 1) / block;
    if (blocks == 0)
        return hipErrorInvalidValue;
    unsigned gx = static_cast<unsigned>((len + 255) / 256 < 16 ? (len + 255) / 256 : 16);
    if (gx == 0)
        gx = 1;
    for (uint64_t b0 = 0; b0 < blocks; b0 += 65535) {
        const uint64_t nb = blocks - b0 < 65535 ? blocks - b0 : 65535;
        hipLaunchKernelGGL(k_block_absmax, dim3(gx, static_cast<unsigned>(nb)), dim3(256), 0, stream,
                           x + b0 * block, block, len, binfo + b0);
    }
    return hipGetLastError();
}

hipError_t launch_block_prefix(const double2 *p, const double2 *pl, uint64_t n, int block, int len, float4 *binfo,
                               float2 *rb, hipStream_t stream)
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
                           p + b0 * block, pl + b0 * block, n - b0 * block, b0 * block, block, len, binfo + b0,
                           rb + b0 * static_cast<uint64_t>(len));
    }
    return hipGetLastError();
}

size_t prefix_work_len(uint64_t n)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    return static_cast<size_t>(4 * chunks + 4);
}

hipError_t launch_prefix(const float *x, uint64_t n, double2 *p, double2 *pl, double *work, size_t work_len,
                         hipStream_t stream)
{
    const uint64_t chunks = (n + kPChunk - 1) / kPChunk;
    if (chunks == 0 || work_len < 4 * chunks)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_chunk_sums, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, work);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(kTopThreads), 0, stream, work, chunks);
    hipLaunchKernelGGL(k_chunk_final, dim3(static_cast<unsigned>(chunks)), dim3(kPThreads), 0, stream, x, n, work,
                       p, pl);
    return hipGetLastError();
}

}  // namespace tmx
