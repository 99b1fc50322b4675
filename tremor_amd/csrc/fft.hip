// Shared parts of the overlap-save FFT (MASS) forms, in the role of the
// reference's FFT fallback (fft_fallback.hpp:4-10, :67-117; fft_fallback.cpp:
// 39-56, 111-176; MASS V3, DMASS_V3.c:805-848, 904-996): the windows are cut
// into pieces of B samples overlapping by Lmax - 1 (step = B - Lmax + 1), the
// spectrum of every centred piece is computed once per waveform (hipFFT r2c,
// FP32, unnormalised), and a template costs one complex product and one
// inverse transform per piece (fft_fused.hip for B = 2^14 and 2^15,
// fft_large.hip for B = 2^16 .. 2^18).  The FFT dot products only filter the
// windows: every window whose distance may be within the bound, allowing for
// the FFT error, is recomputed exactly.  The kernels here prepare the pieces
// and templates and take the spectral norms for the error bound.
//
// Error bound.  With X = F x~ (piece), H = F h (reversed template), y = F^-1
// (X .* H) / B, and forward / inverse transforms of relative 2-norm error
// eps each, the product rounding 2u, every output satisfies
//   |y_fft - y| <= (2 eps + 2u) ||x~||_2 ||H||_inf + eps ||h||_2 ||X||_inf
// (2-norm bounds of the error vector, hence of each element).  The forward
// transforms are hipFFT's, not analysed here: each is allowed eps = 4 x 10
// log2(B) u (measured on adversarial pieces at every B: about 3 u, see the
// probe test).  The inverses are this library's own, with derived bounds
// below that share (fft_fused.hip, fft_large.hip).
#include <hipfft/hipfft.h>

#include "filter.hpp"

namespace tmx {

namespace {

constexpr int kThreads = 256;

// Pieces p0 .. p0 + cnt: x~ = x0 - c_p for samples inside the waveform, 0 past
// its end; sum of squares of each piece into norm2[p] (FP64 atomics) and
// max |x~| into pinfo[p].w (float bits, non-negative: integer max).
__global__ void k_piece_input(const float *__restrict__ x, uint64_t n, int step, int B, uint64_t p0,
                              float4 *__restrict__ pinfo, float *__restrict__ out, double *__restrict__ norm2)
{
    __shared__ double red[kThreads];
    __shared__ float redm[kThreads];
    const uint64_t p = p0 + blockIdx.y;
    const float c = pinfo[p].x;
    const uint64_t base = p * static_cast<uint64_t>(step);
    double s = 0.0;
    float amax = 0.0f;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < B; k += gridDim.x * kThreads) {
        const uint64_t pos = base + static_cast<uint64_t>(k);
        float v = 0.0f;
        if (pos < n) {
            const float r = x[pos];
            v = (r == kMissing ? 0.0f : r) - c;
        }
        out[static_cast<uint64_t>(blockIdx.y) * B + k] = v;
        s += static_cast<double>(v) * static_cast<double>(v);
        amax = fmaxf(amax, fabsf(v));
    }
    red[threadIdx.x] = s;
    redm[threadIdx.x] = amax;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[threadIdx.x] += red[threadIdx.x + o];
            redm[threadIdx.x] = fmaxf(redm[threadIdx.x], redm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&norm2[p], red[0]);
        atomicMax(reinterpret_cast<unsigned int *>(&pinfo[p].w), __float_as_uint(redm[0]));
    }
}

// Template t of the FFT set: the reversed z-normalised template, zero padded.
__global__ void k_tmpl_input(const TmplDev *__restrict__ tmpl, const int32_t *__restrict__ ids, int B,
                             const float *__restrict__ qz, float *__restrict__ out)
{
    const int t = ids[blockIdx.y];
    const int m = tmpl[t].m;
    const float *q = qz + tmpl[t].q_off;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < B; k += gridDim.x * kThreads)
        out[static_cast<uint64_t>(blockIdx.y) * B + k] = k < m ? q[m - 1 - k] : 0.0f;
}

// Largest magnitude of each spectrum (rows of `bins` complex values).
__global__ void k_spec_max(const float2 *__restrict__ spec, int bins, float *__restrict__ out, int stride_out)
{
    __shared__ float red[kThreads];
    const float2 *row = spec + static_cast<uint64_t>(blockIdx.x) * bins;
    float mx = 0.0f;
    for (int k = threadIdx.x; k < bins; k += kThreads) {
        const float2 v = row[k];
        mx = fmaxf(mx, sqrtf(v.x * v.x + v.y * v.y));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[static_cast<uint64_t>(blockIdx.x) * stride_out] = red[0] * (1.0f + 1e-5f);
}

__global__ void k_zero_f64(double *__restrict__ p, uint64_t count)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < count)
        p[i] = 0.0;
}

// ||x~_p||_2 from the accumulated squares (rounded up).
__global__ void k_piece_norm(const double *__restrict__ norm2, float2 *__restrict__ pnorm, uint64_t P)
{
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (p < P)
        pnorm[p].x = static_cast<float>(sqrt(norm2[p]) * (1.0 + 1e-6));
}

}  // namespace

hipError_t launch_piece_input(const float *x, uint64_t n, int step, int B, uint64_t p0, int cnt,
                              float4 *pinfo, float *out, double *norm2, hipStream_t stream)
{
    const int gx = (B + kThreads * 16 - 1) / (kThreads * 16);
    hipLaunchKernelGGL(k_piece_input, dim3(gx, cnt), dim3(kThreads), 0, stream, x, n, step, B, p0, pinfo, out, norm2);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, uint64_t count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_f64, dim3(static_cast<unsigned>((count + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       stream, p, count);
    return hipGetLastError();
}

hipError_t launch_piece_norm(const double *norm2, float2 *pnorm, uint64_t P, hipStream_t stream)
{
    hipLaunchKernelGGL(k_piece_norm, dim3(static_cast<unsigned>((P + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       stream, norm2, pnorm, P);
    return hipGetLastError();
}

hipError_t launch_spec_max(const float2 *spec, int rows, int bins, float *out, int stride_out, hipStream_t stream)
{
    if (rows <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_spec_max, dim3(rows), dim3(kThreads), 0, stream, spec, bins, out, stride_out);
    return hipGetLastError();
}

hipError_t launch_tmpl_input(const TmplDev *tmpl, const int32_t *ids, int count, int B, const float *qz, float *out,
                             hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    const int gx = (B + kThreads * 16 - 1) / (kThreads * 16);
    hipLaunchKernelGGL(k_tmpl_input, dim3(gx, count), dim3(kThreads), 0, stream, tmpl, ids, B, qz, out);
    return hipGetLastError();
}

// kfft of the header comment: 4 (c log2 B + ...) u with c = 10 for each
// transform and 2u for the product.
double fft_error_factor(int B)
{
    double lg = 0.0;
    while ((1 << static_cast<int>(lg)) < B)
        lg += 1.0;
    return 4.0 * (2.0 * 10.0 * lg + 2.0) * 5.9604644775390625e-08;
}

}  // namespace tmx
