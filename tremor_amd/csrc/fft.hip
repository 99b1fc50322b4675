// Overlap-save FFT (MASS) path in FP32 for long templates, in the role of the
// reference's FFT fallback (fft_fallback.hpp:4-10, :67-117; fft_fallback.cpp:
// 39-56, 111-176; MASS V3, DMASS_V3.c:805-848, 904-996): the windows are cut
// into pieces of B samples overlapping by Lmax - 1 (step = B - Lmax + 1), the
// spectrum of every centred piece is computed once per waveform, and a
// template costs one complex product and one inverse transform per piece.
// The FFT dot products only filter the windows: every window whose distance
// may be within the bound, allowing for the FFT error, is recomputed exactly.
//
// Transforms are hipFFT (FP32, unnormalised): r2c forward, and the real
// inverse of length B as a complex inverse of length B / 2 whose input the
// product kernel forms (k_product_z).  The kernels here prepare the pieces
// and templates, take the spectral norms for the error bound, form the
// products and run the fused candidate filter on the inverse transform.
//
// Error bound.  With X = F x~ (piece), H = F h (reversed template), y = F^-1
// (X .* H) / B, and forward / inverse transforms of relative 2-norm error
// eps = c log2(B) u each, the product rounding 2u, every output satisfies
//   |y_fft - y| <= (2 eps + 2u) ||x~||_2 ||H||_inf + eps ||h||_2 ||X||_inf
// (2-norm bounds of the error vector, hence of each element).  c = 10 with
// a further factor 4: the transform library is treated as a black box.  The
// inverse is the Z step (a table twiddle rounded from FP64, one complex
// product and three additions per bin: a few u of relative 2-norm error,
// since (Y[k], conj Y[M-k]) -> (Z[k], conj Z[M-k]) is sqrt 2 times a unitary
// map) followed by a transform of length B / 2, within c log2(B) u =
// c log2(B / 2) u + c u.
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

// The inverse real transform of Y = X[p] .* H[t] (B/2 + 1 bins, B = 2M) as an
// M-point complex inverse transform: with
//   Z[k] = (Y[k] + conj Y[M-k]) + i e^{i pi k / M} (Y[k] - conj Y[M-k]),  k < M,
// z = sum_k Z[k] e^{2 pi i r k / M} gives y[2r] = Re z[r], y[2r+1] = Im z[r],
// the unnormalised real inverse of length B (the layout a c2r writes).  One
// thread per bin pair (k, M - k), so the product is formed once per bin;
// E[k] = e^{i pi k / M} comes from the FP64-rounded table.
__global__ void k_product_z(const FftPair *__restrict__ pairs, const float2 *__restrict__ spectra,
                            const float2 *__restrict__ qspec, const int32_t *__restrict__ slot, int bins,
                            const float2 *__restrict__ E, float2 *__restrict__ Z)
{
    const FftPair pr = pairs[blockIdx.y];
    const int M = bins - 1;
    const float2 *X = spectra + static_cast<uint64_t>(pr.p) * bins;
    const float2 *H = qspec + static_cast<uint64_t>(slot[pr.t]) * bins;
    float2 *out = Z + static_cast<uint64_t>(blockIdx.y) * M;
    const auto cmul = [](float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    // Z[k] from Y[k] and Y[M - k]: A + i E[k] D with A = Yk + conj Ym, D = Yk - conj Ym.
    const auto zval = [&](float2 yk, float2 ym, float2 e) {
        const float2 A = make_float2(yk.x + ym.x, yk.y - ym.y);
        const float2 D = make_float2(yk.x - ym.x, yk.y + ym.y);
        const float2 ED = cmul(e, D);
        return make_float2(A.x - ED.y, A.y + ED.x);
    };
    for (int k = blockIdx.x * kThreads + threadIdx.x; k <= M / 2; k += gridDim.x * kThreads) {
        const int km = M - k;
        const float2 yk = cmul(X[k], H[k]), ym = cmul(X[km], H[km]);
        out[k] = zval(yk, ym, E[k]);
        if (k > 0 && km != k)
            out[km] = zval(ym, yk, E[km]);
    }
}

// Fused candidate filter on one batch of inverse transforms.  Block
// (chunk, pair b): windows w = p * step + i, i in the chunk.
__global__ __launch_bounds__(kThreads) void k_fft_filter(
    Waveform w, const TmplDev *__restrict__ tmpl, const FftPair *__restrict__ pairs, const float *__restrict__ y,
    int B, int step, const float2 *__restrict__ rb, int rb_len, const float4 *__restrict__ pinfo,
    const float2 *__restrict__ pnorm, const float *__restrict__ qmax, const int32_t *__restrict__ slot, double kfft,
    int mode, CandOut co)
{
    __shared__ float s_ub[kThreads / 64];
    __shared__ uint32_t s_j[kThreads / 64];
    const FftPair pr = pairs[blockIdx.y];
    const int t = static_cast<int>(pr.t);
    const int m = tmpl[t].m;
    if (static_cast<uint64_t>(m) > w.n)
        return;
    const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
    const uint64_t base = static_cast<uint64_t>(pr.p) * static_cast<uint64_t>(step);
    const float4 bi = pinfo[pr.p];
    const float c = bi.x;
    const float2 *rbp = rb + static_cast<uint64_t>(pr.p) * static_cast<uint64_t>(rb_len);
    const float *yb = y + static_cast<uint64_t>(blockIdx.y) * B;
    const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
    const double md = static_cast<double>(m);
    const double bound = static_cast<double>(tmpl[t].bound);
    constexpr double u = 5.9604644775390625e-08;
    // Absolute FFT error of every dot product of this (template, piece).
    const double eabs = kfft * (static_cast<double>(pnorm[pr.p].x) * static_cast<double>(qmax[slot[t]]) +
                                sqrt(Q2) * static_cast<double>(pnorm[pr.p].y));
    const double eta = 1.01 * u;  // rounding of x~ = x0 - c_p
    const double eref = ([&] {
        const double rm = sqrt(md), sq = 2.0 * rm + 1.0;
        const double D1 = (u * 1.001) * (3.05 * rm + 1.03 * rm * 1.001e3 + sq);
        const double neff = floor(md / 16.0) + 36.0, gam = neff * u / (1.0 - neff * u);
        return (D1 * (2.0 * sq + D1) + gam * (sq + D1) * (sq + D1)) * 1.02;
    })();
    const float gmin = static_cast<float>((Q2 + md - bound - eref) * 0.5 * (1.0 - 1e-3));
    const float b0 = static_cast<float>(eabs * (1.0 + 1e-3));
    const float eta_sq = static_cast<float>(eta * sqrt(Q2) * (1.0 + 1e-3));
    const float sqrt_m = static_cast<float>(sqrt(md));
    const float rsqrt_m = static_cast<float>(1.0 / sqrt(md)) * (1.0f + 1e-6f);
    const float absS = static_cast<float>(fabs(S) * (1.0 + 1e-3));
    const float mf = static_cast<float>(m), inv_mf = 1.0f / mf;
    const float invB = 1.0f / static_cast<float>(B);
    float best_g = -INFINITY;
    uint32_t best_j = 0xffffffffu;
    float best_dt = 0.0f;
    // Windows i of this block's share of the piece that lie in the waveform
    // (j = base + i < W).  A counted loop: every read below is in bounds for
    // i < step (m - 1 + i < B, i + m <= B), so the reads of several windows
    // are issued together.
    const int i0 = blockIdx.x * (kThreads * 16);
    const int iend = i0 + kThreads * 16 < step ? i0 + kThreads * 16 : step;
    const uint64_t iw = W > base ? W - base : 0;
    const int ilim = static_cast<uint64_t>(iend) < iw ? iend : static_cast<int>(iw);
    // One window ahead: the next window's reads are issued before this
    // window's arithmetic.
    int i = i0 + static_cast<int>(threadIdx.x);
    float ny = 0.0f;
    float2 nra = make_float2(0.0f, 0.0f), nrz = make_float2(0.0f, 0.0f);
    if (i < ilim) {
        ny = yb[m - 1 + i];
        nra = rbp[i];
        nrz = rbp[i + m];
    }
    for (; i < ilim; i += kThreads) {
        const uint64_t j = base + static_cast<uint64_t>(i);
        const float dt = ny * invB;
        const float2 ra = nra, rz = nrz;
        if (i + kThreads < ilim) {
            ny = yb[m - 1 + i + kThreads];
            nra = rbp[i + kThreads];
            nrz = rbp[i + kThreads + m];
        }
        const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
        const float eS1 = 2.0f * bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
        const float eS2 = 2.0f * bi.z + 1.2e-7f * (fabsf(ra.y) + fabsf(rz.y) + fabsf(S2));
        const float V = S2 - S1 * S1 * inv_mf;
        const float eV = eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * inv_mf + 2.4e-7f * (fabsf(S2) + S1 * S1 * inv_mf);
        const float aD = fabsf(S1 * inv_mf) + eS1 * inv_mf;
        const float amu = fabsf(c) + aD;
        const bool judge = V > 1e3f * eV && V > mf * 1e-20f && !(amu * amu * mf > 1e6f * V);
        if (mode == kModeMain) {
            bool survive = !judge || gmin <= 0.0f;
            if (!survive) {
                const float sV = V * __frsqrt_rn(V);
                const float b = b0 + eta_sq * (sV + sqrt_m * aD) + aD * absS;
                survive = dt >= gmin * sV * rsqrt_m * 0.999f - b * 1.001f - 1e-30f;
            }
            if (survive && w.x[j] != kMissing) {
                const DistInterval iv = dist_interval(w, j, m, Q2, S, c, static_cast<double>(dt), eta, eabs);
                if (!(iv.lo > bound))  // NaN (undecidable): a candidate
                    cand_put(co, wave_claim(co.count), static_cast<uint32_t>(t), static_cast<uint32_t>(j),
                             static_cast<float>(iv.est), iv.lo, iv.hi);
            }
        } else if (judge && V > 0.0f) {
            const float g = dt * __frsqrt_rn(V);
            if (g > best_g) {
                best_g = g;
                best_j = static_cast<uint32_t>(j);
                best_dt = dt;
            }
        }
    }
    if (mode == kModeSeed) {
        float ub = INFINITY;
        if (best_j != 0xffffffffu && w.x[best_j] != kMissing) {
            const DistInterval iv = dist_interval(w, best_j, m, Q2, S, c, static_cast<double>(best_dt), eta, eabs);
            if (iv.hi < 3.0e38)
                ub = static_cast<float>(iv.hi) * (1.0f + 1e-6f);
        }
        uint32_t bj = best_j;
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(ub, off);
            const uint32_t oj = __shfl_xor(bj, off);
            if (ov < ub || (ov == ub && oj < bj)) {
                ub = ov;
                bj = oj;
            }
        }
        if ((threadIdx.x & 63) == 0) {
            s_ub[threadIdx.x >> 6] = ub;
            s_j[threadIdx.x >> 6] = bj;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < kThreads / 64; q++)
                if (s_ub[q] < ub || (s_ub[q] == ub && s_j[q] < bj)) {
                    ub = s_ub[q];
                    bj = s_j[q];
                }
            if (ub < INFINITY)
                cand_put(co, atomicAdd(co.count, 1ULL), static_cast<uint32_t>(t), bj, ub, ub, ub);
        }
    }
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

hipError_t launch_product_z(const FftPair *pairs, int npairs, const float2 *spectra, const float2 *qspec,
                            const int32_t *slot, int bins, const float2 *E, float2 *Z, hipStream_t stream)
{
    if (npairs <= 0)
        return hipSuccess;
    const int half = (bins - 1) / 2 + 1;
    const int gx = (half + kThreads * 8 - 1) / (kThreads * 8);
    hipLaunchKernelGGL(k_product_z, dim3(gx, npairs), dim3(kThreads), 0, stream, pairs, spectra, qspec, slot, bins, E,
                       Z);
    return hipGetLastError();
}

hipError_t launch_fft_filter(const Waveform &w, const TmplDev *tmpl, const FftPair *pairs, int npairs, const float *y,
                             int B, int step, const float2 *rb, int rb_len, const float4 *pinfo, const float2 *pnorm,
                             const float *qmax, const int32_t *slot, double kfft, int mode, const CandOut &co,
                             hipStream_t stream)
{
    if (npairs <= 0)
        return hipSuccess;
    const int gx = (step + kThreads * 16 - 1) / (kThreads * 16);
    hipLaunchKernelGGL(k_fft_filter, dim3(gx, npairs), dim3(kThreads), 0, stream, w, tmpl, pairs, y, B, step, rb,
                       rb_len, pinfo, pnorm, qmax, slot, kfft, mode, co);
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
