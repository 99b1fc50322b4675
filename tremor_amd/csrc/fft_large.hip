// The FFT form for long templates (FFT length B = 2^16, 2^17, 2^18): the
// inverse real transform of every (piece, template) product Y = X_p .* H_t
// and the candidate filter on its outputs, hand-written in two kernels.  In
// the reference's role of the MASS overlap-save pass (fft_fallback.hpp:86-116,
// fft_fallback.cpp:111-176; MASS V3, DMASS_V3.c:944-965).
//
// The real inverse of length B is a complex inverse of length M = B / 2 of
//   Z[k] = (Y[k] + conj Y[M-k]) + i e^{i pi k / M} (Y[k] - conj Y[M-k]),
// z[r] = sum_k Z[k] w_M^{rk} (w_M = e^{2 pi i / M}), y[2r] = Re z[r],
// y[2r+1] = Im z[r].  With M = 256 N2 (N2 = 128, 256, 512), k = k2 + N2 k1 and
// r = r1 + 256 r2:
//   z[r1 + 256 r2] = sum_k2 w_N2^{r2 k2} w_M^{r1 k2} sum_k1 Z[k2 + N2 k1] w_256^{r1 k1}.
// Kernel A: per column k2 (16 per workgroup) the product, Z and the
// 256-point transform over k1, times w_M^{r1 k2}, into U[r1][k2] (one M-point
// array per pair in device memory).  Kernel B: per row r1 (4096 / N2 per
// workgroup) the N2-point transform over k2, then the filter on the row's
// outputs straight from LDS: no B-sized array is written after U.
//
// The small transforms are Stockham-style in LDS: a length-N transform with
// N = 16 Q is a radix-16 pass (16 elements per thread, in registers), a
// twiddle w_N^{n1 k2} and a length-Q pass in registers; the register
// transforms are radix-2 after a bit reversal, their twiddles from LDS copies
// of tables rounded from FP64 (after E in the class's table: w_256^t,
// w_M^t for t < 512, w_M^{512 t}, w_N2^t).
//
// Error of this inverse.  Counting, as in Higham's bound for radix-2
// transforms, every radix-2 level and every twiddle multiplication as one
// level of relative 2-norm error mu + gamma_4 (sqrt 2 + mu) with mu <= 4u
// (table twiddles are rounded once; the four-step twiddle w_M^e is the product
// of two table values): 8 + 1 levels for the 256-point transforms, 1 for the
// four-step twiddle, log2 N2 + 1 <= 10 for the N2-point ones, 20 in all, about
// 194 u, inside the share of the filter's allowance that the inverse has
// (4 x 10 log2 B u, fft.hip).
#include "filter.hpp"

namespace tmx {
namespace {

constexpr int kThreads = 256;
constexpr int kN1 = 256;     // length of the first-step transforms
constexpr int kCols = 16;    // columns (first-step transforms) per workgroup of kernel A
constexpr int kElems = 4096;  // complex values per workgroup of either kernel

// Complex product with fused multiply-adds: one rounded product and one
// rounded fused sum per component (relative error at most 2u of |a||b|, inside
// the per-level term below).
__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Reads through a buffer resource: a 32-bit per-thread byte offset plus a
// per-read constant in a scalar register, no 64-bit address per read.
__device__ inline __amdgpu_buffer_rsrc_t rsrc_of(const float2 *p, uint32_t count)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(p), static_cast<short>(0),
                                             static_cast<int>(count * 8u), 0x00020000);
}
__device__ inline float2 ld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

constexpr int rev_bits(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; b++)
        r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

constexpr int log2c(int n) { return n <= 1 ? 0 : 1 + log2c(n / 2); }

// v[k] <- sum_n v[n] w_R^{nk} in registers (R a power of two, <= 32), with
// w_R^j = tw[j * (S / R)] and tw[t] = e^{2 pi i t / S}.
template <int R, int S>
__device__ __forceinline__ void dft_reg(float2 *v, const float2 *tw)
{
    constexpr int L = log2c(R);
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int j = rev_bits(i, L);
        if (j > i) {
            const float2 t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; j++) {
                const float2 a = v[i + j];
                const float2 x = v[i + j + len / 2];
                // w^{S/4} = i exactly: a swap instead of a table product.
                const float2 b = j == 0 ? x : 4 * j == len ? make_float2(-x.y, x.x) : cmul(tw[j * (S / len)], x);
                v[i + j] = cadd(a, b);
                v[i + j + len / 2] = csub(a, b);
            }
        }
    }
}

// Work item of this workgroup: workgroups are dispatched round-robin over
// the 8 XCDs, so workgroup L runs on XCD L mod 8; giving XCD x the
// consecutive items [x chunk, (x + 1) chunk) keeps the pairs of one piece
// (consecutive: pairs are piece-major) on one XCD, where the piece's
// spectrum and prefix pairs are read once into its L2.  -1: past the end.
__device__ inline int xcd_item(int total)
{
    const int chunk = (total + 7) / 8;
    const int L = static_cast<int>(blockIdx.x);
    const int item = (L & 7) * chunk + (L >> 3);
    return item < total ? item : -1;
}

// Z[k] from Y[k] and Y[M - k]: A + i E[k] D, A = Yk + conj Ym, D = Yk - conj Ym.
__device__ inline float2 zval(float2 yk, float2 ym, float2 e)
{
    const float2 A = make_float2(yk.x + ym.x, yk.y - ym.y);
    const float2 D = make_float2(yk.x - ym.x, yk.y + ym.y);
    const float2 ED = cmul(e, D);
    return make_float2(A.x - ED.y, A.y + ED.x);
}

// Kernel A.  Work item (pair, column block): columns k2 = 16 block + c.
// SPEC: Z formed from the piece and template spectra; else read from zin
// (M values per pair: the diagnostic probe).
template <int N2, bool SPEC>
__global__ __launch_bounds__(kThreads) void k_large_a(const FftPair *__restrict__ pairs, int npairs,
                                                      const float2 *__restrict__ spectra,
                                                      const float2 *__restrict__ qspec,
                                                      const int32_t *__restrict__ slot, const float2 *__restrict__ E,
                                                      const float2 *__restrict__ zin, float2 *__restrict__ U)
{
    constexpr int M = kN1 * N2;
    constexpr int kPad = kN1 + 1;
    constexpr int kBlocks = N2 / kCols;  // workgroups per pair
    __shared__ float2 d[kCols * kPad];
    __shared__ float2 tw[kN1];                 // e^{2 pi i t / 256}
    __shared__ float2 ta[512], tb[M / 512];    // w_M^j, w_M^{512 j}
    const int tid = threadIdx.x;
    const int item = xcd_item(npairs * kBlocks);
    if (item < 0)
        return;
    const int pair = item / kBlocks;
    const int k20 = (item % kBlocks) * kCols;
    // The compact twiddle tables after E (E has B = 2 M entries), read into
    // registers first so that their reads overlap the spectra's.
    static_assert(kN1 == kThreads && M / 512 <= kThreads, "one table entry per thread");
    const float2 *lt = E + 2 * M;
    const float2 t_tw = lt[tid], t_ta0 = lt[kN1 + tid], t_ta1 = lt[kN1 + kThreads + tid];
    const float2 t_tb = lt[kN1 + 512 + min(tid, M / 512 - 1)];  // no branch around a read
    if (SPEC) {
        const FftPair pr = pairs[pair];
        const __amdgpu_buffer_rsrc_t Xr = rsrc_of(spectra + static_cast<uint64_t>(pr.p) * (M + 1), M + 1);
        const __amdgpu_buffer_rsrc_t Hr = rsrc_of(qspec + static_cast<uint64_t>(slot[pr.t]) * (M + 1), M + 1);
        const __amdgpu_buffer_rsrc_t Er = rsrc_of(E, M);
        // Element (c, k1), c = tid % 16, k1 = tid / 16 + 16 q: bin k = k20 + c
        // + N2 k1 at byte offset vk + q 128 N2, and its mirror M - k (bin M,
        // the last, for k = 0) at vm + (15 - q) 128 N2.
        const int c = tid % kCols, k1_0 = tid / kCols;
        const uint32_t vk = static_cast<uint32_t>(k20 + c + N2 * k1_0) * 8u;
        const uint32_t vm = static_cast<uint32_t>(16 * N2 - k20 - c - N2 * k1_0) * 8u;
        // Two rounds of eight elements: all forty reads of a round are issued
        // before the first is used.
#pragma unroll
        for (int half = 0; half < 2; half++) {
#pragma unroll
            for (int q = 8 * half; q < 8 * half + 8; q++) {
                const uint32_t sk = static_cast<uint32_t>(q) * 128u * N2;
                const uint32_t sm = static_cast<uint32_t>(15 - q) * 128u * N2;
                const float2 z = zval(cmul(ld2(Xr, vk, sk), ld2(Hr, vk, sk)), cmul(ld2(Xr, vm, sm), ld2(Hr, vm, sm)),
                                      ld2(Er, vk, sk));
                d[c * kPad + k1_0 + 16 * q] = z;
            }
        }
    } else {
        for (int e = tid; e < kCols * kN1; e += kThreads) {
            const int c = e % kCols, k1 = e / kCols;
            d[c * kPad + k1] = zin[static_cast<uint64_t>(pair) * M + k20 + c + N2 * k1];
        }
    }
    tw[tid] = t_tw;
    ta[tid] = t_ta0;
    ta[kThreads + tid] = t_ta1;
    if (tid < M / 512)
        tb[tid] = t_tb;
    __syncthreads();
    // 256 = 16 x 16: radix-16 over n2 for each (column, n1), element n1 + 16 n2.
    {
        const int c = tid / 16, n1 = tid % 16;
        float2 v[16];
#pragma unroll
        for (int s = 0; s < 16; s++)
            v[s] = d[c * kPad + n1 + 16 * s];
        dft_reg<16, kN1>(v, tw);
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++)
            d[c * kPad + n1 + 16 * k2] = k2 == 0 ? v[0] : cmul(tw[n1 * k2], v[k2]);
    }
    __syncthreads();
    // 16-point transforms over n1 for each (column, k2): outputs r1 = k2 + 16 k1.
    {
        const int c = tid % kCols, kk = tid / kCols;
        float2 v[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; n1++)
            v[n1] = d[c * kPad + n1 + 16 * kk];
        dft_reg<16, kN1>(v, tw);
        const int col = k20 + c;
        float2 *out = U + static_cast<uint64_t>(pair) * M;
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) {
            const int r1 = kk + 16 * k1;
            const int ex = r1 * col;  // < 256 N2 = M
            const float2 wm = cmul(tb[ex >> 9], ta[ex & 511]);
            out[r1 * N2 + col] = cmul(wm, v[k1]);
        }
    }
}

// Per-pair constants of the filter (the per-window filter of the FFT forms).
struct PairFilter {
    int t, m;
    uint64_t W, base;
    float4 bi;
    const float2 *rbp;
    double Q2, S, bound, eabs, eta;
    float gmin, b0, eta_sq, sqrt_m, rsqrt_m, absS, mf, inv_mf;
};

// m sigma^2 of a window from its prefix pairs ra (start) and rz (end), and
// whether it is judged (the FP32 interval of V is narrow and the window's
// mean small enough next to its spread for the level-1 bound to hold).
__device__ inline float window_var(const PairFilter &F, float c, float2 ra, float2 rz, bool *judge)
{
    const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
    const float eS1 = 2.0f * F.bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
    const float eS2 = 2.0f * F.bi.z + 1.2e-7f * (fabsf(ra.y) + fabsf(rz.y) + fabsf(S2));
    const float V = S2 - S1 * S1 * F.inv_mf;
    const float eV =
        eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * F.inv_mf + 2.4e-7f * (fabsf(S2) + S1 * S1 * F.inv_mf);
    const float aD = fabsf(S1 * F.inv_mf) + eS1 * F.inv_mf;
    const float amu = fabsf(c) + aD;
    *judge = V > 1e3f * eV && V > F.mf * 1e-20f && !(amu * amu * F.mf > 1e6f * V);
    return V;
}

// The level-1 test of the main pass on one window (dt: its FFT dot product):
// false only for a window whose distance provably exceeds the bound.
__device__ inline bool window_survives(const PairFilter &F, float c, float2 ra, float2 rz, float dt)
{
    bool judge;
    const float V = window_var(F, c, ra, rz, &judge);
    if (!judge || F.gmin <= 0.0f)
        return true;
    const float S1 = rz.x - ra.x;
    const float eS1 = 2.0f * F.bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
    const float aD = fabsf(S1 * F.inv_mf) + eS1 * F.inv_mf;
    const float sV = V * __frsqrt_rn(V);
    const float b = F.b0 + F.eta_sq * (sV + F.sqrt_m * aD) + aD * F.absS;
    return dt >= F.gmin * sV * F.rsqrt_m * 0.999f - b * 1.001f - 1e-30f;
}

// Kernel B.  Work item (pair, row block): rows r1 = RB block + row, RB =
// 4096 / N2.  FILTER: the filter on the outputs y[2r], y[2r+1] (the windows
// i = n - (m - 1) of the piece); else z written to zout (the probe).
template <int N2, bool FILTER>
__global__ __launch_bounds__(kThreads) void k_large_b(Waveform w, const TmplDev *__restrict__ tmpl,
                                                      const FftPair *__restrict__ pairs, int npairs,
                                                      const float2 *__restrict__ U, float2 *__restrict__ zout,
                                                      int step, const float2 *__restrict__ rb, int rb_len,
                                                      const float4 *__restrict__ pinfo,
                                                      const float2 *__restrict__ pnorm,
                                                      const float *__restrict__ qmax,
                                                      const int32_t *__restrict__ slot,
                                                      const float2 *__restrict__ E, double kfft, int mode,
                                                      CandOut co)
{
    constexpr int M = kN1 * N2;
    constexpr int Q = N2 / 16;
    constexpr int RB = kElems / N2;
    // Row stride and in-row slot of element idx: one spare slot per 16
    // elements, so that the second pass's reads (16 elements apart across
    // lanes) fall in different LDS banks.
    constexpr int kPad = N2 + N2 / 16 + 1;
    const auto at = [](int idx) { return idx + (idx >> 4); };
    constexpr int B = 2 * M;
    constexpr int kItems2 = RB * 16;                          // second-pass transforms
    constexpr int kPer2 = (kItems2 + kThreads - 1) / kThreads;  // per thread
    __shared__ float2 d[RB * kPad];
    __shared__ float2 tw[N2];  // e^{2 pi i t / N2}
    __shared__ float s_ub[kThreads / 64];
    __shared__ uint32_t s_j[kThreads / 64];
    constexpr int kBlocks = kN1 / RB;  // workgroups per pair
    const int tid = threadIdx.x;
    const int item = xcd_item(npairs * kBlocks);
    if (item < 0)
        return;  // the whole workgroup: no barrier is skipped by part of it
    const int pair = item / kBlocks;
    const int r10 = (item % kBlocks) * RB;
    // The row block is RB N2 = 4096 consecutive values of U: sixteen reads
    // per thread (and the twiddle table's), all issued before any is used.
    constexpr int kPer = RB * N2 / kThreads;
    const __amdgpu_buffer_rsrc_t Ur = rsrc_of(U + static_cast<uint64_t>(pair) * M + r10 * N2, RB * N2);
    float2 uv[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++)
        uv[q] = ld2(Ur, static_cast<uint32_t>(tid) * 8u, static_cast<uint32_t>(q) * kThreads * 8u);
    const float2 *lt = E + 2 * M + kN1 + 512 + M / 512;
    float2 tv[(N2 + kThreads - 1) / kThreads];
#pragma unroll
    for (int q = 0; q < (N2 + kThreads - 1) / kThreads; q++)
        tv[q] = lt[min(tid + q * kThreads, N2 - 1)];  // no branch around a read
#pragma unroll
    for (int q = 0; q < (N2 + kThreads - 1) / kThreads; q++)
        if (tid + q * kThreads < N2)
            tw[tid + q * kThreads] = tv[q];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const int e = tid + q * kThreads;
        d[(e / N2) * kPad + at(e % N2)] = uv[q];
    }
    __syncthreads();
    // N2 = 16 Q: radix-16 over n2 for each (row, n1 < Q): RB Q = 256 items.
    {
        const int row = tid / Q, n1 = tid % Q;
        float2 v[16];
#pragma unroll
        for (int s = 0; s < 16; s++)
            v[s] = d[row * kPad + at(n1 + Q * s)];
        dft_reg<16, N2>(v, tw);
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++)
            d[row * kPad + at(n1 + Q * k2)] = k2 == 0 ? v[0] : cmul(tw[n1 * k2], v[k2]);
    }
    __syncthreads();
    // Q-point transforms over n1 for each (row, k2 < 16): outputs r2 = k2 + 16 k1.
    float2 v[kPer2][Q];
#pragma unroll
    for (int q = 0; q < kPer2; q++) {
        const int item = tid + q * kThreads;
        if (item < kItems2) {
            const int row = item / 16, kk = item % 16;
#pragma unroll
            for (int n1 = 0; n1 < Q; n1++)
                v[q][n1] = d[row * kPad + at(n1 + Q * kk)];
            dft_reg<Q, N2>(v[q], tw);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer2; q++) {
        const int item = tid + q * kThreads;
        if (item < kItems2) {
            const int row = item / 16, kk = item % 16;
#pragma unroll
            for (int k1 = 0; k1 < Q; k1++)
                d[row * kPad + at(kk + 16 * k1)] = v[q][k1];
        }
    }
    __syncthreads();
    // d[row][r2] = z[r10 + row + 256 r2]
    if (!FILTER) {
        float2 *dst = zout + static_cast<uint64_t>(pair) * M;
        for (int e = tid; e < RB * N2; e += kThreads) {
            const int row = e % RB, r2 = e / RB;
            dst[r10 + row + kN1 * r2] = d[row * kPad + at(r2)];
        }
        return;
    }
    const FftPair pr = pairs[pair];
    PairFilter F;
    F.t = static_cast<int>(pr.t);
    F.m = tmpl[F.t].m;
    if (static_cast<uint64_t>(F.m) > w.n)
        return;  // uniform over the workgroup: no barrier follows for it
    F.W = w.n - static_cast<uint64_t>(F.m) + 1;
    F.base = static_cast<uint64_t>(pr.p) * static_cast<uint64_t>(step);
    F.bi = pinfo[pr.p];
    F.rbp = rb + static_cast<uint64_t>(pr.p) * static_cast<uint64_t>(rb_len);
    F.Q2 = tmpl[F.t].Q2;
    F.S = tmpl[F.t].S;
    const double md = static_cast<double>(F.m);
    F.bound = static_cast<double>(tmpl[F.t].bound);
    constexpr double u = 5.9604644775390625e-08;
    // Absolute FFT error of every dot product of this (template, piece).
    F.eabs = kfft * (static_cast<double>(pnorm[pr.p].x) * static_cast<double>(qmax[slot[F.t]]) +
                     sqrt(F.Q2) * static_cast<double>(pnorm[pr.p].y));
    F.eta = 1.01 * u;  // rounding of x~ = x0 - c_p
    const double eref = ([&] {
        const double rm = sqrt(md), sq = 2.0 * rm + 1.0;
        const double D1 = (u * 1.001) * (3.05 * rm + 1.03 * rm * 1.001e3 + sq);
        const double neff = floor(md / 16.0) + 36.0, gam = neff * u / (1.0 - neff * u);
        return (D1 * (2.0 * sq + D1) + gam * (sq + D1) * (sq + D1)) * 1.02;
    })();
    F.gmin = static_cast<float>((F.Q2 + md - F.bound - eref) * 0.5 * (1.0 - 1e-3));
    F.b0 = static_cast<float>(F.eabs * (1.0 + 1e-3));
    F.eta_sq = static_cast<float>(F.eta * sqrt(F.Q2) * (1.0 + 1e-3));
    F.sqrt_m = static_cast<float>(sqrt(md));
    F.rsqrt_m = static_cast<float>(1.0 / sqrt(md)) * (1.0f + 1e-6f);
    F.absS = static_cast<float>(fabs(F.S) * (1.0 + 1e-3));
    F.mf = static_cast<float>(F.m);
    F.inv_mf = 1.0f / F.mf;
    const float c = F.bi.x;
    const float invB = 1.0f / static_cast<float>(B);
    // Element (row, r2) holds outputs n = 2 r and 2 r + 1, r = r10 + row +
    // 256 r2; consecutive threads take consecutive rows (runs of 2 RB
    // consecutive windows).  Window i = n - (m - 1) of the piece, j = base + i,
    // valid for 0 <= i < iend.  Four elements (eight windows) at a time: their
    // prefix pairs are read first (indices clamped into the piece's windows,
    // so every read is in bounds: i + m <= step - 1 + Lmax = B < rb_len), then
    // filtered.
    static_assert((RB * N2) % (4 * kThreads) == 0, "whole rounds of four elements");
    const int iend = F.base >= F.W ? 0 : static_cast<int>(min(static_cast<uint64_t>(step), F.W - F.base));
    const auto load8 = [&](int e0, float2 (&ra4)[4][2], float2 (&rz4)[4][2]) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * kThreads;
            const int r = r10 + e % RB + kN1 * (e / RB);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                int i = 2 * r + h - (F.m - 1);
                i = i < 0 ? 0 : (i >= step ? step - 1 : i);
                ra4[u][h] = F.rbp[i];
                rz4[u][h] = F.rbp[i + F.m];
            }
        }
    };
    if (mode == kModeMain) {
        // A cheap rejection shared by a thread's eight windows of a round: the
        // level-1 test with its error terms (eS1, eS2, eV, aD, amu) replaced by
        // upper bounds from the largest prefix pair magnitudes among the eight
        // (|S1| <= |ra.x| + |rz.x|, |S2| <= |ra.y| + |rz.y|), with relative
        // slack of at least 1e-5 over the FP32 roundings of both evaluations.
        // Larger error terms make every step of that test more conservative,
        // so a window rejected here fails window_survives too; the others take
        // window_survives itself, and the survivors are exactly its survivors.
        const double Kd = static_cast<double>(F.gmin) * static_cast<double>(F.rsqrt_m) * 0.999 * (1.0 - 1e-5) -
                          1.001 * static_cast<double>(F.eta_sq) * (1.0 + 1e-5);
        const bool cheap = F.gmin > 0.0f && Kd > 0.0;
        const float Kc = static_cast<float>(Kd * (1.0 - 1e-6));  // slope in sqrt(V), rounded down
        const float slope_aD = F.eta_sq * F.sqrt_m + F.absS;
        for (int e0 = tid; e0 < RB * N2; e0 += 4 * kThreads) {
            float2 ra4[4][2], rz4[4][2];
            load8(e0, ra4, rz4);
            float ax = 0.0f, bx = 0.0f, ay = 0.0f, by = 0.0f;
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    ax = fmaxf(ax, fabsf(ra4[u][h].x));
                    bx = fmaxf(bx, fabsf(rz4[u][h].x));
                    ay = fmaxf(ay, fabsf(ra4[u][h].y));
                    by = fmaxf(by, fabsf(rz4[u][h].y));
                }
            const float Mx = (ax + bx) * 1.0001f, My = (ay + by) * 1.0001f;
            const float eS1 = 2.0001f * F.bi.y + 2.41e-7f * Mx;
            const float eS2 = 2.0001f * F.bi.z + 2.41e-7f * My;
            const float aD = (Mx + eS1) * F.inv_mf;
            const float eV = eS2 + (2.0f * Mx * eS1 + eS1 * eS1) * F.inv_mf + 2.41e-7f * (My + Mx * Mx * F.inv_mf);
            const float amu = fabsf(c) + aD;
            const float VT =
                cheap ? fmaxf(fmaxf(1e3f * eV, F.mf * 1e-20f), amu * amu * F.mf * 1e-6f) * 1.00001f : INFINITY;
            const float Cg = (1.001f * (F.b0 + slope_aD * aD) + 1e-30f) * 1.00001f;
            uint32_t need = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * kThreads;
                const int row = e % RB, r2 = e / RB;
                const float2 zz = d[row * kPad + at(r2)];
                const int r = r10 + row + kN1 * r2;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int i = 2 * r + h - (F.m - 1);
                    const float2 ra = ra4[u][h], rz = rz4[u][h];
                    const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
                    const float V = S2 - S1 * S1 * F.inv_mf;
                    const float dt = (h == 0 ? zz.x : zz.y) * invB;
                    // Evaluated for every window (no branch); a NaN compares false.
                    const bool keep = !(V > VT) | !(dt < V * __frsqrt_rn(V) * Kc - Cg);
                    need |= static_cast<uint32_t>(keep & (i >= 0) & (i < iend)) << (2 * u + h);
                }
            }
            // The undecided windows one at a time: the full test, then the
            // FP64 interval (level 2) for its survivors.
            while (need) {
                const int bit = __builtin_ctz(need);
                need &= need - 1;
                const int e = e0 + (bit >> 1) * kThreads, h = bit & 1;
                const int row = e % RB, r2 = e / RB;
                const int i = 2 * (r10 + row + kN1 * r2) + h - (F.m - 1);
                const float2 zz = d[row * kPad + at(r2)];
                const float dt = (h == 0 ? zz.x : zz.y) * invB;
                const uint64_t j = F.base + static_cast<uint64_t>(i);
                if (window_survives(F, c, F.rbp[i], F.rbp[i + F.m], dt) && w.x[j] != kMissing) {
                    const DistInterval iv =
                        dist_interval(w, j, F.m, F.Q2, F.S, c, static_cast<double>(dt), F.eta, F.eabs);
                    if (!(iv.lo > F.bound))  // NaN (undecidable): a candidate
                        cand_put(co, wave_claim(co.count), static_cast<uint32_t>(F.t), static_cast<uint32_t>(j),
                                 static_cast<float>(iv.est), iv.lo, iv.hi);
                }
            }
        }
        return;
    }
    // Seeding: per thread the judged window of the largest dt / sqrt(V).
    float best_g = -INFINITY;
    uint32_t best_j = 0xffffffffu;
    float best_dt = 0.0f;
    for (int e0 = tid; e0 < RB * N2; e0 += 4 * kThreads) {
        float2 ra4[4][2], rz4[4][2];
        load8(e0, ra4, rz4);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * kThreads;
            const int row = e % RB, r2 = e / RB;
            const float2 zz = d[row * kPad + at(r2)];
            const int r = r10 + row + kN1 * r2;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = 2 * r + h - (F.m - 1);
                if (i < 0 || i >= iend)
                    continue;
                bool judge;
                const float V = window_var(F, c, ra4[u][h], rz4[u][h], &judge);
                if (judge && V > 0.0f) {
                    const float dt = (h == 0 ? zz.x : zz.y) * invB;
                    const float g = dt * __frsqrt_rn(V);
                    if (g > best_g) {
                        best_g = g;
                        best_j = static_cast<uint32_t>(F.base + static_cast<uint64_t>(i));
                        best_dt = dt;
                    }
                }
            }
        }
    }
    {
        float ub = INFINITY;
        if (best_j != 0xffffffffu && w.x[best_j] != kMissing) {
            const DistInterval iv =
                dist_interval(w, best_j, F.m, F.Q2, F.S, c, static_cast<double>(best_dt), F.eta, F.eabs);
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
        if ((tid & 63) == 0) {
            s_ub[tid >> 6] = ub;
            s_j[tid >> 6] = bj;
        }
        __syncthreads();
        if (tid == 0) {
            for (int q = 1; q < kThreads / 64; q++)
                if (s_ub[q] < ub || (s_ub[q] == ub && s_j[q] < bj)) {
                    ub = s_ub[q];
                    bj = s_j[q];
                }
            if (ub < INFINITY)
                cand_put(co, atomicAdd(co.count, 1ULL), static_cast<uint32_t>(F.t), bj, ub, ub, ub);
        }
    }
}

// Grid of 8 chunks of ceil(total / 8) workgroups (xcd_item).
inline unsigned grid8(int total) { return static_cast<unsigned>(8 * ((total + 7) / 8)); }

template <int N2>
void launch_pair(const Waveform &w, const TmplDev *tmpl, const FftPair *pairs, int npairs, const float2 *spectra,
                 const float2 *qspec, const int32_t *slot, const float2 *E, float2 *U, int step, const float2 *rb,
                 int rb_len, const float4 *pinfo, const float2 *pnorm, const float *qmax, double kfft, int mode,
                 const CandOut &co, hipStream_t stream)
{
    hipLaunchKernelGGL((k_large_a<N2, true>), dim3(grid8(npairs * (N2 / kCols))), dim3(kThreads), 0, stream, pairs,
                       npairs, spectra, qspec, slot, E, nullptr, U);
    hipLaunchKernelGGL((k_large_b<N2, true>), dim3(grid8(npairs * (kN1 / (kElems / N2)))), dim3(kThreads), 0, stream,
                       w, tmpl, pairs, npairs, U, nullptr, step, rb, rb_len, pinfo, pnorm, qmax, slot, E, kfft, mode,
                       co);
}

template <int N2>
void launch_probe(const float2 *E, const float2 *zin, int count, float2 *U, float2 *zout, hipStream_t stream)
{
    const Waveform w{nullptr, nullptr, nullptr, 0};
    hipLaunchKernelGGL((k_large_a<N2, false>), dim3(grid8(count * (N2 / kCols))), dim3(kThreads), 0, stream,
                       nullptr, count, nullptr, nullptr, nullptr, E, zin, U);
    hipLaunchKernelGGL((k_large_b<N2, false>), dim3(grid8(count * (kN1 / (kElems / N2)))), dim3(kThreads), 0, stream,
                       w, nullptr, nullptr, count, U, zout, 0, nullptr, 0, nullptr, nullptr, nullptr, nullptr, E, 0.0,
                       0, CandOut{});
}

}  // namespace

bool fft_large_supports(int B) { return B == (1 << 16) || B == (1 << 17) || B == (1 << 18); }

hipError_t launch_fft_large(const Waveform &w, const TmplDev *tmpl, const FftPair *pairs, int npairs,
                            const float2 *spectra, const float2 *qspec, const int32_t *slot, int B, const float2 *E,
                            float2 *U, int step, const float2 *rb, int rb_len, const float4 *pinfo,
                            const float2 *pnorm, const float *qmax, double kfft, int mode, const CandOut &co,
                            hipStream_t stream)
{
    if (npairs <= 0)
        return hipSuccess;
    if (npairs > 1024 || rb_len < B + 1)
        return hipErrorInvalidValue;
    switch (B) {
    case 1 << 16:
        launch_pair<128>(w, tmpl, pairs, npairs, spectra, qspec, slot, E, U, step, rb, rb_len, pinfo, pnorm, qmax,
                         kfft, mode, co, stream);
        break;
    case 1 << 17:
        launch_pair<256>(w, tmpl, pairs, npairs, spectra, qspec, slot, E, U, step, rb, rb_len, pinfo, pnorm, qmax,
                         kfft, mode, co, stream);
        break;
    case 1 << 18:
        launch_pair<512>(w, tmpl, pairs, npairs, spectra, qspec, slot, E, U, step, rb, rb_len, pinfo, pnorm, qmax,
                         kfft, mode, co, stream);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The same inverse on given Z (count transforms of B / 2 values, zin) into
// zout, with U as the intermediate (count B / 2 values): the diagnostic probe.
hipError_t launch_fft_large_probe(int B, const float2 *E, const float2 *zin, int count, float2 *U, float2 *zout,
                                  hipStream_t stream)
{
    if (count <= 0 || count > 1024)
        return hipErrorInvalidValue;
    switch (B) {
    case 1 << 16:
        launch_probe<128>(E, zin, count, U, zout, stream);
        break;
    case 1 << 17:
        launch_probe<256>(E, zin, count, U, zout, stream);
        break;
    case 1 << 18:
        launch_probe<512>(E, zin, count, U, zout, stream);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tmx
