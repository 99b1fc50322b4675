// Fused overlap-save FFT candidate pass for FFT lengths B = 2^15 and 2^14
// (pieces of B samples every step = B - Lmax + 1 windows, templates up to
// Lmax samples, Lmax from B / 16 to B / 4), in the role of the
// reference's FFT fallback pass (fft_fallback.hpp:67-117) with its
// filter-then-exact-recheck contract.  B = 2^15: one workgroup of 1024
// threads per CU; B = 2^14: two of 512 (the LDS holds two transforms, so one
// workgroup's barriers overlap the other's work).  Each workgroup takes
// pieces in turn.  (Constants below are for B = 2^15; B = 2^14 halves the
// thread count and the transform, and its last pass is radix 2.)
// Per piece and template of the batch (the piece spectrum X_p, hipFFT r2c
// cached per waveform, comes from HBM once per batch and from L2 after):
//
//   1. Y = X_p .* H_t (H_t the template's spectrum, shared by all pieces:
//      L2-resident), written to LDS;
//   2. the real inverse transform of Y as one 16384-point complex transform:
//      Z[k] = (Y[k] + conj Y[M-k]) + i e^{i pi k / M} (Y[k] - conj Y[M-k]),
//      z = sum_k Z[k] e^{2 pi i r k / M} gives y[2r] = Re z[r], y[2r+1] =
//      Im z[r] (unnormalised, like a c2r of length B), computed in LDS by
//      Stockham passes of radix 16, 16, 16, 4 (the first one in registers;
//      the second and third passes' twiddles from small LDS tables);
//   3. the window dot products dt = y[m - 1 + i] / B are read from LDS by a
//      chunk filter: for the windows of each 64 consecutive outputs, a
//      rigorous lower bound of their reference distances from the largest dt
//      and the statistics of the windows' common core (two FP64 prefix reads
//      per chunk); only chunks that may hold a window within the bound are
//      examined window by window (filter.hpp's interval) and emit candidates.
//
// Nothing of size B leaves the chip: per (piece, template) HBM traffic is the
// piece spectrum (amortised over the batch) and the candidates.  The spectra
// of the next (piece, template) pair are loaded into registers while the
// current pair's chunks are filtered, so their latency is hidden.
//
// Error bound.  The inverse transform here is a sequence of 14 radix-2 levels
// (grouped as radix 16/16/16/4) with twiddles from an FP64-rounded table
// (at most one complex product each: error <= 5u), plus the Z formation; by
// the standard analysis of Cooley-Tukey transforms (relative 2-norm error <=
// levels * (mu + gamma_4 (sqrt 2 + mu)) for twiddle error mu) its relative
// error is below 200u, within the 10 log2(B) u = 150u... x4 allowance of
// fft_error_factor, which budgets 4 * (2 * 10 log2 B + 2) u = 1208u for the
// two forward transforms, the product and the inverse together:
// forward (600u, black box) + product (2u) + inverse (<= 400u with a factor 2
// margin) <= 1208u on ||x~||_2 ||H||_inf, and 600u on ||h||_2 ||X||_inf.
#include "filter.hpp"

#pragma clang fp contract(fast)

namespace tmx {
namespace {

// Geometry of the fused pass for B = 2M = 2^(LOG2M + 1): 16 points per
// thread; B = 2^15 (templates up to 8192) runs one workgroup of 1024 threads
// per CU, B = 2^14 (templates up to 4096) two of 512 (half the LDS each).
template <int LOG2M>
struct Geom {
    static constexpr int kM = 1 << LOG2M;                    // complex points of the half-length transform
    static constexpr int kB = 2 * kM;                        // real length B
    static constexpr int kT = kM / 16;                       // threads per workgroup
    static constexpr int kStepMax = kB - kB / 16 + 1;        // windows per piece, templates up to B / 16
    static constexpr int kLds = kM + (kM >> 3) + 4;          // padded complex slots (index M included)
    static constexpr int kPerCU = LOG2M == 14 ? 1 : 2;       // workgroups per CU
    static_assert((kStepMax + 63) / 64 + 1 <= kT, "one thread per 64-output chunk of a piece");
    static_assert(LOG2M == 13 || LOG2M == 14, "B = 2^14 or 2^15");
};

__device__ __forceinline__ int pad(int i) { return i + 2 * (i >> 4); }

// Outputs per chunk of the filter (two padded LDS rows), and dot-product
// slots per queued chunk record.
constexpr int kChunkW = 64;

// A copy of x the compiler cannot see through: index arithmetic derived from
// it is redone where it is used instead of being hoisted out of the loops
// (and spilled: the transform needs nearly every register).
__device__ __forceinline__ int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Global loads through buffer resources: one 32-bit per-thread offset, the
// per-element part in a scalar register (no 64-bit address per load).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), static_cast<short>(0), static_cast<int>(bytes),
                                             0x00020000);
}
__device__ __forceinline__ float2 bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// e^{i pi r / 16}, r = 0..31 (FP64-rounded): the Z factors (r < 16) and the
// last pass's e^{i pi s / 8}, e^{i pi s / 4}, e^{3 i pi s / 8} (r = 2s, 4s, 6s).
__constant__ float kZc[32] = {
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f, 0.70710678118654757f,
    0.55557023301960229f, 0.38268343236508984f, 0.19509032201612833f, 0.0f, -0.19509032201612819f,
    -0.38268343236508973f, -0.55557023301960196f, -0.70710678118654746f, -0.83146961230254535f,
    -0.92387953251128674f, -0.98078528040323043f, -1.0f, -0.98078528040323043f, -0.92387953251128685f,
    -0.83146961230254546f, -0.70710678118654768f, -0.55557023301960218f, -0.38268343236509034f,
    -0.19509032201612866f, 0.0f, 0.1950903220161283f, 0.38268343236509f, 0.55557023301960184f,
    0.70710678118654735f, 0.83146961230254524f, 0.92387953251128652f, 0.98078528040323032f};
__constant__ float kZs[32] = {
    0.0f, 0.19509032201612825f, 0.38268343236508978f, 0.55557023301960218f, 0.70710678118654746f,
    0.83146961230254524f, 0.92387953251128674f, 0.98078528040323043f, 1.0f, 0.98078528040323043f,
    0.92387953251128674f, 0.83146961230254546f, 0.70710678118654757f, 0.55557023301960218f,
    0.38268343236508989f, 0.19509032201612861f, 0.0f, -0.19509032201612836f, -0.38268343236508967f,
    -0.55557023301960196f, -0.70710678118654746f, -0.83146961230254524f, -0.92387953251128652f,
    -0.98078528040323032f, -1.0f, -0.98078528040323043f, -0.92387953251128663f, -0.83146961230254546f,
    -0.70710678118654768f, -0.55557023301960218f, -0.38268343236509039f, -0.19509032201612872f};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }

// Inverse-sign 4-point DFT in place: X_k = sum_n v_n i^{nk}.
__device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d)
{
    const float2 t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = mul_i(csub(b, d));
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

// Inverse-sign 16-point DFT in place, natural order in and out (4 x 4:
// n = 4 n1 + n2, k = k1 + 4 k2).
__device__ __forceinline__ void dft16(float2 *v)
{
    constexpr float c8 = 0.92387953251128674f, s8 = 0.38268343236508977f, h = 0.70710678118654752f;
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++)
        dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // v[n2 + 4 k1] now holds T[n2][k1]; twiddle by e^{i pi n2 k1 / 8}.
    v[1 + 4] = cmul(v[1 + 4], make_float2(c8, s8));
    v[1 + 8] = cmul(v[1 + 8], make_float2(h, h));
    v[1 + 12] = cmul(v[1 + 12], make_float2(s8, c8));
    v[2 + 4] = cmul(v[2 + 4], make_float2(h, h));
    v[2 + 8] = mul_i(v[2 + 8]);
    v[2 + 12] = cmul(v[2 + 12], make_float2(-h, h));
    v[3 + 4] = cmul(v[3 + 4], make_float2(s8, c8));
    v[3 + 8] = cmul(v[3 + 8], make_float2(-h, h));
    v[3 + 12] = cmul(v[3 + 12], make_float2(-c8, -s8));
    float2 o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 a = v[0 + 4 * k1], b = v[1 + 4 * k1], c = v[2 + 4 * k1], d = v[3 + 4 * k1];
        dft4(a, b, c, d);
        o[k1] = a;
        o[k1 + 4] = b;
        o[k1 + 8] = c;
        o[k1 + 12] = d;
    }
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = o[r];
}

// Twiddles of the second and third passes, from LDS: T[j] = e^{2 pi i j / 256}
// (j < 256) and the two-level pair A[j] = e^{2 pi i j / 64}, Bf[j] =
// e^{2 pi i j / 4096} (j < 64), all entries of the FP64-rounded table E.
constexpr int kTwT = 0, kTwA = 256, kTwB = 320, kTw = 384;

// E[t] = e^{i pi t / M} has 2M entries; T, A and Bf are its entries at
// strides 2M / 256, 2M / 64 and 2M / 4096.
__device__ __forceinline__ void fill_twiddles(float2 *tw, __amdgpu_buffer_rsrc_t Er, int M, int tid, int nthreads)
{
    for (int j = tid; j < kTw; j += nthreads) {
        const int t = j < kTwA ? j * (2 * M / 256) : j < kTwB ? (j - kTwA) * (2 * M / 64) : (j - kTwB) * (2 * M / 4096);
        tw[j] = bload(Er, static_cast<uint32_t>(t) * 8u, 0);
    }
}

// Pass 2 (Ns = 16, w = e^{2 pi i / 256}): v[r] *= w^{r k} = T[r k], r k <= 225.
__device__ __forceinline__ void twiddle_pass2(float2 *v, const float2 *tw, int k)
{
#pragma unroll
    for (int r = 1; r < 16; r++)
        v[r] = cmul(v[r], tw[kTwT + r * k]);
}

// Pass 3 (Ns = 256, w = e^{2 pi i / 4096}): v[r] *= w^{r k} = A[r k / 64]
// Bf[r k mod 64] (one product of two table entries), r k <= 3825.
__device__ __forceinline__ void twiddle_pass3(float2 *v, const float2 *tw, int k)
{
#pragma unroll
    for (int r = 1; r < 16; r++) {
        const int j = r * k;
        v[r] = cmul(v[r], cmul(tw[kTwA + (j >> 6)], tw[kTwB + (j & 63)]));
    }
}

// Upper bound on sqrt(v), v >= 0, without an FP64 square root: the FP32 one
// (V_SQRT_F32, 1 ulp) of v rounded to float, inflated by 2^-20 to cover both
// roundings, plus 1e-18 for values below the float range; +inf from 1e37.
__device__ __forceinline__ double sqrt_up(double v)
{
    if (!(v < 1e37))
        return INFINITY;
    return static_cast<double>(__builtin_amdgcn_sqrtf(static_cast<float>(v))) * (1.0 + 0x1p-20) + 1e-18;
}

// Per-template constants of chunk_dist_lo: 1 / m, an upper bound on sqrt(m),
// the upper bound of gamma_{m/16 + 36} (1 / (1 - x) <= 1 + 2x for x <= 1/2),
// and 1 / (m - 63), the core length of a full chunk.
struct ChunkConst {
    double invm, rm_up, gam, inv_full;
};

__device__ __forceinline__ ChunkConst chunk_const(int m)
{
    constexpr double u = 5.9604644775390625e-08;
    const double md = static_cast<double>(m);
    const double neff = floor(md / 16.0) + 36.0;
    return ChunkConst{1.0 / md, sqrt_up(md), neff * u * (1.0 + 2.0 * neff * u) * (1.0 + 1e-12),
                      m > kChunkW ? 1.0 / static_cast<double>(m - (kChunkW - 1)) : 0.0};
}

// Lower bound on the reference distance of every window j in [j0, j0 + cnt)
// of a template of length m, given that each window's
// N_j = dt_j - (mu_j - c) S is at most nmax: with the windows' common core
// C = [j0 + cnt - 1, j0 + m), m sigma_j^2 = sum_window (x0 - mu_j)^2 >=
// sum_C (x0 - mu_C)^2 = V_C, so sigma_j >= sqrt(V_C / m).  Then
// d_exact = Q2 + m - 2 N_j / sigma_j >= Q2 + m - 2 max(nmax, 0) / sigma_lo,
// and the reference's FP32 value d_ref >= d_exact - E_ref(d_exact) >= dlo -
// E_ref(dlo) (d - E_ref(d) increases wherever the result can exceed 0), with
// E_ref as in filter.hpp and |mu_j| / sigma_j <= mu_abs / sigma_lo.
// pa, pb are the high parts of the prefix pairs at the core's ends a = j0 +
// cnt - 1 and b = j0 + m (Lc = b - a, invL = the rounded 1 / Lc); where they
// cannot bound the core (after a loud stretch) the chunk is queued and its
// windows are judged with the full double-double sums.  The differences and
// V_C are FP64; the square roots and 1 / sigma_lo are FP32 estimates made
// one-sided (sigma_lo rounded down, the error terms up), so the result is
// still a lower bound.  -inf when the core cannot bound sigma.
__device__ double chunk_dist_lo(double2 pa, double2 pb, uint64_t a, uint64_t b, int Lc, double invL,
                                const ChunkConst &K, double Q2, int m, double nmax, double mu_abs)
{
    constexpr double u = 5.9604644775390625e-08;
    constexpr double u64 = 1.1102230246251565e-16;
    if (Lc < 2 || !(nmax <= 1e300))
        return -INFINITY;  // no core, or a non-finite dot product: examine every window
    const WindowSums ws = window_sums_hi(pa, pb, a, b);
    const double S1 = ws.S1, S2 = ws.S2, eS1 = ws.eS1, eS2 = ws.eS2;
    // S1^2 / Lc as S1^2 invL: one rounding more than a division (10 u64).
    const double S1L = S1 * S1 * invL;
    const double V = S2 - S1L;
    const double eV = eS2 + (2.0 * fabs(S1) * eS1 + eS1 * eS1) * invL * (1.0 + 4.0 * u64) + 10.0 * u64 * (fabs(S2) + S1L);
    const double vlo = V - eV;
    if (!(vlo > 0.0))
        return -INFINITY;
    // sigma_lo <= sqrt(vlo / m) and isig >= 1 / sigma_lo (V_RCP_F32, 1 ulp).
    const float vm = static_cast<float>(vlo * K.invm);
    if (!(vm > 1e-30f))
        return -INFINITY;
    const float sig_f = __builtin_amdgcn_sqrtf(fminf(vm, 1e38f)) * (1.0f - 0x1p-20f);
    const double isig = static_cast<double>(__builtin_amdgcn_rcpf(sig_f)) * (1.0 + 0x1p-20);
    const double gmax = nmax > 0.0 ? nmax * isig * (1.0 + 2.0 * u64) : 0.0;
    const double md = static_cast<double>(m);
    const double base = Q2 + md;
    const double dlo = base - 2.0 * gmax - 8.0 * u64 * (base + 2.0 * gmax);
    if (!(dlo > 0.0))
        return dlo;
    const double sq = sqrt_up(dlo);
    const double D1 = (u * 1.001) * (3.05 * K.rm_up + 1.03 * K.rm_up * mu_abs * isig + sq);
    const double Eref = (D1 * (2.0 * sq + D1) + K.gam * (sq + D1) * (sq + D1)) * 1.01;
    return dlo - Eref;
}

// Steps 1 and 2 for one (piece, template): on return data[pad(r)] holds
// z[r] = y[2r] + i y[2r+1], the unnormalised circular correlation of the
// piece with the template (all threads synchronised).
template <int LOG2M>
__device__ __forceinline__ void correlate_piece(float2 *data, const float2 *X, float2 Xn, const float2 *H, float2 Hn,
                                                const float2 *tw, __amdgpu_buffer_rsrc_t Er)
{
    using G = Geom<LOG2M>;
    constexpr int kM = G::kM, kT = G::kT;
    const int tid = opaque(threadIdx.x);
    const uint32_t toff = static_cast<uint32_t>(tid) * 8u;
    // 1. Y = X .* H into LDS (X, H: bins tid + kT r, and M for thread 0).
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        v[r] = cmul(X[r], H[r]);
        data[pad(tid + kT * r)] = v[r];
    }
    if (tid == 0)
        data[pad(kM)] = cmul(Xn, Hn);
    __syncthreads();
    // 2a. Z formation (registers) = the first pass's inputs; the
    // factor e^{i pi k / M} for k = tid + kT r is E[tid] e^{i pi r / 16}.
    {
        const float2 e0 = bload(Er, toff, 0);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int k = tid + kT * r;
            const float2 ym = data[pad(kM - k)];
            const float2 cy = make_float2(ym.x, -ym.y);
            const float2 A = cadd(v[r], cy);
            const float2 D = csub(v[r], cy);
            const float2 wk = r == 0 ? e0 : cmul(e0, make_float2(kZc[r], kZs[r]));
            v[r] = cadd(A, mul_i(cmul(wk, D)));
        }
    }
    __syncthreads();
    // Pass 1: radix 16, Ns = 1.
    dft16(v);
#pragma unroll
    for (int r = 0; r < 16; r++)
        data[pad(16 * tid + r)] = v[r];
    __syncthreads();
    // Pass 2: radix 16, Ns = 16.
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = data[pad(tid + kT * r)];
    twiddle_pass2(v, tw, tid & 15);
    dft16(v);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++)
        data[pad((tid >> 4) * 256 + (tid & 15) + 16 * r)] = v[r];
    __syncthreads();
    // Pass 3: radix 16, Ns = 256.
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = data[pad(tid + kT * r)];
    twiddle_pass3(v, tw, tid & 255);
    dft16(v);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++)
        data[pad((tid >> 8) * 4096 + (tid & 255) + 256 * r)] = v[r];
    __syncthreads();
    // Pass 4, B = 2^15: radix 4, Ns = 4096; each thread owns its four
    // butterflies' slots, so no barrier between read and write.  The
    // twiddles e^{2 pi i r jj / M} for jj = tid + 1024 s are
    // E[2 r tid] e^{i pi r s / 8}.
    if constexpr (LOG2M == 14) {
        const float2 e1 = bload(Er, 2 * toff, 0), e2 = bload(Er, 4 * toff, 0), e3 = bload(Er, 6 * toff, 0);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int jj = tid + 1024 * s;
            float2 a = data[pad(jj)], b = data[pad(jj + 4096)], cc = data[pad(jj + 8192)],
                   d = data[pad(jj + 12288)];
            b = cmul(b, s == 0 ? e1 : cmul(e1, make_float2(kZc[2 * s], kZs[2 * s])));
            cc = cmul(cc, s == 0 ? e2 : cmul(e2, make_float2(kZc[4 * s], kZs[4 * s])));
            d = cmul(d, s == 0 ? e3 : cmul(e3, make_float2(kZc[6 * s], kZs[6 * s])));
            dft4(a, b, cc, d);
            data[pad(jj)] = a;
            data[pad(jj + 4096)] = b;
            data[pad(jj + 8192)] = cc;
            data[pad(jj + 12288)] = d;
        }
    } else {
        // Pass 4, B = 2^14: radix 2, Ns = 4096, butterflies jj = tid + 512 s
        // with factor e^{2 pi i jj / M} = E[2 tid] e^{i pi s / 8}.
        const float2 e1 = bload(Er, 2 * toff, 0);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int jj = tid + 512 * s;
            const float2 a = data[pad(jj)];
            const float2 b = cmul(data[pad(jj + 4096)], s == 0 ? e1 : cmul(e1, make_float2(kZc[2 * s], kZs[2 * s])));
            data[pad(jj)] = cadd(a, b);
            data[pad(jj + 4096)] = csub(a, b);
        }
    }
    __syncthreads();
}


// A chunk that may hold a window within the bound (main mode: its windows'
// dot products follow in the dt queue, kChunkW slots per record), or a seed
// window (cnt = 1: the chunk's most promising window).
struct ChunkRec {
    uint32_t t;   // template
    uint32_t j0;  // first window
    int32_t cnt;  // windows
    float c;      // the piece's centre
    float eabs;   // FFT error bound of the dot products (rounded up)
    uint32_t pad0;
};

template <int LOG2M>
__global__ __launch_bounds__(Geom<LOG2M>::kT, 4) void k_fft_fused(  // 4 waves per SIMD: kPerCU workgroups per CU
    Waveform w, const TmplDev *__restrict__ tmpl, int t0, int F, const float2 *__restrict__ spectra, int bins,
    const float2 *__restrict__ qspec, const float *__restrict__ qmax, const float4 *__restrict__ pinfo,
    const float2 *__restrict__ pnorm, const float2 *__restrict__ E, uint64_t P, int step, int stride, double kfft,
    int mode, ChunkRec *__restrict__ rec, float *__restrict__ rec_dt, unsigned long long *__restrict__ nrec,
    uint64_t rec_cap)
{
    using G = Geom<LOG2M>;
    constexpr int kM = G::kM, kB = G::kB, kT = G::kT;
    __shared__ float2 data[G::kLds];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const float *fdata = reinterpret_cast<const float *>(data);
    constexpr double u = 5.9604644775390625e-08;
    const uint64_t nwork = (P + stride - 1) / stride;
    const __amdgpu_buffer_rsrc_t Er = rsrc(E, 2u * kM * 8u);
    __shared__ float2 tw[kTw];
    fill_twiddles(tw, Er, kM, tid, kT);
    __syncthreads();

    // Spectrum operands (bins tid + kT r, and M for thread 0) of the pair
    // (piece, template) about to be transformed.  The first pair's are loaded
    // here; every later pair's while the previous pair's chunks are filtered,
    // so their latency hides behind the filter.
    float2 X[16], H[16];
    float2 Xn = make_float2(0.0f, 0.0f), Hn = make_float2(0.0f, 0.0f);
    const auto load_spec = [&](float2 *dst, float2 &dst_n, const float2 *src) {
        const __amdgpu_buffer_rsrc_t r = rsrc(src, bins * 8u);
        const uint32_t toff = static_cast<uint32_t>(opaque(tid)) * 8u;
#pragma unroll
        for (int k = 0; k < 16; k++)
            dst[k] = bload(r, toff, 8u * kT * k);
        if (tid == 0)
            dst_n = bload(r, 0, 8u * kM);
    };
    if (blockIdx.x < nwork) {
        load_spec(X, Xn, spectra + static_cast<uint64_t>(blockIdx.x) * stride * static_cast<uint64_t>(bins));
        load_spec(H, Hn, qspec);
    }

    for (uint64_t it = blockIdx.x; it < nwork; it += gridDim.x) {
        const uint64_t p = it * static_cast<uint64_t>(stride);
        const uint64_t base = p * static_cast<uint64_t>(step);
        const float4 pi = pinfo[p];
        const float c = pi.x;
        const double amax = static_cast<double>(pi.w) * (1.0 + 1e-6);
        const float2 pn = pnorm[p];

        for (int ti = 0; ti < F; ti++) {
            const int t = t0 + ti;
            const int m = tmpl[t].m;
            const uint64_t W = static_cast<uint64_t>(m) <= w.n ? w.n - static_cast<uint64_t>(m) + 1 : 0;
            const bool active = W > 0 && !tmpl[t].ea && base < W;
            const uint64_t nwin64 = !active ? 0 : W - base < static_cast<uint64_t>(step) ? W - base
                                                                                        : static_cast<uint64_t>(step);
            const int nwin = static_cast<int>(nwin64);
            // Thread tid owns the chunk of outputs [64 (c0 + tid), + 64): the
            // windows i = n - (m - 1) of its outputs n, clipped to the piece's
            // [0, nwin).  Its core's prefix pairs are loaded ahead of the
            // transform.
            // (The chunk's bounds are recomputed after the transform rather
            // than held across it.)
            const auto chunk = [&](int &n_first, int &n_end) {
                const int nlo = kChunkW * (((m - 1) >> 6) + opaque(tid));
                n_first = nlo > m - 1 ? nlo : m - 1;
                n_end = nlo + kChunkW < m - 1 + nwin ? nlo + kChunkW : m - 1 + nwin;
            };
            double2 pa = make_double2(0.0, 0.0), pb = make_double2(0.0, 0.0);
            if (active) {
                int nf, ne;
                chunk(nf, ne);
                const int cnt = nf < ne ? ne - nf : 0;
                const int i_first = nf - (m - 1);
                if (cnt > 0 && m - (cnt - 1) >= 2) {
                    pa = w.p[base + static_cast<uint64_t>(i_first + cnt - 1)];
                    pb = w.p[base + static_cast<uint64_t>(i_first + m)];
                }
                correlate_piece<LOG2M>(data, X, Xn, H, Hn, tw, Er);
            }
            // The next pair's operands: the next template's spectrum (the
            // first one's after the last) and, after the last template, the
            // next piece's.
            load_spec(H, Hn, qspec + static_cast<uint64_t>(ti + 1 < F ? ti + 1 : 0) * static_cast<uint64_t>(bins));
            if (ti + 1 == F && it + gridDim.x < nwork)
                load_spec(X, Xn, spectra + (it + gridDim.x) * static_cast<uint64_t>(stride) * static_cast<uint64_t>(bins));
            if (!active)
                continue;

            int n_first, n_end;
            chunk(n_first, n_end);
            const int c0 = (m - 1) >> 6;
            const bool has_ch = n_first < n_end;
            const int cnt = has_ch ? n_end - n_first : 0;
            const int i_first = n_first - (m - 1);
            const int Lc = m - (cnt - 1);
            const uint64_t ca = base + static_cast<uint64_t>(i_first + cnt - 1);
            const uint64_t cb = base + static_cast<uint64_t>(i_first + m);
            // 3. Filter.  Output n (dt of window base + n - (m - 1), times B)
            // is at float n + 4 (n >> 5) of the padded array.  The chunk's
            // largest output, and a sum that is non-finite iff an output is.
            const float invB = 1.0f / static_cast<float>(kB);
            float vmax = -INFINITY, vsum = 0.0f;
            if (cnt == kChunkW) {
                // Two padded rows of 32 floats; half the lanes take them in
                // the other order (fewer LDS bank conflicts).
                const int f0 = 72 * (c0 + tid);
                const int sw = (tid >> 4) & 1;
#pragma unroll 1
                for (int hr = 0; hr < 2; hr++) {
                    const float4 *row = reinterpret_cast<const float4 *>(fdata + f0 + 36 * (hr ^ sw));
#pragma unroll 8
                    for (int q = 0; q < 8; q++) {
                        const float4 v = row[q];
                        vmax = fmaxf(vmax, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
                        vsum += v.x;
                        vsum += v.y;
                        vsum += v.z;
                        vsum += v.w;
                    }
                }
            } else {
                for (int n = n_first; n < n_end; n++) {
                    const float v = fdata[n + 4 * (n >> 5)];
                    vmax = fmaxf(vmax, v);
                    vsum += v;
                }
            }
            const bool nonfinite = !(vsum - vsum == 0.0f);
            const float dmax = vmax * invB;
            // Error terms of this (template, piece): FFT error of every dot
            // product plus the rounding of x~ = x0 - c (eta).
            const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
            const double sQ2 = sqrt(Q2);
            const double eabs = kfft * (static_cast<double>(pn.x) * static_cast<double>(qmax[ti]) +
                                        sQ2 * static_cast<double>(pn.y));
            const double eta = 1.01 * u;
            if (mode == kModeMain) {
                if (has_ch) {
                    const double edot = eta * sQ2 * static_cast<double>(pn.x) * (1.0 + 1e-6) + eabs;
                    const double nslack = edot + amax * fabs(S);
                    const double mu_abs = fabs(static_cast<double>(c)) + amax;
                    const double bound = static_cast<double>(tmpl[t].bound);
                    const ChunkConst K = chunk_const(m);
                    const double invL = cnt == kChunkW ? K.inv_full : 1.0 / static_cast<double>(Lc);
                    const double lo = nonfinite ? -INFINITY
                                                : chunk_dist_lo(pa, pb, ca, cb, Lc, invL, K, Q2, m,
                                                                static_cast<double>(dmax) + nslack, mu_abs);
                    if (!(lo > bound)) {
                        const unsigned long long slot = wave_claim(nrec);
                        if (slot < rec_cap) {
                            rec[slot] = ChunkRec{static_cast<uint32_t>(t), static_cast<uint32_t>(base + i_first), cnt,
                                                 c, round_up_f(eabs), 0u};
                            float *dst = rec_dt + slot * kChunkW;
                            if (cnt == kChunkW) {
                                // Two padded rows of 32 outputs, 16 bytes per store.
                                const float2 *row = reinterpret_cast<const float2 *>(fdata + 72 * (c0 + tid));
                                float4 *d4 = reinterpret_cast<float4 *>(dst);
#pragma unroll 4
                                for (int q = 0; q < 16; q++) {
                                    const float2 a = row[(q >> 3) * 18 + 2 * (q & 7)];
                                    const float2 b = row[(q >> 3) * 18 + 2 * (q & 7) + 1];
                                    d4[q] = make_float4(a.x * invB, a.y * invB, b.x * invB, b.y * invB);
                                }
                            } else {
                                for (int l = 0; l < cnt; l++) {
                                    const int n = n_first + l;
                                    dst[l] = fdata[n + 4 * (n >> 5)] * invB;
                                }
                            }
                        }
                    }
                }
            } else {
                // Seed: per wave, the chunk whose largest dt is largest
                // relative to its core's spread; its best window is queued
                // for a rigorous upper bound.
                float g = -INFINITY;
                if (has_ch && !nonfinite && Lc >= 2) {
                    const double S1 = pb.x - pa.x;
                    const double V = (pb.y - pa.y) - S1 * S1 / static_cast<double>(Lc);
                    if (V > 0.0)
                        g = dmax * static_cast<float>(1.0 / sqrt(V));
                }
                int who = lane;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const float og = __shfl_xor(g, off);
                    const int ow = __shfl_xor(who, off);
                    if (og > g || (og == g && ow < who)) {
                        g = og;
                        who = ow;
                    }
                }
                if (lane == who && g > -INFINITY) {
                    int arg = 0;
                    for (int n = n_end - 1; n >= n_first; n--)
                        if (fdata[n + 4 * (n >> 5)] == vmax)
                            arg = n - n_first;
                    const unsigned long long slot = atomicAdd(nrec, 1ULL);
                    if (slot < rec_cap) {
                        rec[slot] = ChunkRec{static_cast<uint32_t>(t),
                                             static_cast<uint32_t>(base + i_first + arg), 1, c, round_up_f(eabs), 0u};
                        rec_dt[slot * kChunkW] = dmax;
                    }
                }
            }
            __syncthreads();  // LDS is rewritten by the next template
        }
    }
}

// The queued chunks, one wavefront each: the rigorous interval of every
// window (filter.hpp) decides the candidates (main), or gives the seed
// window's upper bound (seed).
__global__ __launch_bounds__(256) void k_chunk_windows(Waveform w, const TmplDev *__restrict__ tmpl,
                                                       const ChunkRec *__restrict__ rec,
                                                       const float *__restrict__ rec_dt,
                                                       const unsigned long long *__restrict__ nrec, uint64_t rec_cap,
                                                       int mode, CandOut co)
{
    constexpr double eta = 1.01 * 5.9604644775390625e-08;
    const uint64_t n = *nrec < rec_cap ? *nrec : rec_cap;
    const int lane = threadIdx.x & 63;
    const uint64_t wv = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * 256) >> 6;
    for (uint64_t r = wv; r < n; r += nw) {
        const ChunkRec R = rec[r];
        const int t = static_cast<int>(R.t);
        const int m = tmpl[t].m;
        const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
        const uint64_t j = static_cast<uint64_t>(R.j0) + static_cast<uint64_t>(lane);
        bool emit = false;
        float est = 0.0f;
        double ilo = 0.0, ihi = 0.0;
        if (lane < R.cnt && w.x[j] != kMissing) {
            const double dt = static_cast<double>(rec_dt[r * kChunkW + static_cast<uint64_t>(lane)]);
            const DistInterval iv = dist_interval(w, j, m, Q2, S, R.c, dt, eta, static_cast<double>(R.eabs));
            if (mode == kModeMain) {
                emit = !(iv.lo > static_cast<double>(tmpl[t].bound));  // NaN (undecidable): a candidate
                est = static_cast<float>(iv.est);
                ilo = iv.lo;
                ihi = iv.hi;
            } else if (iv.hi < 3.0e38) {
                emit = true;
                est = static_cast<float>(iv.hi) * (1.0f + 1e-6f);
                ilo = ihi = est;
            }
        }
        const unsigned long long mask = __ballot(emit);
        if (mask) {
            const int leader = __ffsll(static_cast<long long>(mask)) - 1;
            unsigned long long slot0 = 0;
            if (lane == leader)
                slot0 = atomicAdd(co.count, static_cast<unsigned long long>(__popcll(mask)));
            slot0 = __shfl(slot0, leader);
            if (emit)
                cand_put(co, slot0 + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull))), R.t,
                         static_cast<uint32_t>(j), est, ilo, ihi);
        }
    }
}

__global__ void k_zero_nrec(unsigned long long *p) { *p = 0ull; }

// nrec[1] = max(nrec[1], nrec[0]): the largest queue count of a pass made of
// several launches.
__global__ void k_note_max(unsigned long long *nrec)
{
    if (nrec[0] > nrec[1])
        nrec[1] = nrec[0];
}

}  // namespace

// Table E[t] = e^{i pi t / M}, t in [0, 2M), for the fused pass.
// The fused pass serves B = 2^14 and 2^15; its table E[t] = e^{i pi t / M},
// t in [0, 2M), has B entries.
bool fft_fused_supports(int B) { return B == Geom<13>::kB || B == Geom<14>::kB; }

size_t fft_chunk_rec_bytes() { return sizeof(ChunkRec); }
int fft_chunk_width() { return kChunkW; }

// rec / rec_dt: queue of rec_cap records (rec_dt: kChunkW floats each); nrec[0]
// gets the number the launch produced, nrec[1] the largest such number since
// it was last cleared (more than rec_cap: the pass must be rerun with a larger
// queue in main mode; in seed mode the entries kept suffice).
hipError_t launch_fft_fused(const Waveform &w, const TmplDev *tmpl, int t0, int F, const float2 *spectra, int bins,
                            const float2 *qspec, const float *qmax, const float4 *pinfo, const float2 *pnorm,
                            const float2 *table, uint64_t P, int step, int stride, double kfft, int mode, void *rec,
                            float *rec_dt, unsigned long long *nrec, uint64_t rec_cap, const CandOut &co, int cus,
                            hipStream_t stream)
{
    if (F <= 0 || P == 0)
        return hipSuccess;
    const bool small = bins == Geom<13>::kM + 1;
    if (!small && bins != Geom<14>::kM + 1)
        return hipErrorInvalidValue;
    // Pieces of step windows: every 64-output chunk of one needs a thread.
    if (step < 1 || step > (small ? Geom<13>::kStepMax : Geom<14>::kStepMax))
        return hipErrorInvalidValue;
    const int grid = cus * (small ? Geom<13>::kPerCU : Geom<14>::kPerCU);
    const uint64_t nwork = (P + stride - 1) / stride;
    const unsigned g = static_cast<unsigned>(nwork < static_cast<uint64_t>(grid) ? nwork : grid);
    hipLaunchKernelGGL(k_zero_nrec, dim3(1), dim3(1), 0, stream, nrec);
    if (small)
        hipLaunchKernelGGL(k_fft_fused<13>, dim3(g), dim3(Geom<13>::kT), 0, stream, w, tmpl, t0, F, spectra, bins,
                           qspec, qmax, pinfo, pnorm, table, P, step, stride, kfft, mode,
                           static_cast<ChunkRec *>(rec), rec_dt, nrec, rec_cap);
    else
        hipLaunchKernelGGL(k_fft_fused<14>, dim3(g), dim3(Geom<14>::kT), 0, stream, w, tmpl, t0, F, spectra, bins,
                           qspec, qmax, pinfo, pnorm, table, P, step, stride, kfft, mode,
                           static_cast<ChunkRec *>(rec), rec_dt, nrec, rec_cap);
    hipLaunchKernelGGL(k_chunk_windows, dim3(static_cast<unsigned>(4 * cus)), dim3(256), 0, stream, w, tmpl,
                       static_cast<const ChunkRec *>(rec), rec_dt, nrec, rec_cap, mode, co);
    hipLaunchKernelGGL(k_note_max, dim3(1), dim3(1), 0, stream, nrec);
    return hipGetLastError();
}

}  // namespace tmx
