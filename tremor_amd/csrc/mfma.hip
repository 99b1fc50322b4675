// Sliding dot products of many templates against the waveform as a GEMM on
// the gfx950 matrix cores (v_mfma_f32_32x32x16_{bf16,f16}), with the fused
// candidate filter in the epilogue.
//
// Formulation.  For template t (length m) and a waveform block starting at
// J0, write window j = J0 + 32 c + s (s in [0, 32)).  Then
//     dot(t, j) = sum_l A_t[s][l] * B[l][c],
//     A_t[s][l] = q_t[l - s]      (Toeplitz: 32 shifted copies of q_t),
//     B[l][c]   = x~[J0 + 32 c + l]  (strided Hankel view of the waveform),
// with l over K_t = roundup(m + 31, 16).  One 32x32 output tile holds 1024
// consecutive windows of ONE template, so all its window statistics are for
// one length, and every MFMA operand fragment is 8 consecutive samples: a
// B fragment starts at a multiple of 8 (aligned 16-byte LDS reads, no
// shuffles), an A fragment is 8 consecutive template values, read with one
// unaligned 16-byte load from a zero-padded split copy of the template
// (a few KB, L1-resident: no 32-fold expanded operand streams through the
// caches).  B is staged per workgroup in LDS from the float32 waveform,
// centred on the block mean c_b and split on the fly.
//
// Precision forms (FP32 accumulation in both):
//   split-BF16  x = hi + lo, q = hi + lo, dot ~ hi*hi + hi*lo + lo*hi
//               |error| <= (3.01 * 2^-16 + gamma_acc) sum |q| |x~|
//   FP16        the block is scaled by a power of two into FP16 range (max
//               |x~| -> 2^14), the template by 2^7;
//               |error| <= (2.01 * 2^-11 + gamma_acc) sum |q| |x~| + floor terms,
//               the floor being 2^-14 per operand in scaled units (values the
//               matrix core might flush to zero)
// gamma_acc counts every accumulation rounding at 2^-23 (no assumption on the
// matrix core's internal rounding mode).
//
// Epilogue.  A level-1 test in FP32 (per output: the window's m*sigma^2 from
// the FP64 prefix sums, one rsqrt) rejects windows that cannot reach the
// template's bound with generous slack; survivors get the rigorous FP64
// interval of filter.hpp and are emitted as candidates.  In seed mode each
// lane evaluates its most promising window rigorously and the wavefront
// emits the minimum upper bound.
//
// Workgroup: 4 wavefronts x 4 column tiles x 1024 windows = 16384 windows;
// templates of a group (sorted by length) are processed two at a time.
#include "filter.hpp"

namespace tmx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kNM = 2;                        // templates per pass
constexpr int kNN = 4;                        // 32-column tiles per wavefront
constexpr int kTileWin = 1024;                // windows per 32x32 output tile
constexpr int kNW = kWaves * kNN * kTileWin;  // windows per workgroup
constexpr int kKC = 2048;                     // K samples staged per chunk
constexpr int kGroup = 8;                     // templates per workgroup
constexpr int kStage = kNW + kKC;             // staged samples per chunk

// LDS layout of the staged tile (16-bit elements): rows of 32 elements (one
// column of the Hankel view), 8-element chunks permuted within a row by
// (row / 2) mod 4.  The 16-byte reads of 8 consecutive lanes (consecutive
// rows) then land in 8 distinct 16-byte bank groups, without padding.
__device__ inline int swz(int e) { return e ^ (((e >> 6) & 3) << 3); }

// Epilogue scratch, aliased over the tile once a template pair's K loop is
// done: per wavefront the window statistics of one 1024-window tile (padded
// rows of 33) and a queue of windows that passed the level-1 test, decided
// by the rigorous FP64 filter once the accumulators are no longer live.
constexpr int kStatLen = kTileWin + kTileWin / 32;
constexpr int kQ = 384;  // queue entries per wavefront (>= kNM * 64 for the seed mode)
constexpr int kStatBatch = 2;  // tile-statistics loads in flight per lane
constexpr int kTileBytesBf16 = 2 * kStage * 2;
constexpr int kTileBytesF16 = kStage * 2;
constexpr int kEpiBytes = kWaves * kStatLen * 4 + kWaves * kQ * 12 + kWaves * 16 * 64 * 4;
static_assert(kQ >= kNM * 64, "seed entries need one queue slot per lane and template");
constexpr int cmax(int a, int b) { return a > b ? a : b; }

struct alignas(16) Frag {
    uint32_t w[4];
};

__device__ inline uint16_t bf16_rne(float v)
{
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ inline uint16_t f16_bits(float v)
{
    const _Float16 h = static_cast<_Float16>(v);
    uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    return b;
}

// ---- operand packing --------------------------------------------------------

// Per template, the split template values in two 16-bit arrays (hi, then
// lo) of P = kpad + 64 entries, entry 32 + i holding q[i] and every other
// entry zero.  The lane (s, h) operand fragment of k-step kappa,
// A[s][16 kappa + 8 h + e] = q[16 kappa + 8 h + e - s], e < 8, is then the 16
// bytes at entry 32 + 16 kappa + 8 h - s: one (unaligned) vector load from an
// array small enough to stay in the L1 cache.  One thread per entry.
__global__ void k_pack(const TmplDev *__restrict__ tmpl, int count, const float *__restrict__ qz, int precision,
                       Frag *__restrict__ a)
{
    const int t = blockIdx.y;
    if (t >= count)
        return;
    const int m = tmpl[t].m;
    const int P = tmpl[t].kpad + 64;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P)
        return;
    const int i = e - 32;
    const float v = (i >= 0 && i < m) ? qz[tmpl[t].q_off + i] : 0.0f;
    uint16_t hi, lo;
    if (precision == TMX_PREC_FP16) {
        hi = f16_bits(v * 128.0f);  // |q| * 2^7 <= 2^15: no overflow, tiny values stay normal
        lo = 0;
    } else {
        hi = bf16_rne(v);
        lo = bf16_rne(v - bf16_to_f(hi));
    }
    uint16_t *base = reinterpret_cast<uint16_t *>(a + tmpl[t].a_off);
    base[e] = hi;
    base[P + e] = lo;
}

// ---- level-1 filter -------------------------------------------------------------

// Cap on the reference's FP32 rounding E_ref for windows with |mu|/sigma <=
// 1e3 (larger offsets always go to the rigorous filter).
__device__ inline double eref_cap(int m)
{
    constexpr double u = 5.9604644775390625e-08;
    const double md = m, rm = sqrt(md);
    const double sq = 2.0 * rm + 1.0;
    const double D1 = (u * 1.001) * (3.05 * rm + 1.03 * rm * 1.001e3 + sq);
    const double neff = floor(md / 16.0) + 36.0;
    const double gam = neff * u / (1.0 - neff * u);
    return (D1 * (2.0 * sq + D1) + gam * (sq + D1) * (sq + D1)) * 1.02;
}

__device__ inline double eta_mfma(int precision, int kpad)
{
    const double acc = (16.0 + (precision == TMX_PREC_FP16 ? 1.0 : 3.0) * kpad / 16.0) * 1.1920928955078125e-07;
    const double rep = precision == TMX_PREC_FP16 ? 2.01 * 4.8828125e-04 : 3.01 * 1.52587890625e-05;
    return (rep + acc) * 1.001;
}

// ---- main kernel ------------------------------------------------------------

// Orders LDS traffic among the lanes of one wavefront (its private scratch).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int PREC>
__global__ __launch_bounds__(kThreads, 2) void k_scan_mfma(Waveform w, const TmplDev *__restrict__ tmpl,
                                                           const int32_t *__restrict__ group_begin,
                                                           const Frag *__restrict__ afrag, int mode,
                                                           int block_stride, int groups, uint32_t nblocks,
                                                           const float2 *__restrict__ rb,
                                                           const float4 *__restrict__ binfo, int rb_len,
                                                           CandOut co)
{
    constexpr int kLds = cmax(PREC == TMX_PREC_FP16 ? kTileBytesF16 : kTileBytesBf16, kEpiBytes);
    __shared__ __attribute__((aligned(16))) unsigned char lds[kLds];
    __shared__ float s_red[kWaves];
    uint16_t *const lhi = reinterpret_cast<uint16_t *>(lds);
    uint16_t *const llo = lhi + kStage;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int col = lane & 31, h = lane >> 5;
    float *const stats = reinterpret_cast<float *>(lds) + wave * kStatLen;
    uint32_t *const qj = reinterpret_cast<uint32_t *>(lds + kWaves * kStatLen * 4) + wave * kQ;
    float *const qdt = reinterpret_cast<float *>(lds + kWaves * kStatLen * 4 + kWaves * kQ * 4) + wave * kQ;
    int32_t *const qi = reinterpret_cast<int32_t *>(lds + kWaves * kStatLen * 4 + kWaves * kQ * 8) + wave * kQ;
    float *const sdt = reinterpret_cast<float *>(lds + kWaves * kStatLen * 4 + kWaves * kQ * 12) + wave * 16 * 64;
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs;
    // on each XCD, consecutive workgroups take every template group of one
    // waveform block, so the block's samples and prefix sums are shared in
    // that XCD's L2 instead of being fetched once per group.
    const uint32_t xcd = blockIdx.x & 7u, k8 = blockIdx.x >> 3;
    const int g = static_cast<int>(k8 % static_cast<uint32_t>(groups));
    const uint32_t bidx = (k8 / static_cast<uint32_t>(groups)) * 8u + xcd;
    if (bidx >= nblocks)
        return;
    const uint64_t blk = static_cast<uint64_t>(bidx) * static_cast<uint64_t>(block_stride);
    const uint64_t J0 = blk * kNW;
    if (J0 >= w.n)
        return;
    const int tb = group_begin[g], te = group_begin[g + 1];
    const float4 bi = binfo[blk];  // block centre and error bounds of the rebased prefix pairs
    const float c = bi.x;
    const float2 *const rbb = rb + blk * static_cast<uint64_t>(rb_len);

    // Largest K of the group (templates sorted by length: the last one).
    const int kmax = tmpl[te - 1].kpad;

    // FP16: power-of-two scale so the staged block fits FP16.  The waveform
    // buffer carries zero padding past n, so tile reads need no guard.
    // Largest |x~| over the block's samples: bounds every window's |mu - c|
    // for the tile test (and sets the FP16 scale).
    float mx = 0.0f;
    for (int e = tid; e < kNW + kmax; e += kThreads) {
        float v = w.x[J0 + e];
        v = v == kMissing ? 0.0f : v;
        mx = fmaxf(mx, fabsf(v - c));
    }
    for (int off = 32; off > 0; off >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lane == 0)
        s_red[wave] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3]));
    const float aD_max = mx * 1.0001f;  // >= |x0 - c| over the block, hence >= |mu_j - c|
    float scale = 1.0f, inv_scale = 1.0f;
    if (PREC == TMX_PREC_FP16) {
        // max |x~| * 2^-e <= 2^14
        int ex = 0;
        if (mx > 0.0f) {
            (void)frexpf(mx, &ex);  // mx < 2^ex
            ex -= 14;
        }
        inv_scale = ldexpf(1.0f, -ex);  // staged = x~ * 2^-ex
        scale = ldexpf(1.0f, ex - 7);   // real dot = accumulated * 2^ex * 2^-7
    }

    for (int tp = tb; tp < te; tp += kNM) {
        const int nt = te - tp < kNM ? te - tp : kNM;
        // Operand arrays through buffer resources: the lane's entry offset
        // (shared by both templates) in one register, the k-step and the lo
        // part's offset in scalar registers.
        int ks[kNM];
        uint32_t lo_off[kNM];
        __amdgpu_buffer_rsrc_t ar[kNM];
#pragma unroll
        for (int i = 0; i < kNM; i++) {
            const int ti = tp + (i < nt ? i : 0);
            ks[i] = i < nt && !tmpl[ti].ea ? tmpl[ti].kpad / 16 : 0;
            lo_off[i] = static_cast<uint32_t>(tmpl[ti].kpad + 64) * 2u;
            ar[i] = __builtin_amdgcn_make_buffer_rsrc(const_cast<Frag *>(afrag + tmpl[ti].a_off), static_cast<short>(0),
                                                      static_cast<int>(2u * lo_off[i]), 0x00020000);
        }
        const uint32_t avoff = static_cast<uint32_t>(32 + 8 * h - col) * 2u;
        const int kt = (ks[0] > ks[kNM - 1] ? ks[0] : ks[kNM - 1]) * 16;
        f32x16 acc[kNM][kNN];
#pragma unroll
        for (int i = 0; i < kNM; i++)
#pragma unroll
            for (int jn = 0; jn < kNN; jn++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    acc[i][jn][r] = 0.0f;

        for (int k0 = 0; k0 < kt; k0 += kKC) {
            const int kc = kt - k0 < kKC ? kt - k0 : kKC;
            __syncthreads();
            // Stage x~ = (x0 - c) [* 2^-e] for [J0 + k0, J0 + k0 + kNW + kc), 8 samples per item.
            for (int e8 = tid; e8 < (kNW + kc) / 8; e8 += kThreads) {
                const int e = e8 * 8;
                const float4 *src = reinterpret_cast<const float4 *>(w.x + J0 + k0 + e);
                const float4 v0 = src[0], v1 = src[1];
                const float xv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                float xt[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    xt[u] = (xv[u] == kMissing ? 0.0f : xv[u]) - c;
                Frag fh, fl;
                if (PREC == TMX_PREC_FP16) {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        fh.w[q] = static_cast<uint32_t>(f16_bits(xt[2 * q] * inv_scale)) |
                                  (static_cast<uint32_t>(f16_bits(xt[2 * q + 1] * inv_scale)) << 16);
                } else {
                    // Two samples per conversion (v_cvt_pk_bf16_f32, round to
                    // nearest even); x~ - hi is exact in FP32.
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const f32x2 p = {xt[2 * q], xt[2 * q + 1]};
                        const bf16x2 hp = __builtin_convertvector(p, bf16x2);
                        const bf16x2 lp = __builtin_convertvector(p - __builtin_convertvector(hp, f32x2), bf16x2);
                        fh.w[q] = __builtin_bit_cast(uint32_t, hp);
                        fl.w[q] = __builtin_bit_cast(uint32_t, lp);
                    }
                }
                *reinterpret_cast<Frag *>(&lhi[swz(e)]) = fh;
                if (PREC != TMX_PREC_FP16)
                    *reinterpret_cast<Frag *>(&llo[swz(e)]) = fl;
            }
            __syncthreads();
            // Template operands: loaded at each k-step (the small operand
            // arrays stay in the L1 cache, and the other wavefront of the
            // SIMD covers the latency).
            Frag nh[kNM], nl[kNM];
            auto load_a = [&](int kappa) {
#pragma unroll
                for (int i = 0; i < kNM; i++) {
                    if (kappa < ks[i]) {
                        const uint32_t so = 32u * static_cast<uint32_t>(kappa);
                        nh[i] = __builtin_bit_cast(Frag, __builtin_amdgcn_raw_buffer_load_b128(ar[i], avoff, so, 0));
                        if (PREC != TMX_PREC_FP16)
                            nl[i] = __builtin_bit_cast(
                                Frag, __builtin_amdgcn_raw_buffer_load_b128(ar[i], avoff, so + lo_off[i], 0));
                    } else {
                        nh[i] = Frag{{0u, 0u, 0u, 0u}};
                        nl[i] = Frag{{0u, 0u, 0u, 0u}};
                    }
                }
            };
            for (int kk = 0; kk < kc; kk += 16) {
                const int kappa = (k0 + kk) >> 4;
                load_a(kappa);
                Frag ah[kNM], al[kNM];
#pragma unroll
                for (int i = 0; i < kNM; i++) {
                    ah[i] = nh[i];
                    al[i] = nl[i];
                }
#pragma unroll
                for (int jn = 0; jn < (PREC == TMX_PREC_FP16 ? kNN : 0); jn++) {
                    const int cc = (wave * kNN + jn) * 32 + col;  // column within the block
                    const int e = swz(32 * cc + kk + 8 * h);
                    const Frag bh = *reinterpret_cast<const Frag *>(&lhi[e]);
                    if (PREC == TMX_PREC_FP16) {
                        f16x8 bv;
                        __builtin_memcpy(&bv, &bh, 16);
#pragma unroll
                        for (int i = 0; i < kNM; i++) {
                            f16x8 av;
                            __builtin_memcpy(&av, &ah[i], 16);
                            acc[i][jn] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc[i][jn], 0, 0, 0);
                        }
                    }
                }
                if (PREC != TMX_PREC_FP16) {
                    // Two column tiles at a time: the three products of each
                    // accumulator (same order as always) interleaved with the
                    // other three accumulators' products.
#pragma unroll
                    for (int jp = 0; jp < kNN; jp += 2) {
                        bf16x8 bvh[2], bvl[2];
#pragma unroll
                        for (int q = 0; q < 2; q++) {
                            const int cc = (wave * kNN + jp + q) * 32 + col;
                            const int e = swz(32 * cc + kk + 8 * h);
                            const Frag bh = *reinterpret_cast<const Frag *>(&lhi[e]);
                            const Frag bl = *reinterpret_cast<const Frag *>(&llo[e]);
                            __builtin_memcpy(&bvh[q], &bh, 16);
                            __builtin_memcpy(&bvl[q], &bl, 16);
                        }
                        bf16x8 avh[kNM], avl[kNM];
#pragma unroll
                        for (int i = 0; i < kNM; i++) {
                            __builtin_memcpy(&avh[i], &ah[i], 16);
                            __builtin_memcpy(&avl[i], &al[i], 16);
                        }
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] =
                                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(avl[i], bvh[q], acc[i][jp + q], 0, 0, 0);
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] =
                                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(avh[i], bvl[q], acc[i][jp + q], 0, 0, 0);
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] =
                                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(avh[i], bvh[q], acc[i][jp + q], 0, 0, 0);
                    }
                }
            }
        }

        // ---- epilogue (the tile area of LDS is free from here on) -------------
        // Level 1, while the accumulators are live: FP32 only.  Windows that
        // pass go to the wavefront's queue (or, if it is full, straight to
        // the candidate list: extra candidates cost verification time only).
        int qn = 0;         // wavefront-uniform queue length
        uint32_t ovf = 0;   // tiles (i * kNN + jn) whose survivors overflowed the queue
        __syncthreads();  // the epilogue scratch overlays the tile other wavefronts were reading
#pragma unroll
        for (int i = 0; i < kNM; i++) {
            if (i >= nt)
                continue;
            const int t = tp + i;
            const int m = tmpl[t].m;
            if (static_cast<uint64_t>(m) > w.n || tmpl[t].ea)
                continue;
            const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
            const double Q2 = tmpl[t].Q2;
            const double md = static_cast<double>(m);
            const double eta = eta_mfma(PREC, tmpl[t].kpad);
            // FP16 floor terms: staged |x~| * 2^-ex below 2^-14 and |q| * 2^7
            // below 2^-14 may be flushed.  x side: sum |q| 2^-14 2^ex <=
            // 2^-14 2^ex sqrt(m Q2) (absolute); q side: sum 2^-21 |x~| <=
            // 2^-21 sqrt(m) sqrt(sum x~^2) (relative to the window).
            const double abs_unit = PREC == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;  // 2^-21
            const float gmin =
                static_cast<float>((Q2 + md - static_cast<double>(tmpl[t].bound) - eref_cap(m)) * 0.5 * (1.0 - 1e-3));
            const float eta_sq = static_cast<float>(eta * sqrt(Q2) * (1.0 + 1e-3));
            const float sqrt_m = static_cast<float>(sqrt(md));
            const float mf = static_cast<float>(m);
            const float inv_mf = 1.0f / mf;
            const float rsqrt_m = static_cast<float>(1.0 / sqrt(md)) * (1.0f + 1e-6f);
            const float absS = static_cast<float>(fabs(tmpl[t].S) * (1.0 + 1e-3));
            const float eabs_q = static_cast<float>(abs_unit * sqrt(md) * 1.01);  // * sqrt(sum x~^2)
            const float eabs_x = PREC == TMX_PREC_FP16
                                     ? static_cast<float>(6.103515625e-05 * static_cast<double>(scale) * 128.0 *
                                                          sqrt(md * Q2) * 1.01)
                                     : 0.0f;  // absolute
            float best_g = -INFINITY;
            uint32_t best_j = 0xffffffffu;
            float best_dt = 0.0f;
            // Tile test constants: survive iff dt >= K sqrt(V) - rest, V the
            // window's m sigma^2 (the per-window threshold below, rewritten).
            const float Kc = gmin * rsqrt_m * 0.999f - (eta_sq + eabs_q) * 1.001f;
            const float rest = ((eta_sq + eabs_q) * sqrt_m * aD_max + aD_max * absS + eabs_x) * 1.001f + 1e-30f;
            const float amu_max = fabsf(c) + aD_max;
#pragma unroll
            for (int jn = 0; jn < kNN; jn++) {
                const uint64_t tile0 = J0 + 32ull * static_cast<uint64_t>((wave * kNN + jn) * 32);
                if (mode == kModeMain) {
                    // Tile test.  Column col holds windows j0 + s, s < 32 (this
                    // lane and lane ^ 32 hold 16 each); their common core
                    // [j0 + cnt - 1, j0 + m) bounds every window's m sigma^2
                    // from below, |x~| bounds |mu - c|, and the column's
                    // largest dot product bounds every dt.  A tile none of
                    // whose columns can reach the threshold is skipped.
                    float cm = -INFINITY;
                    bool nonfinite = false;
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const float v = acc[i][jn][r];
                        nonfinite |= !(v - v == 0.0f);
                        cm = fmaxf(cm, v);
                    }
                    cm = fmaxf(cm, __shfl_xor(cm, 32)) * scale;
                    nonfinite = __shfl_xor(static_cast<int>(nonfinite), 32) != 0 || nonfinite;
                    const uint64_t j0 = tile0 + 32ull * static_cast<uint64_t>(col);
                    bool surv = false;
                    if (j0 < W) {
                        surv = true;
                        const int cnt = W - j0 < 32 ? static_cast<int>(W - j0) : 32;
                        const int Lc = m - (cnt - 1);
                        if (!nonfinite && Lc >= 2 && gmin > 0.0f && Kc > 0.0f) {
                            const float2 ra = rbb[static_cast<int>(j0 - J0) + cnt - 1];
                            const float2 rz = rbb[static_cast<int>(j0 - J0) + m];
                            const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
                            const float eS1 = 2.0f * bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
                            const float eS2 = 2.0f * bi.z + 1.2e-7f * (fabsf(ra.y) + fabsf(rz.y) + fabsf(S2));
                            const float invL = 1.0f / static_cast<float>(Lc);
                            const float V = S2 - S1 * S1 * invL;
                            const float eV = eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * invL +
                                             2.4e-7f * (fabsf(S2) + S1 * S1 * invL);
                            const float Vlo = V - eV;
                            if (Vlo > 1e3f * eV && Vlo > mf * 1e-20f && !(amu_max * amu_max * mf > 1e6f * Vlo)) {
                                const float thr = Kc * (Vlo * __frsqrt_rn(Vlo)) * 0.999f - rest;
                                surv = !(cm < thr);
                            }
                        }
                    }
                    if (__ballot(surv) == 0ull)
                        continue;
                }
                // Per window of the tile, coalesced, from the block-rebased
                // prefix pairs: main mode, the threshold the dot product must
                // reach to survive level 1 (-inf: cannot be judged - stats
                // swamped by rounding or |mu| / sigma > 1e3 - so it survives;
                // +inf: past the last window); seed mode, 1 / sqrt(m sigma^2)
                // (or -1).
                wave_sync();
#pragma unroll 1
                for (int it0 = 0; it0 < 16; it0 += kStatBatch) {
                    float2 ra[kStatBatch], rz[kStatBatch];
#pragma unroll
                    for (int u = 0; u < kStatBatch; u++) {
                        const int i0 = static_cast<int>(tile0 - J0) + (it0 + u) * 64 + lane;
                        ra[u] = rbb[i0];
                        rz[u] = rbb[i0 + m];
                    }
#pragma unroll
                    for (int u = 0; u < kStatBatch; u++) {
                        const int wl = (it0 + u) * 64 + lane;
                        const uint64_t j = tile0 + static_cast<uint64_t>(wl);
                        float thr = INFINITY;
                        if (j < W) {
                            // Window sums of x0 - c and (x0 - c)^2 with their error bounds.
                            const float S1 = rz[u].x - ra[u].x, S2 = rz[u].y - ra[u].y;
                            const float eS1 = 2.0f * bi.y + 1.2e-7f * (fabsf(ra[u].x) + fabsf(rz[u].x) + fabsf(S1));
                            const float eS2 = 2.0f * bi.z + 1.2e-7f * (fabsf(ra[u].y) + fabsf(rz[u].y) + fabsf(S2));
                            const float V = S2 - S1 * S1 * inv_mf;
                            const float eV = eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * inv_mf +
                                             2.4e-7f * (fabsf(S2) + S1 * S1 * inv_mf);
                            const float aD = fabsf(S1 * inv_mf) + eS1 * inv_mf;  // >= |mu - c|
                            const float amu = fabsf(c) + aD;
                            const bool judge = V > 1e3f * eV && V > mf * 1e-20f && !(amu * amu * mf > 1e6f * V);
                            if (mode == kModeMain) {
                                if (!judge || gmin <= 0.0f) {
                                    thr = -INFINITY;
                                } else {
                                    const float sV = V * __frsqrt_rn(V);
                                    const float b = eta_sq * (sV + sqrt_m * aD) + aD * absS +
                                                    eabs_q * (sV + sqrt_m * aD) + eabs_x;
                                    thr = gmin * sV * rsqrt_m * 0.999f - b * 1.001f - 1e-30f;
                                }
                            } else {
                                thr = judge ? __frsqrt_rn(V) : -1.0f;
                            }
                        }
                        stats[wl + (wl >> 5)] = thr;
                    }
                }
                // The tile's dot products go to LDS (this tile's accumulators
                // die here) and the level-1 test runs as a loop.
#pragma unroll
                for (int r = 0; r < 16; r++)
                    sdt[r * 64 + lane] = acc[i][jn][r] * scale;
                wave_sync();
#pragma unroll 2
                for (int r = 0; r < 16; r++) {
                    const int wl = 32 * col + 8 * (r >> 2) + 4 * h + (r & 3);
                    const float thr = stats[wl + (wl >> 5)];
                    const float dt = sdt[r * 64 + lane];
                    if (mode == kModeMain) {
                        const bool survive = dt >= thr;
                        const uint64_t mask = __ballot(survive);
                        if (mask != 0ull) {
                            // Survivors go to the queue; a tile whose survivors do
                            // not fit is marked and, once the accumulators are
                            // dead, all its windows become candidates (extra
                            // candidates cost verification time only).
                            const int idx = qn + __popcll(mask & ((1ull << lane) - 1ull));
                            if (survive && idx < kQ) {
                                qj[idx] = static_cast<uint32_t>(tile0 + static_cast<uint64_t>(wl));
                                qdt[idx] = dt;
                                qi[idx] = i * kNN + jn;
                            }
                            qn += __popcll(mask);
                            if (qn > kQ)
                                ovf |= 1u << (i * kNN + jn);
                        }
                    } else if (thr > 0.0f && thr < INFINITY) {
                        const float gapprox = dt * thr;
                        if (gapprox > best_g) {
                            best_g = gapprox;
                            best_j = static_cast<uint32_t>(tile0 + static_cast<uint64_t>(wl));
                            best_dt = dt;
                        }
                    }
                }
            }
            if (mode == kModeSeed) {
                qj[i * 64 + lane] = best_j;
                qdt[i * 64 + lane] = best_dt;
            }
        }
        __syncthreads();

        // Level 2: the rigorous FP64 interval of filter.hpp, accumulators dead.
        if (mode == kModeMain) {
            // Overflowed tiles: every valid window is a candidate.
            for (int b = 0; b < kNM * kNN; b++) {
                if (!((ovf >> b) & 1u))
                    continue;
                const int t = tp + b / kNN;
                const int m = tmpl[t].m;
                const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
                const uint64_t tile0 = J0 + 32ull * static_cast<uint64_t>((wave * kNN + b % kNN) * 32);
                for (int wl = lane; wl < kTileWin; wl += 64) {
                    const uint64_t j = tile0 + static_cast<uint64_t>(wl);
                    if (j >= W || w.x[j] == kMissing)
                        continue;
                    cand_put(co, wave_claim(co.count), static_cast<uint32_t>(t), static_cast<uint32_t>(j), NAN,
                             -INFINITY, INFINITY);
                }
            }
            const int qe = qn < kQ ? qn : kQ;
            for (int e = lane; e < qe; e += 64) {
                if ((ovf >> qi[e]) & 1u)
                    continue;
                const int t = tp + qi[e] / kNN;
                const uint64_t j = qj[e];
                if (w.x[j] == kMissing)
                    continue;
                const int m = tmpl[t].m;
                const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
                const double md = static_cast<double>(m);
                const double abs_unit = PREC == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;
                const double ea = eta_mfma(PREC, tmpl[t].kpad) + abs_unit * sqrt(md / Q2) * 1.01;
                DistInterval iv = dist_interval(w, j, m, Q2, S, c, static_cast<double>(qdt[e]), ea);
                if (PREC == TMX_PREC_FP16 && iv.lo > -INFINITY) {
                    const WindowSums ws = window_sums_fast(w, j, j + static_cast<uint64_t>(m));
                    const double S1 = ws.S1;
                    const double V = ws.S2 - S1 * S1 / md;
                    const double sig = sqrt(fmax(V, 1e-300) / md);
                    const double ex = 6.103515625e-05 * static_cast<double>(scale) * 128.0 * sqrt(md * Q2) * 1.01;
                    iv.lo -= 2.0 * ex / (sig * (1.0 - 1e-3)) * 1.01;
                }
                if (iv.lo <= static_cast<double>(tmpl[t].bound))
                    cand_put(co, wave_claim(co.count), static_cast<uint32_t>(t), static_cast<uint32_t>(j),
                             static_cast<float>(iv.est), iv.lo, iv.hi);
            }
        } else {
            for (int i = 0; i < nt; i++) {
                const int t = tp + i;
                const int m = tmpl[t].m;
                float ub = INFINITY;
                const uint32_t bj0 = qj[i * 64 + lane];
                if (static_cast<uint64_t>(m) <= w.n && !tmpl[t].ea && bj0 != 0xffffffffu && w.x[bj0] != kMissing) {
                    const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
                    const double md = static_cast<double>(m);
                    const double abs_unit = PREC == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;
                    const double ea = eta_mfma(PREC, tmpl[t].kpad) + abs_unit * sqrt(md / Q2) * 1.01;
                    DistInterval iv =
                        dist_interval(w, bj0, m, Q2, S, c, static_cast<double>(qdt[i * 64 + lane]), ea);
                    if (PREC == TMX_PREC_FP16 && iv.hi < INFINITY) {
                        const WindowSums ws = window_sums_fast(w, bj0, bj0 + static_cast<uint64_t>(m));
                        const double S1 = ws.S1;
                        const double V = ws.S2 - S1 * S1 / md;
                        const double sig = sqrt(fmax(V, 1e-300) / md);
                        const double ex =
                            6.103515625e-05 * static_cast<double>(scale) * 128.0 * sqrt(md * Q2) * 1.01;
                        iv.hi += 2.0 * ex / (sig * (1.0 - 1e-3)) * 1.01;
                    }
                    if (iv.hi < 3.0e38)
                        ub = static_cast<float>(iv.hi) * (1.0f + 1e-6f);
                }
                uint32_t bj = bj0;
                for (int off = 32; off > 0; off >>= 1) {
                    const float ov = __shfl_xor(ub, off);
                    const uint32_t oj = __shfl_xor(bj, off);
                    if (ov < ub || (ov == ub && oj < bj)) {
                        ub = ov;
                        bj = oj;
                    }
                }
                if (lane == 0 && ub < INFINITY)
                    cand_put(co, atomicAdd(co.count, 1ULL), static_cast<uint32_t>(t), bj, ub, ub, ub);
            }
        }
    }
}

}  // namespace

int mfma_group_size() { return kGroup; }
uint64_t mfma_block_windows() { return kNW; }
int mfma_kpad(int m) { return ((m + 31 + 15) / 16) * 16; }
// Two 16-bit arrays (hi, lo) of kpad + 64 entries, in 16-byte units.
int64_t mfma_frag_count(int kpad) { return static_cast<int64_t>(kpad + 64) / 4; }

hipError_t launch_mfma_pack(const TmplDev *tmpl, int count, const float *qz, int precision, void *a_frag,
                            hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    // Longest template: kpad = 65536 + 32 (+ 64 padding entries).
    const int per_t = 65536 + 32 + 64;
    const dim3 grid((per_t + 255) / 256, count);
    hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, stream, tmpl, count, qz, precision, static_cast<Frag *>(a_frag));
    return hipGetLastError();
}

hipError_t launch_scan_mfma(const Waveform &w, const TmplDev *tmpl, int count, const int32_t *group_begin, int groups,
                            int precision, const void *a_frag, const float2 *rb, const float4 *binfo, int rb_len,
                            int mode, int block_stride, const CandOut &co, hipStream_t stream)
{
    (void)count;
    if (groups <= 0 || w.n == 0)
        return hipSuccess;
    const uint64_t blocks = (w.n + kNW - 1) / kNW;
    const uint64_t launched = (blocks + block_stride - 1) / block_stride;
    const uint64_t total = ((launched + 7) / 8) * 8 * static_cast<uint64_t>(groups);
    if (total >= (1ull << 32))
        return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>(total));
    const Frag *a = static_cast<const Frag *>(a_frag);
    if (precision == TMX_PREC_FP16)
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_FP16>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a, mode,
                           block_stride, groups, static_cast<uint32_t>(launched), rb, binfo, rb_len, co);
    else
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_BF16X3>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a,
                           mode, block_stride, groups, static_cast<uint32_t>(launched), rb, binfo, rb_len, co);
    return hipGetLastError();
}

}  // namespace tmx
