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
// one length, and every MFMA operand fragment is 8 consecutive samples
// starting at a multiple of 8: aligned 16-byte loads, no shuffles.  A_t is
// packed once per call in the exact per-lane fragment order (1 KB per
// k-step), B is staged per workgroup in LDS from the float32 waveform,
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
// Workgroup: 4 wavefronts x 2 column tiles x 1024 windows = 8192 windows;
// templates of a group (sorted by length) are processed two at a time.
#include "filter.hpp"

namespace tmx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kNM = 2;                        // templates per pass
constexpr int kNN = 2;                        // 32-column tiles per wavefront
constexpr int kTileWin = 1024;                // windows per 32x32 output tile
constexpr int kNW = kWaves * kNN * kTileWin;  // windows per workgroup
constexpr int kKC = 1024;                     // K samples staged per chunk
constexpr int kGroup = 8;                     // templates per workgroup
constexpr int kStage = kNW + kKC;             // staged samples per chunk

// LDS swizzle in 16-bit elements: 8 elements of padding after every 32, so
// the 16-byte reads of 8 consecutive lanes (columns 64 bytes apart) hit
// distinct banks.  8-element groups never straddle a 32-element row.
__device__ inline int lds_at(int e) { return e + ((e >> 5) << 3); }
constexpr int kLdsLen = kStage + (kStage >> 2);

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

// One thread per (template, k-step, lane): the lane's 8 Toeplitz entries
// q[16 kappa + 8 h + e - s], split into parts.
__global__ void k_pack(const TmplDev *__restrict__ tmpl, int count, const float *__restrict__ qz, int precision,
                       Frag *__restrict__ a)
{
    const int t = blockIdx.y;
    if (t >= count)
        return;
    const int m = tmpl[t].m;
    const int ksteps = tmpl[t].kpad / 16;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int kappa = idx >> 6, lane = idx & 63;
    if (kappa >= ksteps)
        return;
    const float *q = qz + tmpl[t].q_off;
    const int s = lane & 31, h = lane >> 5;
    uint16_t hi[8], lo[8];
    for (int e = 0; e < 8; e++) {
        const int i = 16 * kappa + 8 * h + e - s;
        const float v = (i >= 0 && i < m) ? q[i] : 0.0f;
        if (precision == TMX_PREC_FP16) {
            hi[e] = f16_bits(v * 128.0f);  // |q| * 2^7 <= 2^15: no overflow, tiny values stay normal
            lo[e] = 0;
        } else {
            hi[e] = bf16_rne(v);
            lo[e] = bf16_rne(v - bf16_to_f(hi[e]));
        }
    }
    Frag fh, fl;
    for (int w = 0; w < 4; w++) {
        fh.w[w] = static_cast<uint32_t>(hi[2 * w]) | (static_cast<uint32_t>(hi[2 * w + 1]) << 16);
        fl.w[w] = static_cast<uint32_t>(lo[2 * w]) | (static_cast<uint32_t>(lo[2 * w + 1]) << 16);
    }
    const int64_t base = tmpl[t].a_off + static_cast<int64_t>(kappa) * 2 * 64;
    a[base + lane] = fh;
    a[base + 64 + lane] = fl;
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

template <int PREC>
__global__ __launch_bounds__(kThreads, 2) void k_scan_mfma(Waveform w, const TmplDev *__restrict__ tmpl,
                                                           const int32_t *__restrict__ group_begin,
                                                           const Frag *__restrict__ afrag, int mode,
                                                           int block_stride, Cand *__restrict__ cand,
                                                           float *__restrict__ cand_val,
                                                           unsigned long long *__restrict__ cand_count, uint64_t cap)
{
    __shared__ __attribute__((aligned(16))) uint16_t lhi[kLdsLen];
    __shared__ __attribute__((aligned(16))) uint16_t llo[PREC == TMX_PREC_FP16 ? 8 : kLdsLen];
    __shared__ float s_red[kThreads / 64];
    __shared__ float s_dt[kWaves][16][64];  // dot products of a tile with filter survivors
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int col = lane & 31, h = lane >> 5;
    const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * static_cast<uint64_t>(block_stride);
    const uint64_t J0 = blk * kNW;
    if (J0 >= w.n)
        return;
    const int g = blockIdx.y;
    const int tb = group_begin[g], te = group_begin[g + 1];
    const uint64_t jend = J0 + kNW < w.n ? J0 + kNW : w.n;
    const float c = static_cast<float>((w.p1[jend] - w.p1[J0]) / static_cast<double>(jend - J0));

    // Largest K of the group (templates sorted by length: the last one).
    const int kmax = tmpl[te - 1].kpad;

    // FP16: power-of-two scale so the staged block fits FP16.
    float scale = 1.0f, inv_scale = 1.0f;
    if (PREC == TMX_PREC_FP16) {
        float mx = 0.0f;
        for (uint64_t e = tid; e < static_cast<uint64_t>(kNW + kmax); e += kThreads) {
            const uint64_t p = J0 + e;
            float v = p < w.n ? w.x[p] : 0.0f;
            v = v == kMissing ? 0.0f : v;
            mx = fmaxf(mx, fabsf(v - c));
        }
        for (int off = 32; off > 0; off >>= 1)
            mx = fmaxf(mx, __shfl_xor(mx, off));
        if (lane == 0)
            s_red[wave] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3]));
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
        int ks[kNM];
        int64_t aoff[kNM];
#pragma unroll
        for (int i = 0; i < kNM; i++) {
            const int ti = tp + (i < nt ? i : 0);
            ks[i] = i < nt ? tmpl[ti].kpad / 16 : 0;
            aoff[i] = tmpl[ti].a_off;
        }
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
                uint16_t hv[8], lv[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint64_t p = J0 + static_cast<uint64_t>(k0 + e + u);
                    float v = p < w.n ? w.x[p] : 0.0f;
                    v = v == kMissing ? 0.0f : v;
                    const float xt = v - c;
                    if (PREC == TMX_PREC_FP16) {
                        hv[u] = f16_bits(xt * inv_scale);
                    } else {
                        hv[u] = bf16_rne(xt);
                        lv[u] = bf16_rne(xt - bf16_to_f(hv[u]));
                    }
                }
                Frag fh;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    fh.w[q] = static_cast<uint32_t>(hv[2 * q]) | (static_cast<uint32_t>(hv[2 * q + 1]) << 16);
                *reinterpret_cast<Frag *>(&lhi[lds_at(e)]) = fh;
                if (PREC != TMX_PREC_FP16) {
                    Frag fl;
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        fl.w[q] = static_cast<uint32_t>(lv[2 * q]) | (static_cast<uint32_t>(lv[2 * q + 1]) << 16);
                    *reinterpret_cast<Frag *>(&llo[lds_at(e)]) = fl;
                }
            }
            __syncthreads();
            for (int kk = 0; kk < kc; kk += 16) {
                const int kappa = (k0 + kk) >> 4;
                Frag ah[kNM], al[kNM];
#pragma unroll
                for (int i = 0; i < kNM; i++) {
                    if (kappa < ks[i]) {
                        const Frag *src = afrag + aoff[i] + static_cast<int64_t>(kappa) * 128;
                        ah[i] = src[lane];
                        al[i] = src[64 + lane];
                    } else {
                        ah[i] = Frag{{0u, 0u, 0u, 0u}};
                        al[i] = Frag{{0u, 0u, 0u, 0u}};
                    }
                }
#pragma unroll
                for (int jn = 0; jn < kNN; jn++) {
                    const int cc = (wave * kNN + jn) * 32 + col;  // column within the block
                    const int e = 32 * cc + kk + 8 * h;
                    const Frag bh = *reinterpret_cast<const Frag *>(&lhi[lds_at(e)]);
                    if (PREC == TMX_PREC_FP16) {
                        f16x8 bv;
                        __builtin_memcpy(&bv, &bh, 16);
#pragma unroll
                        for (int i = 0; i < kNM; i++) {
                            f16x8 av;
                            __builtin_memcpy(&av, &ah[i], 16);
                            acc[i][jn] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc[i][jn], 0, 0, 0);
                        }
                    } else {
                        const Frag bl = *reinterpret_cast<const Frag *>(&llo[lds_at(e)]);
                        bf16x8 bvh, bvl;
                        __builtin_memcpy(&bvh, &bh, 16);
                        __builtin_memcpy(&bvl, &bl, 16);
#pragma unroll
                        for (int i = 0; i < kNM; i++) {
                            bf16x8 avh, avl;
                            __builtin_memcpy(&avh, &ah[i], 16);
                            __builtin_memcpy(&avl, &al[i], 16);
                            acc[i][jn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(avl, bvh, acc[i][jn], 0, 0, 0);
                            acc[i][jn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(avh, bvl, acc[i][jn], 0, 0, 0);
                            acc[i][jn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(avh, bvh, acc[i][jn], 0, 0, 0);
                        }
                    }
                }
            }
        }

        // ---- epilogue -------------------------------------------------------
#pragma unroll
        for (int i = 0; i < kNM; i++) {
            if (i >= nt)
                continue;
            const int t = tp + i;
            const int m = tmpl[t].m;
            if (static_cast<uint64_t>(m) > w.n)
                continue;
            const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
            const double Q2 = tmpl[t].Q2, S = tmpl[t].S;
            const double bound = static_cast<double>(tmpl[t].bound);
            const double eta = eta_mfma(PREC, tmpl[t].kpad);
            const double md = static_cast<double>(m);
            const double inv_m = 1.0 / md;
            // FP16 floor terms: staged |x~| * 2^-ex below 2^-14 and |q| * 2^7
            // below 2^-14 may be flushed.  x side: sum |q| 2^-14 2^ex <=
            // 2^-14 2^ex sqrt(m Q2) (absolute); q side: sum 2^-21 |x~| <=
            // 2^-21 sqrt(m) sqrt(sum x~^2) (relative to the window).
            const double abs_unit = PREC == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;  // 2^-21
            const double ea = eta + abs_unit * sqrt(md / Q2) * 1.01;
            const double eref = eref_cap(m);
            const float gmin = static_cast<float>((Q2 + md - bound - eref) * 0.5 * (1.0 - 1e-3));
            const float eta_sq = static_cast<float>(eta * sqrt(Q2) * (1.0 + 1e-3));
            const float sqrt_m = static_cast<float>(sqrt(md));
            const float rsqrt_m = static_cast<float>(1.0 / sqrt(md)) * (1.0f + 1e-6f);
            const float absS = static_cast<float>(fabs(S) * (1.0 + 1e-3));
            const float eabs_q = static_cast<float>(abs_unit * sqrt(md) * 1.01);  // * sqrt(sum x~^2)
            const float eabs_x = PREC == TMX_PREC_FP16
                                     ? static_cast<float>(6.103515625e-05 * static_cast<double>(scale) * 128.0 *
                                                          sqrt(md * Q2) * 1.01)
                                     : 0.0f;  // absolute
            float best_g = -INFINITY;
            uint32_t best_j = 0xffffffffu;
            float best_dt = 0.0f;
#pragma unroll
            for (int jn = 0; jn < kNN; jn++) {
                const uint64_t jcol = J0 + 32 * static_cast<uint64_t>((wave * kNN + jn) * 32 + col);
                uint32_t surv = 0;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const uint64_t j = jcol + static_cast<uint64_t>(8 * (r >> 2) + 4 * h + (r & 3));
                    if (j >= W)
                        continue;
                    const float dt = acc[i][jn][r] * scale;
                    const double p1a = w.p1[j], p1b = w.p1[j + m];
                    const double p2a = w.p2[j], p2b = w.p2[j + m];
                    const double S1 = p1b - p1a, S2 = p2b - p2a;
                    const double V = S2 - S1 * S1 * inv_m;
                    const double eVc = kPrefixGamma * (p2b + p2a) * 4.0 + 1e-15 * fabs(S2);
                    const bool shaky = !(V > 1e3 * eVc) || !(V > md * 1e-20);
                    const double mu = S1 * inv_m;
                    // |mu| / sigma > 1e3: beyond the cap of the level-1 E_ref bound.
                    const bool offset = mu * mu * md > 1e6 * V;
                    const float Vf = static_cast<float>(V);
                    if (mode == kModeMain) {
                        bool survive = shaky || offset || gmin <= 0.0f;
                        if (!survive) {
                            const float D = static_cast<float>(mu) - c;
                            const float sV = Vf * __frsqrt_rn(Vf);
                            const float rhs = gmin * sV * rsqrt_m;
                            const float b = eta_sq * (sV + sqrt_m * fabsf(D)) + fabsf(D) * absS +
                                            eabs_q * (sV + sqrt_m * fabsf(D)) + eabs_x;
                            survive = dt + b * 1.001f + 1e-30f >= rhs * 0.999f;
                        }
                        surv |= survive ? (1u << r) : 0u;
                    } else if (!shaky && !offset && Vf > 0.0f) {
                        const float gapprox = dt * __frsqrt_rn(Vf);
                        if (gapprox > best_g) {
                            best_g = gapprox;
                            best_j = static_cast<uint32_t>(j);
                            best_dt = dt;
                        }
                    }
                }
                if (mode == kModeMain && __any(surv != 0u)) {
                    // Rare: park the tile's dot products in LDS and walk the
                    // survivors with one instance of the rigorous filter.
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        s_dt[wave][r][lane] = acc[i][jn][r] * scale;
                    while (surv != 0u) {
                        const int r = __builtin_ctz(surv);
                        surv &= surv - 1u;
                        const uint64_t j = jcol + static_cast<uint64_t>(8 * (r >> 2) + 4 * h + (r & 3));
                        if (w.x[j] == kMissing)
                            continue;
                        const float dt = s_dt[wave][r][lane];
                        DistInterval iv = dist_interval(w, j, m, Q2, S, c, static_cast<double>(dt), ea);
                        if (PREC == TMX_PREC_FP16 && iv.lo > -INFINITY) {
                            const double V = (w.p2[j + m] - w.p2[j]) -
                                             (w.p1[j + m] - w.p1[j]) * (w.p1[j + m] - w.p1[j]) * inv_m;
                            const double sig = sqrt(fmax(V, 1e-300) * inv_m);
                            iv.lo -= 2.0 * static_cast<double>(eabs_x) / (sig * (1.0 - 1e-3)) * 1.01;
                        }
                        if (iv.lo <= bound) {
                            const unsigned long long slot = atomicAdd(cand_count, 1ULL);
                            if (slot < cap) {
                                cand[slot].t = static_cast<uint32_t>(t);
                                cand[slot].j = static_cast<uint32_t>(j);
                                cand_val[slot] = static_cast<float>(iv.est);
                            }
                        }
                    }
                }
            }
            if (mode == kModeSeed) {
                float ub = INFINITY;
                if (best_j != 0xffffffffu && w.x[best_j] != kMissing) {
                    DistInterval iv = dist_interval(w, best_j, m, Q2, S, c, static_cast<double>(best_dt), ea);
                    if (PREC == TMX_PREC_FP16 && iv.hi < INFINITY) {
                        const double p1a = w.p1[best_j], p1b = w.p1[best_j + m];
                        const double p2a = w.p2[best_j], p2b = w.p2[best_j + m];
                        const double V = (p2b - p2a) - (p1b - p1a) * (p1b - p1a) * inv_m;
                        const double sig = sqrt(fmax(V, 1e-300) * inv_m);
                        iv.hi += 2.0 * static_cast<double>(eabs_x) / (sig * (1.0 - 1e-3)) * 1.01;
                    }
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
                if (lane == 0 && ub < INFINITY) {
                    const unsigned long long slot = atomicAdd(cand_count, 1ULL);
                    if (slot < cap) {
                        cand[slot].t = static_cast<uint32_t>(t);
                        cand[slot].j = bj;
                        cand_val[slot] = ub;
                    }
                }
            }
        }
    }
}

}  // namespace

int mfma_group_size() { return kGroup; }
uint64_t mfma_block_windows() { return kNW; }
int mfma_kpad(int m) { return ((m + 31 + 15) / 16) * 16; }
// Two parts (hi, lo) of 64 lanes per k-step, in 16-byte fragments.
int64_t mfma_frag_count(int kpad) { return static_cast<int64_t>(kpad / 16) * 128; }

hipError_t launch_mfma_pack(const TmplDev *tmpl, int count, const float *qz, int precision, void *a_frag,
                            hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    // Longest template: 65536 + 31 -> 4098 k-steps * 64 lanes.
    const int per_t = (65536 + 47) / 16 * 64;
    const dim3 grid((per_t + 255) / 256, count);
    hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, stream, tmpl, count, qz, precision, static_cast<Frag *>(a_frag));
    return hipGetLastError();
}

hipError_t launch_scan_mfma(const Waveform &w, const TmplDev *tmpl, int count, const int32_t *group_begin, int groups,
                            int precision, const void *a_frag, int mode, int block_stride, Cand *cand, float *cand_val,
                            unsigned long long *cand_count, uint64_t cand_cap, hipStream_t stream)
{
    (void)count;
    if (groups <= 0 || w.n == 0)
        return hipSuccess;
    const uint64_t blocks = (w.n + kNW - 1) / kNW;
    const uint64_t launched = (blocks + block_stride - 1) / block_stride;
    const dim3 grid(static_cast<unsigned>(launched), static_cast<unsigned>(groups));
    const Frag *a = static_cast<const Frag *>(a_frag);
    if (precision == TMX_PREC_FP16)
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_FP16>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a, mode,
                           block_stride, cand, cand_val, cand_count, cand_cap);
    else
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_BF16X3>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a,
                           mode, block_stride, cand, cand_val, cand_count, cand_cap);
    return hipGetLastError();
}

}  // namespace tmx
