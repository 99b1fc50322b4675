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
// the block-rebased prefix sums, one rsqrt) rejects windows that cannot reach
// the template's bound with generous slack.  It runs on the accumulators in
// place (registers and global loads only), and its survivors go to a global
// queue; k_level2 gives each the rigorous FP64 interval of filter.hpp and
// emits the candidates.  In seed mode each lane evaluates its most promising
// window rigorously and the wavefront emits the minimum upper bound.
//
// Workgroup (one per CU): 8 wavefronts x 4 column tiles x 1024 windows =
// 32768 windows; templates of a group (sorted by length) are processed two at
// a time.  When the group's longest template fits one staged chunk (K <=
// 2048) the block is staged once and the wavefronts then run through every
// template pair with no further barrier: the epilogue leaves LDS untouched,
// and the two wavefronts of each SIMD take the pairs in opposite orders, so
// one's epilogue overlaps the other's matrix-core work.  Inside a K chunk the template operands are loaded one k-step ahead and the
// waveform fragments of each column-tile pair are reloaded for the next
// k-step right after their last product.
#include "filter.hpp"

#include <type_traits>

namespace tmx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = 8;  // two per SIMD, one workgroup per CU
constexpr int kThreads = 64 * kWaves;
constexpr int kNM = 2;                        // templates per pass
constexpr int kNN = 4;                        // 32-column tiles per wavefront
constexpr int kTileWin = 1024;                // windows per 32x32 output tile
constexpr int kNW = kWaves * kNN * kTileWin;  // windows per workgroup
constexpr int kKC = 2048;                     // K samples staged per chunk
constexpr int kGroup = 8;                     // templates per workgroup
constexpr int kStage = kNW + kKC;             // staged samples per chunk
constexpr uint32_t kTileB = 4;                // blocks per template group in the dispatch order (per XCD)

// LDS layout of the staged tile (16-bit elements): rows of 32 elements (one
// column of the Hankel view), the four 16-byte chunks of a row permuted by
// (row / 4) mod 4.  A 16-byte read by 16 consecutive lanes (16 consecutive
// rows, any chunk) then covers all 64 banks once: each row residue mod 16
// takes its own 16-byte slot of the 256 bytes the banks serve per cycle.
// The permutation touches bits 3-4 only and reads bits 7-8, so it commutes
// with adding multiples of 1024: the column tiles of a wavefront share one
// address plus immediate offsets.
__device__ inline int swz(int e) { return e ^ (((e >> 7) & 3) << 3); }

constexpr int kTileBytesBf16 = 2 * kStage * 2;
constexpr int kTileBytesF16 = kStage * 2;
constexpr int kLoBytes = kStage * 2;  // byte offset of the lo half of the tile

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

// Per template, the split template values (hi, then lo).  Let P be the
// zero-padded template of L = kpad + 64 entries, P[32 + i] = q[i] and every
// other entry zero.  The lane (s, h) operand fragment of k-step kappa,
// A[s][16 kappa + 8 h + e] = q[16 kappa + 8 h + e - s], e < 8, is the 8
// entries of P from E = 32 + 16 kappa + 8 h - s, which start at every
// residue mod 8 across the lanes.  So each half holds P in 8 shifted copies,
// interleaved by 8-entry chunk: entry (g, c, e) = P[8 g + c + e] (0 past L),
// g < L / 8, c < 8.  A lane reads copy c = (-s) mod 8 at chunk g = (E - c) / 8:
// one aligned 16-byte load, and a wavefront's 64 loads of one k-step fall in
// 640 consecutive bytes.  One thread per entry.
__global__ void k_pack(const TmplDev *__restrict__ tmpl, int count, const float *__restrict__ qz, int precision,
                       Frag *__restrict__ a)
{
    const int t = blockIdx.y;
    if (t >= count)
        return;
    const int m = tmpl[t].m;
    const int L = tmpl[t].kpad + 64;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    uint16_t *const base = reinterpret_cast<uint16_t *>(a + tmpl[t].a_off);
    // The 64 fragments after the last template are zero too: no read of the
    // operand arrays meets uninitialised memory.
    if (t == count - 1 && idx < 512)
        base[16 * L + idx] = 0;
    if (idx >= 8 * L)
        return;
    const int g = idx >> 6, c = (idx >> 3) & 7, e = idx & 7;
    const int src = 8 * g + c + e;  // entry of P
    const int i = src - 32;
    const float v = (src < L && i >= 0 && i < m) ? qz[tmpl[t].q_off + i] : 0.0f;
    uint16_t hi, lo;
    if (precision == TMX_PREC_FP16) {
        hi = f16_bits(v * 128.0f);  // |q| * 2^7 <= 2^15: no overflow, tiny values stay normal
        lo = 0;
    } else {
        hi = bf16_rne(v);
        lo = bf16_rne(v - bf16_to_f(hi));
    }
    base[idx] = hi;
    base[8 * L + idx] = lo;
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


// Level-1 constants of every matrix-core template from its final bound (one
// thread per template; FP64 once here instead of in every workgroup).
__global__ void k_mfma_consts(TmplDev *__restrict__ tmpl, int count, int precision)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    const int m = tmpl[t].m;
    const double Q2 = tmpl[t].Q2;
    const double md = static_cast<double>(m);
    const double eta = eta_mfma(precision, tmpl[t].kpad);
    // FP16 floor terms: staged |x~| * 2^-ex below 2^-14 and |q| * 2^7 below
    // 2^-14 may be flushed.  x side: sum |q| 2^-14 2^ex <= 2^-14 2^ex
    // sqrt(m Q2) (absolute; the block's scale is 2^ex 2^-7); q side: sum
    // 2^-21 |x~| <= 2^-21 sqrt(m) sqrt(sum x~^2) (relative to the window).
    const double abs_unit = precision == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;  // 2^-21
    L1Consts k;
    k.gmin = static_cast<float>((Q2 + md - static_cast<double>(tmpl[t].bound) - eref_cap(m)) * 0.5 * (1.0 - 1e-3));
    k.eta_sq = static_cast<float>(eta * sqrt(Q2) * (1.0 + 1e-3));
    k.sqrt_m = static_cast<float>(sqrt(md));
    k.inv_mf = 1.0f / static_cast<float>(m);
    k.rsqrt_m = static_cast<float>(1.0 / sqrt(md)) * (1.0f + 1e-6f);
    k.absS = static_cast<float>(fabs(tmpl[t].S) * (1.0 + 1e-3));
    k.eabs_q = static_cast<float>(abs_unit * sqrt(md) * 1.01);
    k.exq = precision == TMX_PREC_FP16 ? static_cast<float>(6.103515625e-05 * 128.0 * sqrt(md * Q2) * 1.011) : 0.0f;
    k.Kc = k.gmin * k.rsqrt_m * 0.999f - (k.eta_sq + k.eabs_q) * 1.001f;
    k.pad = 0.0f;
    tmpl[t].l1 = k;
}

// ---- main kernel ------------------------------------------------------------

// The rigorous interval of one level-1 survivor (main mode) or seed entry.
template <int PREC>
__device__ inline DistInterval survivor_interval(const Waveform &w, uint64_t j, const TmplDev &tp, float c, float dt,
                                                 float scale)
{
    const int m = tp.m;
    const double md = static_cast<double>(m);
    const double abs_unit = PREC == TMX_PREC_FP16 ? 4.76837158203125e-07 : 0.0;
    const double ea = eta_mfma(PREC, tp.kpad) + abs_unit * sqrt(md / tp.Q2) * 1.01;
    DistInterval iv = dist_interval(w, j, m, tp.Q2, tp.S, c, static_cast<double>(dt), ea);
    if (PREC == TMX_PREC_FP16 && iv.lo > -INFINITY) {
        // Staged samples below 2^-14 in scaled units may be flushed: at most
        // 2^-14 2^ex sqrt(m Q2) on the dot product, divided by sigma.
        const WindowSums ws = window_sums_fast(w, j, j + static_cast<uint64_t>(m));
        const double V = ws.S2 - ws.S1 * ws.S1 / md;
        const double sig = sqrt(fmax(V, 1e-300) / md);
        const double ex = 6.103515625e-05 * static_cast<double>(scale) * 128.0 * sqrt(md * tp.Q2) * 1.01;
        const double d = 2.0 * ex / (sig * (1.0 - 1e-3)) * 1.01;
        iv.lo -= d;
        iv.hi += d;
    }
    return iv;
}

// Work item k8 of XCD xcd's share: on each XCD, consecutive items take
// kTileB consecutive blocks (of that XCD's share) for one template group, then
// the next group, so the items in flight on an XCD (one per CU) hold a few
// blocks' samples and prefix sums and a few groups' operands, both
// L2-resident.
__device__ inline void block_of_item(uint32_t xcd, uint32_t k8, int groups, uint32_t *bidx, int *g)
{
    const uint32_t per_tile = kTileB * static_cast<uint32_t>(groups);
    const uint32_t rem = k8 % per_tile;
    *g = static_cast<int>(rem / kTileB);
    *bidx = ((k8 / per_tile) * kTileB + rem % kTileB) * 8u + xcd;
}

// The matrix-core pass over one (block, template group) item; lds: the
// workgroup's staged tile.
template <int PREC>
__device__ __forceinline__ void scan_item(unsigned char *lds, const Waveform &w, const TmplDev *__restrict__ tmpl,
                                          const int32_t *__restrict__ group_begin, const Frag *__restrict__ afrag,
                                          int mode, int block_stride, uint32_t nblocks, const float2 *__restrict__ rb,
                                          const float4 *__restrict__ binfo, int rb_len, const SurvOut &so,
                                          const CandOut &co, uint32_t bidx, int g)
{
    constexpr bool kSplit = PREC != TMX_PREC_FP16;
    uint16_t *const lhi = reinterpret_cast<uint16_t *>(lds);
    uint16_t *const llo = lhi + kStage;
    // Opaque per item: in the persistent form nothing derived from the
    // thread index is hoisted out of the item loop into registers the K loop
    // needs.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int col = lane & 31, h = lane >> 5;
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
    // The block's rebased prefix pairs through a buffer resource: the lane's
    // part of every address is one register, the rest scalar or immediate.
    const __amdgpu_buffer_rsrc_t rbr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2 *>(rbb), static_cast<short>(0), rb_len * 8, 0x00020000);
    auto rbl = [&](uint32_t voff, uint32_t soff) {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rbr, voff, soff, 0));
    };

    // Largest K of the group (templates sorted by length: the last one).
    const int kmax = tmpl[te - 1].kpad;

    // Largest |x~| over the block's samples (resident, over at least the
    // samples staged here): bounds every window's |mu - c| for the level-1
    // test (and sets the FP16 scale).
    const float mx = bi.w;
    const float aD_max = mx * 1.0001f;  // >= |x0 - c| over the block, hence >= |mu_j - c|
    float scale = 1.0f, inv_scale = 1.0f;
    if (PREC == TMX_PREC_FP16) {
        // FP16: power-of-two scale so the staged block fits (max |x~| * 2^-e <= 2^14).
        int ex = 0;
        if (mx > 0.0f) {
            (void)frexpf(mx, &ex);  // mx < 2^ex
            ex -= 14;
        }
        inv_scale = ldexpf(1.0f, -ex);  // staged = x~ * 2^-ex
        scale = ldexpf(1.0f, ex - 7);   // real dot = accumulated * 2^ex * 2^-7
    }

    // Stages x~ = (x0 - c) [* 2^-e] for [J0 + k0, J0 + k0 + kNW + kc), 8
    // samples per item, NB items per thread in flight.
    auto stage = [&](int k0, int kc, auto nb) {
        constexpr int NB = decltype(nb)::value;
        const int items = (kNW + kc) / 8;
        for (int e0 = tid; e0 < items; e0 += NB * kThreads) {
            float4 v[NB][2];
#pragma unroll
            for (int u = 0; u < NB; u++) {
                const int e8 = e0 + u * kThreads;
                if (e8 < items) {
                    const float4 *src = reinterpret_cast<const float4 *>(w.x + J0 + k0 + 8 * e8);
                    v[u][0] = src[0];
                    v[u][1] = src[1];
                }
            }
#pragma unroll
            for (int u = 0; u < NB; u++) {
                const int e8 = e0 + u * kThreads;
                if (e8 >= items)
                    continue;
                const int e = e8 * 8;
                const float xv[8] = {v[u][0].x, v[u][0].y, v[u][0].z, v[u][0].w,
                                     v[u][1].x, v[u][1].y, v[u][1].z, v[u][1].w};
                float xt[8];
#pragma unroll
                for (int q = 0; q < 8; q++)
                    xt[q] = (xv[q] == kMissing ? 0.0f : xv[q]) - c;
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
                if (kSplit)
                    *reinterpret_cast<Frag *>(&llo[swz(e)]) = fl;
            }
        }
    };
    // The group's longest template fits one chunk: stage the block once for
    // every template pair.
    const bool once = kmax <= kKC;
    if (once) {
        stage(0, kmax, std::integral_constant<int, (kStage / 8 + kThreads - 1) / kThreads>{});
        __syncthreads();
    }
    // Byte offset of the wavefront's first column tile in the staged tile.
    const uint32_t wbase = 2048u * static_cast<uint32_t>(wave * kNN);

    // The two wavefronts of a SIMD (wave and wave + 4) take the template
    // pairs in opposite orders when nothing makes them wait for each other
    // (one staged chunk): pairs differ in length, so one wavefront's epilogue
    // falls in the other's matrix-core work instead of both idling the
    // matrix core at once.
    const int npairs = (te - tb + kNM - 1) / kNM;
    const bool rev = once && wave >= kWaves / 2;
    for (int pi = 0; pi < npairs; pi++) {
        const int tp = tb + kNM * (rev ? npairs - 1 - pi : pi);
        const int nt = te - tp < kNM ? te - tp : kNM;
        // Template operands through buffer resources (hi and lo halves of the
        // shifted copies).  Past a template's own K the k-step is clamped to
        // its K, whose fragments read only zero padding, so the shorter
        // template of a pair needs no branch; a template the early-abandoning
        // scan takes, or the missing second template of an odd group, has no
        // records (every read is zero).
        __amdgpu_buffer_rsrc_t rh[kNM], rl[kNM];
        int ks[kNM];
        int kt = 0;
#pragma unroll
        for (int i = 0; i < kNM; i++) {
            const int ti = tp + (i < nt ? i : 0);
            const bool live = i < nt && !tmpl[ti].ea;
            const int L = tmpl[ti].kpad + 64;
            const int bytes = live ? 16 * L : 0;
            uint16_t *base = const_cast<uint16_t *>(reinterpret_cast<const uint16_t *>(afrag + tmpl[ti].a_off));
            rh[i] = __builtin_amdgcn_make_buffer_rsrc(base, static_cast<short>(0), bytes, 0x00020000);
            rl[i] = __builtin_amdgcn_make_buffer_rsrc(base + 8 * L, static_cast<short>(0), bytes, 0x00020000);
            ks[i] = tmpl[ti].kpad / 16;
            if (live)
                kt = tmpl[ti].kpad > kt ? tmpl[ti].kpad : kt;
        }
        // The lane's copy c = (-col) mod 8 and chunk at k-step 0.
        const int cs = (8 - (col & 7)) & 7;
        const uint32_t avoff = static_cast<uint32_t>(((32 + 8 * h - col - cs) / 8) * 8 + cs) * 16u;
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
            if (!once) {
                __syncthreads();
                stage(k0, kc, std::integral_constant<int, 3>{});
                __syncthreads();
            }
            // A fragments of k-step kappa (absolute): 8 consecutive template
            // values at entry 32 + 16 kappa + 8 h - col of each half.
            auto load_a = [&](int kappa, bf16x8 (&oh)[kNM], bf16x8 (&ol)[kNM]) {
#pragma unroll
                for (int i = 0; i < kNM; i++) {
                    const uint32_t so = 256u * static_cast<uint32_t>(kappa < ks[i] ? kappa : ks[i]);
                    oh[i] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rh[i], avoff, so, 0));
                    if (kSplit)
                        ol[i] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rl[i], avoff, so, 0));
                }
            };
            // B fragments of column tile jn at chunk-relative k-step offset kk.
            auto lds_addr = [&](int kk) { return wbase + 2u * static_cast<uint32_t>(swz(32 * col + 8 * h + kk)); };
            auto load_b = [&](uint32_t ad, int jn, bf16x8 &bh, bf16x8 &bl) {
                bh = *reinterpret_cast<const bf16x8 *>(lds + ad + 2048u * jn);
                if (kSplit)
                    bl = *reinterpret_cast<const bf16x8 *>(lds + ad + 2048u * jn + kLoBytes);
            };
            // Two sets of template operands, alternating by k-step, so the
            // next k-step's loads need no register copies.
            bf16x8 a0h[kNM], a0l[kNM], a1h[kNM], a1l[kNM], bh[kNN], bl[kNN];
            {
                const uint32_t ad = lds_addr(0);
#pragma unroll
                for (int jn = 0; jn < kNN; jn++)
                    load_b(ad, jn, bh[jn], bl[jn]);
                load_a(k0 >> 4, a0h, a0l);
            }
            // One k-step on operands (ah, al); with PF the operands of the
            // next are loaded meanwhile (template operands into (nh, nl)
            // first, each column-tile pair's waveform fragments after its
            // last product).
            auto kstep = [&](int kk, bf16x8 (&ah)[kNM], bf16x8 (&al)[kNM], bf16x8 (&nh)[kNM], bf16x8 (&nl)[kNM],
                             auto pf) __attribute__((always_inline)) {
                constexpr bool PF = decltype(pf)::value;
                uint32_t adn = 0;
                if constexpr (PF) {
                    load_a(((k0 + kk) >> 4) + 1, nh, nl);
                    adn = lds_addr(kk + 16);
                }
#pragma unroll
                for (int jp = 0; jp < kNN; jp += 2) {
                    if constexpr (kSplit) {
                        // The three products of each accumulator in the same
                        // order always (lo*hi, hi*lo, hi*hi), interleaved
                        // over four accumulators.
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[jp + q],
                                                                                         acc[i][jp + q], 0, 0, 0);
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[jp + q],
                                                                                         acc[i][jp + q], 0, 0, 0);
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++)
                                acc[i][jp + q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[jp + q],
                                                                                         acc[i][jp + q], 0, 0, 0);
                    } else {
#pragma unroll
                        for (int q = 0; q < 2; q++)
#pragma unroll
                            for (int i = 0; i < kNM; i++) {
                                f16x8 av, bv;
                                __builtin_memcpy(&av, &ah[i], 16);
                                __builtin_memcpy(&bv, &bh[jp + q], 16);
                                acc[i][jp + q] =
                                    __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc[i][jp + q], 0, 0, 0);
                            }
                    }
                    if constexpr (PF) {
#pragma unroll
                        for (int q = 0; q < 2; q++)
                            load_b(adn, jp + q, bh[jp + q], bl[jp + q]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            // kc is a multiple of 32 (padded K and chunk length): k-steps in pairs.
            using Pf = std::integral_constant<bool, true>;
            using NoPf = std::integral_constant<bool, false>;
            int kk = 0;
            for (; kk + 32 < kc; kk += 32) {
                kstep(kk, a0h, a0l, a1h, a1l, Pf{});
                kstep(kk + 16, a1h, a1l, a0h, a0l, Pf{});
            }
            kstep(kk, a0h, a0l, a1h, a1l, Pf{});
            kstep(kk + 16, a1h, a1l, a0h, a0l, NoPf{});
        }

        // ---- epilogue: registers and global memory only ----------------------
        // Accumulator layout: lane (col, h), entry r holds window
        // tile0 + jl + 8 (r / 4) + r % 4, jl = 32 col + 4 h.  (jl is rebuilt
        // opaquely per pair so that its 16 offsets are not hoisted out of the
        // template loop into registers the K loop needs.)
        int jl = 32 * col + 4 * h;
        asm volatile("" : "+v"(jl));
        const uint32_t vl = static_cast<uint32_t>(jl) * 8u;  // lane's byte offset in the prefix pairs
        float best_g[kNM], best_dt[kNM];
        uint32_t best_j[kNM];
        // One template's four tiles.
        auto epilogue = [&](int i, const f32x16 &a0, const f32x16 &a1, const f32x16 &a2,
                            const f32x16 &a3) __attribute__((always_inline)) {
            best_g[i] = -INFINITY;
            best_dt[i] = 0.0f;
            best_j[i] = 0xffffffffu;
            if (i >= nt)
                return;
            const int t = tp + i;
            const int m = tmpl[t].m;
            if (static_cast<uint64_t>(m) > w.n || tmpl[t].ea)
                return;
            const uint64_t W = w.n - static_cast<uint64_t>(m) + 1;
            const L1Consts k1 = tmpl[t].l1;
            const float gmin = k1.gmin, eta_sq = k1.eta_sq, sqrt_m = k1.sqrt_m, inv_mf = k1.inv_mf;
            const float rsqrt_m = k1.rsqrt_m, absS = k1.absS, eabs_q = k1.eabs_q, Kc = k1.Kc;
            const float mf = static_cast<float>(m);
            const float eabs_x = k1.exq * scale;  // FP16 only (absolute)
            // Per window the level-1 test is dt >= Kc sqrt(V) - rest(aD),
            // V the window's m sigma^2 and aD >= |mu - c|; rest uses the
            // block's largest |x~| for aD.
            const float rest = ((eta_sq + eabs_q) * sqrt_m * aD_max + aD_max * absS + eabs_x) * 1.001f + 1e-30f;
            const float amu_max = fabsf(c) + aD_max;
            // Window sums of x0 - c and (x0 - c)^2 over a window from its
            // block-rebased prefix pairs, with their error bounds: the
            // window's m sigma^2, a bound on |mu - c|, and whether the FP32
            // statistics can judge it (else rounding swamps them, or
            // |mu| / sigma > 1e3).
            struct Win {
                float V, aD;
                bool judge;
            };
            auto window = [&](float2 ra, float2 rz) __attribute__((always_inline)) {
                const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
                const float eS1 = 2.0f * bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
                const float eS2 = 2.0f * bi.z + 1.2e-7f * (fabsf(ra.y) + fabsf(rz.y) + fabsf(S2));
                Win r;
                r.V = S2 - S1 * S1 * inv_mf;
                const float eV = eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * inv_mf +
                                 2.4e-7f * (fabsf(S2) + S1 * S1 * inv_mf);
                r.aD = fabsf(S1 * inv_mf) + eS1 * inv_mf;
                const float amu = fabsf(c) + r.aD;
                r.judge = r.V > 1e3f * eV && r.V > mf * 1e-20f && !(amu * amu * mf > 1e6f * r.V);
                return r;
            };
            // Windows of tile jn that exist, [0, Wt) relative to its first, and
            // the column's core prefix pairs (loaded for all four tiles at once).
            auto tile_wt = [&](int jn) __attribute__((always_inline)) {
                const uint64_t tile0 = J0 + static_cast<uint64_t>(1024 * (wave * kNN + jn));
                return W > tile0 ? (W - tile0 < 1024 ? static_cast<int>(W - tile0) : 1024) : 0;
            };
            const int c0 = jl - 4 * h;  // the lane's column's first window, relative to its tile
            float2 cra[kNN], crz[kNN];
            if (mode == kModeMain) {
#pragma unroll
                for (int jn = 0; jn < kNN; jn++) {
                    const int Wt = tile_wt(jn);
                    const int cnt = Wt - c0 < 32 ? (Wt - c0 > 1 ? Wt - c0 : 1) : 32;
                    const uint32_t vc = static_cast<uint32_t>(1024 * (wave * kNN + jn) + c0) * 8u;
                    cra[jn] = rbl(vc + 8u * static_cast<uint32_t>(cnt - 1), 0u);
                    crz[jn] = rbl(vc, 8u * static_cast<uint32_t>(m));
                }
            }
            auto tile = [&](const f32x16 &av, int jn) __attribute__((always_inline)) {
                const int toff = 1024 * (wave * kNN + jn);  // tile's first window, relative to J0
                const uint64_t tile0 = J0 + static_cast<uint64_t>(toff);
                const int Wt = tile_wt(jn);
                if (mode == kModeMain) {
                    // Column threshold.  Column col holds windows j0 + s, s <
                    // 32 (this lane and lane ^ 32 hold 16 each); their common
                    // core [j0 + cnt - 1, j0 + m) bounds every window's
                    // m sigma^2 from below and |x~| bounds |mu - c|, so tc is
                    // at most the level-1 threshold of each of the column's
                    // windows: a window whose dot product is below tc cannot
                    // survive.
                    float tc = INFINITY;
                    if (c0 < Wt) {
                        tc = -INFINITY;
                        const int cnt = Wt - c0 < 32 ? Wt - c0 : 32;
                        const int Lc = m - (cnt - 1);
                        if (Lc >= 2 && gmin > 0.0f && Kc > 0.0f) {
                            const float2 ra = cra[jn], rz = crz[jn];
                            const float S1 = rz.x - ra.x, S2 = rz.y - ra.y;
                            const float eS1 = 2.0f * bi.y + 1.2e-7f * (fabsf(ra.x) + fabsf(rz.x) + fabsf(S1));
                            const float eS2 = 2.0f * bi.z + 1.2e-7f * (fabsf(ra.y) + fabsf(rz.y) + fabsf(S2));
                            const float invL = 1.0f / static_cast<float>(Lc);
                            const float V = S2 - S1 * S1 * invL;
                            const float eV = eS2 + (2.0f * fabsf(S1) * eS1 + eS1 * eS1) * invL +
                                             2.4e-7f * (fabsf(S2) + S1 * S1 * invL);
                            const float Vlo = V - eV;
                            if (Vlo > 1e3f * eV && Vlo > mf * 1e-20f && !(amu_max * amu_max * mf > 1e6f * Vlo))
                                tc = Kc * (Vlo * __frsqrt_rn(Vlo)) * 0.999f - rest;
                        }
                    }
                    uint32_t pre = 0u;  // entries that pass the column threshold
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int wl = jl + 8 * (r >> 2) + (r & 3);
                        if (wl < Wt && av[r] * scale >= tc)
                            pre |= 1u << r;
                    }
                    if (__ballot(pre != 0u) == 0ull)
                        return;
                    // The window's own threshold, only for entries some lane
                    // still holds (-inf: cannot be judged, so it survives).
                    uint32_t sm = 0u;
#pragma unroll 1
                    for (int r = 0; r < 16; r++) {
                        if (__ballot((pre >> r) & 1u) == 0ull)
                            continue;
                        const float dt = av[r] * scale;
                        if ((pre >> r) & 1u) {
                            const uint32_t so0 = static_cast<uint32_t>(toff + 8 * (r >> 2) + (r & 3)) * 8u;
                            const Win wv = window(rbl(vl, so0), rbl(vl, so0 + 8u * static_cast<uint32_t>(m)));
                            float thr = -INFINITY;
                            if (wv.judge) {
                                const float sV = wv.V * __frsqrt_rn(wv.V);
                                const float b = eta_sq * (sV + sqrt_m * wv.aD) + wv.aD * absS +
                                                eabs_q * (sV + sqrt_m * wv.aD) + eabs_x;
                                thr = gmin * sV * rsqrt_m * 0.999f - b * 1.001f - 1e-30f;
                            }
                            if (dt >= thr)
                                sm |= 1u << r;
                        }
                    }
                    if (__ballot(sm != 0u) == 0ull)
                        return;
                    // Survivors to the queue, one atomic per wavefront and
                    // tile: the lane's exclusive prefix of the survivor
                    // counts (at most 16, five bits) from one ballot per bit.
                    const int cnt = __popc(sm);
                    uint32_t excl = 0u, total = 0u;
#pragma unroll
                    for (int b = 0; b < 5; b++) {
                        const unsigned long long bm = __ballot((cnt >> b) & 1);
                        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(bm >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bm), 0u));
                        excl += below << b;
                        total += static_cast<uint32_t>(__popcll(bm)) << b;
                    }
                    unsigned long long base = 0ull;
                    if (lane == 0)
                        base = atomicAdd(so.count, static_cast<unsigned long long>(total));
                    base = __shfl(base, 0);
                    unsigned long long slot = base + excl;
                    const uint32_t jb = static_cast<uint32_t>(tile0) + static_cast<uint32_t>(jl);
#pragma unroll 1
                    for (int r = 0; r < 16; r++) {
                        const float dt = av[r] * scale;
                        if ((sm >> r) & 1u) {
                            if (slot < so.cap)
                                so.rec[slot] = SurvRec{static_cast<uint32_t>(t),
                                                       jb + static_cast<uint32_t>(8 * (r >> 2) + (r & 3)), dt, scale};
                            slot++;
                        }
                    }
                } else {
                    // Seed mode: the lane's most promising window by
                    // dt / sqrt(m sigma^2), over the windows it can judge.
#pragma unroll 1
                    for (int r = 0; r < 16; r++) {
                        const int wl = jl + 8 * (r >> 2) + (r & 3);
                        const uint32_t so0 = static_cast<uint32_t>(toff + 8 * (r >> 2) + (r & 3)) * 8u;
                        const Win wv = window(rbl(vl, so0), rbl(vl, so0 + 8u * static_cast<uint32_t>(m)));
                        const float rs = __frsqrt_rn(wv.V);
                        const float dt = av[r] * scale;
                        const float gapprox = dt * rs;
                        if (wl < Wt && wv.judge && rs > 0.0f && rs < INFINITY && gapprox > best_g[i]) {
                            best_g[i] = gapprox;
                            best_j[i] = static_cast<uint32_t>(tile0) + static_cast<uint32_t>(wl);
                            best_dt[i] = dt;
                        }
                    }
                }
            };
            tile(a0, 0);
            tile(a1, 1);
            tile(a2, 2);
            tile(a3, 3);
        };
        static_assert(kNM == 2 && kNN == 4, "the epilogue calls name every accumulator");
        epilogue(0, acc[0][0], acc[0][1], acc[0][2], acc[0][3]);
        epilogue(1, acc[1][0], acc[1][1], acc[1][2], acc[1][3]);

        // Seed mode: each lane's most promising window of each template,
        // evaluated rigorously (accumulators dead); the wavefront emits the
        // smallest upper bound.
        if (mode == kModeSeed) {
#pragma unroll
            for (int i = 0; i < kNM; i++) {
                if (i >= nt)
                    continue;
                const int t = tp + i;
                float ub = INFINITY;
                const uint32_t bj0 = best_j[i];
                if (bj0 != 0xffffffffu && w.x[bj0] != kMissing) {
                    const DistInterval iv = survivor_interval<PREC>(w, bj0, tmpl[t], c, best_dt[i], scale);
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

__global__ void k_zero_queue(uint32_t *queue) { queue[threadIdx.x] = 0u; }

// One workgroup per (block, group) item: workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup L runs item L / 8 of XCD L mod 8.
template <int PREC>
__global__ __launch_bounds__(kThreads, 1) void k_scan_mfma(Waveform w, const TmplDev *__restrict__ tmpl,
                                                           const int32_t *__restrict__ group_begin,
                                                           const Frag *__restrict__ afrag, int mode,
                                                           int block_stride, int groups, uint32_t nblocks,
                                                           const float2 *__restrict__ rb,
                                                           const float4 *__restrict__ binfo, int rb_len,
                                                           SurvOut so, CandOut co)
{
    constexpr int kLds = PREC != TMX_PREC_FP16 ? kTileBytesBf16 : kTileBytesF16;
    __shared__ __attribute__((aligned(16))) unsigned char lds[kLds];
    uint32_t bidx;
    int g;
    block_of_item(blockIdx.x & 7u, blockIdx.x >> 3, groups, &bidx, &g);
    scan_item<PREC>(lds, w, tmpl, group_begin, afrag, mode, block_stride, nblocks, rb, binfo, rb_len, so, co, bidx,
                    g);
}

// Persistent form: one workgroup per CU, each claiming items from its XCD's
// queue (queue[x] counts the items of XCD x claimed so far; items beyond
// `items` mean that queue is empty) and, once that is empty, from the other
// XCDs' queues in turn, until every queue is empty.  Every workgroup ends
// when all eight counters have passed `items`.
template <int PREC>
__global__ __launch_bounds__(kThreads, 1) void k_scan_mfma_queue(Waveform w, const TmplDev *__restrict__ tmpl,
                                                                 const int32_t *__restrict__ group_begin,
                                                                 const Frag *__restrict__ afrag, int mode,
                                                                 int block_stride, int groups, uint32_t nblocks,
                                                                 const float2 *__restrict__ rb,
                                                                 const float4 *__restrict__ binfo, int rb_len,
                                                                 SurvOut so, CandOut co, uint32_t *queue,
                                                                 uint32_t items)
{
    constexpr int kLds = PREC != TMX_PREC_FP16 ? kTileBytesBf16 : kTileBytesF16;
    __shared__ __attribute__((aligned(16))) unsigned char lds[kLds];
    __shared__ uint32_t s_item[2];  // alternate slots: a slot is rewritten two barriers after it was read
    const uint32_t xcd = blockIdx.x & 7u;
    for (int it = 0;; it++) {
        if (threadIdx.x == 0) {
            uint32_t got = 0xffffffffu;
            for (uint32_t s = 0; s < 8u; s++) {
                const uint32_t x = (xcd + s) & 7u;
                if (__hip_atomic_load(&queue[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= items)
                    continue;
                const uint32_t k8 = atomicAdd(&queue[x], 1u);
                if (k8 < items) {
                    got = k8 << 3 | x;
                    break;
                }
            }
            s_item[it & 1] = got;
        }
        // Also keeps the next item's staging from overwriting the tile while
        // a wavefront still reads the previous item's.
        __syncthreads();
        const uint32_t got = s_item[it & 1];
        if (got == 0xffffffffu)
            return;
        uint32_t bidx;
        int g;
        block_of_item(got & 7u, got >> 3, groups, &bidx, &g);
        scan_item<PREC>(lds, w, tmpl, group_begin, afrag, mode, block_stride, nblocks, rb, binfo, rb_len, so, co,
                        bidx, g);
    }
}

// Level 2: the rigorous FP64 interval of every level-1 survivor; windows
// whose lower bound reaches the template's bound become candidates.
template <int PREC>
__global__ __launch_bounds__(256) void k_level2(Waveform w, const TmplDev *__restrict__ tmpl,
                                                const float4 *__restrict__ binfo, SurvOut so, CandOut co)
{
    const unsigned long long cnt = *so.count;
    const uint64_t n = cnt < so.cap ? cnt : so.cap;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) {
        const SurvRec s = so.rec[e];
        const uint64_t j = s.j;
        if (w.x[j] == kMissing)
            continue;
        const TmplDev &tp = tmpl[s.t];
        // The block that produced it: its centre is the one the dot product used.
        const float c = binfo[j / static_cast<uint64_t>(kNW)].x;
        const DistInterval iv = survivor_interval<PREC>(w, j, tp, c, s.dt, s.scale);
        if (iv.lo <= static_cast<double>(tp.bound))
            cand_put(co, wave_claim(co.count), s.t, s.j, static_cast<float>(iv.est), iv.lo, iv.hi);
    }
}

}  // namespace

int mfma_group_size() { return kGroup; }
uint64_t mfma_block_windows() { return kNW; }
int mfma_block_waves() { return kWaves; }
// K of a template: its m + 31 Toeplitz columns, padded to whole pairs of
// k-steps (the K loop runs k-steps in pairs).
int mfma_kpad(int m) { return ((m + 31 + 31) / 32) * 32; }
static_assert(kKC % 32 == 0, "chunks hold whole pairs of k-steps");
// Two halves (hi, lo) of 8 shifted copies of kpad + 64 16-bit entries, in
// 16-byte units.
int64_t mfma_frag_count(int kpad) { return 2 * static_cast<int64_t>(kpad + 64); }

hipError_t launch_mfma_pack(const TmplDev *tmpl, int count, int max_kpad, const float *qz, int precision,
                            void *a_frag, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    // Entries per half of the longest template: 8 copies of kpad + 64.
    const int per_t = 8 * (max_kpad + 64);
    const dim3 grid((per_t + 255) / 256, count);
    hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, stream, tmpl, count, qz, precision, static_cast<Frag *>(a_frag));
    return hipGetLastError();
}

hipError_t launch_scan_mfma(const Waveform &w, const TmplDev *tmpl, int count, const int32_t *group_begin, int groups,
                            int precision, const void *a_frag, const float2 *rb, const float4 *binfo, int rb_len,
                            int mode, int block_stride, const SurvOut &so, const CandOut &co, int persistent,
                            uint32_t *queue, hipStream_t stream)
{
    (void)count;
    if (groups <= 0 || w.n == 0)
        return hipSuccess;
    const uint64_t blocks = (w.n + kNW - 1) / kNW;
    const uint64_t launched = (blocks + block_stride - 1) / block_stride;
    const uint64_t per_xcd = ((launched + 7) / 8 + kTileB - 1) / kTileB * kTileB;
    const uint64_t total = per_xcd * 8 * static_cast<uint64_t>(groups);
    if (total >= (1ull << 32))
        return hipErrorInvalidConfiguration;
    const Frag *a = static_cast<const Frag *>(a_frag);
    if (persistent > 0) {
        // One workgroup per CU (a multiple of 8, so every XCD has its share);
        // the queues start empty of claims.
        if (!queue)
            return hipErrorInvalidValue;
        const unsigned wgs = static_cast<unsigned>((persistent + 7) / 8 * 8);
        const uint32_t items = static_cast<uint32_t>(per_xcd * static_cast<uint64_t>(groups));
        hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(8), 0, stream, queue);
        if (precision == TMX_PREC_FP16)
            hipLaunchKernelGGL(k_scan_mfma_queue<TMX_PREC_FP16>, dim3(wgs), dim3(kThreads), 0, stream, w, tmpl,
                               group_begin, a, mode, block_stride, groups, static_cast<uint32_t>(launched), rb, binfo,
                               rb_len, so, co, queue, items);
        else
            hipLaunchKernelGGL(k_scan_mfma_queue<TMX_PREC_BF16X3>, dim3(wgs), dim3(kThreads), 0, stream, w, tmpl,
                               group_begin, a, mode, block_stride, groups, static_cast<uint32_t>(launched), rb, binfo,
                               rb_len, so, co, queue, items);
        return hipGetLastError();
    }
    const dim3 grid(static_cast<unsigned>(total));
    if (precision == TMX_PREC_FP16)
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_FP16>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a, mode,
                           block_stride, groups, static_cast<uint32_t>(launched), rb, binfo, rb_len, so, co);
    else
        hipLaunchKernelGGL(k_scan_mfma<TMX_PREC_BF16X3>, grid, dim3(kThreads), 0, stream, w, tmpl, group_begin, a,
                           mode, block_stride, groups, static_cast<uint32_t>(launched), rb, binfo, rb_len, so, co);
    return hipGetLastError();
}

hipError_t launch_mfma_consts(TmplDev *tmpl, int count, int precision, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_mfma_consts, dim3((count + 63) / 64), dim3(64), 0, stream, tmpl, count, precision);
    return hipGetLastError();
}

hipError_t launch_mfma_level2(const Waveform &w, const TmplDev *tmpl, int precision, const float4 *binfo,
                              const SurvOut &so, const CandOut &co, int cus, hipStream_t stream)
{
    const dim3 grid(static_cast<unsigned>((cus > 0 ? cus : 256) * 8));
    if (precision == TMX_PREC_FP16)
        hipLaunchKernelGGL(k_level2<TMX_PREC_FP16>, grid, dim3(256), 0, stream, w, tmpl, binfo, so, co);
    else
        hipLaunchKernelGGL(k_level2<TMX_PREC_BF16X3>, grid, dim3(256), 0, stream, w, tmpl, binfo, so, co);
    return hipGetLastError();
}

}  // namespace tmx
