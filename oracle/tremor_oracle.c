/*
 * CPU oracle for the exact template-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the tremor_amd package, its
 * HIP kernels or its C-ABI library) links, loads or calls this file.  It is the
 * checker that tests/, the smoke entry point and the CPU-baseline leg of the
 * benchmark compare the GPU path against.
 *
 * It restates, in plain C, the semantics of the reference's sequential scan:
 *   - template z-normalisation            ZnormSearch.hpp:33-62
 *   - block visiting order                ZnormSearch.hpp:64-90
 *   - squared z-normalised distance       ZnormSearch.hpp:92-165 (AVX2 lanes emulated)
 *   - cc <-> ed conversions               ZnormSearch.hpp:23-31
 *   - canonical kNN with exclusion        ZnormSearch.hpp:167-189
 *   - bound set (pqueue_bsf)              iSAXPqueue.cpp:328-351, 471-549
 *   - two-pass FP64 window statistics     SkipSequential.cpp:154-190
 *   - threshold scan + merge              SkipSequential.cpp:223-294
 *   - kNN scan (bound set + candidates)   SkipSequential.cpp:296-393
 *   - iSAX words and lower bound          SkipSequential.cpp:155-211, SAX.cpp:24-85,112-180,362-490
 *
 * Floating-point contraction.  The reference is built with GCC at -O2 with FMA
 * available (Makefile:5-6); GCC contracts every a*b+c on this path into one
 * fused multiply-add (C++ defaults to -ffp-contract=fast).  This file is
 * compiled with -ffp-contract=off and writes each of those fusions explicitly
 * with fma()/fmaf(), so its results do not depend on the compiler that builds
 * it.  The fusions are pinned against the reference header compiled with the
 * reference's flags (oracle/_ref, tests/test_oracle_ref.py).
 */
#define _GNU_SOURCE /* M_SQRT1_2 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "tremor_oracle.h"

/* ZnormSearch.hpp:21 */
#define ORC_MISSING (-100000.0f)
/* ZnormSearch.hpp:65 */
#define ORC_BLOCK 32

/* ZnormSearch.hpp:23-26 */
float orc_cc_to_ed(float cc, int m) { return 2.0f * (float)m * (1.0f - cc); }

/* ZnormSearch.hpp:28-31 */
float orc_ed_to_cc(float ed, int m) { return 1.0f - (ed / (2.0f * (float)m)); }

/* ZnormSearch.hpp:33-62.  Single pass FP64 sum and sum of squares (the square
 * fused into the accumulation), population variance with mean*mean fused into
 * the subtraction, clamp at 0, sigma <= 1e-12 gives all zeros. */
void orc_znormalize(float *v, int len)
{
    if (v == NULL || len <= 0)
        return;
    double sum = 0.0, sum_sq = 0.0;
    for (int i = 0; i < len; i++) {
        const double x = (double)v[i];
        sum += x;
        sum_sq = fma(x, x, sum_sq);
    }
    const double mean = sum / (double)len;
    double variance = fma(-mean, mean, sum_sq / (double)len);
    if (variance < 0.0)
        variance = 0.0;
    const double sd = sqrt(variance);
    if (sd <= 1e-12) {
        for (int i = 0; i < len; i++)
            v[i] = 0.0f;
        return;
    }
    for (int i = 0; i < len; i++)
        v[i] = (float)(((double)v[i] - mean) / sd);
}

/* ZnormSearch.hpp:72-90.  Block energies accumulate with fused q*q; a stable
 * sort by descending energy (ties keep ascending block index).  Insertion sort
 * is stable and blocks are few (m/32). */
int orc_block_order(const float *q, int m, int *order)
{
    const int blocks = m / ORC_BLOCK;
    float *energy = (float *)malloc(sizeof(float) * (size_t)(blocks > 0 ? blocks : 1));
    for (int b = 0; b < blocks; b++) {
        float e = 0.0f;
        for (int i = b * ORC_BLOCK; i < (b + 1) * ORC_BLOCK; i++)
            e = fmaf(q[i], q[i], e);
        energy[b] = e;
        order[b] = b * ORC_BLOCK;
    }
    for (int i = 1; i < blocks; i++) {
        const float e = energy[i];
        const int o = order[i];
        int j = i - 1;
        while (j >= 0 && energy[j] < e) { /* move strictly smaller energies right */
            energy[j + 1] = energy[j];
            order[j + 1] = order[j];
            j--;
        }
        energy[j + 1] = e;
        order[j + 1] = o;
    }
    free(energy);
    return blocks;
}

/* SkipSequential.cpp:157-179.  Window `id` of raw[0..raw_len): skipped (left at
 * 0, 0) when its first sample is MISSING; MISSING samples count as 0; two-pass
 * FP64 with the square fused into the second pass; sigma <= 1e-12 stored as 0. */
void orc_window_stats_one(const float *w, int m, float *mean_out, float *std_out)
{
    if (w[0] == ORC_MISSING) {
        *mean_out = 0.0f;
        *std_out = 0.0f;
        return;
    }
    double sum = 0.0;
    for (int d = 0; d < m; d++)
        sum += w[d] == ORC_MISSING ? 0.0 : (double)w[d];
    const double mean = sum / m;
    double squares = 0.0;
    for (int d = 0; d < m; d++) {
        const double v = (w[d] == ORC_MISSING ? 0.0 : (double)w[d]) - mean;
        squares = fma(v, v, squares);
    }
    const double sd = sqrt(squares / m);
    *mean_out = (float)mean;
    *std_out = (float)(sd <= 1e-12 ? 0.0 : sd);
}

void orc_window_stats(const float *raw, int64_t n, int m, int threads, float *mean, float *std)
{
    const int64_t windows = n - m + 1;
    if (windows <= 0)
        return;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t id = 0; id < windows; id++)
        orc_window_stats_one(raw + id, m, &mean[id], &std[id]);
}

/* hsum256 of ZnormSearch.hpp:92-97 applied to acc0 + acc1:
 * ((v0+v4)+(v2+v6)) + ((v1+v5)+(v3+v7)). */
static float hsum8(const float *a0, const float *a1)
{
    float v[8];
    for (int l = 0; l < 8; l++)
        v[l] = a0[l] + a1[l];
    const float s0 = v[0] + v[4], s1 = v[1] + v[5], s2 = v[2] + v[6], s3 = v[3] + v[7];
    return (s0 + s2) + (s1 + s3);
}

/* ZnormSearch.hpp:103-165.  Two 8-lane accumulators: acc0 lane l takes sample
 * off+16s+l, acc1 lane l takes off+16s+8+l, each delta = q - (r - mean)*inv
 * with the product fused into the subtraction, accumulated by fma(delta,
 * delta, acc).  After each block the lanes are summed and compared with the
 * bound; the tail (m mod 32 samples) is added in index order, fused. */
float orc_dist(const float *q, const float *raw, int64_t start, int m, float mean, float sd, float bound,
               const int *order, int *blocks_visited)
{
    int dummy = 0;
    if (blocks_visited == NULL)
        blocks_visited = &dummy;
    if (q == NULL || raw == NULL || m <= 0)
        return FLT_MAX;
    if (sd <= 1e-12f)
        return FLT_MAX;
    const float *r = raw + start;
    const float inv = 1.0f / sd;
    float acc0[8] = {0}, acc1[8] = {0};
    const int blocks = m / ORC_BLOCK;
    for (int b = 0; b < blocks; b++) {
        const int off = order[b];
        for (int i = off; i < off + ORC_BLOCK; i += 16) {
            for (int l = 0; l < 8; l++) {
                const float r0 = r[i + l] == ORC_MISSING ? 0.0f : r[i + l];
                const float r1 = r[i + 8 + l] == ORC_MISSING ? 0.0f : r[i + 8 + l];
                const float d0 = fmaf(-(r0 - mean), inv, q[i + l]);
                const float d1 = fmaf(-(r1 - mean), inv, q[i + 8 + l]);
                acc0[l] = fmaf(d0, d0, acc0[l]);
                acc1[l] = fmaf(d1, d1, acc1[l]);
            }
        }
        const float partial = hsum8(acc0, acc1);
        if (partial > bound) {
            *blocks_visited = b + 1;
            return partial;
        }
    }
    *blocks_visited = blocks;
    float dist = hsum8(acc0, acc1);
    for (int i = blocks * ORC_BLOCK; i < m; i++) {
        const float rv = r[i] == ORC_MISSING ? 0.0f : r[i];
        const float diff = fmaf(-(rv - mean), inv, q[i]);
        dist = fmaf(diff, diff, dist);
    }
    return dist;
}

/* Full distances (bound = FLT_MAX) of one z-normalised template against every
 * window; windows whose first sample is MISSING get FLT_MAX as well (they are
 * never searched: SkipSequential.cpp:256,321). */
void orc_dist_all(const float *qz, int m, const float *raw, int64_t n, const float *mean, const float *sd,
                  int threads, float *out)
{
    const int64_t windows = n - m + 1;
    if (windows <= 0)
        return;
    int *order = (int *)malloc(sizeof(int) * (size_t)(m / ORC_BLOCK + 1));
    orc_block_order(qz, m, order);
#pragma omp parallel for num_threads(threads) schedule(static, 4096)
    for (int64_t id = 0; id < windows; id++) {
        if (raw[id] == ORC_MISSING) {
            out[id] = FLT_MAX;
            continue;
        }
        out[id] = orc_dist(qz, raw, id, m, mean[id], sd[id], FLT_MAX, order, NULL);
    }
    free(order);
}

/* ---------- selection ---------------------------------------------------- */

typedef struct {
    uint64_t pos;
    float d;
} orc_cand;

static int cmp_d_pos(const void *a, const void *b)
{
    const orc_cand *x = (const orc_cand *)a, *y = (const orc_cand *)b;
    if (x->d != y->d)
        return x->d < y->d ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos ? 1 : 0);
}

static int cmp_pos_d(const void *a, const void *b)
{
    const orc_cand *x = (const orc_cand *)a, *y = (const orc_cand *)b;
    if (x->pos != y->pos)
        return x->pos < y->pos ? -1 : 1;
    return x->d < y->d ? -1 : (x->d > y->d ? 1 : 0);
}

/* ZnormSearch.hpp:170-189.  Sort by (d, pos); keep a candidate unless a kept
 * position lies strictly closer than merge_offset on either side; stop at k.
 * merge_offset = 0 keeps everything (distinct positions).  Kept positions are
 * few (<= k), so a linear check replaces the ordered set. */
int orc_select_knn(const uint64_t *pos, const float *d, int64_t count, int k, int merge_offset, uint64_t *out_pos,
                   float *out_d)
{
    orc_cand *c = (orc_cand *)malloc(sizeof(orc_cand) * (size_t)(count > 0 ? count : 1));
    for (int64_t i = 0; i < count; i++) {
        c[i].pos = pos[i];
        c[i].d = d[i];
    }
    qsort(c, (size_t)count, sizeof(orc_cand), cmp_d_pos);
    int kept = 0;
    for (int64_t i = 0; i < count && kept < k; i++) {
        int blocked = 0;
        for (int j = 0; j < kept; j++) {
            const uint64_t p = out_pos[j];
            const uint64_t gap = p > c[i].pos ? p - c[i].pos : c[i].pos - p;
            if (gap < (uint64_t)merge_offset) {
                blocked = 1;
                break;
            }
        }
        if (blocked)
            continue;
        out_pos[kept] = c[i].pos;
        out_d[kept] = c[i].d;
        kept++;
    }
    free(c);
    return kept;
}

/* iSAXPqueue.cpp:328-351 */
void orc_bsf_init(orc_bsf *q, int k)
{
    q->k = k;
    q->pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
    q->knn = (float *)malloc(sizeof(float) * (size_t)k);
    for (int i = 0; i < k; i++) {
        q->knn[i] = FLT_MAX;
        q->pos[i] = -1;
    }
}

void orc_bsf_free(orc_bsf *q)
{
    free(q->pos);
    free(q->knn);
}

/* iSAXPqueue.cpp:522-549: first slot with data <= knn[i] and a different
 * position takes the entry; the tail shifts right. */
void orc_bsf_insert(orc_bsf *q, float data, int64_t position)
{
    for (int i = 0; i < q->k; i++) {
        if (data <= q->knn[i] && q->pos[i] != position) {
            for (int j = q->k - 1; j > i; j--) {
                q->knn[j] = q->knn[j - 1];
                q->pos[j] = q->pos[j - 1];
            }
            q->knn[i] = data;
            q->pos[i] = position;
            return;
        }
    }
}

/* iSAXPqueue.cpp:471-520: reject when an entry at least as good lies closer
 * than merge_offset; otherwise drop every entry closer than merge_offset,
 * refill the tail with (FLT_MAX, -1) and insert in distance order. */
void orc_bsf_insert_offset(orc_bsf *q, float data, int64_t position, int merge_offset)
{
    if (merge_offset == 0) {
        orc_bsf_insert(q, data, position);
        return;
    }
    for (int i = 0; i < q->k; i++) {
        const int64_t gap = q->pos[i] > position ? q->pos[i] - position : position - q->pos[i];
        if (q->knn[i] < FLT_MAX && gap < merge_offset && q->knn[i] <= data)
            return;
    }
    int kept = 0;
    for (int i = 0; i < q->k; i++) {
        const int64_t gap = q->pos[i] > position ? q->pos[i] - position : position - q->pos[i];
        if (q->knn[i] < FLT_MAX && gap < merge_offset)
            continue;
        q->knn[kept] = q->knn[i];
        q->pos[kept] = q->pos[i];
        kept++;
    }
    for (int i = kept; i < q->k; i++) {
        q->knn[i] = FLT_MAX;
        q->pos[i] = -1;
    }
    for (int i = 0; i < q->k; i++) {
        if (data <= q->knn[i]) {
            for (int j = q->k - 1; j > i; j--) {
                q->knn[j] = q->knn[j - 1];
                q->pos[j] = q->pos[j - 1];
            }
            q->knn[i] = data;
            q->pos[i] = position;
            return;
        }
    }
}

/* ---------- scans -------------------------------------------------------- */

/* iSAX breakpoints of cardinality 256, the sub-table of SAXBreakpoints.hpp at
 * offset (c - 1)(c - 2) / 2: the standard normal quantiles Phi^-1(i / 256),
 * i = 1..255, rounded to float.  The lower half is found by bisection on
 * Phi(x) = erfc(-x / sqrt 2) / 2 in FP64 down to adjacent doubles; the median
 * is 0 and the upper half mirrors the lower (Phi^-1(1 - p) = -Phi^-1(p)). */
void orc_sax_breakpoints(float *bp)
{
    bp[127] = 0.0f;
    for (int i = 1; i < 128; i++) {
        const double p = (double)i / 256.0;
        double lo = -10.0, hi = 10.0;
        for (;;) {
            const double mid = 0.5 * (lo + hi);
            if (!(mid > lo && mid < hi))
                break;
            if (0.5 * erfc(-mid * M_SQRT1_2) < p)
                lo = mid;
            else
                hi = mid;
        }
        bp[i - 1] = (float)(0.5 * (lo + hi));
        bp[255 - i] = -bp[i - 1];
    }
}

/* SAX.cpp:73-85 (paa_from_ts): FP32 sums in index order, then one division. */
void orc_paa(const float *z, int m, int segments, float *paa)
{
    const int len = m / segments;
    for (int s = 0; s < segments; s++) {
        float acc = 0.0f;
        for (int i = 0; i < len; i++)
            acc += z[s * len + i];
        paa[s] = acc / (float)len;
    }
}

/* SAX.cpp:46-66 (the bsearch with its compare of SAX.cpp:7-22): the symbol of
 * a PAA value is the number of breakpoints strictly below it (255 above the
 * last; 0 for NaN). */
static uint8_t sax_symbol(const float *bp, float v)
{
    int c = 0;
    for (int i = 0; i < 255; i++)
        c += bp[i] < v;
    return (uint8_t)c;
}

/* SkipSequential.cpp:155-190 with lower_bounds: for window id (first sample
 * not MISSING), the two-pass FP64 mean and deviation (sigma <= 1e-12 -> 1),
 * the samples (v - mean) / sigma rounded to float, their PAA and its symbols
 * (sax_from_ts, SAX.cpp:24-70).  Windows starting with MISSING keep zeros. */
void orc_sax_words(const float *raw, int64_t n, int m, int segments, int threads, uint8_t *out)
{
    const int64_t windows = n - m + 1;
    if (windows <= 0 || segments <= 0 || m % segments != 0)
        return;
    float bp[255];
    orc_sax_breakpoints(bp);
#pragma omp parallel num_threads(threads)
    {
        float *z = (float *)malloc(sizeof(float) * (size_t)m);
        float *paa = (float *)malloc(sizeof(float) * (size_t)segments);
#pragma omp for schedule(static)
        for (int64_t id = 0; id < windows; id++) {
            const float *w = raw + id;
            uint8_t *o = out + (size_t)id * segments;
            memset(o, 0, (size_t)segments);
            if (w[0] == ORC_MISSING)
                continue;
            double sum = 0.0;
            for (int d = 0; d < m; d++)
                sum += w[d] == ORC_MISSING ? 0.0 : (double)w[d];
            const double mean = sum / m;
            double squares = 0.0;
            for (int d = 0; d < m; d++) {
                const double v = (w[d] == ORC_MISSING ? 0.0 : (double)w[d]) - mean;
                squares = fma(v, v, squares);
            }
            const double sd = sqrt(squares / m);
            const double div = sd <= 1e-12 ? 1.0 : sd;
            for (int d = 0; d < m; d++) {
                const double v = w[d] == ORC_MISSING ? 0.0 : (double)w[d];
                z[d] = (float)((v - mean) / div);
            }
            orc_paa(z, m, segments, paa);
            for (int s = 0; s < segments; s++)
                o[s] = sax_symbol(bp, paa[s]);
        }
        free(z);
        free(paa);
    }
}

/* The lower bound of SkipSequential.cpp:200-211: per segment the distance of
 * the template's PAA value from the word's symbol region [bp[v - 1], bp[v]]
 * (MINVAL / MAXVAL = -/+2e6 at the ends, iSAXTypes.hpp:17-18), all at full
 * cardinality.  16 segments (minidist_paa_to_isax_rawa_SIMD, SAX.cpp:362-490):
 * FP32 squares summed as two DPPS of each 8-lane half, ((s0 + s1) + (s2 +
 * s3)), the halves' sums added lane-wise, then lanes 0 and 4; otherwise
 * (minidist_paa_to_isax, SAX.cpp:112-180) the FP32 difference widened to FP64
 * and accumulated with one fused multiply-add per segment into the FP32 sum.
 * Then times m / segments in FP32. */
float orc_lower_bound(const float *bp, const float *paa, const uint8_t *word, int m, int segments)
{
    float d[64];
    for (int s = 0; s < segments && s < 64; s++) {
        const int v = word[s];
        const float lo = v == 0 ? -2000000.0f : bp[v - 1];
        const float hi = v == 255 ? 2000000.0f : bp[v];
        d[s] = lo > paa[s] ? lo - paa[s] : (hi < paa[s] ? hi - paa[s] : 0.0f);
    }
    const float ratio = (float)m / (float)segments;
    if (segments == 16) {
        float q[16];
        for (int s = 0; s < 16; s++)
            q[s] = d[s] * d[s];
        float h[4];
        for (int g = 0; g < 4; g++)
            h[g] = (q[4 * g] + q[4 * g + 1]) + (q[4 * g + 2] + q[4 * g + 3]);
        const float lane0 = h[0] + h[2], lane4 = h[1] + h[3];
        return (lane0 + lane4) * ratio;
    }
    float dist = 0.0f;
    for (int s = 0; s < segments; s++) {
        const double e = (double)d[s];
        dist = (float)fma(e, e, (double)dist);
    }
    return ratio * dist;
}

/* orc_lower_bound for `count` consecutive words. */
void orc_lower_bounds(const float *bp, const float *paa, const uint8_t *words, int64_t count, int m, int segments,
                      float *out)
{
    for (int64_t i = 0; i < count; i++)
        out[i] = orc_lower_bound(bp, paa, words + (size_t)i * segments, m, segments);
}

/* Per-template lower-bound state of the skip-sequential scan (NULL words: no
 * bound, the sequential method). */
typedef struct {
    const uint8_t *words;
    float paa[64];
    float bp[255];
    int m, segments;
} orc_lb;

static int lb_setup(orc_lb *lb, const float *raw, int64_t n, int m, int segments, int threads, const float *qz)
{
    lb->words = NULL;
    if (segments <= 0 || segments > 64 || m % segments != 0 || n - m + 1 <= 0)
        return 0;
    uint8_t *w = (uint8_t *)malloc((size_t)(n - m + 1) * (size_t)segments);
    orc_sax_words(raw, n, m, segments, threads, w);
    orc_sax_breakpoints(lb->bp);
    orc_paa(qz, m, segments, lb->paa);
    lb->words = w;
    lb->m = m;
    lb->segments = segments;
    return 1;
}

static int lb_skip(const orc_lb *lb, int64_t id, float bound)
{
    return lb->words &&
           orc_lower_bound(lb->bp, lb->paa, lb->words + (size_t)id * lb->segments, lb->m, lb->segments) > bound;
}

/* Prepared template: SkipSequential.cpp:213-221 (z-norm + block order; the PAA
 * only feeds the lower bounds, which the sequential method does not use). */
static void prepare(const float *t, int m, float *qz, int *order)
{
    memcpy(qz, t, sizeof(float) * (size_t)m);
    orc_znormalize(qz, m);
    orc_block_order(qz, m, order);
}

/* SkipSequential.cpp:296-393 on one rank with lower_bounds = false, for
 * templates of individual lengths.  Windows are scanned in position order in
 * chunks of 2^16 claimed dynamically; the bound set holds k entries (2k-1 with
 * a merge offset); candidates are windows with distance <= the bound when
 * offered, filtered by the final bound, then selected canonically.  Results
 * are k slots per template, unfilled slots (0, FLT_MAX). */
int orc_scan_knn(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                 const int64_t *offsets, int count, int k, int merge_offset, int threads, uint64_t *out_pos,
                 float *out_d, double *seconds)
{
    return orc_scan_knn_lb(raw, n, templates, lengths, offsets, count, k, merge_offset, threads, 0, out_pos, out_d,
                           seconds);
}

/* The same with lower_bounds = true and `segments` PAA segments (the
 * skip-sequential method, SkipSequential.cpp:321): a window is skipped when
 * its lower bound exceeds the bound of the moment; templates whose length is
 * not a multiple of `segments` are scanned without bounds. */
int orc_scan_knn_lb(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                    const int64_t *offsets, int count, int k, int merge_offset, int threads, int segments,
                    uint64_t *out_pos, float *out_d, double *seconds)
{
    const int bound_size = merge_offset > 0 ? 2 * k - 1 : k;
    for (int t = 0; t < count; t++) {
        const int m = lengths[t];
        const int64_t windows = n - m + 1;
        uint64_t *op = out_pos + (size_t)t * k;
        float *od = out_d + (size_t)t * k;
        for (int j = 0; j < k; j++) {
            op[j] = 0;
            od[j] = FLT_MAX;
        }
        if (windows <= 0 || m <= 0)
            continue;
        float *mean = (float *)malloc(sizeof(float) * (size_t)windows);
        float *sd = (float *)malloc(sizeof(float) * (size_t)windows);
        orc_window_stats(raw, n, m, threads, mean, sd);
        float *qz = (float *)malloc(sizeof(float) * (size_t)m);
        int *order = (int *)malloc(sizeof(int) * (size_t)(m / ORC_BLOCK + 1));
        prepare(templates + offsets[t], m, qz, order);
        orc_lb lb;
        lb_setup(&lb, raw, n, m, segments, threads, qz);

        double t0 = 0.0;
#ifdef _OPENMP
        t0 = omp_get_wtime();
#endif
        orc_bsf best;
        orc_bsf_init(&best, bound_size);
        size_t cap = 1024, nc = 0;
        orc_cand *cand = (orc_cand *)malloc(sizeof(orc_cand) * cap);
#pragma omp parallel num_threads(threads)
        {
            int blocks = 0;
#pragma omp for schedule(dynamic, 65536)
            for (int64_t id = 0; id < windows; id++) {
                float bound;
#pragma omp atomic read
                bound = best.knn[bound_size - 1];
                if (raw[id] == ORC_MISSING || lb_skip(&lb, id, bound))
                    continue;
                const float dist = orc_dist(qz, raw, id, m, mean[id], sd[id], bound, order, &blocks);
                if (dist > bound)
                    continue;
#pragma omp critical(orc_knn)
                {
                    if (dist <= best.knn[bound_size - 1]) {
                        if (nc == cap) {
                            cap *= 2;
                            cand = (orc_cand *)realloc(cand, sizeof(orc_cand) * cap);
                        }
                        cand[nc].pos = (uint64_t)id;
                        cand[nc].d = dist;
                        nc++;
                    }
                    if (dist < best.knn[bound_size - 1])
                        orc_bsf_insert_offset(&best, dist, id, merge_offset);
                }
            }
        }
        const float fb = best.knn[bound_size - 1];
        size_t kept = 0;
        for (size_t i = 0; i < nc; i++)
            if (cand[i].d <= fb)
                cand[kept++] = cand[i];
        uint64_t *pp = (uint64_t *)malloc(sizeof(uint64_t) * (kept + 1));
        float *dd = (float *)malloc(sizeof(float) * (kept + 1));
        for (size_t i = 0; i < kept; i++) {
            pp[i] = cand[i].pos;
            dd[i] = cand[i].d;
        }
        orc_select_knn(pp, dd, (int64_t)kept, k, merge_offset, op, od);
#ifdef _OPENMP
        if (seconds)
            seconds[t] = omp_get_wtime() - t0;
#else
        if (seconds)
            seconds[t] = 0.0;
#endif
        free(pp);
        free(dd);
        free(cand);
        free((void *)lb.words);
        orc_bsf_free(&best);
        free(qz);
        free(order);
        free(mean);
        free(sd);
    }
    return 0;
}

/* SkipSequential.cpp:223-294 on one rank with lower_bounds = false.  A window
 * is a hit when its distance <= cc_to_ed(rho, m); hits are sorted by (pos, d)
 * and a hit closer than merge_offset + 1 to the last kept one is dropped.
 * Writes at most `cap` rows (template, pos, d) and returns the number of rows
 * of the full answer. */
int64_t orc_scan_threshold(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                           const int64_t *offsets, int count, float rho, int merge_offset, int threads,
                           int64_t cap, int32_t *out_t, uint64_t *out_pos, float *out_d)
{
    return orc_scan_threshold_lb(raw, n, templates, lengths, offsets, count, rho, merge_offset, threads, 0, cap,
                                 out_t, out_pos, out_d);
}

/* The same with lower_bounds = true (SkipSequential.cpp:256): a window whose
 * lower bound exceeds cc_to_ed(rho, m) is skipped. */
int64_t orc_scan_threshold_lb(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                              const int64_t *offsets, int count, float rho, int merge_offset, int threads,
                              int segments, int64_t cap, int32_t *out_t, uint64_t *out_pos, float *out_d)
{
    int64_t rows = 0;
    for (int t = 0; t < count; t++) {
        const int m = lengths[t];
        const int64_t windows = n - m + 1;
        if (windows <= 0 || m <= 0)
            continue;
        const float bound = orc_cc_to_ed(rho, m);
        float *mean = (float *)malloc(sizeof(float) * (size_t)windows);
        float *sd = (float *)malloc(sizeof(float) * (size_t)windows);
        orc_window_stats(raw, n, m, threads, mean, sd);
        float *qz = (float *)malloc(sizeof(float) * (size_t)m);
        int *order = (int *)malloc(sizeof(int) * (size_t)(m / ORC_BLOCK + 1));
        prepare(templates + offsets[t], m, qz, order);
        orc_lb lb;
        lb_setup(&lb, raw, n, m, segments, threads, qz);
        size_t cap_h = 1024, nh = 0;
        orc_cand *hits = (orc_cand *)malloc(sizeof(orc_cand) * cap_h);
#pragma omp parallel num_threads(threads)
        {
            int blocks = 0;
            size_t lcap = 64, ln = 0;
            orc_cand *local = (orc_cand *)malloc(sizeof(orc_cand) * lcap);
#pragma omp for schedule(dynamic, 65536)
            for (int64_t id = 0; id < windows; id++) {
                if (raw[id] == ORC_MISSING || lb_skip(&lb, id, bound))
                    continue;
                const float dist = orc_dist(qz, raw, id, m, mean[id], sd[id], bound, order, &blocks);
                if (dist <= bound) {
                    if (ln == lcap) {
                        lcap *= 2;
                        local = (orc_cand *)realloc(local, sizeof(orc_cand) * lcap);
                    }
                    local[ln].pos = (uint64_t)id;
                    local[ln].d = dist;
                    ln++;
                }
            }
#pragma omp critical(orc_thr)
            {
                if (nh + ln > cap_h) {
                    while (nh + ln > cap_h)
                        cap_h *= 2;
                    hits = (orc_cand *)realloc(hits, sizeof(orc_cand) * cap_h);
                }
                memcpy(hits + nh, local, sizeof(orc_cand) * ln);
                nh += ln;
            }
            free(local);
        }
        qsort(hits, nh, sizeof(orc_cand), cmp_pos_d);
        int has_last = 0;
        uint64_t last = 0;
        for (size_t i = 0; i < nh; i++) {
            if (has_last && merge_offset > 0 && hits[i].pos < last + (uint64_t)(merge_offset + 1))
                continue;
            if (rows < cap) {
                out_t[rows] = t;
                out_pos[rows] = hits[i].pos;
                out_d[rows] = hits[i].d;
            }
            rows++;
            last = hits[i].pos;
            has_last = 1;
        }
        free(hits);
        free((void *)lb.words);
        free(qz);
        free(order);
        free(mean);
        free(sd);
    }
    return rows;
}
