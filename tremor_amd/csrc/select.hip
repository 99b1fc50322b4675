// Grouping of candidate windows by template, kNN bound seeding and the
// canonical kNN selection on the device.
//
//   seed bound     the k-th (merge 0) or (2k-1)-th pairwise-separated (merge
//                  > 0) smallest distance upper bound among sampled windows:
//                  the role of the reference's bound set pqueue_bsf
//                  (iSAXPqueue.cpp:471-549, bound size SkipSequential.cpp:299)
//   kNN select     select_knn_with_exclusion (ZnormSearch.hpp:170-189): sort by
//                  (d, pos), keep a window unless one kept lies closer than
//                  merge_offset, stop at k, fill (0, FLT_MAX)
#include <hipcub/hipcub.hpp>

#include "tmx_internal.hpp"

namespace tmx {

// Predicates of the grouping pass.
enum GroupPred : int {
    kPredAll = 0,        // every entry
    kPredHit = 1,        // exact distance <= the template's bound (threshold hits)
    kPredFinite = 2,     // value < +inf (seed entries)
    kPredQuery = 3,      // per template: every entry for kNN, hits for threshold
};

// Sort keys: (value bits << 32) | window, or (window << 32) | value bits.
enum GroupKey : int {
    kKeyValuePos = 0,
    kKeyPosValue = 1,
    kKeyQuery = 2,       // per template: (value, pos) for kNN, (pos, value) for threshold
};

namespace {

constexpr int kThreads = 256;

__device__ inline bool pred_ok(int pred, const TmplDev *tmpl, uint32_t t, float v)
{
    if (pred == kPredHit)
        return v <= tmpl[t].bound;
    if (pred == kPredFinite)
        return v < INFINITY;
    if (pred == kPredQuery)
        return tmpl[t].k > 0 || v <= tmpl[t].bound;
    return true;
}

__global__ void k_group_count(const Cand *__restrict__ cand, const float *__restrict__ val, uint64_t count,
                              const TmplDev *__restrict__ tmpl, int pred, uint32_t *__restrict__ counts)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= count)
        return;
    const Cand c = cand[i];
    if (pred_ok(pred, tmpl, c.t, val[i]))
        (void)wave_add_by_key(counts, c.t);
}

// Exclusive scan of T counts into offsets[0..T], one workgroup; also resets
// the scatter cursors.
__global__ __launch_bounds__(1024) void k_group_offsets(const uint32_t *__restrict__ counts, int T,
                                                        uint32_t *__restrict__ offsets, uint32_t *__restrict__ cursor)
{
    __shared__ uint32_t part[1024];
    const int per = (T + 1023) / 1024;
    const int lo = per * threadIdx.x;
    const int hi = lo + per < T ? lo + per : T;
    uint32_t s = 0;
    for (int t = lo; t < hi; t++)
        s += counts[t];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
    for (int t = lo; t < hi; t++) {
        offsets[t] = run;
        cursor[t] = 0u;
        run += counts[t];
    }
    if (threadIdx.x == 1023)
        offsets[T] = part[1023];
}

__device__ inline uint64_t make_key(int keykind, const TmplDev *tmpl, const Cand &c, float v)
{
    const uint64_t vb = static_cast<uint64_t>(__float_as_uint(v));
    const bool value_first = keykind == kKeyValuePos || (keykind == kKeyQuery && tmpl[c.t].k > 0);
    return value_first ? ((vb << 32) | c.j) : ((static_cast<uint64_t>(c.j) << 32) | vb);
}

__global__ void k_group_scatter(const Cand *__restrict__ cand, const float *__restrict__ val, uint64_t count,
                                const TmplDev *__restrict__ tmpl, int pred, int keykind,
                                const uint32_t *__restrict__ offsets, uint32_t *__restrict__ cursor,
                                uint64_t *__restrict__ keys)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= count)
        return;
    const Cand c = cand[i];
    const float v = val[i];
    if (!pred_ok(pred, tmpl, c.t, v))
        return;
    const uint32_t slot = offsets[c.t] + wave_add_by_key(cursor, c.t);
    keys[slot] = make_key(keykind, tmpl, c, v);
}

__global__ void k_zero_counts(uint32_t *__restrict__ counts, int T)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < T)
        counts[i] = 0u;
}


// Large inputs: the entries satisfying pred as (key, template) pairs in any
// order, behind a prefill of sentinels (key ~0, template T) up to count.
__global__ void k_fill_sentinel(uint64_t *__restrict__ keys, uint32_t *__restrict__ tv, uint64_t count, uint32_t T)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < count) {
        keys[i] = ~0ull;
        tv[i] = T;
    }
}

__global__ void k_group_compact(const Cand *__restrict__ cand, const float *__restrict__ val, uint64_t count,
                                const TmplDev *__restrict__ tmpl, int pred, int keykind,
                                unsigned long long *__restrict__ total, uint64_t *__restrict__ keys,
                                uint32_t *__restrict__ tv)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    bool ok = false;
    Cand c{0u, 0u};
    float v = 0.0f;
    if (i < count) {
        c = cand[i];
        v = val[i];
        ok = pred_ok(pred, tmpl, c.t, v);
    }
    const unsigned long long mask = __ballot(ok);
    if (!mask)
        return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader)
        base = atomicAdd(total, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    if (ok) {
        const unsigned long long slot = base + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull)));
        keys[slot] = make_key(keykind, tmpl, c, v);
        tv[slot] = c.t;
    }
}

// Candidates of kNN templates whose lower bound exceeds the (tightened)
// bound cannot be answers: the rest are compacted into the second set.
__global__ void k_refine(const Cand *__restrict__ cand, const float *__restrict__ est, const float *__restrict__ lo,
                         const float *__restrict__ hi, uint64_t count, const TmplDev *__restrict__ tmpl,
                         Cand *__restrict__ cand2, float *__restrict__ est2, float *__restrict__ lo2,
                         float *__restrict__ hi2, unsigned long long *__restrict__ total)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    bool ok = false;
    if (i < count) {
        const uint32_t t = cand[i].t;
        ok = !(tmpl[t].k > 0 && lo[i] > tmpl[t].bound);
    }
    const unsigned long long mask = __ballot(ok);
    if (!mask)
        return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader)
        base = atomicAdd(total, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    if (ok) {
        const unsigned long long slot = base + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull)));
        cand2[slot] = cand[i];
        est2[slot] = est[i];
        lo2[slot] = lo[i];
        hi2[slot] = hi[i];
    }
}

__global__ void k_zero_total(unsigned long long *p) { *p = 0ull; }

__device__ inline float key_val(uint64_t k) { return __uint_as_float(static_cast<uint32_t>(k >> 32)); }
__device__ inline uint32_t key_pos(uint64_t k) { return static_cast<uint32_t>(k); }

// Greedy pass over (value, pos)-sorted keys of one template by one wavefront:
// keep an entry unless a kept one lies closer than `merge` (strictly), until
// `want` are kept.  Kept entries go to LDS.  Returns the number kept.
__device__ int greedy_separated(const uint64_t *keys, uint32_t cnt, int want, int merge, uint64_t *kept)
{
    const int lane = threadIdx.x & 63;
    int nk = 0;
    for (uint32_t base = 0; base < cnt && nk < want; base += 64) {
        // 64 keys at a time in registers, broadcast one by one.
        const uint64_t mine = base + lane < cnt ? keys[base + lane] : ~0ull;
        const int span = cnt - base < 64u ? static_cast<int>(cnt - base) : 64;
        for (int l = 0; l < span && nk < want; l++) {
            const uint64_t k = __shfl(mine, l);
            const uint32_t p = key_pos(k);
            bool blocked = false;
            for (int kb = 0; kb < nk; kb += 64) {
                const int i = kb + lane;
                bool near = false;
                if (i < nk) {
                    const uint32_t q = key_pos(kept[i]);
                    const uint32_t gap = q > p ? q - p : p - q;
                    near = gap < static_cast<uint32_t>(merge);
                }
                if (__any(near)) {
                    blocked = true;
                    break;
                }
            }
            if (!blocked) {
                if (lane == 0)
                    kept[nk] = k;
                nk++;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
    return nk;
}

// Seed bound per kNN template (k > 0): one wavefront per template; the
// bound set holds k entries (merge 0) or 2k - 1 (SkipSequential.cpp:299).
__global__ __launch_bounds__(64) void k_seed_bound(const uint64_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ offsets, TmplDev *tmpl)
{
    extern __shared__ uint64_t kept[];
    const int t = blockIdx.x;
    const int k = tmpl[t].k, merge = tmpl[t].merge;
    if (k <= 0)
        return;
    const int want = merge > 0 ? 2 * k - 1 : k;
    const uint32_t b = offsets[t], e = offsets[t + 1];
    float U = FLT_MAX;
    if (merge == 0) {
        if (e - b >= static_cast<uint32_t>(want))
            U = key_val(keys[b + want - 1]);
    } else {
        const int nk = greedy_separated(keys + b, e - b, want, merge, kept);
        __syncthreads();
        if (nk >= want)
            U = key_val(kept[want - 1]);
    }
    // Only ever tightens: a bound from any sample of upper bounds is valid.
    if (threadIdx.x == 0)
        tmpl[t].bound = fminf(tmpl[t].bound, U < FLT_MAX ? U : FLT_MAX);
}

// Canonical kNN answer per kNN template from (d, pos)-sorted exact
// candidates, into k slots from the template's output offset.
__global__ __launch_bounds__(64) void k_knn_select(const uint64_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ offsets,
                                                   const TmplDev *__restrict__ tmpl, uint64_t *__restrict__ out_pos,
                                                   float *__restrict__ out_d)
{
    extern __shared__ uint64_t kept[];
    const int t = blockIdx.x;
    const int k = tmpl[t].k, merge = tmpl[t].merge;
    if (k <= 0)
        return;
    const uint32_t b = offsets[t], e = offsets[t + 1];
    int nk;
    if (merge == 0) {
        nk = static_cast<int>(e - b) < k ? static_cast<int>(e - b) : k;
        for (int i = threadIdx.x; i < nk; i += 64)
            kept[i] = keys[b + i];
    } else {
        nk = greedy_separated(keys + b, e - b, k, merge, kept);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += 64) {
        const size_t o = static_cast<size_t>(tmpl[t].out_off) + i;
        if (i < nk) {
            out_pos[o] = key_pos(kept[i]);
            out_d[o] = key_val(kept[i]);
        } else {
            out_pos[o] = 0;
            out_d[o] = FLT_MAX;
        }
    }
}

}  // namespace

// Groups the entries satisfying `pred` by template into `keys` (sorted within
// each template); offsets has T + 1 slots.  Many small segments: a scatter
// into the template segments and a segmented sort.  Otherwise: the pairs
// (key, template) compacted behind sentinels and sorted by two stable
// device-wide radix sorts, by key and then by template (one sort for a single
// template).  A segmented sort gives one workgroup to each segment, which
// serialises a large one (22K keys of one template: 0.7 ms).
constexpr uint64_t kSegmentedMax = 1u << 16;
constexpr uint64_t kSegmentedAvg = 1024;  // entries per template, on average

static bool use_segmented(uint64_t count, int T)
{
    return count <= kSegmentedMax && count <= kSegmentedAvg * static_cast<uint64_t>(T);
}

static int template_bits(int T)
{
    int b = 1;
    while ((1u << b) <= static_cast<uint32_t>(T))
        b++;
    return b;
}

hipError_t launch_group_sort(const Cand *cand, const float *val, uint64_t count, const TmplDev *tmpl, int T,
                             int pred, int keykind, uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                             uint64_t *keys_tmp, uint64_t *keys, uint32_t *tv_a, uint32_t *tv_b,
                             unsigned long long *total, void *sort_tmp, size_t sort_tmp_bytes, hipStream_t stream)
{
    // Counters cleared by a kernel, not a memset node: the sequence stays
    // capturable in a HIP graph.
    hipLaunchKernelGGL(k_zero_counts, dim3((T + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, counts, T);
    const unsigned grid = static_cast<unsigned>((count + kThreads - 1) / kThreads);
    if (count > 0)
        hipLaunchKernelGGL(k_group_count, dim3(grid), dim3(kThreads), 0, stream, cand, val, count, tmpl, pred, counts);
    hipLaunchKernelGGL(k_group_offsets, dim3(1), dim3(1024), 0, stream, counts, T, offsets, cursor);
    if (count == 0)
        return hipGetLastError();
    size_t bytes = sort_tmp_bytes;
    if (use_segmented(count, T)) {
        hipLaunchKernelGGL(k_group_scatter, dim3(grid), dim3(kThreads), 0, stream, cand, val, count, tmpl, pred,
                           keykind, offsets, cursor, keys_tmp);
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess)
            return err;
        return hipcub::DeviceSegmentedRadixSort::SortKeys(sort_tmp, bytes, keys_tmp, keys, static_cast<int>(count),
                                                         T, offsets, offsets + 1, 0, 64, stream);
    }
    hipLaunchKernelGGL(k_fill_sentinel, dim3(grid), dim3(kThreads), 0, stream, keys_tmp, tv_a, count,
                       static_cast<uint32_t>(T));
    hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(1), 0, stream, total);
    hipLaunchKernelGGL(k_group_compact, dim3(grid), dim3(kThreads), 0, stream, cand, val, count, tmpl, pred, keykind,
                       total, keys_tmp, tv_a);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return err;
    // By key (keys_tmp -> keys, templates along), then stably by template
    // (keys -> keys_tmp), then back into keys.
    err = hipcub::DeviceRadixSort::SortPairs(sort_tmp, bytes, keys_tmp, keys, tv_a, tv_b, static_cast<int>(count), 0,
                                             64, stream);
    if (err != hipSuccess || T == 1)
        return err;
    bytes = sort_tmp_bytes;
    err = hipcub::DeviceRadixSort::SortPairs(sort_tmp, bytes, tv_b, tv_a, keys, keys_tmp, static_cast<int>(count), 0,
                                             template_bits(T), stream);
    if (err != hipSuccess)
        return err;
    return hipMemcpyAsync(keys, keys_tmp, sizeof(uint64_t) * count, hipMemcpyDeviceToDevice, stream);
}

size_t group_sort_tmp_bytes(uint64_t count, int T)
{
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, a, static_cast<const uint64_t *>(nullptr),
                                                     static_cast<uint64_t *>(nullptr),
                                                     static_cast<int>(count < kSegmentedMax ? count : kSegmentedMax),
                                                     T, static_cast<const uint32_t *>(nullptr),
                                                     static_cast<const uint32_t *>(nullptr), 0, 64);
    if (!use_segmented(count, T)) {
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<const uint64_t *>(nullptr),
                                                 static_cast<uint64_t *>(nullptr),
                                                 static_cast<const uint32_t *>(nullptr),
                                                 static_cast<uint32_t *>(nullptr), static_cast<int>(count), 0, 64);
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, c, static_cast<const uint32_t *>(nullptr),
                                                 static_cast<uint32_t *>(nullptr),
                                                 static_cast<const uint64_t *>(nullptr),
                                                 static_cast<uint64_t *>(nullptr), static_cast<int>(count), 0,
                                                 template_bits(T));
    }
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// Drops the kNN candidates whose lower bound exceeds their template's bound;
// the kept ones go to the second set, their number to *total.
hipError_t launch_refine(const Cand *cand, const float *est, const float *lo, const float *hi, uint64_t count,
                         const TmplDev *tmpl, Cand *cand2, float *est2, float *lo2, float *hi2,
                         unsigned long long *total, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(1), 0, stream, total);
    if (count > 0)
        hipLaunchKernelGGL(k_refine, dim3(static_cast<unsigned>((count + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           stream, cand, est, lo, hi, count, tmpl, cand2, est2, lo2, hi2, total);
    return hipGetLastError();
}

hipError_t launch_seed_bound(const uint64_t *keys, const uint32_t *offsets, TmplDev *tmpl, int T, int max_want,
                             hipStream_t stream)
{
    if (T <= 0 || max_want <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_seed_bound, dim3(T), dim3(64), sizeof(uint64_t) * static_cast<size_t>(max_want), stream,
                       keys, offsets, tmpl);
    return hipGetLastError();
}

// max_k: the largest k over the templates.
hipError_t launch_knn_select(const uint64_t *keys, const uint32_t *offsets, const TmplDev *tmpl, int T, int max_k,
                             uint64_t *out_pos, float *out_d, hipStream_t stream)
{
    if (T <= 0 || max_k <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_knn_select, dim3(T), dim3(64), sizeof(uint64_t) * static_cast<size_t>(max_k), stream, keys,
                       offsets, tmpl, out_pos, out_d);
    return hipGetLastError();
}

}  // namespace tmx
