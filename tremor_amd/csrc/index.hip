// The iSAX index on the device (TREMOR's index, indexing.cpp:24-281 build;
// Tremor.cpp:1195-1518 leaf-order search with node mindist and the record
// pass): windows sorted by an interleaved iSAX key, cut top-down into leaves
// of at most `leaf_size` windows, and a search that visits only the leaves
// whose region's lower bound does not exceed a template's bound, then only
// their windows whose own word's bound does not.
//
// Key of a window (16 PAA segments, symbols at cardinality 256): bit plane p
// (0 = most significant) of segment s's symbol at key bit 63 - (16 p + s), for
// p < 4.  A node at depth d holds the windows whose keys share the top d bits:
// segment s is then known to ceil((d - s) / 16) bits (clamped to [0, 4]).  The
// root level is depth 16 (one bit per segment, the reference's 2^16 root
// nodes); a node splits on its next key bit while it holds more than
// leaf_size windows (the reference's informed split picks the segment by the
// data; here the segments split in turn, which keeps the tree a pure function
// of the keys).  A node that cannot split (64 bits known) stays a leaf.
#include <hipcub/hipcub.hpp>

#include "ctx.hpp"
#include "ea.hpp"

namespace tmx {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_ix_keys(const uint8_t *__restrict__ words, uint64_t W,
                                                      uint64_t *__restrict__ keys, uint32_t *__restrict__ ids)
{
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (j >= W)
        return;
    const uint4 wv = *reinterpret_cast<const uint4 *>(words + 16 * j);
    const uint32_t wd[4] = {wv.x, wv.y, wv.z, wv.w};
    uint64_t key = 0;
#pragma unroll
    for (int p = 0; p < kIxPlanes; p++)
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const uint32_t sym = (wd[s >> 2] >> (8 * (s & 3))) & 0xffu;
            key |= static_cast<uint64_t>((sym >> (7 - p)) & 1u) << (63 - (16 * p + s));
        }
    keys[j] = key;
    ids[j] = static_cast<uint32_t>(j);
}

// One level of the top-down split: node i covers sorted keys [start, start +
// count) sharing their top `depth` bits.  A leaf is appended to the leaf list;
// a node that splits appends its non-empty children (the first index whose
// next bit is set found by binary search: within a node that bit is
// monotone in the sorted order).
// counters: [0] next nodes, [1] leaves, [2] set when a list would exceed its
// capacity (the entry is dropped and the host fails the build).
__global__ __launch_bounds__(kThreads) void k_ix_split(const uint64_t *__restrict__ keys,
                                                       const uint4 *__restrict__ nodes, uint32_t nn,
                                                       uint4 *__restrict__ next, uint32_t next_cap,
                                                       uint4 *__restrict__ leaves, uint32_t leaf_cap,
                                                       uint32_t *__restrict__ counters, uint32_t leaf_size)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= nn)
        return;
    const uint4 nd = nodes[i];
    const uint32_t start = nd.x, count = nd.y, depth = nd.z;
    const auto put = [&](uint4 *list, uint32_t cap, uint32_t *n, uint4 v) {
        const uint32_t slot = atomicAdd(n, 1u);
        if (slot < cap)
            list[slot] = v;
        else
            atomicOr(counters + 2, 1u);
    };
    if (!(depth < 64u && (depth < kIxRootDepth || count > leaf_size))) {
        put(leaves, leaf_cap, counters + 1, nd);
        return;
    }
    const uint64_t bit = 1ull << (63 - depth);
    uint32_t lo = start, hi = start + count;  // first index with the bit set
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] & bit)
            hi = mid;
        else
            lo = mid + 1;
    }
    if (lo > start)
        put(next, next_cap, counters, make_uint4(start, lo - start, depth + 1, 0));
    if (lo < start + count)
        put(next, next_cap, counters, make_uint4(lo, start + count - lo, depth + 1, 0));
}

// The words and the window sums in key order (one gather at build): a
// leaf's are then consecutive for the search, which reads them coalesced
// instead of two prefix-sum lines and the first sample per window.  The sums
// are window_sums_fast's, exactly as the scan computes them; eS1 = -1 marks
// a window whose first sample is missing.
__global__ __launch_bounds__(kThreads) void k_ix_gather(Waveform w, int m, const uint8_t *__restrict__ words,
                                                        uint64_t W, const uint32_t *__restrict__ ids,
                                                        uint4 *__restrict__ out, WindowSums *__restrict__ sums)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= W)
        return;
    const uint64_t j = ids[i];
    out[i] = *reinterpret_cast<const uint4 *>(words + 16ull * j);
    WindowSums ws;
    if (w.x[j] == kMissing) {
        ws.S1 = ws.S2 = ws.eS2 = 0.0;
        ws.eS1 = -1.0;
    } else {
        ws = window_sums_fast(w, j, j + static_cast<uint64_t>(m));
    }
    sums[i] = ws;
}

__global__ __launch_bounds__(kThreads) void k_ix_leaf_keys(const uint4 *__restrict__ leaves, uint32_t n,
                                                           uint32_t *__restrict__ order_key,
                                                           uint32_t *__restrict__ order_val)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        order_key[i] = leaves[i].x;
        order_val[i] = i;
    }
}

__global__ __launch_bounds__(kThreads) void k_ix_leaf_gather(const uint4 *__restrict__ leaves, uint32_t n,
                                                             const uint32_t *__restrict__ order,
                                                             const uint64_t *__restrict__ keys,
                                                             IxLeaf *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const uint4 nd = leaves[order[i]];
        out[i] = IxLeaf{keys[nd.x], nd.x, nd.y, static_cast<int32_t>(nd.z), 0};
    }
}

// The reference's routing sample (Tremor.cpp:2045-2059): for template
// ids[blockIdx.y], the windows j * stride, j < samples, whose word bound does
// not exceed the template's bound, counted into passed[template].
__global__ __launch_bounds__(kThreads) void k_ix_sample(const TmplDev *__restrict__ tmpl,
                                                        const int32_t *__restrict__ ids, const float *__restrict__ bp_g,
                                                        const float *__restrict__ paa_g, uint64_t n, int samples,
                                                        unsigned long long *__restrict__ passed)
{
    __shared__ float bp[256];
    __shared__ float paa[16];
    const int tid = threadIdx.x;
    const int d = ids[blockIdx.y];
    const TmplDev &t = tmpl[d];
    for (int i = tid; i < 255; i += kThreads)
        bp[i] = bp_g[i];
    if (tid < 16)
        paa[tid] = paa_g[t.lb_paa + tid];
    __syncthreads();
    const uint64_t W = n - static_cast<uint64_t>(t.m) + 1;
    const uint64_t stride = W / static_cast<uint64_t>(samples) > 0 ? W / static_cast<uint64_t>(samples) : 1;
    unsigned long long c = 0;
    for (int j = blockIdx.x * kThreads + tid; j < samples; j += gridDim.x * kThreads) {
        const uint64_t id = static_cast<uint64_t>(j) * stride;
        if (id < W && sax_lower_bound(bp, paa, t.lb_words + 16ull * id, t.m, 16) <= t.bound)
            c++;
    }
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_xor(c, off);
    if ((tid & 63) == 0 && c > 0)
        atomicAdd(passed + d, c);
}

// Templates ids[blockIdx.y]: per leaf (one per thread), its bound against the
// template's bound; the windows of the leaves that pass counted into
// passed[template] (the auto rule's cost of the exact search).
__global__ __launch_bounds__(kThreads) void k_ix_count(const TmplDev *__restrict__ tmpl,
                                                       const int32_t *__restrict__ ids,
                                                       const float *__restrict__ bp_g, const float *__restrict__ paa_g,
                                                       unsigned long long *__restrict__ passed)
{
    __shared__ float bp[256];
    __shared__ float paa[16];
    const int tid = threadIdx.x;
    const int d = ids[blockIdx.y];
    const TmplDev &t = tmpl[d];
    for (int i = tid; i < 255; i += kThreads)
        bp[i] = bp_g[i];
    if (tid < 16)
        paa[tid] = paa_g[t.lb_paa + tid];
    __syncthreads();
    const uint32_t leaf = blockIdx.x * kThreads + tid;
    if (leaf < t.ix_nleaves) {
        const IxLeaf L = t.ix_leaves[leaf];
        if (ix_pass(ix_leaf_bound(bp, paa, L.key, L.depth, t.m), t.bound))
            atomicAdd(passed + d, static_cast<unsigned long long>(L.count));
    }
}

constexpr int kIxPer = 8;  // windows of passing leaves per thread and step of the exact search

// The running bounds of the exact search start at the templates' bounds.
__global__ __launch_bounds__(kThreads) void k_ix_bound_init(const TmplDev *__restrict__ tmpl,
                                                            const int32_t *__restrict__ ids, int count,
                                                            unsigned int *__restrict__ dyn)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count)
        dyn[ids[i]] = __float_as_uint(tmpl[ids[i]].bound);
}

// The exact search (Tremor.cpp:1195-1518 visits the leaves whose bound
// passes and computes the distance of each of their windows whose word's
// bound passes too, with early abandoning), template ids[blockIdx.y].
// Workgroup b of G = gridDim.x takes the leaves b, b + G, b + 2G, ...: the
// leaves that pass lie close together in key order (near the template's own
// key), and the stride spreads them evenly over the workgroups.  Per round of
// 256 leaves (one per thread) the passing ones' windows are numbered through
// a workgroup-wide prefix sum, and the workgroup takes 256 kIxPer numbers per
// step, so every lane has windows while any are left.  Their
// counts go to passed[template]; each window whose word bound passes goes
// through the early-abandoning window kernel of the scan pass (ea.hpp), which
// drops it once its running sum proves the reference's distance above the
// bound and otherwise makes it a candidate with an interval of that distance.
// kNN templates with k = 1 tighten their bound as they go, as the reference's
// best-so-far does: dyn[template] (float bits; non-negative floats order as
// their bits) holds the smallest upper bound of a candidate's reference
// distance found so far, which bounds the nearest window's distance too.
template <bool DYN>
__global__ __launch_bounds__(kThreads) void k_ix_search(Waveform w, const TmplDev *__restrict__ tmpl,
                                                        const int32_t *__restrict__ ids,
                                                        const float *__restrict__ qz, const int32_t *__restrict__ order,
                                                        const float *__restrict__ bp_g, const float *__restrict__ paa_g,
                                                        unsigned long long *__restrict__ passed,
                                                        unsigned int *__restrict__ dyn, CandOut co)
{
    __shared__ float bp[256];
    __shared__ float paa[16];
    __shared__ uint32_t s_first[kThreads];    // the round's leaves: first sorted index
    __shared__ uint32_t s_off[kThreads + 1];  // exclusive prefix of the passing leaves' counts (0 if failing)
    __shared__ uint32_t s_wave[kThreads / 64];
    __shared__ uint32_t s_sorted[kIxPer][kThreads];  // each thread's windows of the step (sorted positions)
    __shared__ float gap[16 * 256];  // gap[256 s + v] = sax_gap(bp, paa[s], v): a word's bound is 16 reads
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int d = ids[blockIdx.y];
    const TmplDev &t = tmpl[d];
    for (int i = tid; i < 255; i += kThreads)
        bp[i] = bp_g[i];
    if (tid < 16)
        paa[tid] = paa_g[t.lb_paa + tid];
    __syncthreads();
    for (int e = tid; e < 16 * 256; e += kThreads)
        gap[e] = sax_gap(bp, paa[e >> 8], e & 255);
    const uint64_t G = gridDim.x;
    const uint64_t nl = static_cast<uint64_t>(t.ix_nleaves);
    // DYN: a separate instance for batches with k = 1 templates, so that the
    // others keep the plain loop.
    const bool dynamic = DYN && t.k == 1;
    const auto current = [&]() {
        return dynamic ? fminf(t.bound, __uint_as_float(__hip_atomic_load(dyn + d, __ATOMIC_RELAXED,
                                                                           __HIP_MEMORY_SCOPE_AGENT)))
                       : t.bound;
    };
    for (uint64_t r = 0; r * G * kThreads < nl; r++) {  // uniform
        __syncthreads();  // the gap table written; the last round's s_first / s_off read
        const uint64_t leaf = (r * kThreads + static_cast<uint64_t>(tid)) * G + blockIdx.x;
        uint32_t cnt = 0, start = 0;
        if (leaf < nl) {
            const IxLeaf L = t.ix_leaves[leaf];
            if (ix_pass(ix_leaf_bound(bp, paa, L.key, L.depth, t.m), current())) {
                cnt = L.count;
                start = L.start;
            }
        }
        // Workgroup-wide exclusive prefix of cnt (the leaves partition the W <
        // 2^31 windows, so every sum of distinct leaves' counts fits 31 bits).
        uint32_t incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off);
            if (lane >= off)
                incl += v;
        }
        if (lane == 63)
            s_wave[wv] = incl;
        s_first[tid] = start;
        __syncthreads();
        uint32_t before = 0;
        for (int i = 0; i < wv; i++)
            before += s_wave[i];
        s_off[tid] = before + incl - cnt;
        uint32_t total = 0;
        for (int i = 0; i < kThreads / 64; i++)
            total += s_wave[i];
        if (tid == kThreads - 1)
            s_off[kThreads] = total;
        if (tid == 0 && total > 0)
            atomicAdd(passed + d, static_cast<unsigned long long>(total));
        __syncthreads();
        for (uint32_t base = 0; base < total; base += static_cast<uint32_t>(kThreads * kIxPer)) {
            // Window numbers base + k 256 + tid: consecutive lanes take
            // consecutive windows in key order (mostly neighbouring
            // positions, so the window kernel's sample reads coalesce).  The
            // leaf of number f is the last entry whose offset is <= f
            // (entries of failing leaves share the next one's offset).
            float bnd = current();
            uint32_t keep = 0;
#pragma unroll
            for (int k = 0; k < kIxPer; k++) {
                const uint32_t f = base + static_cast<uint32_t>(k * kThreads + tid);
                int e = 0;
#pragma unroll
                for (int step = kThreads / 2; step > 0; step >>= 1)
                    if (s_off[e + step] <= f)
                        e += step;
                const uint32_t sorted = s_first[e] + (f - s_off[e]);
                s_sorted[k][tid] = sorted;
                if (f < total && ix_pass(sax_lower_bound16_table(gap, t.ix_words + 16ull * sorted, t.m), bnd))
                    keep |= 1u << k;
            }
            // One copy of the window kernel, the lanes on the same k at once,
            // from the window's sums in key order (coalesced like the words).
#pragma unroll 1
            for (int k = 0; k < kIxPer; k++) {
                if (keep & (1u << k)) {
                    const uint32_t sorted = s_sorted[k][tid];
                    const WindowSums ws = t.ix_sums[sorted];
                    if (ws.eS1 < 0.0)
                        continue;  // its first sample is missing: never searched (as in the scan)
                    const float hi =
                        scan_window_sums(w, tmpl, 0, d, t.m, qz, order, kModeMain, co, nullptr, t.ix_ids[sorted], ws, bnd);
                    if (dynamic && hi < bnd) {
                        atomicMin(dyn + d, __float_as_uint(hi));
                        bnd = hi;
                    }
                }
            }
        }
    }
}

// The approximate search that seeds the reference's exact index search (the
// query's own leaf first): for template ids[blockIdx.x], its word from its
// PAA (symbols as the windows'), its key, and the `per` windows nearest that
// key in key order (its leaf and the neighbouring ones) as candidates whose
// exact distances tighten the kNN bound.  Windows starting with a missing
// sample are left out.
__global__ __launch_bounds__(kThreads) void k_ix_home(Waveform w, const TmplDev *__restrict__ tmpl,
                                                      const int32_t *__restrict__ ids,
                                                      const float *__restrict__ bp_g, const float *__restrict__ paa_g,
                                                      int per, CandOut co)
{
    __shared__ float bp[256];
    __shared__ uint32_t s_pos;
    const int tid = threadIdx.x;
    const int d = ids[blockIdx.x];
    const TmplDev &t = tmpl[d];
    for (int i = tid; i < 255; i += kThreads)
        bp[i] = bp_g[i];
    __syncthreads();
    const uint64_t W = w.n - static_cast<uint64_t>(t.m) + 1;
    const uint64_t *keys = t.ix_keys;
    if (tid == 0) {
        uint64_t key = 0;
        for (int s = 0; s < 16; s++) {
            const uint32_t sym = static_cast<uint32_t>(sax_symbol(bp, paa_g[t.lb_paa + s]));
            for (int p = 0; p < kIxPlanes; p++)
                key |= static_cast<uint64_t>((sym >> (7 - p)) & 1u) << (63 - (16 * p + s));
        }
        uint64_t lo = 0, hi = W;  // first sorted key >= the template's
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (keys[mid] < key)
                lo = mid + 1;
            else
                hi = mid;
        }
        s_pos = static_cast<uint32_t>(lo);
    }
    __syncthreads();
    const uint64_t n = static_cast<uint64_t>(per) < W ? static_cast<uint64_t>(per) : W;
    const uint64_t half = n / 2;
    uint64_t first = s_pos > half ? s_pos - half : 0;
    if (first + n > W)
        first = W - n;
    for (uint64_t e = tid; e < n; e += kThreads) {
        const uint32_t j = t.ix_ids[first + e];
        if (w.x[j] != kMissing)
            cand_put(co, atomicAdd(co.count, 1ULL), static_cast<uint32_t>(d), j, NAN, -INFINITY, INFINITY);
    }
}

}  // namespace

hipError_t build_isax_index(const Waveform &w, int m, const uint8_t *words, uint64_t W, int leaf_size, IxIndex &ix,
                            hipStream_t stream)
{
    if (W == 0 || W > 0x7fffffffull || leaf_size < 1 || m < 1 || W != w.n - static_cast<uint64_t>(m) + 1)
        return hipErrorInvalidValue;
    const int n = static_cast<int>(W);
    hipError_t e;
    // Keys and ids, then sorted by key (ids along).
    DevMem keys_in, ids_in, tmp;
    if ((e = keys_in.reserve(8 * W)) || (e = ids_in.reserve(4 * W)) || (e = ix.keys.reserve(8 * W)) ||
        (e = ix.ids.reserve(4 * W)) || (e = ix.words.reserve(16 * W)) || (e = ix.sums.reserve(sizeof(WindowSums) * W)))
        return e;
    const unsigned grid = static_cast<unsigned>((W + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_ix_keys, dim3(grid), dim3(kThreads), 0, stream, words, W, keys_in.as<uint64_t>(),
                       ids_in.as<uint32_t>());
    size_t bytes = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in.as<uint64_t>(), ix.keys.as<uint64_t>(),
                                                ids_in.as<uint32_t>(), ix.ids.as<uint32_t>(), n, 0, 64, stream)))
        return e;
    if ((e = tmp.reserve(bytes)))
        return e;
    if ((e = hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys_in.as<uint64_t>(), ix.keys.as<uint64_t>(),
                                                ids_in.as<uint32_t>(), ix.ids.as<uint32_t>(), n, 0, 64, stream)))
        return e;
    hipLaunchKernelGGL(k_ix_gather, dim3(grid), dim3(kThreads), 0, stream, w, m, words, W, ix.ids.as<uint32_t>(),
                       ix.words.as<uint4>(), ix.sums.as<WindowSums>());
    // Top-down split, one level per launch, from the root node (depth 0).
    // The nodes of one level are disjoint and non-empty: at most 2^16 of them
    // down to the root depth, then at most twice the nodes over leaf_size
    // windows above; the leaves, at depth >= 16, at most 2^16 plus that many
    // per depth.  Capacities from these counts (never above W), checked by the
    // kernel.
    const uint64_t per_level = std::min<uint64_t>(W, std::max<uint64_t>(1ull << 16, 2 * W / leaf_size)) + 1;
    const uint64_t leaf_cap = std::min<uint64_t>(W, (1ull << 16) + 96 * (W / leaf_size) + 1) + 1;
    DevMem nodes_a, nodes_b, leaves, counters;
    if ((e = nodes_a.reserve(16 * per_level)) || (e = nodes_b.reserve(16 * per_level)) ||
        (e = leaves.reserve(16 * leaf_cap)) || (e = counters.reserve(16)))
        return e;
    uint4 *cur = nodes_a.as<uint4>(), *nxt = nodes_b.as<uint4>();
    const uint4 root = make_uint4(0u, static_cast<uint32_t>(W), 0u, 0u);
    if ((e = hipMemcpyAsync(cur, &root, sizeof(root), hipMemcpyHostToDevice, stream)) ||
        (e = hipMemsetAsync(counters.p, 0, 12, stream)))
        return e;
    uint32_t nn = 1, cnt[3] = {0, 0, 0};
    for (int level = 0; nn > 0; level++) {
        if (level > 64)
            return hipErrorUnknown;  // a node's depth grows by one per level: at most 65 levels
        if ((e = hipMemsetAsync(counters.as<uint32_t>(), 0, 4, stream)))
            return e;
        hipLaunchKernelGGL(k_ix_split, dim3((nn + kThreads - 1) / kThreads), dim3(kThreads), 0, stream,
                           ix.keys.as<uint64_t>(), cur, nn, nxt, static_cast<uint32_t>(per_level), leaves.as<uint4>(),
                           static_cast<uint32_t>(leaf_cap), counters.as<uint32_t>(),
                           static_cast<uint32_t>(leaf_size));
        if ((e = hipMemcpyAsync(cnt, counters.p, 12, hipMemcpyDeviceToHost, stream)) ||
            (e = hipStreamSynchronize(stream)))
            return e;
        if (cnt[2])
            return hipErrorUnknown;  // capacity exceeded: nothing past it was written
        nn = cnt[0];
        std::swap(cur, nxt);
    }
    const uint32_t nl = cnt[1];
    // Leaves in key order (by their first sorted index).
    DevMem ok_in, ov_in, ok_out, ov_out;
    if ((e = ok_in.reserve(4 * nl)) || (e = ov_in.reserve(4 * nl)) || (e = ok_out.reserve(4 * nl)) ||
        (e = ov_out.reserve(4 * nl)) || (e = ix.leaves.reserve(sizeof(IxLeaf) * nl)))
        return e;
    const unsigned lgrid = (nl + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_ix_leaf_keys, dim3(lgrid), dim3(kThreads), 0, stream, leaves.as<uint4>(), nl,
                       ok_in.as<uint32_t>(), ov_in.as<uint32_t>());
    bytes = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ok_in.as<uint32_t>(), ok_out.as<uint32_t>(),
                                                ov_in.as<uint32_t>(), ov_out.as<uint32_t>(), static_cast<int>(nl), 0,
                                                32, stream)))
        return e;
    if ((e = tmp.reserve(bytes)))
        return e;
    if ((e = hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, ok_in.as<uint32_t>(), ok_out.as<uint32_t>(),
                                                ov_in.as<uint32_t>(), ov_out.as<uint32_t>(), static_cast<int>(nl), 0,
                                                32, stream)))
        return e;
    hipLaunchKernelGGL(k_ix_leaf_gather, dim3(lgrid), dim3(kThreads), 0, stream, leaves.as<uint4>(), nl,
                       ov_out.as<uint32_t>(), ix.keys.as<uint64_t>(), ix.leaves.as<IxLeaf>());
    ix.nleaves = nl;
    ix.leaf_size = leaf_size;
    if ((e = hipGetLastError()))
        return e;
    return hipStreamSynchronize(stream);  // the scratch above is freed on return
}

hipError_t launch_ix_count(const TmplDev *tmpl, const int32_t *ids, int count, int64_t max_leaves, const float *bp,
                           const float *paa, unsigned long long *passed, hipStream_t stream)
{
    if (count <= 0 || max_leaves <= 0)
        return hipSuccess;
    if (count > 65535 || max_leaves > 0x7fffffffll || !passed)
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>((max_leaves + kThreads - 1) / kThreads), static_cast<unsigned>(count));
    hipLaunchKernelGGL(k_ix_count, grid, dim3(kThreads), 0, stream, tmpl, ids, bp, paa, passed);
    return hipGetLastError();
}

hipError_t launch_ix_search(const Waveform &w, const TmplDev *tmpl, const int32_t *ids, int count, int64_t max_leaves,
                            const float *qz, const int32_t *order, const float *bp, const float *paa,
                            unsigned long long *passed, unsigned int *dyn, const CandOut &co, int cus,
                            hipStream_t stream)
{
    if (count <= 0 || max_leaves <= 0)
        return hipSuccess;
    if (count > 65535 || max_leaves > 0x7fffffffll || !passed)
        return hipErrorInvalidValue;
    if (dyn)
        hipLaunchKernelGGL(k_ix_bound_init, dim3((count + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, tmpl,
                           ids, count, dyn);
    // Workgroups per template: about 4 per compute unit (all resident at once
    // at the window kernel's ~100 registers), and no more than a round of
    // leaves each.  The workgroups run template after template (x before y):
    // every template's search has the whole device, and the windows in
    // flight are one template's (16 templates at once, each on 1/16 of the
    // device, took 3.6x as long at n = 10^8 on their scattered reads).
    const int64_t want = 4ll * std::max(cus, 1);
    const int64_t G = std::max<int64_t>(1, std::min<int64_t>(want, (max_leaves + kThreads - 1) / kThreads));
    const dim3 grid(static_cast<unsigned>(G), static_cast<unsigned>(count));
    if (dyn)
        hipLaunchKernelGGL(k_ix_search<true>, grid, dim3(kThreads), 0, stream, w, tmpl, ids, qz, order, bp, paa, passed,
                           dyn, co);
    else
        hipLaunchKernelGGL(k_ix_search<false>, grid, dim3(kThreads), 0, stream, w, tmpl, ids, qz, order, bp, paa, passed,
                           dyn, co);
    return hipGetLastError();
}

hipError_t launch_ix_home(const Waveform &w, const TmplDev *tmpl, const int32_t *ids, int count, const float *bp,
                          const float *paa, int per, const CandOut &co, hipStream_t stream)
{
    if (count <= 0 || per <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ix_home, dim3(static_cast<unsigned>(count)), dim3(kThreads), 0, stream, w, tmpl, ids, bp, paa,
                       per, co);
    return hipGetLastError();
}

hipError_t launch_ix_sample(const TmplDev *tmpl, const int32_t *ids, int count, const float *bp, const float *paa,
                            uint64_t n, int samples, unsigned long long *passed, hipStream_t stream)
{
    if (count <= 0 || samples <= 0)
        return hipSuccess;
    if (count > 65535)
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(std::min(64, (samples + kThreads - 1) / kThreads)),
                    static_cast<unsigned>(count));
    hipLaunchKernelGGL(k_ix_sample, grid, dim3(kThreads), 0, stream, tmpl, ids, bp, paa, n, samples, passed);
    return hipGetLastError();
}

}  // namespace tmx
