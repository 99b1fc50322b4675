// Early-abandoning scan as a candidate pass over every window of its
// templates (the window kernel and its error bound: ea.hpp).
#include "ea.hpp"

namespace tmx {
namespace {

constexpr int kThreads = 256;
constexpr int kTableSegments = 16;  // lower bounds over at most this many segments use a gap table
constexpr uint32_t kBlocksPerTemplate = 2048;  // main mode: workgroups per template (windows in a stride loop)

// ids[0..count): the templates of this pass (blockIdx.y).  Main mode: the
// windows [256 blockIdx.x, + 256) of each; sample mode: sample_n windows
// spread evenly over the template's windows, visited[t] += blocks visited,
// visited[T + t] += windows sampled.  A template with a lower bound (lb_seg >
// 0) skips every window whose bound exceeds its own bound before reading a
// sample (SkipSequential.cpp:256, :321); main mode counts them in *skipped.
// Main mode with dyn: kNN templates with k = 1 tighten their bound as they
// go, as the reference's bound set does (SkipSequential.cpp:299-330):
// dyn[template] (float bits; non-negative floats order as their bits) holds
// the smallest upper bound of a candidate's reference distance found so far,
// which bounds the nearest window's distance too.
template <bool DYN>
__global__ __launch_bounds__(kThreads) void k_scan_ea(Waveform w, const TmplDev *__restrict__ tmpl, int T,
                                                      const float *__restrict__ qz, const int32_t *__restrict__ order,
                                                      const int32_t *__restrict__ ids, int mode, uint64_t sample_n,
                                                      CandOut co, unsigned long long *__restrict__ visited,
                                                      const float *__restrict__ bp_g, const float *__restrict__ paa_g,
                                                      unsigned long long *__restrict__ skipped,
                                                      unsigned int *__restrict__ dyn)
{
    const int t = ids[blockIdx.y];
    const int m = tmpl[t].m;
    if (static_cast<uint64_t>(m) > w.n)
        return;
    const int lb_seg = tmpl[t].lb_seg;
    // Lower bound: the breakpoints and the template's PAA; up to 16 segments
    // also every (segment, symbol) gap of the template's PAA value from the
    // symbol's region, so a window's bound is 16 table reads.
    __shared__ float bp[256];
    __shared__ float paa[kMaxSegments];
    __shared__ float gap[kTableSegments * 256];
    if (lb_seg > 0) {
        for (int i = threadIdx.x; i < 255; i += kThreads)
            bp[i] = bp_g[i];
        if (threadIdx.x < lb_seg)
            paa[threadIdx.x] = paa_g[tmpl[t].lb_paa + threadIdx.x];
        __syncthreads();
        if (lb_seg <= kTableSegments) {
            for (int e = threadIdx.x; e < lb_seg * 256; e += kThreads)
                gap[e] = sax_gap(bp, paa[e >> 8], e & 255);
            __syncthreads();
        }
    }
    const uint64_t stride = mode == kModeSample ? 0 : static_cast<uint64_t>(gridDim.x) * kThreads;
    // DYN: a separate instance, so that launches without k = 1 templates
    // keep the plain loop (the running bound's code cost 15% there).
    const bool dynamic = DYN && mode != kModeSample && tmpl[t].k == 1;
    unsigned long long nskip = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;;) {
        const float bnd = dynamic ? fminf(tmpl[t].bound, __uint_as_float(__hip_atomic_load(
                                                               dyn + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
                                  : tmpl[t].bound;
        float hi;
        nskip += scan_window(w, tmpl, T, t, m, qz, order, mode, sample_n, co, visited, bp, paa, gap, lb_seg, i, bnd,
                             &hi)
                     ? 1
                     : 0;
        if (dynamic && hi < bnd)
            atomicMin(dyn + t, __float_as_uint(hi));
        if (mode == kModeSample)
            break;
        i += stride;
        if (i >= w.n - static_cast<uint64_t>(m) + 1)
            break;
    }
    // Windows the bound skipped: one atomic per wavefront.
    if (skipped && lb_seg > 0) {
        for (int off = 32; off > 0; off >>= 1)
            nskip += __shfl_xor(nskip, off);
        if ((threadIdx.x & 63) == 0 && nskip)
            atomicAdd(skipped, nskip);
    }
}

// The running bounds start at the templates' bounds.
__global__ void k_ea_bound_init(const TmplDev *__restrict__ tmpl, const int32_t *__restrict__ ids, int count,
                                unsigned int *__restrict__ dyn)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count)
        dyn[ids[i]] = __float_as_uint(tmpl[ids[i]].bound);
}

__global__ void k_mark_ea(TmplDev *__restrict__ tmpl, const int32_t *__restrict__ ids, int count, int form)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < count)
        tmpl[ids[i]].ea = form;
}

}  // namespace

// Main pass (mode kModeMain) over the templates ids[0..count) (all windows),
// or a sample pass (sample_n windows each) accumulating the blocks visited.
// dyn (main pass, may be null): one word per template of the batch, the
// running bound of k = 1 templates (pass null when none is among ids).
hipError_t launch_scan_ea(const Waveform &w, const TmplDev *tmpl, int T, const float *qz, const int32_t *order,
                          const int32_t *ids, int count, uint64_t max_windows, int sample, uint64_t sample_n,
                          const CandOut &co, unsigned long long *visited, const float *bp, const float *paa,
                          unsigned long long *skipped, unsigned int *dyn, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    const uint64_t win = sample ? sample_n : max_windows;
    uint64_t gx = (win + kThreads - 1) / kThreads;
    if (!sample && gx > kBlocksPerTemplate)
        gx = kBlocksPerTemplate;  // the windows in a stride loop
    if (gx == 0)
        return hipSuccess;
    if (gx > 0x7fffffffull || count > 65535)
        return hipErrorInvalidValue;
    if (sample)
        dyn = nullptr;
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(count));
    const int mode = sample ? kModeSample : static_cast<int>(kModeMain);
    if (dyn) {
        hipLaunchKernelGGL(k_ea_bound_init, dim3((count + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, tmpl,
                           ids, count, dyn);
        hipLaunchKernelGGL(k_scan_ea<true>, grid, dim3(kThreads), 0, stream, w, tmpl, T, qz, order, ids, mode,
                           sample_n, co, visited, bp, paa, skipped, dyn);
    } else {
        hipLaunchKernelGGL(k_scan_ea<false>, grid, dim3(kThreads), 0, stream, w, tmpl, T, qz, order, ids, mode,
                           sample_n, co, visited, bp, paa, skipped, dyn);
    }
    return hipGetLastError();
}

// form: 1 the early-abandoning scan, 2 the iSAX index.
hipError_t launch_mark_ea(TmplDev *tmpl, const int32_t *ids, int count, int form, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_mark_ea, dim3((count + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, tmpl, ids,
                       count, form);
    return hipGetLastError();
}

}  // namespace tmx
