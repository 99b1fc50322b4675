// Allocation-free variant of the exact distance batch (tmx_distance_batch_noalloc).
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The caller
// passes the raw template, the window starts and the output in device memory
// and a stream; the prepared template (its record, z-normalised samples and
// block order) lives in `scratch`, a fixed array of this file that every
// device holds from the library's first context on that device until the
// process ends.  Nothing here allocates, frees or synchronises.
//
// The array's name is a host-side stand-in in host code, not a device
// address: host code uses only the address hipGetSymbolAddress returns for
// the current device, kept per device by noalloc_setup.
//
// Every call on one device shares that device's scratch and leaves its
// contents behind; the record is rewritten by each call before use.  Calls
// whose work overlaps on one device (different streams, no ordering between
// them) race on it.
#include <mutex>

#include "tmx_internal.hpp"

namespace tmx {

constexpr int kNoallocMaxTemplate = 65536;                 // longest template the scratch holds
constexpr size_t kNoallocRecBytes = 256;                   // the template record (TmplDev), padded
constexpr size_t kNoallocQzBytes = sizeof(float) * kNoallocMaxTemplate;
constexpr size_t kNoallocOrderBytes = sizeof(int32_t) * (kNoallocMaxTemplate / kBlock);
constexpr size_t kNoallocBytes = kNoallocRecBytes + kNoallocQzBytes + kNoallocOrderBytes;
static_assert(sizeof(TmplDev) <= kNoallocRecBytes, "template record must fit its scratch slot");

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// Writes the one-template record the preparation and distance kernels read:
// raw samples, z-normalised samples and block order all at offset 0 of their
// arrays.
__global__ void k_noalloc_record(TmplDev *rec, int m)
{
    if (threadIdx.x != 0)
        return;
    TmplDev t;
    t.q_off = 0;
    t.order_off = 0;
    t.raw_off = 0;
    t.m = m;
    t.kpad = 0;
    t.a_off = 0;
    t.Q2 = 0.0;
    t.S = 0.0;
    t.bound = 0.0f;
    t.k = 0;
    t.merge = 0;
    t.ea = 0;
    t.out_off = 0;
    *rec = t;
}

}  // namespace

int noalloc_max_template() { return kNoallocMaxTemplate; }

// Once per device, with that device current: the device address of this
// device's copy of `scratch`.
hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device])
        return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    if (!ptr)
        return hipErrorInvalidSymbol;
    g_scratch[device] = ptr;
    return hipSuccess;
}

// Enqueues record, preparation and distances on `stream`.  The caller has
// checked the arguments (m <= kNoallocMaxTemplate, device pointers non-null).
hipError_t noalloc_distance_batch(int device, const float *x, uint64_t n, const float *d_tmpl, int m,
                                  const uint64_t *d_starts, int count, float *d_out, hipStream_t stream)
{
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (device >= 0 && device < kMaxDevices)
            ptr = g_scratch[device];
    }
    if (!ptr)
        return hipErrorNotInitialized;
    unsigned char *base = static_cast<unsigned char *>(ptr);
    TmplDev *rec = reinterpret_cast<TmplDev *>(base);
    float *qz = reinterpret_cast<float *>(base + kNoallocRecBytes);
    int32_t *order = reinterpret_cast<int32_t *>(base + kNoallocRecBytes + kNoallocQzBytes);
    hipLaunchKernelGGL(k_noalloc_record, dim3(1), dim3(64), 0, stream, rec, m);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = launch_prepare(d_tmpl, rec, 1, qz, order, stream);
    if (e == hipSuccess)
        e = launch_dist_batch(x, n, rec, qz, order, d_starts, count, d_out, stream);
    return e;
}

}  // namespace tmx
