// Small utility kernels: counter reset (a kernel rather than a memset node,
// so the calls stay capturable in HIP graphs) and the maximum correlation
// error of the fast inner product over the verified candidates.
#include "tmx_internal.hpp"

namespace tmx {
namespace {

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// |rho_fast - rho_exact| = |d_fast - d_exact| / (2m); non-negative floats
// order like their bit patterns, so atomicMax on the bits is a float max.
__global__ void k_rho_err(const Cand *__restrict__ cand, const float *__restrict__ est,
                          const float *__restrict__ exact, uint64_t count, const TmplDev *__restrict__ tmpl,
                          unsigned int *__restrict__ out_bits)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (i < count) {
        const float a = est[i], b = exact[i];
        if (a == a && b < FLT_MAX && a < 3.0e38f) {
            const float m2 = 2.0f * static_cast<float>(tmpl[cand[i].t].m);
            e = fabsf(a - b) / m2;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        e = fmaxf(e, __shfl_xor(e, off));
    if ((threadIdx.x & 63) == 0 && e > 0.0f)
        atomicMax(out_bits, __float_as_uint(e));
}

__global__ void k_zero_u32(unsigned int *p) { *p = 0u; }

}  // namespace

hipError_t launch_zero_u64(unsigned long long *p, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_rho_err(const Cand *cand, const float *est, const float *exact, uint64_t count,
                          const TmplDev *tmpl, unsigned int *out_bits, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, out_bits);
    if (count == 0)
        return hipGetLastError();
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(k_rho_err, dim3(grid), dim3(256), 0, stream, cand, est, exact, count, tmpl, out_bits);
    return hipGetLastError();
}

}  // namespace tmx
