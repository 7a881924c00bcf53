// Self-checks of device arithmetic that the pair passes rely on.
#include "device_math.h"
#include "runtime_internal.h"

using sphhip::fail;

namespace {

// Every float bit pattern in [t * per, (t + 1) * per): recip(b) against the
// compiler's correctly rounded double division, bit for bit (NaN against NaN
// equal).  One counter per block, added once.
__global__ void __launch_bounds__(256) k_check_recip(unsigned per, unsigned long long *mismatches) {
    __shared__ unsigned block_bad;
    if (threadIdx.x == 0) block_bad = 0;
    __syncthreads();
    const unsigned long long t = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned bad = 0;
    for (unsigned k = 0; k < per; ++k) {
        const float b = __uint_as_float(static_cast<unsigned>(t * per + k));
        const double fast = sphk::recip(b), slow = sphk::recip_by_division(b);
        const bool same = __double_as_longlong(fast) == __double_as_longlong(slow) ||
                          (__builtin_isnan(fast) && __builtin_isnan(slow));
        bad += same ? 0u : 1u;
    }
    if (bad) atomicAdd(&block_bad, bad);
    __syncthreads();
    if (threadIdx.x == 0 && block_bad) atomicAdd(mismatches, static_cast<unsigned long long>(block_bad));
}

// The same sweep for sqrt_rn(s) against the compiler's sqrtf.
__global__ void __launch_bounds__(256) k_check_sqrt(unsigned per, unsigned long long *mismatches) {
    __shared__ unsigned block_bad;
    if (threadIdx.x == 0) block_bad = 0;
    __syncthreads();
    const unsigned long long t = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned bad = 0;
    for (unsigned k = 0; k < per; ++k) {
        const float s = __uint_as_float(static_cast<unsigned>(t * per + k));
        const float fast = sphk::sqrt_rn(s), slow = sqrtf(s);
        const bool same = __float_as_uint(fast) == __float_as_uint(slow) ||
                          (__builtin_isnan(fast) && __builtin_isnan(slow));
        bad += same ? 0u : 1u;
    }
    if (bad) atomicAdd(&block_bad, bad);
    __syncthreads();
    if (threadIdx.x == 0 && block_bad) atomicAdd(mismatches, static_cast<unsigned long long>(block_bad));
}

// Runs one sweep kernel over all 2^32 bit patterns: 4096 blocks of 256
// threads, 4096 patterns per thread, one device counter.
int sweep_all_floats(void (*kernel)(unsigned, unsigned long long *), const char *what,
                     unsigned long long *mismatches) {
    if (!mismatches) return fail("null argument");
    int dev = 0;
    if (sphhip::current_device(&dev)) return -1;
    SPH_HIP_TRY(hipSetDevice(dev));
    constexpr unsigned kPer = 4096, kThreads = 256;
    constexpr unsigned kBlocks = static_cast<unsigned>((1ull << 32) / kPer / kThreads);   // 4096
    unsigned long long *d = nullptr;
    SPH_HIP_TRY(hipMalloc(&d, sizeof *d));
    hipError_t e = hipMemset(d, 0, sizeof *d);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kernel, dim3(kBlocks), dim3(kThreads), 0, 0, kPer, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(mismatches, d, sizeof *d, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail("%s self-check failed: %s", what, hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" int sph_hip_check_recip(unsigned long long *mismatches) {
    return sweep_all_floats(k_check_recip, "recip", mismatches);
}

extern "C" int sph_hip_check_sqrt(unsigned long long *mismatches) {
    return sweep_all_floats(k_check_sqrt, "sqrt", mismatches);
}
