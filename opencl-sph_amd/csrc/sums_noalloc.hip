// Allocation-free variant of global_sums_hip for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).  Inputs and outputs are the caller's device
// arrays; the reduction's per-block partial sums live in `scratch`, a
// file-scope device array of fixed size that every GPU the library runs on
// holds until the process ends.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"
#include "runtime_internal.h"

using sphhip::fail;

namespace {

// Partial sums of the FP64 reduction: 7 sums x kReduceBlocks blocks (14 KiB).
// k_sums_partial writes every entry on every call, so nothing needs clearing.
constexpr unsigned kScratchDoubles = 7u * sphk::kReduceBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Largest particle count: k_sums_partial indexes position[3 * i] with a
// 32-bit i.
constexpr unsigned kMaxParticles = 0xffffffffu / 3u;

constexpr int kMaxDevices = 64;
std::mutex g_lock;
double *g_scratch[kMaxDevices] = {};   // device address of `scratch` on each GPU

}  // namespace

namespace sphhip {

// Looks up the device address of `scratch` on the current device, once per
// device.  Called by init_hip for the selected device and by the entry point
// on any other device at its first call there.
int noalloc_bind_current_device(double **out) {
    int dev = 0;
    SPH_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail("device %d: at most %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    if (out) *out = g_scratch[dev];
    return 0;
}

}  // namespace sphhip

extern "C" int global_sums_noalloc_hip(const float *position, const float *velocity, unsigned n, float mass,
                                       double *out7, void *stream) {
    if (!position || !out7) return fail("global_sums_noalloc_hip: position and out7 must be device pointers");
    if (n > kMaxParticles)
        return fail("global_sums_noalloc_hip: n = %u exceeds the maximum of %u particles", n, kMaxParticles);
    double *partials = nullptr;
    if (sphhip::noalloc_bind_current_device(&partials)) return -1;
    sphk::StepParams p{};
    p.n = n;
    p.mass = mass;
    sphk::FieldPtrs f{};
    f.position = const_cast<float *>(position);
    f.velocity = const_cast<float *>(velocity);
    SPH_HIP_TRY(sphk::launch_global_sums(p, f, partials, out7, static_cast<hipStream_t>(stream)));
    return 0;
}
