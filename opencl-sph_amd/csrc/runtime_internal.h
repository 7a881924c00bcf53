// Declarations shared by the device-side host code of libsph_hip.so: the
// per-psdata runtime (runtime.hip) and the scene batch (batch.hip).  Not
// part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "host_common.h"
#include "kernels.h"

namespace sphhip {

// Records a formatted error message for sph_hip_last_error (thread-local),
// prints it at log level 2 and returns -1.
int fail(const char *fmt, ...);

#define SPH_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t err_ = (expr);                                                               \
        if (err_ != hipSuccess) return ::sphhip::fail("%s: %s", #expr, hipGetErrorString(err_)); \
    } while (0)

// Kernel files named in a create_psdata call (kernels/<name>.cl in the
// reference); they select which kernels a program has.
struct KernelSet {
    bool common = false, fluids = false, solids = false, corot = false, svk = false,
         constraints = false, rigidbody = false;
};
KernelSet parse_file_list(const char *list);

// Device selected by init_hip (initialising on first use).
int current_device(int *device);

// Step parameters of a psdata whose field values are `blob` (the host blob,
// or a copy of it with the same layout): scalars, grid, planes, constraints,
// equation-of-state options.  n is the number of active particles.  p0,
// flags, grid.c0 and body_base are left 0; compensated and polar_fp64 are
// copied from the arguments.
int params_from_blob(const psdata &meta, const char *blob, const KernelSet &k, unsigned n, bool has_grid,
                     int compensated, int polar_fp64, sphk::StepParams *out);

// Device address of the allocation-free entry points' scratch array on the
// current device (looked up once per device); *out may be NULL.
int noalloc_bind_current_device(double **out);

}  // namespace sphhip
