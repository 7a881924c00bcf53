// Scripted rigid bodies and initial-condition helpers: the lattice
// initialiser and the whole-scene rotation (kernels/common.cl:533-604), body
// translation and rotation (kernels/rigidbody.cl:1-120).
//
// All four are elementwise over slots (one lane per particle or lattice
// point).  Uniform matrices (the rotation of rotate_particles, the inverse
// body matrix of set_body_rotation) are formed once on the host and passed by
// value instead of per lane; the per-lane arithmetic is the reference's,
// operation for operation (no contraction).
#include "kernels.h"

namespace sphk {
namespace {

constexpr unsigned kBodyBlock = 256;

unsigned blocks_for(unsigned long long count) {
    return static_cast<unsigned>((count + kBodyBlock - 1) / kBodyBlock);
}

__device__ __forceinline__ float3 mat_vec(const Mat3 &m, float3 v) {
    return make_float3(m.m[0] * v.x + m.m[1] * v.y + m.m[2] * v.z, m.m[3] * v.x + m.m[4] * v.y + m.m[5] * v.z,
                       m.m[6] * v.x + m.m[7] * v.y + m.m[8] * v.z);
}

__device__ __forceinline__ float3 ld3(const float *a, unsigned i) {
    return make_float3(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
}

__device__ __forceinline__ void st3(float *a, unsigned i, float3 v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}

// Lane p is particle p = i + sx j + sx sy k of the lattice; its coordinate
// along each axis is (index / (size - 1)) (corner2 - corner1) + corner1.
__global__ void __launch_bounds__(kBodyBlock)
    k_populate_cuboid(float *__restrict__ pos, unsigned count, unsigned sx, unsigned sy, unsigned sz, Box c) {
    const unsigned p = blockIdx.x * kBodyBlock + threadIdx.x;
    if (p >= count) return;
    const unsigned i = p % sx;
    const unsigned j = (p / sx) % sy;
    const unsigned k = static_cast<unsigned>(p / (static_cast<unsigned long long>(sx) * sy));
    float3 v;
    v.x = (static_cast<float>(i) / static_cast<float>(sx - 1)) * (c.hi[0] - c.lo[0]) + c.lo[0];
    v.y = (static_cast<float>(j) / static_cast<float>(sy - 1)) * (c.hi[1] - c.lo[1]) + c.lo[1];
    v.z = (static_cast<float>(k) / static_cast<float>(sz - 1)) * (c.hi[2] - c.lo[2]) + c.lo[2];
    st3(pos, p, v);
}

__global__ void __launch_bounds__(kBodyBlock) k_rotate(float *__restrict__ pos, unsigned n, Mat3 m) {
    const unsigned s = blockIdx.x * kBodyBlock + threadIdx.x;
    if (s >= n) return;
    st3(pos, s, mat_vec(m, ld3(pos, s)));
}

// Particles tagged `tag` (= body + 1) move by target - stored and take
// velocity (target - stored) / dt; every lane reads the stored body position,
// which a later launch overwrites.
__global__ void __launch_bounds__(kBodyBlock)
    k_set_body_position(float *__restrict__ pos, float *__restrict__ vel, const unsigned *__restrict__ body_number,
                        unsigned n, unsigned tag, const float *__restrict__ stored, float3 target, float dt) {
    const unsigned s = blockIdx.x * kBodyBlock + threadIdx.x;
    if (s >= n || body_number[s] != tag) return;
    const float3 old = make_float3(stored[0], stored[1], stored[2]);
    const float3 x = ld3(pos, s);
    const float3 rel = make_float3(x.x - old.x, x.y - old.y, x.z - old.z);
    st3(pos, s, make_float3(rel.x + target.x, rel.y + target.y, rel.z + target.z));
    st3(vel, s, make_float3((target.x - old.x) / dt, (target.y - old.y) / dt, (target.z - old.z) / dt));
}

// T = Bt P (Bt: transposed body matrix, P: the particle's rotation field);
// position <- T x, rotation <- T P.
__global__ void __launch_bounds__(kBodyBlock)
    k_set_body_rotation(float *__restrict__ pos, float *__restrict__ rot, const unsigned *__restrict__ body_number,
                        unsigned n, unsigned tag, Mat3 bt) {
    const unsigned s = blockIdx.x * kBodyBlock + threadIdx.x;
    if (s >= n || body_number[s] != tag) return;
    float *r = rot + 9 * static_cast<size_t>(s);
    Mat3 p, t;
#pragma unroll
    for (int q = 0; q < 9; ++q) p.m[q] = r[q];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            t.m[3 * a + b] = bt.m[3 * a] * p.m[b] + bt.m[3 * a + 1] * p.m[3 + b] + bt.m[3 * a + 2] * p.m[6 + b];
    st3(pos, s, mat_vec(t, ld3(pos, s)));
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            r[3 * a + b] = t.m[3 * a] * p.m[b] + t.m[3 * a + 1] * p.m[3 + b] + t.m[3 * a + 2] * p.m[6 + b];
}

__global__ void k_store3(float *__restrict__ dst, float3 v) {
    if (threadIdx.x == 0) {
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
    }
}

}  // namespace

hipError_t launch_populate_cuboid(float *pos, unsigned count, unsigned sx, unsigned sy, unsigned sz, const Box &c,
                                  hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_populate_cuboid, dim3(blocks_for(count)), dim3(kBodyBlock), 0, s, pos, count, sx, sy, sz,
                       c);
    return hipGetLastError();
}

hipError_t launch_rotate(float *pos, unsigned n, const Mat3 &m, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rotate, dim3(blocks_for(n)), dim3(kBodyBlock), 0, s, pos, n, m);
    return hipGetLastError();
}

hipError_t launch_set_body_position(float *pos, float *vel, const unsigned *body_number, unsigned n, unsigned body,
                                    bool move, float *body_position, const float *target, float dt,
                                    hipStream_t s) {
    const float3 t = make_float3(target[0], target[1], target[2]);
    float *stored = body_position + 3 * static_cast<size_t>(body);
    if (move && n) {
        hipLaunchKernelGGL(k_set_body_position, dim3(blocks_for(n)), dim3(kBodyBlock), 0, s, pos, vel, body_number,
                           n, body + 1, stored, t, dt);
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_store3, dim3(1), dim3(64), 0, s, stored, t);
    return hipGetLastError();
}

hipError_t launch_set_body_rotation(float *pos, float *rot, const unsigned *body_number, unsigned n, unsigned body,
                                    bool move, const Mat3 &bt, float *body_euler, const float *angles,
                                    hipStream_t s) {
    if (move && n) {
        hipLaunchKernelGGL(k_set_body_rotation, dim3(blocks_for(n)), dim3(kBodyBlock), 0, s, pos, rot, body_number,
                           n, body + 1, bt);
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_store3, dim3(1), dim3(64), 0, s, body_euler + 3 * static_cast<size_t>(body),
                       make_float3(angles[0], angles[1], angles[2]));
    return hipGetLastError();
}

}  // namespace sphk
