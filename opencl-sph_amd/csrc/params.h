// Where a per-particle pass finds its parameters: one scene given by value
// (a psdata handle), or a table of scenes indexed by each slot's scene (a
// batch).  Kernels are written once against this interface, so a scene's
// results are the same in a batch as on its own.
#pragma once

#include "kernels.h"

namespace sphk {

struct SingleScene {
    StepParams p;
    __device__ __forceinline__ bool active(unsigned i) const { return i < p.n; }
    __device__ __forceinline__ const StepParams &at(unsigned) const { return p; }
    __host__ __device__ __forceinline__ unsigned slots() const { return p.n; }
};

struct SceneTable {
    BatchTable t;
    SceneFilter sel;
    __device__ __forceinline__ bool active(unsigned i) const {
        if (i >= t.capacity) return false;
        const unsigned s = t.scene_of[i];
        if (s == kNoScene) return false;
        const StepParams &p = t.scenes[s];
        return i < p.p0 + p.n && t.steps_left[s] > 0u && ((sel.system_mask >> p.system) & 1u) &&
               (p.flags & sel.flag_mask) == sel.flag_mask;
    }
    __device__ __forceinline__ const StepParams &at(unsigned i) const { return t.scenes[t.scene_of[i]]; }
    __host__ __device__ __forceinline__ unsigned slots() const { return t.capacity; }
};

}  // namespace sphk
