// Where a per-particle pass finds its parameters: one scene given by value
// (a psdata handle), or a table of scenes indexed by each slot's scene (a
// batch).  Kernels are written once against this interface, so a scene's
// results are the same in a batch as on its own.
#pragma once

#include "kernels.h"

namespace sphk {

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each XCD with its own L2.  Neighbour passes launch a multiple of 8
// blocks and turn the hardware block id into a logical one so that XCD x
// sweeps the contiguous eighth [x * per, (x + 1) * per) of the work, and the
// neighbour rows its blocks share stay in its own L2.
constexpr unsigned kXcds = 8;

__host__ __device__ __forceinline__ unsigned xcd_grid(unsigned blocks) {
    return (blocks + kXcds - 1) / kXcds * kXcds;
}

__device__ __forceinline__ unsigned xcd_block() {
    const unsigned per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// slot_xcd / slot_linear: the slot of this thread in a launch of
// `block_threads`-thread blocks sized by slots(), with the XCD-aware or the
// plain block order; inactive threads get a slot that active() rejects.
struct SingleScene {
    StepParams p;
    __device__ __forceinline__ bool active(unsigned i) const { return i < p.n; }
    __device__ __forceinline__ const StepParams &at(unsigned) const { return p; }
    __host__ __device__ __forceinline__ unsigned slots() const { return p.n; }
    __device__ __forceinline__ unsigned slot_xcd(unsigned block_threads) const {
        return xcd_block() * block_threads + threadIdx.x;
    }
    __device__ __forceinline__ unsigned slot_linear(unsigned block_threads) const {
        return blockIdx.x * block_threads + threadIdx.x;
    }
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
    __host__ __device__ __forceinline__ unsigned slots() const {
        return t.chunk_slot ? t.nchunks * kChunk : t.capacity;
    }
    __device__ __forceinline__ unsigned from_block(unsigned block, unsigned block_threads) const {
        if (!t.chunk_slot) return block * block_threads + threadIdx.x;
        const unsigned k = block * (block_threads / kChunk) + threadIdx.x / kChunk;
        if (k >= t.nchunks) return kNoScene;
        const unsigned i = t.chunk_slot[k] + threadIdx.x % kChunk;
        return i < t.chunk_end[k] ? i : kNoScene;
    }
    __device__ __forceinline__ unsigned slot_xcd(unsigned block_threads) const {
        return from_block(xcd_block(), block_threads);
    }
    __device__ __forceinline__ unsigned slot_linear(unsigned block_threads) const {
        return from_block(blockIdx.x, block_threads);
    }
};

}  // namespace sphk
