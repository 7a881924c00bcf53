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

// slot_xcd / slot_linear: the slot of this thread, with the XCD-aware or the
// plain block order; threads without work get a slot that active() rejects.
// launch_blocks: the grid of a pass (a multiple of 8 blocks for the XCD
// order).  A single scene launches one thread per slot; a scene table covers
// grid_units slots or chunks, at least what its layout holds.
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
    __host__ bool has_work() const { return p.n > 0; }
    __host__ dim3 launch_blocks(unsigned block_threads, bool xcd) const {
        const unsigned b = (p.n + block_threads - 1) / block_threads;
        return dim3(xcd ? xcd_grid(b) : b);
    }
};

struct SceneTable {
    BatchTable t;
    SceneFilter sel;
    __device__ __forceinline__ unsigned capacity() const { return t.layout->used_slots; }
    __device__ __forceinline__ bool active(unsigned i) const {
        if (i >= capacity()) return false;
        const unsigned s = t.scene_of[i];
        if (s == kNoScene) return false;
        const StepParams &p = t.scenes[s];
        return i < p.p0 + p.n && t.steps_left[s] > 0u && ((sel.system_mask >> p.system) & 1u) &&
               (p.flags & sel.flag_mask) == sel.flag_mask;
    }
    __device__ __forceinline__ const StepParams &at(unsigned i) const { return t.scenes[t.scene_of[i]]; }
    // Chunk k of the map: slots [chunk_slot[k], min(chunk_slot[k] + 64, chunk_end[k])).
    __device__ __forceinline__ unsigned chunk_lane(unsigned k) const {
        if (k >= t.layout->nchunks[t.map]) return kNoScene;
        const unsigned i = t.chunk_slot[k] + threadIdx.x % kChunk;
        return i < t.chunk_end[k] ? i : kNoScene;
    }
    // XCD x takes the contiguous eighth [x * per, (x + 1) * per) of the
    // chunks the layout holds, whatever the grid's size: neighbouring chunks
    // share neighbour rows in one L2.  With a segment size that divides the
    // grid's L chunk positions per XCD, the map is instead cut into segments
    // of xcd_segment chunks dealt round-robin to the XCDs (segment q to XCD
    // q % 8), covering chunks [0, 8 L) once: each XCD then meets the same mix
    // of scenes, which balances passes whose cost per chunk varies with the
    // scene.
    __device__ __forceinline__ unsigned slot_xcd(unsigned block_threads) const {
        if (!t.chunk_slot) return slot_linear(block_threads);
        const unsigned cpb = block_threads / kChunk;
        const unsigned local = (blockIdx.x / kXcds) * cpb + threadIdx.x / kChunk;
        const unsigned x = blockIdx.x % kXcds;
        const unsigned L = (gridDim.x / kXcds) * cpb;
        const unsigned seg = t.xcd_segment;
        if (seg != 0u && L % seg == 0u) return chunk_lane(((local / seg) * kXcds + x) * seg + local % seg);
        const unsigned per = (t.layout->nchunks[t.map] + kXcds - 1u) / kXcds;
        return local < per ? chunk_lane(x * per + local) : kNoScene;
    }
    __device__ __forceinline__ unsigned slot_linear(unsigned block_threads) const {
        if (!t.chunk_slot) return blockIdx.x * block_threads + threadIdx.x;
        return chunk_lane(blockIdx.x * (block_threads / kChunk) + threadIdx.x / kChunk);
    }
    __host__ bool has_work() const { return t.grid_units > 0; }
    __host__ dim3 launch_blocks(unsigned block_threads, bool xcd) const {
        const unsigned per_block = t.chunk_slot ? block_threads / kChunk : block_threads;
        const unsigned b = (t.grid_units + per_block - 1) / per_block;
        return dim3(xcd ? xcd_grid(b) : b);
    }
};

}  // namespace sphk
