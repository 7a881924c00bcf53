// Cell keys and per-cell arrival ranks, shared by the single-scene binning
// (kernels_grid.hip) and the batch binning (batch.hip).
#pragma once

#include "kernels.h"

#pragma clang fp contract(off)

namespace sphk {

__device__ __forceinline__ unsigned grid_index(float p, float lo, float hi, unsigned res) {
    // (p - lo) / (hi - lo) * res, truncated toward zero, clamped to [0, res-1]
    // (kernels/common.cl:384-388).  Values below 0 and NaN land in cell 0.
    float v = (p - lo) / (hi - lo) * static_cast<float>(res);
    unsigned u;
    if (!(v > 0.0f)) u = 0u;
    else if (v >= 4294967296.0f) u = 0xffffffffu;
    else u = static_cast<unsigned>(v);
    return u < res - 1u ? u : res - 1u;
}

// Key of the scene-local cell of slot s (kernels/common.cl:390).
__device__ __forceinline__ unsigned cell_key(const float *pos, unsigned s, const GridGeom &g) {
    unsigned gx = grid_index(pos[3 * s], g.lo[0], g.hi[0], g.rx);
    unsigned gy = grid_index(pos[3 * s + 1], g.lo[1], g.hi[1], g.ry);
    unsigned gz = grid_index(pos[3 * s + 2], g.lo[2], g.hi[2], g.rz);
    return gz * g.rx * g.ry + gy * g.rx + gx;
}

// Counts key k of every active lane into count[k] and returns the lane's
// arrival rank inside its cell.  Slots are cell-sorted from the previous
// step, so runs of equal keys inside a wavefront are common; each run of
// consecutive active lanes with one key issues one atomic and hands out
// consecutive ranks.  Every lane of the wavefront must call it.
__device__ __forceinline__ unsigned claim_ranks(unsigned k, bool active, unsigned *count) {
    const int lane = threadIdx.x & 63;
    const unsigned prev = __shfl_up(k, 1, 64);
    const bool head = active && (lane == 0 || prev != k);
    const unsigned long long heads = __ballot(head);
    const unsigned long long act = __ballot(active);
    if (!active) return 0u;
    // Run start = highest head lane <= lane; run end = the next lane that
    // starts a run or is inactive (inactive lanes may sit anywhere in the
    // wavefront, e.g. between two scenes of a batch).
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll(heads & upto);
    const unsigned long long breaks = (heads | ~act) & ~upto;
    const int end = breaks ? __ffsll(static_cast<long long>(breaks)) - 1 : 64;
    unsigned base = 0;
    if (lane == start) base = atomicAdd(&count[k], static_cast<unsigned>(end - start));
    base = __shfl(base, start, 64);
    return base + static_cast<unsigned>(lane - start);
}

}  // namespace sphk
