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

// Rank of id `me` among the ids of its cell, ids[start, start + cnt): the
// number of smaller ids, which is me's offset in the stable (ascending id)
// order of the cell.  The ids are loaded 8 at a time, unconditionally (slots
// past the cell's end re-read its last member and are not counted), so a cell
// costs one memory round trip per 8 members rather than one per member.  The
// cell holds me itself, so cnt >= 1.
__device__ __forceinline__ unsigned rank_in_cell(const unsigned *__restrict__ ids, unsigned start, unsigned cnt,
                                                 unsigned me) {
    constexpr unsigned kBatch = 8;
    const unsigned last = start + cnt - 1u;
    unsigned rank = 0;
    for (unsigned k0 = 0; k0 < cnt; k0 += kBatch) {
        unsigned v[kBatch];
#pragma unroll
        for (unsigned k = 0; k < kBatch; ++k) v[k] = ids[min(start + k0 + k, last)];
#pragma unroll
        for (unsigned k = 0; k < kBatch; ++k) rank += (k0 + k < cnt && v[k] < me) ? 1u : 0u;
    }
    return rank;
}

// W consecutive 32-bit words with 4-byte alignment, moved as one wide load
// and one wide store (global_load_dwordx3 for a position, and so on).
template <unsigned W>
struct __attribute__((packed, aligned(4))) Words {
    unsigned w[W];
};
// Record `rec` of an array of W-word records.
template <unsigned W, typename T>
__device__ __forceinline__ Words<W> load_words(const T *__restrict__ p, unsigned rec) {
    static_assert(sizeof(T) == 4, "records of 32-bit words");
    return *reinterpret_cast<const Words<W> *>(p + static_cast<size_t>(rec) * W);
}
template <unsigned W, typename T>
__device__ __forceinline__ void store_words(T *__restrict__ p, unsigned rec, const Words<W> &v) {
    static_assert(sizeof(T) == 4, "records of 32-bit words");
    *reinterpret_cast<Words<W> *>(p + static_cast<size_t>(rec) * W) = v;
}
// dst record `to` = src record `from`.
template <unsigned W, typename T>
__device__ __forceinline__ void copy_words(const T *__restrict__ src, T *__restrict__ dst, unsigned from,
                                           unsigned to) {
    store_words<W>(dst, to, load_words<W>(src, from));
}

}  // namespace sphk
