// Uniform-grid neighbour search: cell keys, per-cell counts, exclusive scan,
// stable placement (ascending particle id inside each cell) and the
// cell-sorted reorder of particle state.
//
// Reference: kernels/common.cl:372-529 (find_particle_bins, zero_gridcount,
// count_particles_in_bins, prefix_sum, copy_celloffset_to_backup,
// insert_particles_in_bin_array) and the host sequence of
// src/opencl/particle_system_host.c:373-454.  The reference leaves the order
// inside a cell to atomic arrival; here it is ascending particle index, which
// is what a sequential run of the reference produces, so gridcell, gridcount,
// celloffset and cellparticles are deterministic and bit-exact against the
// sequential oracle.
#include <algorithm>

#include "cells.h"

#pragma clang fp contract(off)

namespace sphk {
namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kDevGrid = 2048;   // blocks of the launches whose count is read from the device

__global__ void k_fill_u32(unsigned *p, size_t count, unsigned value) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// The count read from device memory (launches whose size must not depend on it).
__global__ void k_fill_u32_dev(unsigned *p, const unsigned *count, unsigned value) {
    const unsigned n = *count;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

__global__ void k_iota_u32(unsigned *p, size_t count) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < count; i += stride) p[i] = static_cast<unsigned>(i);
}

__global__ void k_cell_keys(const float *__restrict__ pos, unsigned n, GridGeom g,
                            unsigned *__restrict__ key) {
    unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    key[s] = cell_key(pos, s, g);
}

// Key and arrival rank of every active slot.
__global__ void k_key_count(const float *__restrict__ pos, unsigned n, GridGeom g,
                            unsigned *__restrict__ key, unsigned *__restrict__ local,
                            unsigned *__restrict__ count) {
    unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < n;
    const unsigned k = active ? cell_key(pos, s, g) : 0xffffffffu;
    const unsigned rank = claim_ranks(k, active, count);
    if (!active) return;
    key[s] = k;
    local[s] = rank;
}

// ---- exclusive scan over cells -------------------------------------------
constexpr unsigned kScanItems = kScanTile / kBlock;   // 16 cells per thread

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned *total) {
    __shared__ unsigned wave_sums[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    unsigned wave_off = 0, all = 0;
    for (int w = 0; w < static_cast<int>(kBlock / 64); ++w) {
        unsigned ws = wave_sums[w];
        if (w < wave) wave_off += ws;
        all += ws;
    }
    __syncthreads();
    *total = all;
    return wave_off + inc - v;
}

// The scan runs in tiles of kScanTile cells: per-tile sums, the exclusive
// scan of the tile sums, then each tile's own scan.  VEC: the arrays are
// 16-byte aligned and full tiles move as 16-byte words.
template <bool VEC>
__device__ __forceinline__ void scan_reduce_tile(const unsigned *__restrict__ in, unsigned count, unsigned tile,
                                                 unsigned *__restrict__ tile_sums) {
    const unsigned base = tile * kScanTile;
    unsigned sum = 0;
    if (VEC && base + kScanTile <= count) {
        const uint4 *v = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (unsigned k = 0; k < kScanTile / 4 / kBlock; ++k) {
            const uint4 q = v[k * kBlock + threadIdx.x];
            sum += (q.x + q.y) + (q.z + q.w);
        }
    } else {
        for (unsigned t = threadIdx.x; t < kScanTile; t += kBlock) {
            unsigned i = base + t;
            if (i < count) sum += in[i];
        }
    }
    unsigned total;
    block_exclusive_scan(sum, &total);
    if (threadIdx.x == 0) tile_sums[tile] = total;
}

template <bool VEC>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const unsigned *__restrict__ in, unsigned count,
                                                        unsigned *__restrict__ tile_sums) {
    scan_reduce_tile<VEC>(in, count, blockIdx.x, tile_sums);
}

__device__ __forceinline__ void scan_tile_sums(unsigned *tile_sums, unsigned ntiles) {
    unsigned carry = 0;
    for (unsigned base = 0; base < ntiles; base += kBlock) {
        unsigned i = base + threadIdx.x;
        unsigned v = i < ntiles ? tile_sums[i] : 0u;
        unsigned total;
        unsigned ex = block_exclusive_scan(v, &total);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kBlock) k_scan_tile_sums(unsigned *tile_sums, unsigned ntiles) {
    scan_tile_sums(tile_sums, ntiles);
}

// Tiles up to which each tile of the final pass sums the tile sums before it
// itself (at most kPrefixTiles values per block) instead of a separate
// one-block scan of the tile sums.
constexpr unsigned kPrefixTiles = 4096;

// PREFIX: tile_sums holds the raw tile sums and each block adds up those of
// the tiles before it; otherwise tile_sums holds their exclusive scan.
template <bool VEC, bool PREFIX>
__device__ __forceinline__ void scan_final_tile(const unsigned *__restrict__ in, unsigned count, unsigned tile,
                                                const unsigned *__restrict__ tile_sums, unsigned *__restrict__ out) {
    unsigned tile_off;
    if (PREFIX) {
        unsigned pre = 0;
        for (unsigned k = threadIdx.x; k < tile; k += kBlock) pre += tile_sums[k];
        block_exclusive_scan(pre, &tile_off);
    } else {
        tile_off = tile_sums[tile];
    }
    const unsigned base = tile * kScanTile + threadIdx.x * kScanItems;
    const bool full = VEC && base + kScanItems <= count;
    unsigned v[kScanItems];
    if (full) {
        const uint4 *src = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (unsigned k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = src[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (unsigned k = 0; k < kScanItems; ++k) v[k] = base + k < count ? in[base + k] : 0u;
    }
    unsigned sum = 0;
#pragma unroll
    for (unsigned k = 0; k < kScanItems; ++k) sum += v[k];
    unsigned total;
    unsigned run = tile_off + block_exclusive_scan(sum, &total);
    unsigned o[kScanItems];
#pragma unroll
    for (unsigned k = 0; k < kScanItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (full) {
        uint4 *dst = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (unsigned k = 0; k < kScanItems / 4; ++k) dst[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (unsigned k = 0; k < kScanItems; ++k)
            if (base + k < count) out[base + k] = o[k];
    }
}

template <bool VEC, bool PREFIX>
__global__ void __launch_bounds__(kBlock) k_scan_final(const unsigned *__restrict__ in, unsigned count,
                                                       const unsigned *__restrict__ tile_sums,
                                                       unsigned *__restrict__ out) {
    scan_final_tile<VEC, PREFIX>(in, count, blockIdx.x, tile_sums, out);
}

// Persistent forms: the count read from device memory, the blocks of a fixed
// grid looping over the tiles (uniformly per block: the tile functions
// synchronise the block).
template <bool VEC>
__global__ void __launch_bounds__(kBlock) k_scan_reduce_dev(const unsigned *__restrict__ in,
                                                            const unsigned *__restrict__ count_p,
                                                            unsigned *__restrict__ tile_sums) {
    const unsigned count = *count_p;
    const unsigned ntiles = (count + kScanTile - 1) / kScanTile;
    for (unsigned tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
        scan_reduce_tile<VEC>(in, count, tile, tile_sums);
}

__global__ void __launch_bounds__(kBlock) k_scan_tile_sums_dev(unsigned *tile_sums, const unsigned *count_p) {
    scan_tile_sums(tile_sums, (*count_p + kScanTile - 1) / kScanTile);
}

template <bool VEC>
__global__ void __launch_bounds__(kBlock) k_scan_final_dev(const unsigned *__restrict__ in,
                                                           const unsigned *__restrict__ count_p,
                                                           const unsigned *__restrict__ tile_sums,
                                                           unsigned *__restrict__ out) {
    const unsigned count = *count_p;
    const unsigned ntiles = (count + kScanTile - 1) / kScanTile;
    for (unsigned tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
        scan_final_tile<VEC, false>(in, count, tile, tile_sums, out);
}

__global__ void k_scatter_unsorted(unsigned n, const unsigned *__restrict__ key,
                                   const unsigned *__restrict__ local,
                                   const unsigned *__restrict__ celloffset,
                                   const unsigned *__restrict__ slot_id, unsigned *__restrict__ tmp_id,
                                   unsigned *__restrict__ tmp_src, unsigned *__restrict__ tmp_key) {
    unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    unsigned k = key[s];
    unsigned t = celloffset[k] + local[s];
    tmp_id[t] = slot_id[s];
    tmp_src[t] = s;
    tmp_key[t] = k;
}

// Moves the position (v0) and, with VEL, the velocity (v1) with the particle.
template <bool VEL>
__global__ void k_rank_place(unsigned n, const unsigned *__restrict__ tmp_id,
                             const unsigned *__restrict__ tmp_src, const unsigned *__restrict__ tmp_key,
                             const unsigned *__restrict__ celloffset,
                             const unsigned *__restrict__ gridcount, unsigned *__restrict__ perm,
                             unsigned *__restrict__ slot_id_new, unsigned *__restrict__ gridcell,
                             unsigned *__restrict__ cellparticles, const float *__restrict__ v0s,
                             float *__restrict__ v0d, const float *__restrict__ v1s,
                             float *__restrict__ v1d) {
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned c = tmp_key[t];
    const unsigned me = tmp_id[t];
    const unsigned src = tmp_src[t];
    // the moved vectors are fetched first, so that their latency overlaps
    // the rank count
    const Words<3> w0 = load_words<3>(v0s, src);
    Words<3> w1{};
    if (VEL) w1 = load_words<3>(v1s, src);
    const unsigned start = celloffset[c], cnt = gridcount[c];
    const unsigned f = start + rank_in_cell(tmp_id, start, cnt, me);
    perm[f] = src;
    slot_id_new[f] = me;
    gridcell[f] = c;
    cellparticles[f] = me;
    store_words<3>(v0d, f, w0);
    if (VEL) store_words<3>(v1d, f, w1);
}

template <unsigned W>
__global__ void k_gather_words(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                               const unsigned *__restrict__ map, unsigned n, unsigned count) {
    unsigned f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= count) return;
    unsigned s = f < n ? map[f] : f;
#pragma unroll
    for (unsigned w = 0; w < W; ++w) dst[static_cast<size_t>(f) * W + w] = src[static_cast<size_t>(s) * W + w];
}

template <unsigned W>
__global__ void k_scatter_words(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                                const unsigned *__restrict__ map, unsigned n, unsigned count) {
    unsigned f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= count) return;
    unsigned d = f < n ? map[f] : f;
#pragma unroll
    for (unsigned w = 0; w < W; ++w) dst[static_cast<size_t>(d) * W + w] = src[static_cast<size_t>(f) * W + w];
}

__global__ void k_gather_words_any(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                                   const unsigned *__restrict__ map, unsigned n, unsigned count,
                                   unsigned words) {
    unsigned f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= count) return;
    unsigned s = f < n ? map[f] : f;
    for (unsigned w = 0; w < words; ++w)
        dst[static_cast<size_t>(f) * words + w] = src[static_cast<size_t>(s) * words + w];
}

__global__ void k_scatter_words_any(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                                    const unsigned *__restrict__ map, unsigned n, unsigned count,
                                    unsigned words) {
    unsigned f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= count) return;
    unsigned d = f < n ? map[f] : f;
    for (unsigned w = 0; w < words; ++w)
        dst[static_cast<size_t>(d) * words + w] = src[static_cast<size_t>(f) * words + w];
}

inline unsigned blocks_for(size_t count) {
    return static_cast<unsigned>((count + kBlock - 1) / kBlock);
}

}  // namespace

hipError_t launch_fill_u32(unsigned *p, size_t count, unsigned value, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned blocks = blocks_for(count);
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(kBlock), 0, s, p, count, value);
    return hipGetLastError();
}

hipError_t launch_fill_u32_dev(unsigned *p, const unsigned *count, unsigned value, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32_dev, dim3(kDevGrid), dim3(kBlock), 0, s, p, count, value);
    return hipGetLastError();
}

hipError_t launch_iota_u32(unsigned *p, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned blocks = blocks_for(count);
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(k_iota_u32, dim3(blocks), dim3(kBlock), 0, s, p, count);
    return hipGetLastError();
}

hipError_t launch_cell_keys(const float *position, unsigned n, const GridGeom &g, unsigned *key_out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(n)), dim3(kBlock), 0, s, position, n, g, key_out);
    return hipGetLastError();
}

hipError_t launch_key_count(const float *position, unsigned n, const GridGeom &g, BinScratch &b,
                            unsigned *gridcount, hipStream_t s) {
    hipError_t e = launch_fill_u32(gridcount, g.ncells, 0u, s);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_key_count, dim3(blocks_for(n)), dim3(kBlock), 0, s, position, n, g, b.key,
                       b.local, gridcount);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan(const unsigned *in, unsigned count, unsigned *out, unsigned *tile_sums,
                                 hipStream_t s) {
    if (count == 0) return hipSuccess;
    const unsigned ntiles = (count + kScanTile - 1) / kScanTile;
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    const dim3 grid(ntiles), block(kBlock);
    if (vec) hipLaunchKernelGGL(k_scan_reduce<true>, grid, block, 0, s, in, count, tile_sums);
    else hipLaunchKernelGGL(k_scan_reduce<false>, grid, block, 0, s, in, count, tile_sums);
    if (ntiles <= kPrefixTiles) {
        if (vec) hipLaunchKernelGGL((k_scan_final<true, true>), grid, block, 0, s, in, count, tile_sums, out);
        else hipLaunchKernelGGL((k_scan_final<false, true>), grid, block, 0, s, in, count, tile_sums, out);
    } else {
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), block, 0, s, tile_sums, ntiles);
        if (vec) hipLaunchKernelGGL((k_scan_final<true, false>), grid, block, 0, s, in, count, tile_sums, out);
        else hipLaunchKernelGGL((k_scan_final<false, false>), grid, block, 0, s, in, count, tile_sums, out);
    }
    return hipGetLastError();
}

hipError_t launch_exclusive_scan_dev(const unsigned *in, const unsigned *count, unsigned max_count, unsigned *out,
                                     unsigned *tile_sums, hipStream_t s) {
    const unsigned max_tiles = (max_count + kScanTile - 1) / kScanTile;
    const dim3 grid(std::max(1u, std::min(max_tiles, kDevGrid))), block(kBlock);
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    if (vec) hipLaunchKernelGGL(k_scan_reduce_dev<true>, grid, block, 0, s, in, count, tile_sums);
    else hipLaunchKernelGGL(k_scan_reduce_dev<false>, grid, block, 0, s, in, count, tile_sums);
    hipLaunchKernelGGL(k_scan_tile_sums_dev, dim3(1), block, 0, s, tile_sums, count);
    if (vec) hipLaunchKernelGGL(k_scan_final_dev<true>, grid, block, 0, s, in, count, tile_sums, out);
    else hipLaunchKernelGGL(k_scan_final_dev<false>, grid, block, 0, s, in, count, tile_sums, out);
    return hipGetLastError();
}

hipError_t launch_scatter_unsorted(unsigned n, const BinScratch &b, const unsigned *celloffset,
                                   const unsigned *slot_id, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_unsorted, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, b.key, b.local,
                       celloffset, slot_id, b.tmp_id, b.tmp_src, b.tmp_key);
    return hipGetLastError();
}

hipError_t launch_rank_place(unsigned n, const BinScratch &b, const unsigned *celloffset,
                             const unsigned *gridcount, unsigned *perm, unsigned *slot_id_new,
                             unsigned *gridcell, unsigned *cellparticles, const float *vec_src0,
                             float *vec_dst0, const float *vec_src1, float *vec_dst1, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!vec_src0 || !vec_dst0) return hipErrorInvalidValue;
    if (vec_src1 && vec_dst1)
        hipLaunchKernelGGL(k_rank_place<true>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, b.tmp_id, b.tmp_src,
                           b.tmp_key, celloffset, gridcount, perm, slot_id_new, gridcell, cellparticles,
                           vec_src0, vec_dst0, vec_src1, vec_dst1);
    else
        hipLaunchKernelGGL(k_rank_place<false>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, b.tmp_id, b.tmp_src,
                           b.tmp_key, celloffset, gridcount, perm, slot_id_new, gridcell, cellparticles,
                           vec_src0, vec_dst0, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_gather_words(const void *src, void *dst, const unsigned *map, unsigned n,
                               unsigned count, unsigned words, hipStream_t s) {
    if (count == 0 || words == 0) return hipSuccess;
    const unsigned *a = static_cast<const unsigned *>(src);
    unsigned *b = static_cast<unsigned *>(dst);
    dim3 grid(blocks_for(count)), block(kBlock);
    switch (words) {
        case 1: hipLaunchKernelGGL(k_gather_words<1>, grid, block, 0, s, a, b, map, n, count); break;
        case 3: hipLaunchKernelGGL(k_gather_words<3>, grid, block, 0, s, a, b, map, n, count); break;
        case 6: hipLaunchKernelGGL(k_gather_words<6>, grid, block, 0, s, a, b, map, n, count); break;
        case 9: hipLaunchKernelGGL(k_gather_words<9>, grid, block, 0, s, a, b, map, n, count); break;
        default:
            hipLaunchKernelGGL(k_gather_words_any, grid, block, 0, s, a, b, map, n, count, words);
    }
    return hipGetLastError();
}

hipError_t launch_scatter_words(const void *src, void *dst, const unsigned *map, unsigned n,
                                unsigned count, unsigned words, hipStream_t s) {
    if (count == 0 || words == 0) return hipSuccess;
    const unsigned *a = static_cast<const unsigned *>(src);
    unsigned *b = static_cast<unsigned *>(dst);
    dim3 grid(blocks_for(count)), block(kBlock);
    switch (words) {
        case 1: hipLaunchKernelGGL(k_scatter_words<1>, grid, block, 0, s, a, b, map, n, count); break;
        case 3: hipLaunchKernelGGL(k_scatter_words<3>, grid, block, 0, s, a, b, map, n, count); break;
        case 6: hipLaunchKernelGGL(k_scatter_words<6>, grid, block, 0, s, a, b, map, n, count); break;
        case 9: hipLaunchKernelGGL(k_scatter_words<9>, grid, block, 0, s, a, b, map, n, count); break;
        default:
            hipLaunchKernelGGL(k_scatter_words_any, grid, block, 0, s, a, b, map, n, count, words);
    }
    return hipGetLastError();
}

}  // namespace sphk
