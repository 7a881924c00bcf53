// Neighbour iteration over the cell-sorted slots and the mirror-plane tests.
//
// Particle state lives in slot order: slots [celloffset[c], celloffset[c] +
// gridcount[c]) hold the particles of cell c in ascending particle index.
// The reference visits the 27 cells around a particle in (z, y, x) offset
// order, skipping cells outside the grid, and each cell's particles in
// cellparticles order (kernels/common.cl:49-68, 354-371).  The three x-cells
// of one (z, y) row are consecutive in slot space, so the same visiting order
// is 9 contiguous slot ranges.  Keys outside the grid visit no cell.
#pragma once

#include "device_math.h"
#include "kernels.h"
#include "params.h"

namespace sphk {

template <typename Body>
__device__ __forceinline__ void for_each_candidate(const GridGeom &g, unsigned lo, unsigned hi, unsigned key,
                                                   const unsigned *__restrict__ celloffset,
                                                   const unsigned *__restrict__ gridcount,
                                                   Body &&body) {
    if (key < g.c0 || key - g.c0 >= g.ncells) return;
    const unsigned local = key - g.c0;
    const unsigned zo = local / g.layer;
    const unsigned rem = local - g.layer * zo;
    const unsigned yo = rem / g.rx;
    const unsigned xo = rem - g.rx * yo;
    const unsigned x0 = xo > 0u ? xo - 1u : 0u;
    const unsigned x1 = xo + 1u < g.rx ? xo + 1u : xo;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
        const int zn = static_cast<int>(zo) + dz;
        if (zn < 0 || zn >= static_cast<int>(g.rz)) continue;
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const int yn = static_cast<int>(yo) + dy;
            if (yn < 0 || yn >= static_cast<int>(g.ry)) continue;
            const unsigned row = g.c0 + static_cast<unsigned>(zn) * g.layer + static_cast<unsigned>(yn) * g.rx;
            // Ranges are clamped to the scene's slots so that grid arrays
            // uploaded from the host can never index past the particle data.
            const unsigned b = min(max(celloffset[row + x0], lo), hi);
            const unsigned e = min(max(celloffset[row + x1] + gridcount[row + x1], lo), hi);
            for (unsigned j = b; j < e; ++j) body(j);
        }
    }
}

// Iterates slot i's pair-cache list in groups of G entries: the G entries
// are loaded, then load(entry) is issued for all G of them, then
// use(entry, data) runs on each in list order.  Issuing a group's loads
// together lets their memory latency overlap; the order in which use()
// sees the entries, and so every sum, is the list order.  Entries past the
// end read slot i's own data (load must accept it) and are not used.
// The next group's entries are loaded while the current group is used, so
// only the first group waits for its list entries.
template <unsigned G, typename Load, typename Use>
__device__ __forceinline__ void for_each_listed(const PairCache &c, unsigned i, unsigned cnt, Load &&load,
                                                Use &&use) {
    if (cnt == 0) return;
    const unsigned *__restrict__ col = c.list + i;
    const size_t stride = c.stride;
    unsigned e[G];
#pragma unroll
    for (unsigned g = 0; g < G; ++g) e[g] = g < cnt ? col[g * stride] : i;
    for (unsigned k0 = 0; k0 < cnt; k0 += G) {
        decltype(load(0u)) d[G];
#pragma unroll
        for (unsigned g = 0; g < G; ++g) d[g] = load(e[g]);
        unsigned nx[G];
#pragma unroll
        for (unsigned g = 0; g < G; ++g) nx[g] = k0 + G + g < cnt ? col[(k0 + G + g) * stride] : i;
#pragma unroll
        for (unsigned g = 0; g < G; ++g)
            if (k0 + g < cnt) use(e[g], d[g]);
#pragma unroll
        for (unsigned g = 0; g < G; ++g) e[g] = nx[g];
    }
}

// for_each_listed over a list of 16-bit entries (the fluid list: entries
// 2m and 2m + 1 of slot i in the low and high half of word m * stride + i),
// each turned into its slot by decode(entry).  G is even.  Words are read
// without a test (the column holds kListCap words; with cnt <= kFluidListMax
// the entries fetched end before cnt + 2 G - 1, inside it); entries at or
// past cnt become i by a select, so that no branch separates the loads.
template <unsigned G, typename Decode, typename Load, typename Use>
__device__ __forceinline__ void for_each_listed16(const unsigned *__restrict__ list, size_t stride, unsigned i,
                                                  unsigned cnt, Decode &&decode, Load &&load, Use &&use) {
    static_assert(G % 2 == 0, "entries are read in pairs");
    static_assert((kFluidListMax + 2 * G - 2) / 2 < kListCap, "fetched words stay in the slot's column");
    if (cnt == 0) return;
    const unsigned *__restrict__ col = list + i;
    auto fetch = [&](unsigned k, unsigned *out) {   // entries k, k + 1 (k even)
        const unsigned w = col[(k / 2u) * stride];
        const unsigned a = decode(w & 0xffffu), b = decode(w >> 16);
        out[0] = k < cnt ? a : i;
        out[1] = k + 1u < cnt ? b : i;
    };
    unsigned e[G];
#pragma unroll
    for (unsigned g = 0; g < G; g += 2) fetch(g, e + g);
    for (unsigned k0 = 0; k0 < cnt; k0 += G) {
        decltype(load(0u)) d[G];
#pragma unroll
        for (unsigned g = 0; g < G; ++g) d[g] = load(e[g]);
        unsigned nx[G];
#pragma unroll
        for (unsigned g = 0; g < G; g += 2) fetch(k0 + G + g, nx + g);
#pragma unroll
        for (unsigned g = 0; g < G; ++g)
            if (k0 + g < cnt) use(e[g], d[g]);
#pragma unroll
        for (unsigned g = 0; g < G; ++g) e[g] = nx[g];
    }
}

__device__ __forceinline__ V3 plane_point(const StepParams &p, unsigned s) {
    return V3{p.planes[s][0], p.planes[s][1], p.planes[s][2]};
}
__device__ __forceinline__ V3 plane_normal(const StepParams &p, unsigned s) {
    return V3{p.planes[s][3], p.planes[s][4], p.planes[s][5]};
}

// Planes (with a non-zero normal) that particle i is within h of.  A plane
// combination can only contribute a mirror image when every plane in it is
// in this set (kernels/common.cl:658-670: any plane failing the test skips
// the combination, and a zero normal always fails the 1e-5 test).
__device__ __forceinline__ unsigned planes_near_i(const StepParams &p, V3 xi) {
    unsigned m = 0;
    for (unsigned s = 0; s < p.nplanes; ++s) {
        if (!((p.active_planes >> s) & 1u)) continue;
        float d = fabsf(dot(plane_normal(p, s), xi - plane_point(p, s)));
        if (!(d > p.h)) m |= 1u << s;
    }
    return m;
}
// Planes of `cand` whose distance to j lies in [1e-5, h].  The 1e-5 bound is
// a double literal in the reference, so the comparison is done in double.
__device__ __forceinline__ unsigned planes_near_j(const StepParams &p, V3 xj, unsigned cand) {
    unsigned m = 0;
    for (unsigned s = 0; s < p.nplanes; ++s) {
        if (!((cand >> s) & 1u)) continue;
        float d = fabsf(dot(plane_normal(p, s), xj - plane_point(p, s)));
        if (!(d > p.h || static_cast<double>(d) < 1.0e-5)) m |= 1u << s;
    }
    return m;
}

// Position reflected in every plane of combination b, lowest plane first.
__device__ __forceinline__ V3 reflect_combo(const StepParams &p, unsigned b, V3 x) {
    for (unsigned s = 0; s < p.nplanes; ++s)
        if ((b >> s) & 1u) x = reflect_point(x, plane_point(p, s), plane_normal(p, s));
    return x;
}

__device__ __forceinline__ KernelConsts kconsts(const StepParams &p) {
    return KernelConsts{p.h, p.h2, p.poly6, p.spiky, p.lap};
}

}  // namespace sphk
