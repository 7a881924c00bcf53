// Density, fluid forces, integrator, plane constraints and the solid setup
// passes that share the density loop.
//
// Reference kernels: compute_density (kernels/common.cl:610-690),
// compute_forces_fluids (kernels/fluids.cl:1-63), step_forward
// (kernels/common.cl:691-722), apply_plane_constraints
// (kernels/constraints.cl:1-30), init_original_position and
// compute_original_density (kernels/solids.cl:21-107).
//
// One thread per active slot; every sum runs over the candidates in the
// reference's visiting order, so with compensated sums off each output is
// the same IEEE operation sequence as the sequential oracle.  Every kernel is
// a template over its parameter source (one scene, or a batch's scene
// table), so a scene computes the same bits alone and in a batch.
#include "neighbours.h"

#pragma clang fp contract(off)

namespace sphk {
namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kPairGroup = 8;   // list entries whose loads are issued together
constexpr unsigned kScanUnroll = 8;  // candidates whose loads are issued together in a scan
constexpr unsigned kSolidScanUnroll = 4;   // the same for the solid scan (two positions per candidate)

template <bool KAHAN>
struct Acc {
    float s = 0.0f;
    __device__ __forceinline__ void add(float v) { s = s + v; }
    __device__ __forceinline__ float value() const { return s; }
};
template <>
struct Acc<true> {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s; }
};

// Equation of state: the reference's linear p = k (rho - rho0)
// (kernels/fluids.cl:28-29), or Tait with exponent 7 when the conf sets
// eos_tait.
__device__ __forceinline__ float pressure(const StepParams &p, float rho) {
    if (!p.eos_tait) return (rho - p.restdens) * p.stiffness;
    const float r = rho / p.restdens;
    const float r2 = r * r, r4 = r2 * r2;
    return p.stiffness * (r4 * r2 * r - 1.0f);
}

// The reference density loop: every candidate, direct term then mirror
// images (kernels/common.cl:634-687).  Returns the kernel sum; counts r < h
// in *nn (COUNT); when `list` is given, records the slots with r <= h in
// list[k * stride + i] (at most kListCap) and their number in *listed.
// The direct term's range test is the square-root-free form of r <= h
// (s < r2_out), so out-of-range candidates take no square root.
template <bool KAHAN, bool COUNT>
__device__ __forceinline__ float density_scan(const StepParams &p, const FieldPtrs &f, unsigned i, V3 xi,
                                              unsigned inear, unsigned *nn, unsigned *list, unsigned stride,
                                              unsigned *listed) {
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const unsigned ncombo = 1u << p.nplanes;
    const float r2_out = p.r2_out;   // read once: with a scene table p is in global memory
    Acc<KAHAN> rho;
    unsigned cnt = 0, k = 0;
    for_each_candidate(p.grid, p.p0, p.p0 + p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 xj = load3(pos, j);
        const float s = sqlength(xj - xi);
        if (!(s >= r2_out)) {
            const float r = sqrtf(s);
            rho.add(w_poly6(r, kc));
            if (list && k < kListCap) list[static_cast<size_t>(k) * stride + i] = j;
            ++k;
            if (COUNT && r < kc.h) ++cnt;
        }
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok)
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const float rr = length(reflect_combo(p, b, xj) - xi);
                    if (rr < kc.h) {
                        rho.add(w_poly6(rr, kc));
                        if (COUNT) ++cnt;
                    }
                }
        }
    });
    *nn = cnt;
    if (listed) *listed = k;
    return rho.value();
}

// ORIGINAL = compute_original_density (density0, no neighbour count).
template <class PS, bool ORIGINAL, bool KAHAN>
__global__ void __launch_bounds__(kBlock) k_density(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const V3 xi = load3(f.position, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    unsigned nn = 0;
    const float rho = density_scan<KAHAN, !ORIGINAL>(p, f, i, xi, inear, &nn, nullptr, 0, nullptr);
    const float out = rho * p.mass;
    if (ORIGINAL) {
        f.density0[i] = out;
    } else {
        f.density[i] = out;
        if (f.num_neighbours) f.num_neighbours[i] = nn;
    }
}

// compute_density for fluid programs, in two phases per lane:
//   1. scan the candidates with the range prefilter (sqlength_fused, a
//      superset of the square-root-free test) and compact the hits into this
//      lane's column of an LDS buffer, as 16-bit entries (dz group << 14 |
//      offset from the group's first slot);
//   2. re-test each hit with sqlength and evaluate the kernel sum over those
//      in range (every lane busy), writing the entries out as the pair
//      cache's neighbour list; a hit out of range is written as an entry of
//      the slot itself, which the force pass skips as it skips j == i.
// The summation order is the candidate order, as in the reference.  Hits
// past the LDS column go straight to the pair cache.  Lanes near a mirror
// plane, with more than kFluidListMax hits or with a dz group spanning more
// than 16384 slots take the reference loop instead.
constexpr unsigned kListBlock = 128;
// Entries of a lane's LDS list column in the density scans (the rest of a
// longer list goes straight to the pair cache).
constexpr unsigned kFluidListCap = 40;
constexpr unsigned kSolidListCap = 40;
static_assert(kFluidListCap % kPairGroup == 0 && kSolidListCap % kPairGroup == 0,
              "the kernel sums take whole groups from the LDS column");
static_assert(kFluidListCap <= kFluidListMax && kSolidListCap <= kListCap,
              "the LDS column is a prefix of the list");

// First slot of neighbour row r (0..8, (dz, dy) = (r / 3 - 1, r % 3 - 1))
// around the cell (xo, yo, zo) of the scene's grid, and one past its last
// slot, both clamped to the scene's slots [lo, hi); false when the row lies
// outside the grid.
__device__ __forceinline__ bool row_range(const GridGeom &g, unsigned lo, unsigned hi, unsigned xo, unsigned yo,
                                          unsigned zo, unsigned r, const unsigned *__restrict__ celloffset,
                                          const unsigned *__restrict__ gridcount, unsigned *b, unsigned *e) {
    const int zn = static_cast<int>(zo) + static_cast<int>(r / 3u) - 1;
    const int yn = static_cast<int>(yo) + static_cast<int>(r % 3u) - 1;
    if (zn < 0 || zn >= static_cast<int>(g.rz) || yn < 0 || yn >= static_cast<int>(g.ry)) return false;
    const unsigned x0 = xo > 0u ? xo - 1u : 0u;
    const unsigned x1 = xo + 1u < g.rx ? xo + 1u : xo;
    const unsigned row = g.c0 + static_cast<unsigned>(zn) * g.layer + static_cast<unsigned>(yn) * g.rx;
    *b = min(max(celloffset[row + x0], lo), hi);
    *e = min(max(celloffset[row + x1] + gridcount[row + x1], lo), hi);
    return true;
}

// Where a lane's list entries past its LDS column go: straight into its
// column of the pair cache.  Entry k of a fluid list is the low (k even) or
// high half of word k / 2; entry k of a solid list is word k (slot | flags).
struct ListSpill {
    unsigned *list;
    size_t stride;
    unsigned i;
    __device__ __forceinline__ unsigned short *fluid_entry(unsigned k) const {
        return reinterpret_cast<unsigned short *>(list + (k / 2u) * stride + i) + (k & 1u);
    }
    __device__ __forceinline__ unsigned *solid_entry(unsigned k) const { return list + k * stride + i; }
};

// One round of the fluid scan: the range prefilter (sqlength_fused against
// r2_wide) on kScanUnroll candidates xs in slot order, appending ent + u to
// the lane's LDS column at w for each hit.  CHECK_LEFT: the round reaches
// past the row's end, and only its first `left` candidates are used.  CLAMP:
// the column may fill up during the round; entries past it go to the pair
// cache (up to kFluidListMax; beyond that the count alone grows, and the
// lane's list is discarded).
template <bool CHECK_LEFT, bool CLAMP>
__device__ __forceinline__ void scan_round(const V3 (&xs)[kScanUnroll], unsigned left, V3 xi, float r2_wide,
                                           unsigned ent, unsigned short *&w, const unsigned short *col,
                                           const ListSpill &spill) {
#pragma unroll
    for (unsigned u = 0; u < kScanUnroll; ++u) {
        const float s = sqlength_fused(xs[u] - xi);
        const bool hit = CHECK_LEFT ? ((u < left) & !(s >= r2_wide)) : !(s >= r2_wide);
        if (hit) {
            const unsigned short v = static_cast<unsigned short>(ent + u);
            if (CLAMP) {
                const unsigned k = static_cast<unsigned>(w - col) / kListBlock;
                if (k < kFluidListCap) *w = v;
                else if (k < kFluidListMax) *spill.fluid_entry(k) = v;
            } else {
                *w = v;
            }
            w += kListBlock;
        }
    }
}
// The candidates of a round: from global memory (the round's offsets fold
// into the loads), or from the wave's LDS copy of the row (below).
template <unsigned N>
__device__ __forceinline__ void candidates_from(const Packed3 *q, V3 (&xs)[N]) {
#pragma unroll
    for (unsigned u = 0; u < N; ++u) {
        const Packed3 t = q[u];
        xs[u] = V3{t.x, t.y, t.z};
    }
}

// Row staging.  The lanes of a wave hold consecutive slots of one scene, so
// their cells are in key order and so are the slot ranges of each of their
// neighbour rows: the union of a row's ranges over the wave is the
// consecutive slots from its first scanning lane's start to its last
// scanning lane's end.  The wave copies those positions into LDS once and
// its lanes read their candidates there, instead of every lane reading its
// 20-30 candidates through the vector cache.  A row whose union exceeds
// kStageCap slots is scanned from global memory.  The LDS of a block (the
// lanes' list columns of kFluidListCap entries and the staged rows) is sized
// so that LDS does not limit the resident waves below what the registers
// allow.
constexpr unsigned kWave = 64;
constexpr unsigned kStageCap = 96;
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First slots of the three dz groups of neighbour rows of a lane (rows 3g
// to 3g + 2: the rows of one z layer of cells, in slot order), kept in
// registers between the scan and the kernel sum and stored with the list.
// A fluid list entry is (g << kGroupBits) | (j - base of group g), 16 bits.
constexpr unsigned kGroupBits = 14;
__device__ __forceinline__ unsigned group_slot(unsigned ent, unsigned b0, unsigned b1, unsigned b2) {
    const unsigned g = ent >> kGroupBits;
    return (g == 0u ? b0 : (g == 1u ? b1 : b2)) + (ent & ((1u << kGroupBits) - 1u));
}
struct GroupBases {
    unsigned b0 = 0, b1 = 0, b2 = 0;
    __device__ __forceinline__ unsigned slot(unsigned ent) const { return group_slot(ent, b0, b1, b2); }
};

template <class PS, bool KAHAN>
__global__ void __launch_bounds__(kListBlock) k_density_listed(PS ps, FieldPtrs f, PairCache c) {
    __shared__ unsigned short buf[kFluidListCap * kListBlock];
    // per wave: one staged row, with room for a round read past its end
    __shared__ Packed3 stage[kListBlock / kWave][kStageCap + kScanUnroll];
    const unsigned i = ps.slot_xcd(kListBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const unsigned lo = p.p0, hi = p.p0 + p.n;
    const float r2_out = p.r2_out;
    const float r2_wide = widened_r2(r2_out);
    const GridGeom g = p.grid;
    const V3 xi = load3(pos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned key = f.gridcell[i];
    unsigned short *col = buf + threadIdx.x;
    const ListSpill spill{c.list, c.stride, i};
    Packed3 *const stg = stage[threadIdx.x / kWave];
    unsigned cnt = 0;
    const bool key_in_grid = key >= g.c0 && key - g.c0 < g.ncells;
    bool fast = !inear && key_in_grid;
    GroupBases gb;
    // Every active lane of the wave takes part in the row loop (to stage
    // rows); the lanes on the fast path scan.  The copy is spread over the
    // active lanes by their rank among them (slots past a scene's end have
    // left the kernel).
    if (__any(fast)) {
        const unsigned long long act = __ballot(1);
        const unsigned copier = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(act >> 32),
                                                          __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(act), 0u));
        const unsigned copiers = static_cast<unsigned>(__builtin_popcountll(act));
        const unsigned local = key_in_grid ? key - g.c0 : 0u;
        const unsigned zo = local / g.layer;
        const unsigned rem = local - g.layer * zo;
        const unsigned yo = rem / g.rx;
        const unsigned xo = rem - g.rx * yo;
        // next free entry of the lane's column; the last position with room
        // for a whole round
        unsigned short *w = col;
        unsigned short *const wroom = col + (kFluidListCap - kScanUnroll) * kListBlock;
        unsigned gbase = 0;   // first slot of the current group's first row inside the grid
        bool gset = false;
#pragma unroll 1
        for (unsigned r = 0; r < 9u; ++r) {
            const unsigned gi = r / 3u;
            if (r == 3u * gi) gset = false;
            unsigned b = 0, e = 0;
            const bool scan = row_range(g, lo, hi, xo, yo, zo, r, f.celloffset, f.gridcount, &b, &e) && fast;
            const unsigned long long m = __ballot(scan);
            if (m == 0) continue;
            // the wave's union of this row's ranges: [first lane's b, last lane's e)
            const unsigned sb = __builtin_amdgcn_readlane(b, __builtin_ctzll(m));
            const unsigned se = __builtin_amdgcn_readlane(e, 63 - __builtin_clzll(m));
            const bool staged = se - sb <= kStageCap;
            if (staged) {
                for (unsigned k = copier; k < se - sb; k += copiers) {
                    const V3 x = load3(pos, sb + k);
                    stg[k] = Packed3{x.x, x.y, x.z};
                }
                wave_sync();
            }
            if (scan) {
                if (!gset) {
                    gbase = b;
                    gset = true;
                    if (gi == 0u) gb.b0 = b;
                    else if (gi == 1u) gb.b1 = b;
                    else gb.b2 = b;
                }
                if (e - gbase > (1u << kGroupBits)) fast = false;
                // kScanUnroll candidates per round, tested in slot order:
                // entry ebase + (j0 - b) + u.  The last round of a row reads
                // past its end (within the buffer's kScanPad padding, or the
                // stage's spare room) and does not use those candidates.
                static_assert(kScanUnroll <= kScanPad, "scan rounds read past the row end");
                const unsigned ebase = (gi << kGroupBits) + (b - gbase);
                // Whole rounds need no end-of-row test; a round needs no cap
                // test when every lane of the wave has room for all its
                // candidates.
                unsigned j0 = b;
                V3 xs[kScanUnroll];
                for (; j0 + kScanUnroll <= e; j0 += kScanUnroll) {
                    const unsigned ent = ebase + (j0 - b);
                    if (staged) candidates_from(stg + (j0 - sb), xs);
                    else candidates_from(reinterpret_cast<const Packed3 *>(pos) + j0, xs);
                    if (__all(w <= wroom)) scan_round<false, false>(xs, 0u, xi, r2_wide, ent, w, col, spill);
                    else scan_round<false, true>(xs, 0u, xi, r2_wide, ent, w, col, spill);
                }
                if (j0 < e) {
                    const unsigned ent = ebase + (j0 - b);
                    if (staged) candidates_from(stg + (j0 - sb), xs);
                    else candidates_from(reinterpret_cast<const Packed3 *>(pos) + j0, xs);
                    if (__all(w <= wroom)) scan_round<true, false>(xs, e - j0, xi, r2_wide, ent, w, col, spill);
                    else scan_round<true, true>(xs, e - j0, xi, r2_wide, ent, w, col, spill);
                }
            }
            // this row's reads before the next row's copy
            if (staged) wave_sync();
        }
        cnt = static_cast<unsigned>(w - col) / kListBlock;
        if (cnt > kFluidListMax) fast = false;
    }
    float rho;
    unsigned nn = 0;
    if (fast) {
        // the list keeps the LDS entries as they are, two per word of the
        // slot's own column (see kFluidPlanes), with the group bases
        static_assert(kPairGroup % 2 == 0, "entries are stored in pairs");
        Acc<KAHAN> acc;
        // the slot's own entry: its cell is in the dz = 0 group, whose span
        // the scan checked
        const unsigned self_ent = (1u << kGroupBits) | (i - gb.b1);
        for (unsigned k0 = 0; k0 < cnt; k0 += kPairGroup) {
            unsigned es[kPairGroup], js[kPairGroup];
            // entries of the LDS column, then (k0 is the same for every
            // lane) those the scan put in the pair cache
#pragma unroll
            for (unsigned q = 0; q < kPairGroup; ++q) es[q] = 0;
            if (k0 < kFluidListCap) {
#pragma unroll
                for (unsigned q = 0; q < kPairGroup; ++q)
                    if (k0 + q < cnt) es[q] = col[(k0 + q) * kListBlock];
            } else {
#pragma unroll
                for (unsigned q = 0; q < kPairGroup; ++q)
                    if (k0 + q < cnt) es[q] = *spill.fluid_entry(k0 + q);
            }
#pragma unroll
            for (unsigned q = 0; q < kPairGroup; ++q) js[q] = k0 + q < cnt ? gb.slot(es[q]) : i;
            V3 xs[kPairGroup];
#pragma unroll
            for (unsigned q = 0; q < kPairGroup; ++q) xs[q] = load3(pos, js[q]);
#pragma unroll
            for (unsigned q = 0; q < kPairGroup; ++q) {
                if (k0 + q >= cnt) break;
                const float sq = sqlength(xs[q] - xi);
                if (!(sq >= r2_out)) {
                    // in range: RN(sqrt(sq)) <= h
                    const float rr = sqrt_rn(sq);
                    acc.add(w_poly6_in_range(rr, kc));
                    nn += rr < kc.h ? 1u : 0u;
                } else {
                    es[q] = self_ent;
                }
            }
#pragma unroll
            for (unsigned q = 0; q < kPairGroup; q += 2)
                if (k0 + q < cnt)
                    c.list[static_cast<size_t>((k0 + q) / 2u) * c.stride + i] = es[q] | (es[q + 1] << 16);
        }
        rho = acc.value();
    } else {
        // the force pass scans for these lanes (no list)
        rho = density_scan<KAHAN, true>(p, f, i, xi, inear, &nn, nullptr, 0, nullptr);
        cnt = kFluidListMax + 1u;
    }
    const float out = rho * p.mass;
    f.density[i] = out;
    if (f.num_neighbours) f.num_neighbours[i] = nn;
    c.count[i] = cnt;
    const V3 vi = load3(f.velocity, i);
    const size_t st = c.stride;
    c.rec[i] = make_float4(xi.x, xi.y, xi.z, out);
    c.rec[st + i] = make_float4(vi.x, vi.y, vi.z, pressure(p, out));
    c.rec[2 * st + i] = make_float4(__uint_as_float(gb.b0), __uint_as_float(gb.b1), __uint_as_float(gb.b2), 0.0f);
}

// compute_density for solid programs, which also records the kSolidList
// neighbour list (current or reference distance within h, with flags) in
// candidate order.  Two phases per lane as in k_density_listed: the scan
// compacts candidates in either range into LDS as 16-bit entries (dz group
// << 14 | current << 13 | reference << 12 | offset from the group's base),
// then the kernel sum runs over the compacted entries (those past the LDS
// column go straight to the pair cache as slot | flags).  The range tests
// are the square-root-free form of length(d) <= h.  Lanes near a mirror
// plane take the reference loop and mark their list as overflowed; lanes
// with more than kListCap entries or a group spanning more than 4096 slots
// scan again without LDS.
constexpr unsigned kSolidOffsetBits = 12;

// One round of the solid scan: kSolidScanUnroll candidates (current
// positions xs, reference positions x0s) tested against the current and the
// reference position of the lane, appending
// (ent + u) | current << 13 | reference << 12 for each candidate in either
// range.  CHECK_LEFT and CLAMP as in scan_round; an entry past the LDS
// column goes to the pair cache as (slot0 + u) | flags, up to kListCap.
template <bool CHECK_LEFT, bool CLAMP>
__device__ __forceinline__ void solid_scan_round(const V3 (&xs)[kSolidScanUnroll], const V3 (&x0s)[kSolidScanUnroll],
                                                 unsigned left, V3 xi, V3 x0i, float r2_out, unsigned ent,
                                                 unsigned slot0, unsigned short *&w, const unsigned short *col,
                                                 const ListSpill &spill) {
#pragma unroll
    for (unsigned u = 0; u < kSolidScanUnroll; ++u) {
        const bool cur = !(sqlength(xs[u] - xi) >= r2_out);
        const bool ref = !(sqlength(x0s[u] - x0i) >= r2_out);
        const bool hit = CHECK_LEFT ? ((u < left) & (cur | ref)) : (cur | ref);
        if (hit) {
            const unsigned short v =
                static_cast<unsigned short>((ent + u) | (cur ? 1u << 13 : 0u) | (ref ? 1u << 12 : 0u));
            if (CLAMP) {
                const unsigned k = static_cast<unsigned>(w - col) / kListBlock;
                if (k < kSolidListCap) *w = v;
                else if (k < kListCap)
                    *spill.solid_entry(k) = (slot0 + u) | (cur ? kListCurrent : 0u) | (ref ? kListReference : 0u);
            } else {
                *w = v;
            }
            w += kListBlock;
        }
    }
}
// The solid scan's LDS: list columns of kSolidListCap entries and per wave
// the current positions of one row of up to kSolidStageCap slots (as for the
// fluid scan).  The reference positions are read from global memory: staging
// them too costs more occupancy than their reads do.
constexpr unsigned kSolidStageCap = 128;

template <class PS, bool KAHAN>
__global__ void __launch_bounds__(kListBlock) k_density_solid_listed(PS ps, FieldPtrs f, PairCache c) {
    __shared__ unsigned short buf[kSolidListCap * kListBlock];
    // per wave: one staged row, with room for a round read past its end
    __shared__ Packed3 stage[kListBlock / kWave][kSolidStageCap + kSolidScanUnroll];
    const unsigned i = ps.slot_xcd(kListBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const float *__restrict__ opos = f.originalpos;
    const unsigned lo = p.p0, hi = p.p0 + p.n;
    const float r2_out = p.r2_out;
    const GridGeom g = p.grid;
    const V3 xi = load3(pos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned key = f.gridcell[i];
    unsigned short *col = buf + threadIdx.x;
    const ListSpill spill{c.list, c.stride, i};
    Packed3 *const stg = stage[threadIdx.x / kWave];
    const V3 x0i = load3(opos, i);
    const bool key_in_grid = key >= g.c0 && key - g.c0 < g.ncells;
    bool fast = !inear && key_in_grid;
    unsigned nn = 0, cnt = 0;
    unsigned b0 = 0, b1 = 0, b2 = 0;   // dz group bases, as in k_density_listed
    // the row loop as in k_density_listed: every active lane stages, the
    // lanes on the fast path scan
    if (__any(fast)) {
        const unsigned long long act = __ballot(1);
        const unsigned copier = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(act >> 32),
                                                          __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(act), 0u));
        const unsigned copiers = static_cast<unsigned>(__builtin_popcountll(act));
        const unsigned local = key_in_grid ? key - g.c0 : 0u;
        const unsigned zo = local / g.layer;
        const unsigned rem = local - g.layer * zo;
        const unsigned yo = rem / g.rx;
        const unsigned xo = rem - g.rx * yo;
        unsigned short *w = col;   // as in k_density_listed
        unsigned short *const wroom = col + (kSolidListCap - kSolidScanUnroll) * kListBlock;
        unsigned gbase = 0;
        bool gset = false;
#pragma unroll 1
        for (unsigned r = 0; r < 9u; ++r) {
            const unsigned gi = r / 3u;
            if (r == 3u * gi) gset = false;
            unsigned b = 0, e = 0;
            const bool scan = row_range(g, lo, hi, xo, yo, zo, r, f.celloffset, f.gridcount, &b, &e) && fast;
            const unsigned long long m = __ballot(scan);
            if (m == 0) continue;
            const unsigned sb = __builtin_amdgcn_readlane(b, __builtin_ctzll(m));
            const unsigned se = __builtin_amdgcn_readlane(e, 63 - __builtin_clzll(m));
            const bool staged = se - sb <= kSolidStageCap;
            if (staged) {
                for (unsigned k = copier; k < se - sb; k += copiers) {
                    const V3 x = load3(pos, sb + k);
                    stg[k] = Packed3{x.x, x.y, x.z};
                }
                wave_sync();
            }
            if (scan) {
                if (!gset) {
                    gbase = b;
                    gset = true;
                    if (gi == 0u) b0 = b;
                    else if (gi == 1u) b1 = b;
                    else b2 = b;
                }
                if (e - gbase > (1u << kSolidOffsetBits)) fast = false;
                static_assert(kSolidScanUnroll <= kScanPad, "scan rounds read past the row end");
                // entry of slot j: (gi << kGroupBits) + (j - gbase), ent + u in a round from j0
                const unsigned ebase = (gi << kGroupBits) + (b - gbase);
                unsigned j0 = b;
                V3 xs[kSolidScanUnroll], x0s[kSolidScanUnroll];
                for (; j0 + kSolidScanUnroll <= e; j0 += kSolidScanUnroll) {
                    const unsigned ent = ebase + (j0 - b);
                    if (staged) candidates_from(stg + (j0 - sb), xs);
                    else candidates_from(reinterpret_cast<const Packed3 *>(pos) + j0, xs);
                    candidates_from(reinterpret_cast<const Packed3 *>(opos) + j0, x0s);
                    if (__all(w <= wroom))
                        solid_scan_round<false, false>(xs, x0s, 0u, xi, x0i, r2_out, ent, j0, w, col, spill);
                    else solid_scan_round<false, true>(xs, x0s, 0u, xi, x0i, r2_out, ent, j0, w, col, spill);
                }
                if (j0 < e) {
                    const unsigned ent = ebase + (j0 - b);
                    if (staged) candidates_from(stg + (j0 - sb), xs);
                    else candidates_from(reinterpret_cast<const Packed3 *>(pos) + j0, xs);
                    candidates_from(reinterpret_cast<const Packed3 *>(opos) + j0, x0s);
                    if (__all(w <= wroom))
                        solid_scan_round<true, false>(xs, x0s, e - j0, xi, x0i, r2_out, ent, j0, w, col, spill);
                    else solid_scan_round<true, true>(xs, x0s, e - j0, xi, x0i, r2_out, ent, j0, w, col, spill);
                }
            }
            if (staged) wave_sync();
        }
        cnt = static_cast<unsigned>(w - col) / kListBlock;
        if (cnt > kListCap) fast = false;
    }
    float rho;
    if (inear) {
        rho = density_scan<KAHAN, true>(p, f, i, xi, inear, &nn, nullptr, 0, nullptr);
        cnt = kListCap + 1u;
    } else {
        Acc<KAHAN> acc;
        if (fast) {
            for (unsigned k0 = 0; k0 < cnt; k0 += kPairGroup) {
                // list entries as slot | flags: from the LDS column, then (k0
                // is the same for every lane) those the scan put in the pair
                // cache
                unsigned es[kPairGroup];
#pragma unroll
                for (unsigned q = 0; q < kPairGroup; ++q) es[q] = i;
                if (k0 < kSolidListCap) {
#pragma unroll
                    for (unsigned q = 0; q < kPairGroup; ++q)
                        if (k0 + q < cnt) {
                            const unsigned ent = col[(k0 + q) * kListBlock];
                            const unsigned gq = ent >> kGroupBits;
                            es[q] = ((gq == 0u ? b0 : (gq == 1u ? b1 : b2)) +
                                     (ent & ((1u << kSolidOffsetBits) - 1u))) |
                                    (((ent >> 13) & 1u) ? kListCurrent : 0u) |
                                    (((ent >> 12) & 1u) ? kListReference : 0u);
                        }
                } else {
#pragma unroll
                    for (unsigned q = 0; q < kPairGroup; ++q)
                        if (k0 + q < cnt) es[q] = *spill.solid_entry(k0 + q);
                }
                V3 xs[kPairGroup];
#pragma unroll
                for (unsigned q = 0; q < kPairGroup; ++q) xs[q] = load3(pos, es[q] & kListSlot);
#pragma unroll
                for (unsigned q = 0; q < kPairGroup; ++q) {
                    if (k0 + q >= cnt) break;
                    if (es[q] & kListCurrent) {
                        const float rr = length(xs[q] - xi);
                        acc.add(w_poly6(rr, kc));
                        if (rr < kc.h) ++nn;
                    }
                    c.list[static_cast<size_t>(k0 + q) * c.stride + i] = es[q];
                }
            }
        } else {
            cnt = 0;
            for_each_candidate(g, lo, hi, key, f.celloffset, f.gridcount, [&](unsigned j) {
                const V3 d = load3(pos, j) - xi;
                const bool cur = !(sqlength(d) >= r2_out);
                const bool ref = !(sqlength(load3(opos, j) - x0i) >= r2_out);
                if (cur) {
                    const float rr = length(d);
                    acc.add(w_poly6(rr, kc));
                    if (rr < kc.h) ++nn;
                }
                if (cur || ref) {
                    if (cnt < kListCap)
                        c.list[static_cast<size_t>(cnt) * c.stride + i] =
                            j | (cur ? kListCurrent : 0u) | (ref ? kListReference : 0u);
                    ++cnt;
                }
            });
        }
        rho = acc.value();
    }
    f.density[i] = rho * p.mass;
    if (f.num_neighbours) f.num_neighbours[i] = nn;
    c.count[i] = cnt;
}

// Penalty walls at +-2 m, evaluated in double as in the reference macro
// (kernels/common.cl:37-43 with the double literals of kernels/fluids.cl:60).
__device__ __forceinline__ float cube_wall(float x, float fc) {
    const double lo = -2.0, hi = 2.0;
    const double xd = static_cast<double>(x);
    if (xd < lo + static_cast<double>(0.2f))
        return static_cast<float>(static_cast<double>(fc) + (1.0 / (-lo + xd) - 5.0));
    if (xd > hi - static_cast<double>(0.2f))
        return static_cast<float>(static_cast<double>(fc) - (1.0 / (hi - xd) - 5.0));
    return fc;
}

// The pressure + viscosity pair term of compute_forces_fluids
// (kernels/fluids.cl:27-50), for a pair already known to be in range, and
// the final scaling and walls (kernels/fluids.cl:56-62).
template <bool KAHAN>
struct FluidForce {
    Acc<KAHAN> fx, fy, fz;
    __device__ __forceinline__ void pair(float viscosity, const KernelConsts &kc, V3 xi, V3 vi, float rhoi,
                                         float pi, V3 xj, V3 vj, float rhoj, float pj) {
        const V3 d = xj - xi;
        const float r = length(d);
        const V3 rv = vj - vi;
        const float din = kc.h - r;
        const float didj = rhoi * rhoj;
        const float vterm = (viscosity * din) * kc.lap;
        // (p_i + p_j) * 0.5 / r * din * din * spiky, in double
        // (the 0.5 literal promotes it, kernels/fluids.cl:48)
        const float pterm = static_cast<float>(static_cast<double>(pi + pj) * 0.5 / static_cast<double>(r) *
                                               static_cast<double>(din) * static_cast<double>(din) *
                                               static_cast<double>(kc.spiky));
        const double rdidj = recip(didj);   // the three quotients below, bit for bit
        fx.add(div_recip(d.x * pterm + rv.x * vterm, rdidj));
        fy.add(div_recip(d.y * pterm + rv.y * vterm, rdidj));
        fz.add(div_recip(d.z * pterm + rv.z * vterm, rdidj));
    }
    __device__ __forceinline__ V3 finish(const StepParams &p, V3 xi) const {
        V3 fi = V3{fx.value(), fy.value(), fz.value()} * p.mass;
        if (p.mass_consistent) fi = fi * p.mass;   // the reference sums m_j only
        fi.x = cube_wall(xi.x, fi.x);
        fi.y = cube_wall(xi.y, fi.y);
        fi.z = cube_wall(xi.z, fi.z);
        return fi;
    }
};

template <class PS, bool KAHAN>
__global__ void __launch_bounds__(kBlock) k_forces_fluids(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float r2o = p.r2_out;   // read once: with a scene table p is in global memory
    const float *__restrict__ pos = f.position;
    const float *__restrict__ vel = f.velocity;
    const float *__restrict__ rho = f.density;
    const V3 xi = load3(pos, i), vi = load3(vel, i);
    const float rhoi = rho[i];
    const float pi = pressure(p, rhoi);
    const float visc = p.viscosity;   // read once: with a scene table p is in global memory
    FluidForce<KAHAN> acc;
    for_each_candidate(p.grid, p.p0, p.p0 + p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        if (j == i) return;
        const V3 xj = load3(pos, j);
        if (sqlength(xj - xi) >= r2o) return;   // length > h, without the square root
        const float rhoj = rho[j];
        acc.pair(visc, kc, xi, vi, rhoi, pi, xj, load3(vel, j), rhoj, pressure(p, rhoj));
    });
    store3(f.force, i, acc.finish(p, xi));
}

__device__ __forceinline__ bool is_static_body(const StepParams &p, const FieldPtrs &f, unsigned i) {
    if (!f.body_number) return false;
    const unsigned b = f.body_number[i];
    if (b == 0u || !f.body_dynamic || b > p.body_count) return false;
    return f.body_dynamic[p.body_base + b - 1u] == 0u;
}

// step_forward of slot i (kernels/common.cl:691-722) from its force and its
// current velocity and position.
__device__ __forceinline__ void integrate(const StepParams &p, const FieldPtrs &f, unsigned i, V3 force, V3 v,
                                          V3 x) {
    if (is_static_body(p, f, i)) return;
    const V3 g = V3{p.gravity[0], p.gravity[1], p.gravity[2]};
    const V3 a = force / p.mass + g;
    if (f.acceleration) store3(f.acceleration, i, a);
    const V3 vn = v + p.dt * a;
    store3(f.position, i, x + p.dt * vn);
    store3(f.velocity, i, vn);
}

// compute_forces_fluids over the pair cache: the in-range list of each slot
// and the (position, density) and (velocity, pressure) planes of the density
// pass, two aligned 16-byte gathers per neighbour (the pressures are
// identical values to evaluating the equation of state per pair).
// Overflowed lists fall back to the scan.  INTEGRATE: also step_forward
// for the slot, in place: every neighbour's position and velocity come from
// the density pass's planes, so no lane reads what another lane updates.
template <class PS, bool KAHAN, unsigned G, bool INTEGRATE>
__global__ void __launch_bounds__(kBlock) k_forces_fluids_listed(PS ps, FieldPtrs f, PairCache c) {
    const unsigned i = ps.slot_xcd(blockDim.x);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const KernelConsts kc = kconsts(p);
    const float r2o = p.r2_out;   // read once: with a scene table p is in global memory
    const float4 *__restrict__ xr = c.rec;
    const float4 *__restrict__ vp = c.rec + c.stride;
    const float4 ai = xr[i], bi = vp[i];
    const V3 xi{ai.x, ai.y, ai.z}, vi{bi.x, bi.y, bi.z};
    const float rhoi = ai.w;
    const float pi = bi.w;
    const float visc = p.viscosity;   // read once: with a scene table p is in global memory
    FluidForce<KAHAN> acc;
    const unsigned cnt = c.count[i];
    if (cnt <= kFluidListMax) {
        // List entries other than slot i are in range by construction (same
        // positions, same distance expression), so the pair terms need no
        // range test here.
        struct Nbr {
            float4 a, b;
        };
        // the slot's list bases, as values (the decode is a select on them)
        const float4 gbf = c.rec[2 * static_cast<size_t>(c.stride) + i];
        const unsigned b0 = __float_as_uint(gbf.x), b1 = __float_as_uint(gbf.y), b2 = __float_as_uint(gbf.z);
        for_each_listed16<G>(
            c.list, c.stride, i, cnt,
            [b0, b1, b2](unsigned ent) { return group_slot(ent, b0, b1, b2); },
            [&](unsigned j) { return Nbr{xr[j], vp[j]}; },
            [&](unsigned j, const Nbr &n) {
                if (j != i)
                    acc.pair(visc, kc, xi, vi, rhoi, pi, V3{n.a.x, n.a.y, n.a.z}, V3{n.b.x, n.b.y, n.b.z}, n.a.w,
                             n.b.w);
            });
    } else {
        for_each_candidate(p.grid, p.p0, p.p0 + p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
            if (j == i) return;
            const float4 a = xr[j];
            const V3 xj{a.x, a.y, a.z};
            if (sqlength(xj - xi) >= r2o) return;   // length > h, without the square root
            const float4 b = vp[j];
            acc.pair(visc, kc, xi, vi, rhoi, pi, xj, V3{b.x, b.y, b.z}, a.w, b.w);
        });
    }
    const V3 fi = acc.finish(p, xi);
    store3(f.force, i, fi);
    if (INTEGRATE) integrate(p, f, i, fi, vi, xi);
}

template <class PS>
__global__ void __launch_bounds__(kBlock) k_step_forward(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_linear(kBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    if (is_static_body(p, f, i)) return;
    integrate(p, f, i, load3(f.force, i), load3(f.velocity, i), load3(f.position, i));
}

template <class PS>
__global__ void __launch_bounds__(kBlock) k_plane_constraints(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_linear(kBlock);
    if (!ps.active(i)) return;
    const StepParams &p = ps.at(i);
    const unsigned mask = f.plane_constraints_particles[i];
    if (mask == 0u) return;
    V3 x = load3(f.position, i), v = load3(f.velocity, i);
    for (unsigned c = 0; c < p.nconstraints; ++c) {
        if (!((mask >> c) & 1u)) continue;
        const V3 pt = V3{p.constraints[c][0], p.constraints[c][1], p.constraints[c][2]};
        const V3 nr = V3{p.constraints[c][3], p.constraints[c][4], p.constraints[c][5]};
        x = x - dot(nr, x - pt) * nr;
        v = v - dot(nr, v) * nr;
    }
    store3(f.position, i, x);
    store3(f.velocity, i, v);
}

template <class PS>
__global__ void __launch_bounds__(kBlock) k_init_original_position(PS ps, FieldPtrs f) {
    const unsigned i = ps.slot_linear(kBlock);
    if (!ps.active(i)) return;
    store3(f.originalpos, i, load3(f.position, i));
}

// Deterministic two-level FP64 reduction (fixed grid, fixed tree order).
__global__ void __launch_bounds__(kBlock) k_sums_partial(StepParams p, FieldPtrs f, double *partials) {
    __shared__ double sh[7][kBlock];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const double m = static_cast<double>(p.mass);
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < p.n; i += kReduceBlocks * kBlock) {
        const double vx = f.velocity ? f.velocity[3 * i] : 0.0;
        const double vy = f.velocity ? f.velocity[3 * i + 1] : 0.0;
        const double vz = f.velocity ? f.velocity[3 * i + 2] : 0.0;
        acc[0] += m * vx;
        acc[1] += m * vy;
        acc[2] += m * vz;
        acc[3] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        acc[4] += m * static_cast<double>(f.position[3 * i]);
        acc[5] += m * static_cast<double>(f.position[3 * i + 1]);
        acc[6] += m * static_cast<double>(f.position[3 * i + 2]);
    }
    for (int k = 0; k < 7; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (unsigned w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 7; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; ++k) partials[k * kReduceBlocks + blockIdx.x] = sh[k][0];
}

__global__ void k_sums_final(const double *partials, double *out7, unsigned n, float mass) {
    const unsigned k = threadIdx.x;
    if (k >= 7) return;
    double s = 0.0;
    for (unsigned b = 0; b < kReduceBlocks; ++b) s += partials[k * kReduceBlocks + b];
    if (k >= 4) {
        const double total = static_cast<double>(mass) * static_cast<double>(n);
        s = total > 0.0 ? s / total : 0.0;
    }
    out7[k] = s;
}

inline dim3 grid_for(unsigned n) { return dim3((n + kBlock - 1) / kBlock); }
// Neighbour passes: a multiple of 8 blocks for the XCD-aware slot mapping.
inline dim3 nbr_grid(unsigned n, unsigned block) { return dim3(xcd_grid((n + block - 1) / block)); }

// ---- launchers shared by both parameter sources --------------------------
template <class PS>
hipError_t density(const PS &ps, const FieldPtrs &f, bool original, int compensated, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    const unsigned nb = neighbour_block();
    const dim3 g = ps.launch_blocks(nb, true), b(nb);
    const bool k = (compensated & kKahanDensity) != 0;
    if (original) {
        if (k) hipLaunchKernelGGL((k_density<PS, true, true>), g, b, 0, s, ps, f);
        else hipLaunchKernelGGL((k_density<PS, true, false>), g, b, 0, s, ps, f);
    } else {
        if (k) hipLaunchKernelGGL((k_density<PS, false, true>), g, b, 0, s, ps, f);
        else hipLaunchKernelGGL((k_density<PS, false, false>), g, b, 0, s, ps, f);
    }
    return hipGetLastError();
}

template <class PS>
hipError_t density_listed(const PS &ps, const FieldPtrs &f, const PairCache &c, int compensated, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    const dim3 g = ps.launch_blocks(kListBlock, true), b(kListBlock);
    if (compensated & kKahanDensity) hipLaunchKernelGGL((k_density_listed<PS, true>), g, b, 0, s, ps, f, c);
    else hipLaunchKernelGGL((k_density_listed<PS, false>), g, b, 0, s, ps, f, c);
    return hipGetLastError();
}

template <class PS>
hipError_t density_solid_listed(const PS &ps, const FieldPtrs &f, const PairCache &c, int compensated,
                                hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    const dim3 g = ps.launch_blocks(kListBlock, true), b(kListBlock);
    if (compensated & kKahanDensity) hipLaunchKernelGGL((k_density_solid_listed<PS, true>), g, b, 0, s, ps, f, c);
    else hipLaunchKernelGGL((k_density_solid_listed<PS, false>), g, b, 0, s, ps, f, c);
    return hipGetLastError();
}

template <class PS, bool INTEGRATE>
hipError_t forces_listed_g(const PS &ps, const FieldPtrs &f, const PairCache &c, bool kahan, hipStream_t s) {
    const unsigned nb = neighbour_block();
    const dim3 g = ps.launch_blocks(nb, true), b(nb);
    switch (fluid_pair_group()) {
    case 12:
        if (kahan) hipLaunchKernelGGL((k_forces_fluids_listed<PS, true, 12, INTEGRATE>), g, b, 0, s, ps, f, c);
        else hipLaunchKernelGGL((k_forces_fluids_listed<PS, false, 12, INTEGRATE>), g, b, 0, s, ps, f, c);
        break;
    case 16:
        if (kahan) hipLaunchKernelGGL((k_forces_fluids_listed<PS, true, 16, INTEGRATE>), g, b, 0, s, ps, f, c);
        else hipLaunchKernelGGL((k_forces_fluids_listed<PS, false, 16, INTEGRATE>), g, b, 0, s, ps, f, c);
        break;
    default:
        if (kahan) hipLaunchKernelGGL((k_forces_fluids_listed<PS, true, kPairGroup, INTEGRATE>), g, b, 0, s, ps, f, c);
        else hipLaunchKernelGGL((k_forces_fluids_listed<PS, false, kPairGroup, INTEGRATE>), g, b, 0, s, ps, f, c);
    }
    return hipGetLastError();
}

template <class PS>
hipError_t forces_listed(const PS &ps, const FieldPtrs &f, const PairCache &c, int compensated, bool integrate,
                         hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    const bool kahan = (compensated & kKahanForces) != 0;
    return integrate ? forces_listed_g<PS, true>(ps, f, c, kahan, s) : forces_listed_g<PS, false>(ps, f, c, kahan, s);
}

template <class PS>
hipError_t step_forward(const PS &ps, const FieldPtrs &f, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    hipLaunchKernelGGL((k_step_forward<PS>), ps.launch_blocks(kBlock, false), dim3(kBlock), 0, s, ps, f);
    return hipGetLastError();
}

template <class PS>
hipError_t init_original_position(const PS &ps, const FieldPtrs &f, hipStream_t s) {
    if (!ps.has_work()) return hipSuccess;
    hipLaunchKernelGGL((k_init_original_position<PS>), ps.launch_blocks(kBlock, false), dim3(kBlock), 0, s, ps, f);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_compute_density(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return density(SingleScene{p}, f, false, p.compensated, s);
}

hipError_t launch_compute_density_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                         hipStream_t s) {
    return density_listed(SingleScene{p}, f, c, p.compensated, s);
}

hipError_t launch_forces_fluids_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c, bool integrate,
                                       hipStream_t s) {
    return forces_listed(SingleScene{p}, f, c, p.compensated, integrate, s);
}

hipError_t launch_compute_density_solid_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                               hipStream_t s) {
    return density_solid_listed(SingleScene{p}, f, c, p.compensated, s);
}

hipError_t launch_compute_original_density(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return density(SingleScene{p}, f, true, p.compensated, s);
}

hipError_t launch_forces_fluids(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    const SingleScene ps{p};
    const unsigned nb = neighbour_block();
    if (p.compensated & kKahanForces)
        hipLaunchKernelGGL((k_forces_fluids<SingleScene, true>), nbr_grid(p.n, nb), dim3(nb), 0, s, ps, f);
    else
        hipLaunchKernelGGL((k_forces_fluids<SingleScene, false>), nbr_grid(p.n, nb), dim3(nb), 0, s, ps, f);
    return hipGetLastError();
}

hipError_t launch_step_forward(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return step_forward(SingleScene{p}, f, s);
}

hipError_t launch_plane_constraints(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_plane_constraints<SingleScene>), grid_for(p.n), dim3(kBlock), 0, s, SingleScene{p}, f);
    return hipGetLastError();
}

hipError_t launch_init_original_position(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    return init_original_position(SingleScene{p}, f, s);
}

hipError_t launch_global_sums(const StepParams &p, const FieldPtrs &f, double *partials, double *out7,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_sums_partial, dim3(kReduceBlocks), dim3(kBlock), 0, s, p, f, partials);
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, s, partials, out7, p.n, p.mass);
    return hipGetLastError();
}

hipError_t launch_batch_density_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                       const PairCache &c, int compensated, hipStream_t s) {
    return density_listed(SceneTable{b, sel}, f, c, compensated, s);
}

hipError_t launch_batch_density(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, int compensated,
                                hipStream_t s) {
    return density(SceneTable{b, sel}, f, false, compensated, s);
}

hipError_t launch_batch_density_solid_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                             const PairCache &c, int compensated, hipStream_t s) {
    return density_solid_listed(SceneTable{b, sel}, f, c, compensated, s);
}

hipError_t launch_batch_original_density(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                         int compensated, hipStream_t s) {
    return density(SceneTable{b, sel}, f, true, compensated, s);
}

hipError_t launch_batch_init_original_position(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                               hipStream_t s) {
    return init_original_position(SceneTable{b, sel}, f, s);
}

hipError_t launch_batch_forces_fluids_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                             const PairCache &c, int compensated, hipStream_t s) {
    return forces_listed(SceneTable{b, sel}, f, c, compensated, false, s);
}

hipError_t launch_batch_step_forward(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, hipStream_t s) {
    return step_forward(SceneTable{b, sel}, f, s);
}

}  // namespace sphk
