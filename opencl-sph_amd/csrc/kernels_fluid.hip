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
// the same IEEE operation sequence as the sequential oracle.
#include "neighbours.h"

#pragma clang fp contract(off)

namespace sphk {
namespace {

constexpr unsigned kBlock = 256;

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

// ORIGINAL = compute_original_density (density0, no neighbour count).
template <bool ORIGINAL, bool KAHAN>
__global__ void __launch_bounds__(kBlock) k_density(StepParams p, FieldPtrs f) {
    const unsigned i = xcd_slot(blockDim.x);
    if (i >= p.n) return;
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const V3 xi = load3(pos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned ncombo = 1u << p.nplanes;
    Acc<KAHAN> rho;
    unsigned nn = 0;
    for_each_candidate(p.grid, p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        const V3 xj = load3(pos, j);
        const float r = length(xj - xi);
        if (!(r > kc.h)) rho.add(w_poly6(r, kc));
        if (!ORIGINAL && r < kc.h) ++nn;
        if (inear) {
            const unsigned ok = planes_near_j(p, xj, inear);
            if (ok) {
                for (unsigned b = 1; b < ncombo; ++b) {
                    if (b & ~ok) continue;
                    const float rr = length(reflect_combo(p, b, xj) - xi);
                    if (rr < kc.h) {
                        rho.add(w_poly6(rr, kc));
                        if (!ORIGINAL) ++nn;
                    }
                }
            }
        }
    });
    const float out = rho.value() * p.mass;
    if (ORIGINAL) {
        f.density0[i] = out;
    } else {
        f.density[i] = out;
        if (f.num_neighbours) f.num_neighbours[i] = nn;
    }
}

__device__ __forceinline__ float pressure(const StepParams &p, float rho);

// compute_density for fluid programs, in two phases per lane:
//   1. scan the candidates with the square-root-free range test and compact
//      the in-range slots into this lane's column of an LDS buffer, as 16-bit
//      entries (row number << 12 | offset from the row's first slot);
//   2. evaluate the kernel sum over the compacted slots (every lane busy),
//      writing them out as the pair cache's neighbour list.
// The summation order is the candidate order, as in the reference.  Lanes
// near a mirror plane, with more than kListCap in-range candidates or with a
// row longer than 4096 slots take the reference loop instead.
constexpr unsigned kListBlock = 128;
constexpr unsigned kRowBits = 12;

// First slot of neighbour row r (0..8, (dz, dy) = (r / 3 - 1, r % 3 - 1))
// around the cell (xo, yo, zo), and one past its last slot; false when the
// row lies outside the grid.
__device__ __forceinline__ bool row_range(const GridGeom &g, unsigned n, unsigned xo, unsigned yo, unsigned zo,
                                          unsigned r, const unsigned *__restrict__ celloffset,
                                          const unsigned *__restrict__ gridcount, unsigned *b, unsigned *e) {
    const int zn = static_cast<int>(zo) + static_cast<int>(r / 3u) - 1;
    const int yn = static_cast<int>(yo) + static_cast<int>(r % 3u) - 1;
    if (zn < 0 || zn >= static_cast<int>(g.rz) || yn < 0 || yn >= static_cast<int>(g.ry)) return false;
    const unsigned x0 = xo > 0u ? xo - 1u : 0u;
    const unsigned x1 = xo + 1u < g.rx ? xo + 1u : xo;
    const unsigned row = static_cast<unsigned>(zn) * g.layer + static_cast<unsigned>(yn) * g.rx;
    *b = min(celloffset[row + x0], n);
    *e = min(celloffset[row + x1] + gridcount[row + x1], n);
    return true;
}

template <bool KAHAN>
__global__ void __launch_bounds__(kListBlock) k_density_listed(StepParams p, FieldPtrs f, PairCache c) {
    __shared__ unsigned short buf[kListCap * kListBlock];
    const unsigned i = xcd_slot(kListBlock);
    if (i >= p.n) return;
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const V3 xi = load3(pos, i);
    const unsigned inear = p.nplanes ? planes_near_i(p, xi) : 0u;
    const unsigned key = f.gridcell[i];
    unsigned short *col = buf + threadIdx.x;
    unsigned cnt = 0;
    bool fast = !inear && key < p.grid.ncells;
    unsigned xo = 0, yo = 0, zo = 0;
    if (fast) {
        zo = key / p.grid.layer;
        const unsigned rem = key - p.grid.layer * zo;
        yo = rem / p.grid.rx;
        xo = rem - p.grid.rx * yo;
#pragma unroll 1
        for (unsigned r = 0; r < 9u; ++r) {
            unsigned b, e;
            if (!row_range(p.grid, p.n, xo, yo, zo, r, f.celloffset, f.gridcount, &b, &e)) continue;
            if (e - b > (1u << kRowBits)) fast = false;
            for (unsigned j = b; j < e; ++j) {
                const float s = sqlength(load3(pos, j) - xi);
                if (!(s >= p.r2_out)) {
                    if (cnt < kListCap)
                        col[cnt * kListBlock] = static_cast<unsigned short>((r << kRowBits) | (j - b));
                    ++cnt;
                }
            }
        }
        if (cnt > kListCap) fast = false;
    }
    Acc<KAHAN> rho;
    unsigned nn = 0;
    if (fast) {
        unsigned cur = 0xffu, base = 0, e = 0;
        for (unsigned k = 0; k < cnt; ++k) {
            const unsigned ent = col[k * kListBlock];
            const unsigned r = ent >> kRowBits;
            if (r != cur) {
                cur = r;
                row_range(p.grid, p.n, xo, yo, zo, r, f.celloffset, f.gridcount, &base, &e);
            }
            const unsigned j = base + (ent & ((1u << kRowBits) - 1u));
            const float rr = length(load3(pos, j) - xi);
            rho.add(w_poly6(rr, kc));
            if (rr < kc.h) ++nn;
            c.list[static_cast<size_t>(k) * p.n + i] = j;
        }
    } else {
        const unsigned ncombo = 1u << p.nplanes;
        cnt = 0;
        for_each_candidate(p.grid, p.n, key, f.celloffset, f.gridcount, [&](unsigned j) {
            const V3 xj = load3(pos, j);
            const float r = length(xj - xi);
            if (!(r > kc.h)) {
                rho.add(w_poly6(r, kc));
                if (cnt < kListCap) c.list[static_cast<size_t>(cnt) * p.n + i] = j;
                ++cnt;
            }
            if (r < kc.h) ++nn;
            if (inear) {
                const unsigned ok = planes_near_j(p, xj, inear);
                if (ok)
                    for (unsigned b = 1; b < ncombo; ++b) {
                        if (b & ~ok) continue;
                        const float rr = length(reflect_combo(p, b, xj) - xi);
                        if (rr < kc.h) {
                            rho.add(w_poly6(rr, kc));
                            ++nn;
                        }
                    }
            }
        });
    }
    const float out = rho.value() * p.mass;
    f.density[i] = out;
    if (f.num_neighbours) f.num_neighbours[i] = nn;
    c.count[i] = cnt;
    c.pressure[i] = pressure(p, out);
}

// Equation of state: the reference's linear p = k (rho - rho0)
// (kernels/fluids.cl:28-29), or Tait with exponent 7 when the conf sets
// eos_tait.
__device__ __forceinline__ float pressure(const StepParams &p, float rho) {
    if (!p.eos_tait) return (rho - p.restdens) * p.stiffness;
    const float r = rho / p.restdens;
    const float r2 = r * r, r4 = r2 * r2;
    return p.stiffness * (r4 * r2 * r - 1.0f);
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

template <bool KAHAN>
__global__ void __launch_bounds__(kBlock) k_forces_fluids(StepParams p, FieldPtrs f) {
    const unsigned i = xcd_slot(blockDim.x);
    if (i >= p.n) return;
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const float *__restrict__ vel = f.velocity;
    const float *__restrict__ rho = f.density;
    const V3 xi = load3(pos, i), vi = load3(vel, i);
    const float rhoi = rho[i];
    const float pi = pressure(p, rhoi);
    Acc<KAHAN> fx, fy, fz;
    for_each_candidate(p.grid, p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
        if (j == i) return;
        const V3 xj = load3(pos, j);
        const V3 d = xj - xi;
        const float r = length(d);
        if (r > kc.h) return;
        const V3 vj = load3(vel, j);
        const V3 rv = vj - vi;
        const float rhoj = rho[j];
        const float pj = pressure(p, rhoj);
        const float din = kc.h - r;
        const float didj = rhoi * rhoj;
        const float vterm = (p.viscosity * din) * kc.lap;
        // (p_i + p_j) * 0.5 / r * din * din * spiky, in double
        // (the 0.5 literal promotes it, kernels/fluids.cl:48)
        const float pterm = static_cast<float>(static_cast<double>(pi + pj) * 0.5 /
                                               static_cast<double>(r) * static_cast<double>(din) *
                                               static_cast<double>(din) * static_cast<double>(kc.spiky));
        fx.add((d.x * pterm + rv.x * vterm) / didj);
        fy.add((d.y * pterm + rv.y * vterm) / didj);
        fz.add((d.z * pterm + rv.z * vterm) / didj);
    });
    V3 fi = V3{fx.value(), fy.value(), fz.value()} * p.mass;
    if (p.mass_consistent) fi = fi * p.mass;   // the reference sums m_j only
    fi.x = cube_wall(xi.x, fi.x);
    fi.y = cube_wall(xi.y, fi.y);
    fi.z = cube_wall(xi.z, fi.z);
    store3(f.force, i, fi);
}

// compute_forces_fluids over the pair cache: the in-range list of each slot
// and the pressures of the density pass (identical values to evaluating the
// equation of state per pair).  Overflowed lists fall back to the scan.
template <bool KAHAN>
__global__ void __launch_bounds__(kBlock) k_forces_fluids_listed(StepParams p, FieldPtrs f, PairCache c) {
    const unsigned i = xcd_slot(blockDim.x);
    if (i >= p.n) return;
    const KernelConsts kc = kconsts(p);
    const float *__restrict__ pos = f.position;
    const float *__restrict__ vel = f.velocity;
    const float *__restrict__ rho = f.density;
    const float *__restrict__ prs = c.pressure;
    const V3 xi = load3(pos, i), vi = load3(vel, i);
    const float rhoi = rho[i];
    const float pi = prs[i];
    Acc<KAHAN> fx, fy, fz;
    auto pair_in_range = [&](unsigned j) {
        const V3 xj = load3(pos, j), vj = load3(vel, j);
        const float rhoj = rho[j];
        const float pj = prs[j];
        const V3 d = xj - xi;
        const float r = length(d);
        const V3 rv = vj - vi;
        const float din = kc.h - r;
        const float didj = rhoi * rhoj;
        const float vterm = (p.viscosity * din) * kc.lap;
        const float pterm = static_cast<float>(static_cast<double>(pi + pj) * 0.5 /
                                               static_cast<double>(r) * static_cast<double>(din) *
                                               static_cast<double>(din) * static_cast<double>(kc.spiky));
        fx.add((d.x * pterm + rv.x * vterm) / didj);
        fy.add((d.y * pterm + rv.y * vterm) / didj);
        fz.add((d.z * pterm + rv.z * vterm) / didj);
    };
    auto pair = [&](unsigned j) {
        if (length(load3(pos, j) - xi) > kc.h) return;
        pair_in_range(j);
    };
    const unsigned cnt = c.count[i];
    if (cnt <= kListCap) {
        // List entries are in range by construction (same positions, same
        // distance expression), so the pair terms need no range test here.
        unsigned j = cnt ? c.list[i] : 0u;
        for (unsigned k = 0; k < cnt; ++k) {
            const unsigned jn = k + 1 < cnt ? c.list[static_cast<size_t>(k + 1) * p.n + i] : 0u;
            if (j != i) pair_in_range(j);
            j = jn;
        }
    } else {
        for_each_candidate(p.grid, p.n, f.gridcell[i], f.celloffset, f.gridcount, [&](unsigned j) {
            if (j != i) pair(j);
        });
    }
    V3 fi = V3{fx.value(), fy.value(), fz.value()} * p.mass;
    if (p.mass_consistent) fi = fi * p.mass;
    fi.x = cube_wall(xi.x, fi.x);
    fi.y = cube_wall(xi.y, fi.y);
    fi.z = cube_wall(xi.z, fi.z);
    store3(f.force, i, fi);
}

__device__ __forceinline__ bool is_static_body(const StepParams &p, const FieldPtrs &f, unsigned i) {
    if (!f.body_number) return false;
    const unsigned b = f.body_number[i];
    if (b == 0u || !f.body_dynamic || b > p.body_count) return false;
    return f.body_dynamic[b - 1u] == 0u;
}

__global__ void __launch_bounds__(kBlock) k_step_forward(StepParams p, FieldPtrs f) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || is_static_body(p, f, i)) return;
    const V3 g = V3{p.gravity[0], p.gravity[1], p.gravity[2]};
    const V3 a = load3(f.force, i) / p.mass + g;
    if (f.acceleration) store3(f.acceleration, i, a);
    const V3 v = load3(f.velocity, i);
    const V3 vn = v + p.dt * a;
    const V3 x = load3(f.position, i);
    store3(f.position, i, x + p.dt * vn);
    store3(f.velocity, i, vn);
}

__global__ void __launch_bounds__(kBlock) k_plane_constraints(StepParams p, FieldPtrs f) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
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

__global__ void __launch_bounds__(kBlock) k_init_original_position(StepParams p, FieldPtrs f) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
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

}  // namespace

hipError_t launch_compute_density(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    if (p.compensated & kKahanDensity)
        hipLaunchKernelGGL((k_density<false, true>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    else
        hipLaunchKernelGGL((k_density<false, false>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_compute_density_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                         hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    const dim3 grid = nbr_grid(p.n, kListBlock), block(kListBlock);
    if (p.compensated & kKahanDensity)
        hipLaunchKernelGGL((k_density_listed<true>), grid, block, 0, s, p, f, c);
    else
        hipLaunchKernelGGL((k_density_listed<false>), grid, block, 0, s, p, f, c);
    return hipGetLastError();
}

hipError_t launch_forces_fluids_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                       hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    if (p.compensated & kKahanForces)
        hipLaunchKernelGGL((k_forces_fluids_listed<true>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f, c);
    else
        hipLaunchKernelGGL((k_forces_fluids_listed<false>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f, c);
    return hipGetLastError();
}

hipError_t launch_compute_original_density(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    if (p.compensated & kKahanDensity)
        hipLaunchKernelGGL((k_density<true, true>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    else
        hipLaunchKernelGGL((k_density<true, false>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_forces_fluids(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    if (p.compensated & kKahanForces)
        hipLaunchKernelGGL((k_forces_fluids<true>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    else
        hipLaunchKernelGGL((k_forces_fluids<false>), nbr_grid(p.n, kBlock), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_step_forward(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_step_forward, grid_for(p.n), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_plane_constraints(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plane_constraints, grid_for(p.n), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_init_original_position(const StepParams &p, const FieldPtrs &f, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_original_position, grid_for(p.n), dim3(kBlock), 0, s, p, f);
    return hipGetLastError();
}

hipError_t launch_global_sums(const StepParams &p, const FieldPtrs &f, double *partials, double *out7,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_sums_partial, dim3(kReduceBlocks), dim3(kBlock), 0, s, p, f, partials);
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, s, partials, out7, p.n, p.mass);
    return hipGetLastError();
}

}  // namespace sphk
