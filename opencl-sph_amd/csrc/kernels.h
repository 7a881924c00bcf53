// Launch interface between the HIP runtime (runtime.hip) and the device
// kernels (kernels_*.hip).  Internal to libsph_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sphk {

constexpr int kMaxSymmetryPlanes = 5;      // 2^5 plane combinations fit one mask word
constexpr int kMaxPlaneConstraints = 32;   // bit c of plane_constraints_particles

// Kahan-compensated neighbour sums, per family of passes.
constexpr int kKahanDensity = 1;           // density, rest density
constexpr int kKahanForces = 2;            // forces, A_pq / A_qq, displacement gradient

// Uniform grid: res and the box [lo, hi] per axis.  Cell keys are global:
// the scene's cells are [c0, c0 + ncells) of the cell arrays (c0 = 0 for a
// single scene).
struct GridGeom {
    unsigned rx, ry, rz, layer, ncells;
    unsigned c0;
    float lo[3], hi[3];
};

enum SystemKind : unsigned { kFluid = 0, kCorotational = 1, kSvk = 2 };
constexpr unsigned kSceneSetup = 1u;
// Position and reference-position buffers hold kScanPad slots past the last
// one: the unrolled candidate scans load up to kScanPad - 1 slots beyond a
// row's end (never used) instead of clamping every index.
constexpr unsigned kScanPad = 8;
// Slot byte offsets of the widest per-slot record loaded with a 32-bit
// offset (a 3x3 rotation, 36 bytes) are 32-bit values.
constexpr unsigned kMaxSlots = 0xffffffffu / 36u - kScanPad;

// Everything the per-particle passes read besides the per-particle fields.
// A single psdata is one scene at slots [0, n); a batch holds one StepParams
// per scene, at slots [p0, p0 + n).
struct StepParams {
    unsigned n;               // active particles
    unsigned p0;              // first slot
    unsigned system;          // SystemKind (batches dispatch passes on it)
    unsigned flags;           // kSceneSetup: reference configuration still to compute
    GridGeom grid;
    float h, h2, poly6, spiky, lap;
    // Smallest float s with RN(sqrt(s)) > h: for s = squared distance,
    // "length > h" holds exactly when s >= r2_out (sqrt is correctly rounded
    // and monotonic), so out-of-range candidates need no square root.
    float r2_out;
    float mass, dt;
    float restdens, stiffness, viscosity;
    float gravity[3];
    float youngs, poisson;
    unsigned nplanes;                              // symmetry planes (0: no field)
    unsigned active_planes;                        // bit s: plane s has a non-zero normal
    float planes[kMaxSymmetryPlanes][6];           // point xyz, normal xyz
    unsigned nconstraints;                         // plane constraints
    float constraints[kMaxPlaneConstraints][6];
    unsigned body_count;                           // 0: no body fields
    unsigned body_base;                            // first entry of body_dynamic for this scene
    int eos_tait;                                  // conf field eos_tait: Tait EOS (exponent 7)
    int mass_consistent;                           // conf field mass_consistent: fluid force times m
    int svk_szz_zero;                              // conf field svk_sigma_zz_zero: the reference's sigma_zz
    int compensated;                               // kKahanDensity | kKahanForces
    int polar_fp64;                                // FP64 polar decomposition / inverse
};

// Device pointers of the per-particle fields in the current slot order.
// Optional fields are nullptr when the conf does not define them.
struct FieldPtrs {
    float *position, *velocity, *force, *acceleration;
    float *density, *density0, *originalpos;
    float *rotation, *strain, *stress, *deformation;
    float *apq_temp, *intermediate;
    float *force_elastic, *force_viscous, *cutoff_force;
    unsigned *num_neighbours, *body_number, *plane_constraints_particles;
    const unsigned *body_dynamic;                  // [body_count], not per particle
    unsigned *gridcell;                            // per-slot cell key
    const unsigned *celloffset, *gridcount;        // per cell
};

// ---- launch tuning ------------------------------------------------------------
// Threads per block of the neighbour passes that take any block size of 64,
// 128 or 256 (the list consumers and the candidate scans; the
// LDS-compaction density passes keep their fixed block).  Chosen by the
// autotuner through sph_hip_set_tuning; 256 by default.
constexpr unsigned kMaxNeighbourBlock = 256;
unsigned neighbour_block();
bool set_neighbour_block(unsigned threads);
// List entries whose neighbour loads the fluid force pass issues together
// (8, 12 or 16): more loads in flight against fewer waves per SIMD.
unsigned fluid_pair_group();
bool set_fluid_pair_group(unsigned entries);

// ---- neighbour grid (kernels_grid.hip) ------------------------------------
struct BinScratch {
    unsigned *key;        // [pnum] cell of each old slot
    unsigned *local;      // [pnum] arrival order inside the cell
    unsigned *tmp_id;     // [pnum] particle id at each unsorted cell position
    unsigned *tmp_src;    // [pnum] old slot at each unsorted cell position
    unsigned *tmp_key;    // [pnum] cell at each unsorted cell position
    unsigned *tile_sums;  // [ceil(ncells / kScanTile) + 1]
};
constexpr unsigned kScanTile = 4096;

hipError_t launch_fill_u32(unsigned *p, size_t count, unsigned value, hipStream_t s);
// The same with the count read from device memory (*count, at most
// max_count), as fixed grids: launches whose size does not depend on it.
hipError_t launch_fill_u32_dev(unsigned *p, const unsigned *count, unsigned value, hipStream_t s);
hipError_t launch_exclusive_scan_dev(const unsigned *in, const unsigned *count, unsigned max_count, unsigned *out,
                                     unsigned *tile_sums, hipStream_t s);
hipError_t launch_iota_u32(unsigned *p, size_t count, hipStream_t s);
hipError_t launch_cell_keys(const float *position, unsigned n, const GridGeom &g, unsigned *key_out,
                            hipStream_t s);
hipError_t launch_key_count(const float *position, unsigned n, const GridGeom &g, BinScratch &b,
                            unsigned *gridcount, hipStream_t s);
hipError_t launch_exclusive_scan(const unsigned *in, unsigned count, unsigned *out,
                                 unsigned *tile_sums, hipStream_t s);
hipError_t launch_scatter_unsorted(unsigned n, const BinScratch &b, const unsigned *celloffset,
                                   const unsigned *slot_id, hipStream_t s);
// Stable placement: ascending particle id inside every cell.  Writes the
// gather map (perm[new] = old slot), new slot ids, per-slot cell keys and the
// cellparticles array, and gathers up to two 3-float fields on the way.
hipError_t launch_rank_place(unsigned n, const BinScratch &b, const unsigned *celloffset,
                             const unsigned *gridcount, unsigned *perm, unsigned *slot_id_new,
                             unsigned *gridcell, unsigned *cellparticles, const float *vec_src0,
                             float *vec_dst0, const float *vec_src1, float *vec_dst1,
                             hipStream_t s);

// Field movement between orders.  `words` is the number of 32-bit words per
// particle.  gather: dst[f] = src[map[f]] for f < n, dst[f] = src[f] for
// n <= f < count.  scatter: dst[map[f]] = src[f] likewise.
hipError_t launch_gather_words(const void *src, void *dst, const unsigned *map, unsigned n,
                               unsigned count, unsigned words, hipStream_t s);
hipError_t launch_scatter_words(const void *src, void *dst, const unsigned *map, unsigned n,
                                unsigned count, unsigned words, hipStream_t s);

// ---- per-particle passes (kernels_fluid.hip, kernels_solid.hip) ------------
// Neighbour lists written by the density pass, entries k < count[i] at
// list[k * stride + i] in the reference's visiting order (self included);
// count[i] above the list's capacity (kListCap solid entries, kFluidListMax
// fluid entries) means the list overflowed (or was not built for slot i)
// and readers scan the candidates instead.  A fluid list may also hold
// extra entries of slot i itself, which the force pass skips as it skips
// j == i.
//   kFluidList: the slots j with length(x_j - x_i) <= h, as 16-bit entries
//     (see kFluidPlanes).
//   kSolidList: the slots j with length(x_j - x_i) <= h (flag
//     kListCurrent) or length(x0_j - x0_i) <= h (flag kListReference), the
//     union of the current and reference neighbourhoods that the solid passes
//     iterate; slot = entry & kListSlot.
constexpr unsigned kListCap = 48;
constexpr unsigned kFluidListMax = 64;   // 16-bit entries, two per word of a column
constexpr int kFluidList = 1;
constexpr int kSolidList = 2;
constexpr unsigned kListCurrent = 1u << 31;
constexpr unsigned kListReference = 1u << 30;
constexpr unsigned kListSlot = kListReference - 1u;   // solid lists need slots < 2^30
// Fluid neighbour records, written by the density pass for the listed force
// pass: per slot, float4 planes (position, density), (velocity, equation of
// state of the density) and the slot's list bases (three slots as bits),
// plane k of slot j at float4 index k * stride + j, so that each gather
// reads one 16-byte-strided array.  A fluid list holds 16-bit entries
// (dz group << 14 | offset from that group's base), entries 2m and 2m + 1
// of slot i in the low and high half of list[m * stride + i]: within the
// words a 32-bit list of slot i would use, so that in a batch the fluid and
// solid lists of different slots never share a word.
constexpr unsigned kFluidPlanes = 3;

struct PairCache {
    unsigned *list;        // [kListCap * stride]
    unsigned *count;       // [stride]
    float4 *rec;           // [kFluidPlanes * stride] fluid neighbour records
    unsigned stride;       // slots per list column (pnum, or the batch capacity)
    int valid;             // kFluidList / kSolidList: built from the current state; 0: stale
};

// Per-scene parameters of a batch: slot s belongs to scene scene_of[s]
// (kNoScene for unused slots) and is active when s < p0 + n of that scene
// and the scene has steps left.
constexpr unsigned kNoScene = 0xffffffffu;
// A launch covers either the slots [0, capacity), or, with a chunk map, the
// 64-slot chunks listed in it: chunk k holds slots [chunk_slot[k],
// min(chunk_slot[k] + 64, chunk_end[k])) of one scene (chunk_end = p0 + n),
// so each wavefront works on one scene and a pass for one system launches
// only over that system's scenes, spread evenly over the XCDs.
constexpr unsigned kChunk = 64;
// The part of a pool's layout that changes with admissions and removals.  It
// lives in device memory and every batch pass reads its counts there:
// threads past a count have no work, so a launch sized for at least the
// current work is correct for every layout, and a step sequence captured with
// grids of a given size class replays unchanged for every layout of that
// class.
struct BatchLayout {
    unsigned used_slots;       // slots [0, used_slots) are tiled by segments
    unsigned used_cells;       // cells [0, used_cells) likewise
    unsigned nsegs;            // segments
    unsigned nchunks[3];       // chunk map lengths (fluid, corotational, svk)
};
struct BatchTable {
    const StepParams *scenes;
    const unsigned *scene_of;
    const unsigned *steps_left;   // per scene
    const BatchLayout *layout;    // device
    const unsigned *chunk_slot;   // nullptr: no chunk map (every used slot)
    const unsigned *chunk_end;
    unsigned map;                 // chunk map index into layout->nchunks
    unsigned grid_units;          // host: slots (no chunk map) or chunks the launch covers
    unsigned xcd_segment;         // chunks per segment dealt round-robin to the XCDs; 0: one
                                  // contiguous eighth of the map per XCD (SceneTable::slot_xcd)
};

hipError_t launch_compute_density(const StepParams &p, const FieldPtrs &f, hipStream_t s);
// Density that also fills the pair cache (fluid programs).
hipError_t launch_compute_density_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                         hipStream_t s);
// integrate: also step_forward, in the same kernel (the pass reads every
// neighbour from the density pass's planes).
hipError_t launch_forces_fluids_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c, bool integrate,
                                       hipStream_t s);
// Density that also fills a kSolidList pair cache (solid programs).
hipError_t launch_compute_density_solid_listed(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                               hipStream_t s);
hipError_t launch_compute_original_density(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_forces_fluids(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_step_forward(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_plane_constraints(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_init_original_position(const StepParams &p, const FieldPtrs &f, hipStream_t s);
// The solid neighbour passes iterate `c` when c.valid == kSolidList.
hipError_t launch_rotations_and_strains(const StepParams &p, const FieldPtrs &f, const PairCache &c,
                                        hipStream_t s);
hipError_t launch_stresses_corotational(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_deformations_svk(const StepParams &p, const FieldPtrs &f, const PairCache &c, hipStream_t s);
hipError_t launch_stresses_svk(const StepParams &p, const FieldPtrs &f, hipStream_t s);
hipError_t launch_forces_solids(const StepParams &p, const FieldPtrs &f, const PairCache &c, bool corotational,
                                hipStream_t s);

// ---- batches: the same kernels with per-scene parameters ------------------
// A pass runs for the scenes whose system is in system_mask (bit SystemKind)
// and whose flags contain all of flag_mask.
struct SceneFilter {
    unsigned system_mask;
    unsigned flag_mask;
};
constexpr SceneFilter kAllScenes = {7u, 0u};

// `compensated` is the batch's precision setting (kKahanDensity |
// kKahanForces), shared by all its scenes.
hipError_t launch_batch_density_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                       const PairCache &c, int compensated, hipStream_t s);
hipError_t launch_batch_density(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, int compensated,
                                hipStream_t s);
hipError_t launch_batch_density_solid_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                             const PairCache &c, int compensated, hipStream_t s);
hipError_t launch_batch_original_density(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                         int compensated, hipStream_t s);
hipError_t launch_batch_init_original_position(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                               hipStream_t s);
hipError_t launch_batch_forces_fluids_listed(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                             const PairCache &c, int compensated, hipStream_t s);
hipError_t launch_batch_step_forward(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, hipStream_t s);
hipError_t launch_batch_rotations_and_strains(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                              const PairCache &c, int compensated, int polar_fp64, hipStream_t s);
hipError_t launch_batch_stresses_corotational(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                              hipStream_t s);
hipError_t launch_batch_deformations_svk(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                         const PairCache &c, int compensated, int polar_fp64, hipStream_t s);
hipError_t launch_batch_stresses_svk(const BatchTable &b, SceneFilter sel, const FieldPtrs &f, hipStream_t s);
hipError_t launch_batch_forces_solids(const BatchTable &b, SceneFilter sel, const FieldPtrs &f,
                                      const PairCache &c, bool corotational, int compensated, hipStream_t s);

// FP64 reductions over slots [0, n): momentum (3), kinetic energy, centre of
// mass (3).  `partials` holds 7 * kReduceBlocks doubles.
constexpr unsigned kReduceBlocks = 256;
hipError_t launch_global_sums(const StepParams &p, const FieldPtrs &f, double *partials,
                              double *out7, hipStream_t s);

// Scripted rigid bodies and initial-condition helpers (kernels_bodies.hip).
// Row-major 3x3 matrix passed by value; Box = two corners.
struct Mat3 {
    float m[9];
};
struct Box {
    float lo[3], hi[3];
};
// Lattice of sx*sy*sz points into particles [0, count) (slot = particle).
hipError_t launch_populate_cuboid(float *pos, unsigned count, unsigned sx, unsigned sy, unsigned sz, const Box &c,
                                  hipStream_t s);
// position <- m position for slots [0, n).
hipError_t launch_rotate(float *pos, unsigned n, const Mat3 &m, hipStream_t s);
// Moves the particles of `body` (when `move`), then stores `target` as the
// body's position.
hipError_t launch_set_body_position(float *pos, float *vel, const unsigned *body_number, unsigned n, unsigned body,
                                    bool move, float *body_position, const float *target, float dt,
                                    hipStream_t s);
// Rotates the particles of `body` by the transposed body matrix `bt` (when
// `move`), then stores `angles` as the body's Euler angles.
hipError_t launch_set_body_rotation(float *pos, float *rot, const unsigned *body_number, unsigned n, unsigned body,
                                    bool move, const Mat3 &bt, float *body_euler, const float *angles,
                                    hipStream_t s);

}  // namespace sphk
