// Scene batch: many independent scenes resident in one pool of particle
// slots and grid cells, stepped together by one launch per pass.
//
// Layout.  The pool is a sequence of segments that tile the used slots and
// the used cells in the same order: segment k owns slots [p0, p0 + cap) and
// cells [c0, c0 + cellcap).  A scene's grid is the first ncells cells of its
// segment (global cell key = c0 + the scene's own key, the same key as on its
// own); the segment's last cell is a padding cell whose count is the number
// of the segment's slots that are not binned (unused tail, or all of them
// for a free segment or a scene with no steps left).  One exclusive scan
// over all used cells then gives celloffset[c0] = p0 for every segment, so
// the stable (cell, particle id) sort of the whole pool keeps each scene in
// its own slots, in exactly the order it has on its own.  Every pass reads
// the scene of each slot from the scene table; neighbour ranges are clamped
// to the scene's slots.  A scene therefore computes the same bits in a batch
// as on its own, at the same precision settings.
//
// Stepping.  One step = bin every resident scene (key/count with the
// padding cells, scan, scatter, rank with the gather of the persistent
// per-slot state), then per system, on parallel streams, the density and
// force passes over that system's chunk map, then the integrator and a
// per-scene step counter.  Scenes stop on their own count (a finished scene
// is still binned, which keeps its slots' order, until it is removed); the
// host mirrors the counters, so it knows which scenes are finished without
// reading the device.  The K-step sequence can be captured into a hipGraph
// and replayed while the pool's layout is unchanged.
//
// Counterpart in the reference: none (the reference runs one psdata per
// OpenCL program, src/opencl/particle_system_host.c:361-528); the per-scene
// passes are the reference kernels as in the single-scene path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cells.h"
#include "params.h"
#include "runtime_internal.h"

using sphhip::fail;

namespace sphk {
namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kBodiesPerScene = 8;   // body_dynamic entries a scene may define

struct Segment {
    unsigned p0, cap;       // slots [p0, p0 + cap)
    unsigned c0, cellcap;   // cells [c0, c0 + cellcap); the last is the padding cell
    unsigned scene;         // kNoScene: free
};

// Per-slot state that persists from step to step and moves with the sort.
struct SlotState {
    float *pos, *vel, *opos, *d0, *rot, *strain;
    unsigned *body, *id;
};

__device__ __forceinline__ unsigned linear_slot() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Slot i holds a particle of a resident scene (finished or not).
__device__ __forceinline__ bool live(const SceneTable &ps, unsigned i) {
    if (i >= ps.capacity()) return false;
    const unsigned s = ps.t.scene_of[i];
    if (s == kNoScene) return false;
    const StepParams &p = ps.t.scenes[s];
    return i < p.p0 + p.n;
}
// ... of a scene with steps left: those are sorted; the slots of finished
// scenes keep their order (their outputs stay where they were computed).
__device__ __forceinline__ bool binned(const SceneTable &ps, unsigned i) {
    return live(ps, i) && ps.t.steps_left[ps.t.scene_of[i]] > 0u;
}

// Cell key and arrival rank of every binned slot; threads k < nsegs also
// set the padding cell of segment k to the number of its slots that are not
// binned.  The grid covers at least max(used slots, segments).
__global__ void __launch_bounds__(kBlock) k_batch_key_count(SceneTable ps, const float *__restrict__ pos,
                                                            const Segment *__restrict__ segs,
                                                            unsigned *__restrict__ key,
                                                            unsigned *__restrict__ local,
                                                            unsigned *__restrict__ count) {
    const unsigned s = linear_slot();
    if (s < ps.t.layout->nsegs) {
        const Segment g = segs[s];
        const unsigned used =
            g.scene != kNoScene && ps.t.steps_left[g.scene] > 0u ? ps.t.scenes[g.scene].n : 0u;
        count[g.c0 + g.cellcap - 1u] = g.cap - used;
    }
    const bool active = binned(ps, s);
    unsigned k = 0xffffffffu;
    if (active) {
        const StepParams &p = ps.at(s);
        k = p.grid.c0 + cell_key(pos, s, p.grid);
    }
    const unsigned rank = claim_ranks(k, active, count);
    if (!active) return;
    key[s] = k;
    local[s] = rank;
}

__global__ void __launch_bounds__(kBlock) k_batch_scatter(SceneTable ps, const unsigned *__restrict__ key,
                                                          const unsigned *__restrict__ local,
                                                          const unsigned *__restrict__ celloffset,
                                                          const unsigned *__restrict__ id,
                                                          unsigned *__restrict__ tmp_id,
                                                          unsigned *__restrict__ tmp_src,
                                                          unsigned *__restrict__ tmp_key) {
    const unsigned s = linear_slot();
    if (!binned(ps, s)) return;
    const unsigned k = key[s];
    const unsigned t = celloffset[k] + local[s];
    tmp_id[t] = id[s];
    tmp_src[t] = s;
    tmp_key[t] = k;
}

// Stable placement (ascending particle id inside each cell) and the move of
// the persistent state of each particle into its new slot.  The unsorted
// positions of a scene's particles after the scatter are exactly its slots
// [p0, p0 + n), so `binned` selects them; slots of finished scenes copy
// their state in place.
__global__ void __launch_bounds__(kBlock) k_batch_rank_gather(SceneTable ps, const unsigned *__restrict__ tmp_id,
                                                              const unsigned *__restrict__ tmp_src,
                                                              const unsigned *__restrict__ tmp_key,
                                                              const unsigned *__restrict__ celloffset,
                                                              const unsigned *__restrict__ gridcount,
                                                              unsigned *__restrict__ gridcell, SlotState a,
                                                              SlotState b) {
    const unsigned t = linear_slot();
    if (!live(ps, t)) return;
    unsigned f = t, from = t;
    if (ps.t.steps_left[ps.t.scene_of[t]] > 0u) {
        const unsigned c = tmp_key[t];
        const unsigned start = celloffset[c], cnt = gridcount[c];
        const unsigned me = tmp_id[t];
        f = start + rank_in_cell(tmp_id, start, cnt, me);
        from = tmp_src[t];
        gridcell[f] = c;
    }
    copy_words<1>(a.id, b.id, from, f);
    const unsigned system = ps.at(t).system;
    copy_words<3>(a.pos, b.pos, from, f);
    copy_words<3>(a.vel, b.vel, from, f);
    copy_words<1>(a.body, b.body, from, f);   // static bodies are frozen in every system
    if (system != kFluid) {
        copy_words<3>(a.opos, b.opos, from, f);
        copy_words<1>(a.d0, b.d0, from, f);
    }
    if (system == kCorotational) {
        copy_words<9>(a.rot, b.rot, from, f);
        copy_words<6>(a.strain, b.strain, from, f);
    }
}

// Clears the setup flag of every scene once the setup step is enqueued (the
// device copy of the scene table; the host mirror is cleared alongside).
__global__ void k_batch_clear_setup(unsigned nscenes, StepParams *scenes) {
    const unsigned s = linear_slot();
    if (s < nscenes) scenes[s].flags &= ~kSceneSetup;
}

__global__ void k_batch_tick(unsigned nscenes, unsigned *steps_left) {
    const unsigned s = linear_slot();
    if (s < nscenes && steps_left[s] > 0u) steps_left[s] -= 1u;
}

// dst[id[f]] = src[f] for the scene's slots f in [p0, p0 + n).
__global__ void k_batch_unpermute(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                                  const unsigned *__restrict__ id, unsigned p0, unsigned n, unsigned words) {
    const unsigned k = linear_slot();
    if (k >= n) return;
    const unsigned f = p0 + k;
    const unsigned d = id[f];
    for (unsigned w = 0; w < words; ++w)
        dst[static_cast<size_t>(d) * words + w] = src[static_cast<size_t>(f) * words + w];
}

__global__ void k_batch_iota(unsigned *p, unsigned count) {
    const unsigned k = linear_slot();
    if (k < count) p[k] = k;
}

inline dim3 blocks(size_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

}  // namespace
}  // namespace sphk

using sphk::Segment;
using sphk::SlotState;
using sphk::StepParams;
using sphk::kNoScene;

struct BatchTiming {
    std::string name;
    hipEvent_t start, stop;
};

struct sph_batch {
    hipStream_t stream = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};   // corotational / svk branches of a step
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    int device = 0;
    unsigned slot_cap = 0, cell_cap = 0, scene_cap = 0;
    int compensated = 0, polar_fp64 = 1;
    std::vector<void *> allocations;

    // device, per slot
    SlotState st[2]{};
    int cur = 0;
    float *force = nullptr, *density = nullptr, *stress = nullptr, *deformation = nullptr;
    unsigned *gridcell = nullptr, *scene_of = nullptr;
    sphk::BinScratch bins{};
    sphk::PairCache pairs{};
    void *staging = nullptr;
    // device, per cell / scene / segment
    unsigned *gridcount = nullptr, *celloffset = nullptr;
    StepParams *d_scenes = nullptr;
    unsigned *d_steps_left = nullptr, *body_dynamic = nullptr;
    Segment *d_segs = nullptr;
    sphk::BatchLayout *d_layout = nullptr;   // what the passes read of the layout

    // per-system chunk maps (fluid, corotational, svk): the 64-slot chunks of
    // the running scenes of each, rebuilt on layout changes
    static constexpr int kMaps = 3;
    unsigned *d_chunk_slot[kMaps] = {}, *d_chunk_end[kMaps] = {};
    size_t max_chunks = 0;
    unsigned long long uploaded_version = 0;   // layout version on the device
    sphk::BatchLayout layout{};                // as uploaded

    // host mirror
    std::vector<StepParams> scenes;
    std::vector<unsigned> steps_left;
    std::vector<unsigned> seg_of;          // segment index of each scene (kNoScene: free entry)
    std::vector<Segment> segs;             // in slot (and cell) order
    unsigned used_slots = 0, used_cells = 0;
    unsigned setup_pending = 0;
    unsigned long long layout_version = 1;   // advances on every admission and removal

    // per-pass timing
    bool profiling = false;
    std::map<std::string, std::pair<unsigned long long, double>> stats;   // launches, total ms
    std::vector<BatchTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> stat_names;

    // graph replay of K steps, one captured sequence per state-buffer parity,
    // systems present, set-up step and size class of the work (the passes
    // read the layout's counts from the device, so a sequence replays for
    // every layout of its class)
    unsigned graph_steps = 0;
    std::map<std::array<unsigned, 4>, hipGraphExec_t> graphs;
    unsigned long long graph_launches = 0, graph_captures = 0;
};

namespace {

template <typename T>
int dev_alloc(sph_batch *b, T **p, size_t count) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 4);
    if (hipMalloc(&q, bytes) != hipSuccess) return fail("out of device memory (batch pool, %zu bytes)", bytes);
    b->allocations.push_back(q);
    *p = static_cast<T *>(q);
    return 0;
}

// Device bytes of a pool (what sph_batch_create allocates).
size_t pool_bytes(size_t slots, size_t cells, size_t scenes) {
    const size_t per_slot = 2 * (3 + 3 + 3 + 1 + 9 + 6 + 1 + 1) * 4   // double-buffered state
                            + (3 + 1 + 6 + 9) * 4                      // force, density, stress, deformation
                            + 2 * 4                                    // gridcell, scene_of
                            + 5 * 4                                    // bin scratch
                            + (sphk::kListCap + 1 + 4 * sphk::kFluidPlanes) * 4   // pair cache, fluid records
                            + 9 * 4                                    // readback staging
                            + 1;                                       // chunk maps (3 x 8 B per 64 slots)
    const size_t per_cell = 2 * 4;
    const size_t tiles = (cells + sphk::kScanTile - 1) / sphk::kScanTile + 1;
    const size_t per_scene = sizeof(StepParams) + 4 + sphk::kBodiesPerScene * 4 + sizeof(Segment) * 2 + 4 * 8;
    const size_t scan_pad = 2 * 2 * 3 * 4 * sphk::kScanPad;               // padding of position, reference position
    return slots * per_slot + cells * per_cell + tiles * 4 + scenes * per_scene + scan_pad;
}

void release(sph_batch *b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);   // no graph launch is still running
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    for (BatchTiming &t : b->pending) {
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    for (hipEvent_t e : b->event_pool) hipEventDestroy(e);
    for (hipEvent_t e : {b->ev_fork, b->ev_join[0], b->ev_join[1]})
        if (e) hipEventDestroy(e);
    for (hipStream_t q : b->side)
        if (q) hipStreamDestroy(q);
    for (void *p : b->allocations) hipFree(p);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

// Size class of a work count: launches cover the next power of two (at
// least `floor`), so that a captured sequence serves every count of its class.
unsigned size_class(unsigned units, unsigned floor, unsigned cap) {
    if (units == 0) return 0;
    unsigned c = floor;
    while (c < units && c < (1u << 31)) c <<= 1;
    return std::max(units, std::min(c, cap));
}

unsigned slot_class(const sph_batch *b) {
    return size_class(std::max(b->layout.used_slots, b->layout.nsegs), 4096u,
                      std::max(b->slot_cap, 2 * b->scene_cap + 1));
}

sphk::BatchTable table(const sph_batch *b) {
    return sphk::BatchTable{b->d_scenes, b->scene_of, b->d_steps_left, b->d_layout, nullptr, nullptr, 0,
                            slot_class(b)};
}

enum MapIndex { kMapFluid = 0, kMapCorot = 1, kMapSvk = 2 };

// Chunks every system's passes cover: the size class of the longest map
// (one class for the three, so that few sequences serve every mix).
unsigned chunk_class(const sph_batch *b) {
    const unsigned most = std::max(b->layout.nchunks[0], std::max(b->layout.nchunks[1], b->layout.nchunks[2]));
    return size_class(most, 64u, static_cast<unsigned>(b->max_chunks));
}

// Chunks per XCD segment of the passes that balance their chunk map over
// the XCDs (SceneTable::slot_xcd).
constexpr unsigned kBalancedSegment = 64;

// The scene table restricted to the chunks of one system's running scenes.
sphk::BatchTable table(const sph_batch *b, int map) {
    sphk::BatchTable t = table(b);
    t.chunk_slot = b->d_chunk_slot[map];
    t.chunk_end = b->d_chunk_end[map];
    t.map = static_cast<unsigned>(map);
    t.grid_units = b->layout.nchunks[map] ? chunk_class(b) : 0u;
    return t;
}

// Uploads what the passes read of the layout, when it changed: the
// segments, the per-system chunk maps (the 64-slot chunks of the scenes with
// steps left) and the counts.  Enqueued on the batch stream, so that it
// takes effect after the steps enqueued before it.
//
// The solid maps list their scenes in slot order.  The fluid map lists
// whole scenes grouped into kXcds bins of nearly equal chunk counts (the
// largest scene first, each into the bin with the fewest chunks so far),
// then bin after bin: with its passes dealing 64-chunk segments to the XCDs
// this spreads large and small scenes (a chunk's cost follows its scene's
// size and surface) more evenly than slot order.  Measured on the served
// stream's profiled pass: fluid density 274 -> 235 us, fluid forces 245 ->
// 209 us per launch; the solid passes were 1-2 % slower with the binned
// order and keep slot order.
int upload_layout(sph_batch *b) {
    if (b->uploaded_version == b->layout_version) return 0;
    std::vector<unsigned> slot[sph_batch::kMaps], end[sph_batch::kMaps];
    std::vector<const StepParams *> members[sph_batch::kMaps];
    for (const Segment &g : b->segs) {
        if (g.scene == kNoScene || b->steps_left[g.scene] == 0) continue;
        const StepParams &p = b->scenes[g.scene];
        const int m = p.system == sphk::kFluid ? kMapFluid : p.system == sphk::kCorotational ? kMapCorot : kMapSvk;
        members[m].push_back(&p);
    }
    for (int m = 0; m < sph_batch::kMaps; ++m) {
        std::vector<const StepParams *> &ms = members[m];
        if (m == kMapFluid)
            std::stable_sort(ms.begin(), ms.end(), [](const StepParams *a, const StepParams *c) { return a->n > c->n; });
        std::vector<const StepParams *> bins[sphk::kXcds];
        size_t load[sphk::kXcds] = {};
        for (const StepParams *p : ms) {
            const size_t x =
                m == kMapFluid ? static_cast<size_t>(std::min_element(load, load + sphk::kXcds) - load) : 0u;
            bins[x].push_back(p);
            load[x] += (p->n + sphk::kChunk - 1u) / sphk::kChunk;
        }
        for (const auto &bin : bins)
            for (const StepParams *p : bin)
                for (unsigned c = 0; c < p->n; c += sphk::kChunk) {
                    slot[m].push_back(p->p0 + c);
                    end[m].push_back(p->p0 + p->n);
                }
    }
    sphk::BatchLayout lay{};
    lay.used_slots = b->used_slots;
    lay.used_cells = b->used_cells;
    lay.nsegs = static_cast<unsigned>(b->segs.size());
    for (int m = 0; m < sph_batch::kMaps; ++m) {
        if (slot[m].size() > b->max_chunks) return fail("internal: chunk map overflow");
        lay.nchunks[m] = static_cast<unsigned>(slot[m].size());
        if (slot[m].empty()) continue;
        SPH_HIP_TRY(hipMemcpyAsync(b->d_chunk_slot[m], slot[m].data(), slot[m].size() * 4, hipMemcpyHostToDevice,
                                   b->stream));
        SPH_HIP_TRY(hipMemcpyAsync(b->d_chunk_end[m], end[m].data(), end[m].size() * 4, hipMemcpyHostToDevice,
                                   b->stream));
    }
    if (!b->segs.empty())
        SPH_HIP_TRY(hipMemcpyAsync(b->d_segs, b->segs.data(), b->segs.size() * sizeof(Segment), hipMemcpyHostToDevice,
                                   b->stream));
    SPH_HIP_TRY(hipMemcpyAsync(b->d_layout, &lay, sizeof(lay), hipMemcpyHostToDevice, b->stream));
    SPH_HIP_TRY(hipStreamSynchronize(b->stream));   // the host copies go out of scope
    b->layout = lay;
    b->uploaded_version = b->layout_version;
    return 0;
}

sphk::FieldPtrs field_ptrs(const sph_batch *b) {
    const SlotState &s = b->st[b->cur];
    sphk::FieldPtrs f{};
    f.position = s.pos;
    f.velocity = s.vel;
    f.force = b->force;
    f.density = b->density;
    f.density0 = s.d0;
    f.originalpos = s.opos;
    f.rotation = s.rot;
    f.strain = s.strain;
    f.stress = b->stress;
    f.deformation = b->deformation;
    f.body_number = s.body;
    f.body_dynamic = b->body_dynamic;
    f.gridcell = b->gridcell;
    f.celloffset = b->celloffset;
    f.gridcount = b->gridcount;
    return f;
}

int upload_scene(sph_batch *b, unsigned s) {
    SPH_HIP_TRY(hipMemcpyAsync(b->d_scenes + s, &b->scenes[s], sizeof(StepParams), hipMemcpyHostToDevice,
                               b->stream));
    SPH_HIP_TRY(hipMemcpyAsync(b->d_steps_left + s, &b->steps_left[s], 4, hipMemcpyHostToDevice, b->stream));
    return 0;
}

// Systems with binned scenes: bit SystemKind.
unsigned systems_present(const sph_batch *b) {
    unsigned m = 0;
    for (unsigned s = 0; s < b->scene_cap; ++s)
        if (b->seg_of[s] != kNoScene && b->steps_left[s] > 0) m |= 1u << b->scenes[s].system;
    return m;
}

// ---- per-pass timing (HIP events on the batch stream, eager steps only) ----
hipEvent_t take_event(sph_batch *b) {
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

int drain_timings(sph_batch *b) {
    if (b->pending.empty()) return 0;
    SPH_HIP_TRY(hipEventSynchronize(b->pending.back().stop));
    for (BatchTiming &t : b->pending) {
        float ms = 0.0f;
        SPH_HIP_TRY(hipEventElapsedTime(&ms, t.start, t.stop));
        auto &st = b->stats[t.name];
        st.first += 1;
        st.second += ms;
        b->event_pool.push_back(t.start);
        b->event_pool.push_back(t.stop);
    }
    b->pending.clear();
    return 0;
}

// Runs `launch` (which returns the launch status); brackets it with events
// when profiling.
template <typename F>
int timed(sph_batch *b, const char *name, F &&launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (b->profiling) {
        e0 = take_event(b);
        e1 = take_event(b);
        if (!e0 || !e1) return fail("hipEventCreate failed");
        SPH_HIP_TRY(hipEventRecord(e0, b->stream));
    }
    const hipError_t e = launch();
    if (e != hipSuccess) return fail("%s launch failed: %s", name, hipGetErrorString(e));
    if (b->profiling) {
        SPH_HIP_TRY(hipEventRecord(e1, b->stream));
        b->pending.push_back(BatchTiming{name, e0, e1});
        if (b->pending.size() >= 4096) return drain_timings(b);
    }
    return 0;
}

// Binning of every scene with steps left; swaps the state buffers.
int enqueue_bins(sph_batch *b) {
    hipStream_t s = b->stream;
    const sphk::BatchTable t = table(b);
    const sphk::SceneTable all{t, sphk::kAllScenes};
    const SlotState a = b->st[b->cur];
    const SlotState n = b->st[b->cur ^ 1];
    const dim3 grid = all.launch_blocks(sphk::kBlock, false), block(sphk::kBlock);
    const unsigned *used_cells = &b->d_layout->used_cells;
    if (timed(b, "batch_key_count", [&] {
            hipError_t e = sphk::launch_fill_u32_dev(b->gridcount, used_cells, 0u, s);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(sphk::k_batch_key_count, grid, block, 0, s, all, a.pos, b->d_segs, b->bins.key,
                               b->bins.local, b->gridcount);
            return hipGetLastError();
        }) ||
        timed(b, "batch_scan", [&] {
            return sphk::launch_exclusive_scan_dev(b->gridcount, used_cells, b->cell_cap, b->celloffset,
                                                   b->bins.tile_sums, s);
        }) ||
        timed(b, "batch_scatter", [&] {
            hipLaunchKernelGGL(sphk::k_batch_scatter, grid, block, 0, s, all, b->bins.key, b->bins.local,
                               b->celloffset, a.id, b->bins.tmp_id, b->bins.tmp_src, b->bins.tmp_key);
            return hipGetLastError();
        }) ||
        timed(b, "batch_rank_gather", [&] {
            hipLaunchKernelGGL(sphk::k_batch_rank_gather, grid, block, 0, s, all, b->bins.tmp_id, b->bins.tmp_src,
                               b->bins.tmp_key, b->celloffset, b->gridcount, b->gridcell, a, n);
            return hipGetLastError();
        }))
        return -1;
    b->cur ^= 1;
    return 0;
}

// One step of every scene with steps left.  `systems` selects the passes to
// launch (bit SystemKind); `setup` adds the reference-configuration passes
// for the scenes flagged kSceneSetup (none, in a replayed graph whose
// scenes were all set up before) and clears the flags after them.  The three
// systems' passes touch disjoint slots, so unless per-pass timing is on they
// run on three streams, forked after binning and joined before the
// integrator (parallel branches of a captured graph).
int enqueue_step(sph_batch *b, unsigned systems, bool setup) {
    hipStream_t s = b->stream;
    const bool branch = !b->profiling;
    hipStream_t sc = branch ? b->side[0] : s, sv = branch ? b->side[1] : s;
    const sphk::BatchTable t = table(b);
    const sphk::BatchTable tf = table(b, kMapFluid), tc = table(b, kMapCorot), tv = table(b, kMapSvk);
    // The density passes and the fluid force pass deal 64-chunk segments
    // round-robin to the XCDs (their cost per chunk follows the scene's size
    // and surface, so contiguous eighths of a map left XCDs idle); the solid
    // sweeps and force passes keep contiguous eighths, whose neighbour rows
    // stay in one L2 (measured on the served stream's profiled pass: fluid
    // density 297 -> 273 us, fluid forces 284 -> 247 us, solid density 233
    // -> 224 us per launch, while the corotational force pass lost 12 %
    // with segments).
    sphk::BatchTable tfb = tf, tcb = tc, tvb = tv;
    tfb.xcd_segment = tcb.xcd_segment = tvb.xcd_segment = kBalancedSegment;
    const int kahan = b->compensated, fp64 = b->polar_fp64;
    const unsigned solid_mask = (1u << sphk::kCorotational) | (1u << sphk::kSvk);
    const sphk::SceneFilter fluid{1u << sphk::kFluid, 0u};
    const sphk::SceneFilter corot{1u << sphk::kCorotational, 0u}, svk{1u << sphk::kSvk, 0u};
    const sphk::SceneFilter setup_sel{solid_mask, sphk::kSceneSetup};
    if (setup && timed(b, "init_original_position", [&] {
            return sphk::launch_batch_init_original_position(t, setup_sel, field_ptrs(b), s);
        }))
        return -1;
    if (enqueue_bins(b)) return -1;
    const sphk::FieldPtrs f = field_ptrs(b);
    if (setup && (timed(b, "compute_original_density",
                        [&] { return sphk::launch_batch_original_density(t, setup_sel, f, kahan, s); }) ||
                  timed(b, "batch_clear_setup", [&] {
                      hipLaunchKernelGGL(sphk::k_batch_clear_setup, sphk::blocks(b->scene_cap), dim3(sphk::kBlock), 0,
                                         s, b->scene_cap, b->d_scenes);
                      return hipGetLastError();
                  })))
        return -1;
    const bool has_c = (systems >> sphk::kCorotational) & 1u, has_v = (systems >> sphk::kSvk) & 1u;
    if (branch) {
        SPH_HIP_TRY(hipEventRecord(b->ev_fork, s));
        if (has_c) SPH_HIP_TRY(hipStreamWaitEvent(sc, b->ev_fork, 0));
        if (has_v) SPH_HIP_TRY(hipStreamWaitEvent(sv, b->ev_fork, 0));
    }
    sphk::PairCache solid_pairs = b->pairs;
    solid_pairs.valid = sphk::kSolidList;
    if (has_c &&
        (timed(b, "compute_density_solid",
               [&] { return sphk::launch_batch_density_solid_listed(tcb, corot, f, solid_pairs, kahan, sc); }) ||
         timed(b, "compute_rotations_and_strains",
               [&] { return sphk::launch_batch_rotations_and_strains(tc, corot, f, solid_pairs, kahan, fp64, sc); }) ||
         timed(b, "compute_stresses_corotational",
               [&] { return sphk::launch_batch_stresses_corotational(tc, corot, f, sc); }) ||
         timed(b, "compute_forces_solids_corotational",
               [&] { return sphk::launch_batch_forces_solids(tc, corot, f, solid_pairs, true, kahan, sc); })))
        return -1;
    if (has_v &&
        (timed(b, "compute_density_solid",
               [&] { return sphk::launch_batch_density_solid_listed(tvb, svk, f, solid_pairs, kahan, sv); }) ||
         timed(b, "compute_deformations",
               [&] { return sphk::launch_batch_deformations_svk(tv, svk, f, solid_pairs, kahan, fp64, sv); }) ||
         timed(b, "compute_stresses_svk", [&] { return sphk::launch_batch_stresses_svk(tv, svk, f, sv); }) ||
         timed(b, "compute_forces_solids_svk",
               [&] { return sphk::launch_batch_forces_solids(tv, svk, f, solid_pairs, false, kahan, sv); })))
        return -1;
    if ((systems & (1u << sphk::kFluid)) &&
        (timed(b, "compute_density_fluid",
               [&] { return sphk::launch_batch_density_listed(tfb, fluid, f, b->pairs, kahan, s); }) ||
         timed(b, "compute_forces_fluids",
               [&] { return sphk::launch_batch_forces_fluids_listed(tfb, fluid, f, b->pairs, kahan, s); })))
        return -1;
    if (branch) {
        if (has_c) {
            SPH_HIP_TRY(hipEventRecord(b->ev_join[0], sc));
            SPH_HIP_TRY(hipStreamWaitEvent(s, b->ev_join[0], 0));
        }
        if (has_v) {
            SPH_HIP_TRY(hipEventRecord(b->ev_join[1], sv));
            SPH_HIP_TRY(hipStreamWaitEvent(s, b->ev_join[1], 0));
        }
    }
    if (timed(b, "step_forward", [&] { return sphk::launch_batch_step_forward(t, sphk::kAllScenes, f, s); }) ||
        timed(b, "batch_tick", [&] {
            hipLaunchKernelGGL(sphk::k_batch_tick, sphk::blocks(b->scene_cap), dim3(sphk::kBlock), 0, s,
                               b->scene_cap, b->d_steps_left);
            return hipGetLastError();
        }))
        return -1;
    return 0;
}

void reindex(sph_batch *b) {
    for (unsigned s = 0; s < b->scene_cap; ++s) b->seg_of[s] = kNoScene;
    for (size_t j = 0; j < b->segs.size(); ++j)
        if (b->segs[j].scene != kNoScene) b->seg_of[b->segs[j].scene] = static_cast<unsigned>(j);
    b->layout_version += 1;
}

// Whether alloc_segment would find room (a free segment or the unused tail).
bool segment_fits(const sph_batch *b, unsigned need_slots, unsigned need_cells) {
    for (const Segment &g : b->segs)
        if (g.scene == kNoScene && g.cap >= need_slots && g.cellcap >= need_cells) return true;
    return static_cast<size_t>(b->used_slots) + need_slots <= b->slot_cap &&
           static_cast<size_t>(b->used_cells) + need_cells <= b->cell_cap;
}

int alloc_segment(sph_batch *b, unsigned need_slots, unsigned need_cells, unsigned scene, unsigned *seg_index) {
    // Best fit among free segments (smallest slot count that holds the scene).
    size_t best = b->segs.size();
    for (size_t k = 0; k < b->segs.size(); ++k) {
        const Segment &g = b->segs[k];
        if (g.scene != kNoScene || g.cap < need_slots || g.cellcap < need_cells) continue;
        if (best == b->segs.size() || g.cap < b->segs[best].cap) best = k;
    }
    if (best < b->segs.size()) {
        Segment g = b->segs[best];
        if (g.cellcap - need_cells >= 1u && (g.cap > need_slots || g.cellcap > need_cells)) {
            // split: the scene takes the front, the rest stays free (with its
            // own padding cell)
            Segment rest{g.p0 + need_slots, g.cap - need_slots, g.c0 + need_cells, g.cellcap - need_cells, kNoScene};
            g.cap = need_slots;
            g.cellcap = need_cells;
            b->segs.insert(b->segs.begin() + static_cast<long>(best) + 1, rest);
        }
        g.scene = scene;
        b->segs[best] = g;
    } else {
        if (static_cast<size_t>(b->used_slots) + need_slots > b->slot_cap ||
            static_cast<size_t>(b->used_cells) + need_cells > b->cell_cap)
            return 1;
        b->segs.push_back(Segment{b->used_slots, need_slots, b->used_cells, need_cells, scene});
        b->used_slots += need_slots;
        b->used_cells += need_cells;
        best = b->segs.size() - 1;
    }
    reindex(b);
    *seg_index = static_cast<unsigned>(best);
    return 0;
}

void free_segment(sph_batch *b, unsigned k) {
    b->segs[k].scene = kNoScene;
    // merge with free neighbours
    if (k + 1 < b->segs.size() && b->segs[k + 1].scene == kNoScene) {
        b->segs[k].cap += b->segs[k + 1].cap;
        b->segs[k].cellcap += b->segs[k + 1].cellcap;
        b->segs.erase(b->segs.begin() + static_cast<long>(k) + 1);
    }
    if (k > 0 && b->segs[k - 1].scene == kNoScene) {
        b->segs[k - 1].cap += b->segs[k].cap;
        b->segs[k - 1].cellcap += b->segs[k].cellcap;
        b->segs.erase(b->segs.begin() + static_cast<long>(k));
        k -= 1;
    }
    // a free segment at the end returns to the unused tail
    if (!b->segs.empty() && b->segs.back().scene == kNoScene) {
        b->used_slots -= b->segs.back().cap;
        b->used_cells -= b->segs.back().cellcap;
        b->segs.pop_back();
    }
    reindex(b);
}

struct FieldSlice {
    const char *name;
    int index;          // in the psdata, -1 when absent
    unsigned words;
};

// Field `name` of `data` as `words` 32-bit values per particle, or -1.
int particle_field(const psdata &data, const char *name, unsigned words, unsigned pnum, int *index) {
    *index = get_field_psdata(data, name);
    if (*index < 0) return 0;
    if (data.entry_sizes[*index] != 4 || data.data_sizes[*index] != static_cast<size_t>(pnum) * words * 4)
        return fail("field %s must hold %u float/uint values per particle", name, words);
    return 0;
}

int upload_or_fill(sph_batch *b, const psdata &data, int index, void *dst, unsigned p0, unsigned n,
                   unsigned words, unsigned fill) {
    unsigned *d = static_cast<unsigned *>(dst) + static_cast<size_t>(p0) * words;
    if (index >= 0) {
        const char *src = static_cast<const char *>(data.data) + data.data_offsets[index];
        SPH_HIP_TRY(hipMemcpyAsync(d, src, static_cast<size_t>(n) * words * 4, hipMemcpyHostToDevice, b->stream));
    } else {
        SPH_HIP_TRY(sphk::launch_fill_u32(d, static_cast<size_t>(n) * words, fill, b->stream));
    }
    return 0;
}

// Device buffer holding field `name` for readback, and its words per slot.
bool readback_source(sph_batch *b, const std::string &name, const unsigned **src, unsigned *words) {
    const SlotState &s = b->st[b->cur];
    struct Entry {
        const char *name;
        const void *ptr;
        unsigned words;
    } table[] = {{"position", s.pos, 3},       {"velocity", s.vel, 3},  {"originalpos", s.opos, 3},
                 {"density0", s.d0, 1},        {"rotation", s.rot, 9},  {"strain", s.strain, 6},
                 {"body_number", s.body, 1},   {"force", b->force, 3},  {"density", b->density, 1},
                 {"stress", b->stress, 6},     {"deformation", b->deformation, 9}};
    for (const Entry &e : table)
        if (name == e.name) {
            *src = static_cast<const unsigned *>(e.ptr);
            *words = e.words;
            return true;
        }
    return false;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" size_t sph_batch_device_bytes(unsigned slots, unsigned cells, unsigned scenes) {
    return pool_bytes(slots, cells, scenes);
}

extern "C" int sph_batch_create(unsigned slots, unsigned cells, unsigned scenes, int compensated_sums, int polar_fp64,
                                sph_batch **out) {
    if (!out) return fail("sph_batch_create: out is NULL");
    *out = nullptr;
    if (slots == 0 || cells == 0 || scenes == 0) return fail("sph_batch_create: empty pool");
    if (slots > sphk::kListSlot || slots > sphk::kMaxSlots)
        return fail("sph_batch_create: at most %u slots", std::min(sphk::kListSlot, sphk::kMaxSlots));
    sph_batch *b = new sph_batch();
    if (sphhip::current_device(&b->device)) {
        delete b;
        return -1;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        delete b;
        return fail("hipSetDevice failed");
    }
    b->slot_cap = slots;
    b->cell_cap = cells;
    b->scene_cap = scenes;
    b->compensated = compensated_sums & (sphk::kKahanDensity | sphk::kKahanForces);
    b->polar_fp64 = polar_fp64 ? 1 : 0;
    const size_t S = slots, C = cells;
    const size_t tiles = (C + sphk::kScanTile - 1) / sphk::kScanTile + 1;
    int rc = 0;
    for (int k = 0; k < 2 && !rc; ++k) {
        SlotState &st = b->st[k];
        rc = dev_alloc(b, &st.pos, 3 * (S + sphk::kScanPad)) || dev_alloc(b, &st.vel, 3 * S) ||
             dev_alloc(b, &st.opos, 3 * (S + sphk::kScanPad)) ||
             dev_alloc(b, &st.d0, S) || dev_alloc(b, &st.rot, 9 * S) || dev_alloc(b, &st.strain, 6 * S) ||
             dev_alloc(b, &st.body, S) || dev_alloc(b, &st.id, S);
    }
    unsigned *pair_list = nullptr, *pair_count = nullptr;
    float4 *pair_rec = nullptr;
    rc = rc || dev_alloc(b, &b->force, 3 * S) || dev_alloc(b, &b->density, S) || dev_alloc(b, &b->stress, 6 * S) ||
         dev_alloc(b, &b->deformation, 9 * S) || dev_alloc(b, &b->gridcell, S) || dev_alloc(b, &b->scene_of, S) ||
         dev_alloc(b, &b->bins.key, S) || dev_alloc(b, &b->bins.local, S) ||
         dev_alloc(b, &b->bins.tmp_id, S) || dev_alloc(b, &b->bins.tmp_src, S) || dev_alloc(b, &b->bins.tmp_key, S) ||
         dev_alloc(b, &pair_list, sphk::kListCap * S) || dev_alloc(b, &pair_count, S) ||
         dev_alloc(b, &pair_rec, sphk::kFluidPlanes * S) || dev_alloc(b, reinterpret_cast<unsigned **>(&b->staging), 9 * S) ||
         dev_alloc(b, &b->gridcount, C) || dev_alloc(b, &b->celloffset, C) ||
         dev_alloc(b, &b->bins.tile_sums, tiles) || dev_alloc(b, &b->d_scenes, scenes) ||
         dev_alloc(b, &b->d_steps_left, scenes) || dev_alloc(b, &b->body_dynamic, sphk::kBodiesPerScene * scenes) ||
         dev_alloc(b, &b->d_segs, 2 * static_cast<size_t>(scenes) + 1) ||
         dev_alloc(b, &b->d_layout, 1);
    b->max_chunks = S / sphk::kChunk + scenes;
    for (int m = 0; m < sph_batch::kMaps && !rc; ++m)
        rc = dev_alloc(b, &b->d_chunk_slot[m], b->max_chunks) || dev_alloc(b, &b->d_chunk_end[m], b->max_chunks);
    if (rc) {
        release(b);
        return -1;
    }
    b->pairs = sphk::PairCache{pair_list, pair_count, pair_rec, slots, 1};
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side[1], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[1], hipEventDisableTiming) != hipSuccess) {
        release(b);
        return fail("stream / event creation failed");
    }
    b->scenes.assign(scenes, StepParams{});
    b->steps_left.assign(scenes, 0u);
    b->seg_of.assign(scenes, kNoScene);
    if (sphk::launch_fill_u32(b->scene_of, S, kNoScene, b->stream) != hipSuccess ||
        sphk::launch_fill_u32(b->d_steps_left, scenes, 0u, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        release(b);
        return fail("batch initialisation failed");
    }
    *out = b;
    return 0;
}

extern "C" int sph_batch_free(sph_batch *b) {
    release(b);
    return 0;
}

extern "C" int sph_batch_add(sph_batch *b, psdata *data, const char *kernel_file_list, unsigned steps,
                             int *scene_out) {
    if (!b || !data || !scene_out) return fail("sph_batch_add: null argument");
    *scene_out = -1;
    const sphhip::KernelSet ks = sphhip::parse_file_list(kernel_file_list);
    if (!ks.common) return fail("the kernel file list must include 'common'");
    if (ks.corot && ks.svk) return fail("solid_corotation and solid_svk both define compute_stresses/compute_forces_solids");
    if (!ks.fluids && !ks.corot && !ks.svk) return fail("the kernel file list names no system (fluids, solid_corotation, solid_svk)");
    if ((ks.corot || ks.svk) && !ks.solids) return fail("solid systems need 'solids' in the kernel file list");
    const int pn = get_field_psdata(*data, "pnum"), nn = get_field_psdata(*data, "n");
    if (pn < 0 || nn < 0) return fail("the psdata needs pnum and n fields");
    unsigned pnum, n;
    std::memcpy(&pnum, static_cast<char *>(data->data) + data->data_offsets[pn], 4);
    std::memcpy(&n, static_cast<char *>(data->data) + data->data_offsets[nn], 4);
    if (n == 0 || n > pnum) return fail("n = %u must be in [1, pnum = %u]", n, pnum);
    const bool has_grid = get_field_psdata(*data, "gridres") >= 0 && get_field_psdata(*data, "gridbounds") >= 0 &&
                          get_field_psdata(*data, "gridcount") >= 0;
    if (!has_grid) return fail("the psdata has no grid arrays (smoothingradius, gridbounds and pnum need values)");
    StepParams p;
    if (sphhip::params_from_blob(*data, static_cast<const char *>(data->data), ks, n, true, b->compensated,
                                 b->polar_fp64, &p))
        return -1;
    if (p.grid.rx == 0 || p.grid.ry == 0 || p.grid.rz == 0)
        return fail("grid resolution %ux%ux%u has an empty axis", p.grid.rx, p.grid.ry, p.grid.rz);
    if (p.body_count > sphk::kBodiesPerScene)
        return fail("a batched scene may define at most %u bodies (body_dynamic has %u)", sphk::kBodiesPerScene,
                    p.body_count);
    int fpos, fvel, fbody, frot, fstrain;
    if (particle_field(*data, "position", 3, pnum, &fpos) || particle_field(*data, "velocity", 3, pnum, &fvel) ||
        particle_field(*data, "body_number", 1, pnum, &fbody) || particle_field(*data, "rotation", 9, pnum, &frot) ||
        particle_field(*data, "strain", 6, pnum, &fstrain))
        return -1;
    if (fpos < 0 || fvel < 0) return fail("a batched scene needs position and velocity fields");
    const unsigned need_cells = p.grid.ncells + 1u;
    if (p.grid.ncells > 0xffffffffu - 1u || need_cells > b->cell_cap || n > b->slot_cap)
        return fail("the scene (%u particles, %u cells) is larger than the pool", n, p.grid.ncells);

    unsigned scene = kNoScene;
    for (unsigned s = 0; s < b->scene_cap; ++s)
        if (b->seg_of[s] == kNoScene) {
            scene = s;
            break;
        }
    if (scene == kNoScene) return 1;   // scene table full
    unsigned k;
    const int room = alloc_segment(b, n, need_cells, scene, &k);
    if (room) return room;             // 1: no room now
    const Segment g = b->segs[k];
    p.p0 = g.p0;
    p.grid.c0 = g.c0;
    p.body_base = scene * sphk::kBodiesPerScene;
    p.flags = p.system == sphk::kFluid ? 0u : sphk::kSceneSetup;
    b->scenes[scene] = p;
    b->steps_left[scene] = steps;

    const SlotState &st = b->st[b->cur];
    int rc = upload_or_fill(b, *data, fpos, st.pos, g.p0, n, 3, 0u) ||
             upload_or_fill(b, *data, fvel, st.vel, g.p0, n, 3, 0u) ||
             upload_or_fill(b, *data, fbody, b->st[0].body, g.p0, n, 1, 0u) ||
             upload_or_fill(b, *data, fbody, b->st[1].body, g.p0, n, 1, 0u) ||
             upload_or_fill(b, *data, frot, st.rot, g.p0, n, 9, 0u) ||
             upload_or_fill(b, *data, fstrain, st.strain, g.p0, n, 6, 0u);
    if (!rc && p.body_count > 0) {
        const int bd = get_field_psdata(*data, "body_dynamic");
        const char *src = static_cast<const char *>(data->data) + data->data_offsets[bd];
        if (hipMemcpyAsync(b->body_dynamic + p.body_base, src, p.body_count * 4, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess)
            rc = fail("body_dynamic upload failed");
    }
    if (!rc) {
        hipLaunchKernelGGL(sphk::k_batch_iota, sphk::blocks(n), dim3(sphk::kBlock), 0, b->stream, st.id + g.p0, n);
        if (sphk::launch_fill_u32(b->scene_of + g.p0, g.cap, scene, b->stream) != hipSuccess)
            rc = fail("scene_of fill failed");
    }
    if (!rc) rc = upload_scene(b, scene);
    if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = fail("scene upload failed");
    if (rc) {
        b->steps_left[scene] = 0;
        free_segment(b, b->seg_of[scene]);
        b->seg_of[scene] = kNoScene;
        return -1;
    }
    if (p.flags & sphk::kSceneSetup) b->setup_pending += 1;
    *scene_out = static_cast<int>(scene);
    return 0;
}

extern "C" int sph_batch_can_admit(sph_batch *b, unsigned slots, unsigned cells) {
    if (!b) return fail("null batch");
    bool entry = false;
    for (unsigned s = 0; s < b->scene_cap && !entry; ++s) entry = b->seg_of[s] == kNoScene;
    return entry && slots > 0 && segment_fits(b, slots, cells) ? 1 : 0;
}

extern "C" int sph_batch_remove(sph_batch *b, int scene) {
    if (!b) return fail("null batch");
    if (scene < 0 || static_cast<unsigned>(scene) >= b->scene_cap || b->seg_of[scene] == kNoScene)
        return fail("scene %d is not in the batch", scene);
    const unsigned s = static_cast<unsigned>(scene);
    const Segment g = b->segs[b->seg_of[s]];
    if (b->scenes[s].flags & sphk::kSceneSetup) b->setup_pending -= 1;
    b->steps_left[s] = 0;
    b->scenes[s] = StepParams{};
    SPH_HIP_TRY(sphk::launch_fill_u32(b->scene_of + g.p0, g.cap, kNoScene, b->stream));
    if (upload_scene(b, s)) return -1;
    SPH_HIP_TRY(hipStreamSynchronize(b->stream));
    free_segment(b, b->seg_of[s]);
    return 0;
}

namespace {

void destroy_graphs(sph_batch *b) {
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

}  // namespace

extern "C" int sph_batch_set_graph_steps(sph_batch *b, unsigned steps_per_graph) {
    if (!b) return fail("null batch");
    if (steps_per_graph % 2u != 0u) return fail("steps per graph must be even (the state buffers alternate)");
    if (steps_per_graph == b->graph_steps) return 0;
    SPH_HIP_TRY(hipStreamSynchronize(b->stream));   // no launch of the old graphs is still running
    destroy_graphs(b);
    b->graph_steps = steps_per_graph;
    return 0;
}

namespace {

constexpr size_t kMaxGraphs = 256;   // cached sequences per batch

// Captures graph_steps batch steps from the current buffer parity for the
// given systems, with the set-up passes in the first step when `setup` (they
// act on the scenes flagged kSceneSetup, if any).
int capture_graph(sph_batch *b, unsigned systems, bool setup, hipGraphExec_t *out) {
    const int cur0 = b->cur;
    SPH_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (unsigned k = 0; k < b->graph_steps && !rc; ++k) rc = enqueue_step(b, systems, setup && k == 0);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->cur = cur0;   // an even number of steps: the parity the graph starts from
    if (rc || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc ? -1 : fail("graph capture failed: %s", hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *out = nullptr;
        return fail("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    b->graph_captures += 1;
    return 0;
}

// The cached sequence for the current parity, systems and size classes,
// captured on first use.  With solids present its first step carries the
// set-up passes, which cost a flag test per slot when no scene is pending.
int graph_for(sph_batch *b, unsigned systems, hipGraphExec_t *out) {
    const unsigned solids = (1u << sphk::kCorotational) | (1u << sphk::kSvk);
    const bool setup = (systems & solids) != 0;
    const std::array<unsigned, 4> key = {static_cast<unsigned>(b->cur), systems, slot_class(b), chunk_class(b)};
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return 0;
    }
    if (b->graphs.size() >= kMaxGraphs) {
        SPH_HIP_TRY(hipStreamSynchronize(b->stream));   // none of them is still running
        destroy_graphs(b);
    }
    hipGraphExec_t g = nullptr;
    if (capture_graph(b, systems, setup, &g)) return -1;
    b->graphs[key] = g;
    *out = g;
    return 0;
}

void clear_setup_mirror(sph_batch *b) {
    for (unsigned s = 0; s < b->scene_cap; ++s) b->scenes[s].flags &= ~sphk::kSceneSetup;
    b->setup_pending = 0;
}

}  // namespace

extern "C" int sph_batch_step(sph_batch *b, unsigned num_steps) {
    if (!b) return fail("null batch");
    if (hipSetDevice(b->device) != hipSuccess) return fail("hipSetDevice failed");
    if (upload_layout(b)) return -1;
    unsigned done = 0;
    while (done < num_steps) {
        if (b->used_slots == 0) break;
        const bool setup = b->setup_pending > 0;
        const unsigned systems = systems_present(b);
        if (b->graph_steps > 0 && !b->profiling && num_steps - done >= b->graph_steps) {
            hipGraphExec_t g = nullptr;
            if (graph_for(b, systems, &g)) return -1;
            SPH_HIP_TRY(hipGraphLaunch(g, b->stream));
            b->graph_launches += 1;
            done += b->graph_steps;
        } else {
            if (enqueue_step(b, systems, setup)) return -1;
            done += 1;
        }
        if (setup) clear_setup_mirror(b);
    }
    // mirror of the device step counters for the steps just enqueued
    for (unsigned s = 0; s < b->scene_cap; ++s)
        if (b->seg_of[s] != kNoScene) b->steps_left[s] = b->steps_left[s] > done ? b->steps_left[s] - done : 0u;
    return 0;
}

extern "C" int sph_batch_synchronize(sph_batch *b) {
    if (!b) return fail("null batch");
    SPH_HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int sph_batch_steps_left(sph_batch *b, int scene, unsigned *steps_left) {
    if (!b || !steps_left) return fail("null argument");
    if (scene < 0 || static_cast<unsigned>(scene) >= b->scene_cap || b->seg_of[scene] == kNoScene)
        return fail("scene %d is not in the batch", scene);
    *steps_left = b->steps_left[scene];
    return 0;
}

extern "C" int sph_batch_read(sph_batch *b, int scene, psdata data, size_t num_fields, const char *const *names) {
    if (!b) return fail("null batch");
    if (scene < 0 || static_cast<unsigned>(scene) >= b->scene_cap || b->seg_of[scene] == kNoScene)
        return fail("scene %d is not in the batch", scene);
    const StepParams &p = b->scenes[scene];
    const int pn = get_field_psdata(data, "pnum");
    unsigned pnum = 0;
    if (pn >= 0) std::memcpy(&pnum, static_cast<char *>(data.data) + data.data_offsets[pn], 4);
    if (pnum < p.n) return fail("the psdata holds %u particles, the scene %u", pnum, p.n);
    const unsigned *ids = b->st[b->cur].id;
    for (size_t k = 0; k < num_fields; ++k) {
        if (!names[k]) continue;
        const unsigned *src;
        unsigned words;
        if (!readback_source(b, names[k], &src, &words)) return fail("field %s is not kept by the batch", names[k]);
        const int fi = get_field_psdata(data, names[k]);
        if (fi < 0) return fail("the psdata has no field %s", names[k]);
        if (data.entry_sizes[fi] != 4 || data.data_sizes[fi] != static_cast<size_t>(pnum) * words * 4)
            return fail("field %s must hold %u float/uint values per particle", names[k], words);
        hipLaunchKernelGGL(sphk::k_batch_unpermute, sphk::blocks(p.n), dim3(sphk::kBlock), 0, b->stream, src,
                           static_cast<unsigned *>(b->staging), ids, p.p0, p.n, words);
        SPH_HIP_TRY(hipGetLastError());
        char *dst = static_cast<char *>(data.data) + data.data_offsets[fi];
        SPH_HIP_TRY(hipMemcpyAsync(dst, b->staging, static_cast<size_t>(p.n) * words * 4, hipMemcpyDeviceToHost,
                                   b->stream));
        SPH_HIP_TRY(hipStreamSynchronize(b->stream));
    }
    return 0;
}

extern "C" int sph_batch_set_profiling(sph_batch *b, int enabled) {
    if (!b) return fail("null batch");
    if (!enabled && drain_timings(b)) return -1;
    b->profiling = enabled != 0;
    return 0;
}

extern "C" int sph_batch_reset_kernel_stats(sph_batch *b) {
    if (!b) return fail("null batch");
    if (drain_timings(b)) return -1;
    b->stats.clear();
    return 0;
}

extern "C" int sph_batch_get_kernel_stats(sph_batch *b, int capacity, const char **names,
                                          unsigned long long *launches, double *total_ms) {
    if (!b) return fail("null batch");
    if (drain_timings(b)) return -1;
    b->stat_names.clear();
    for (auto &kv : b->stats) b->stat_names.push_back(kv.first);
    int k = 0;
    for (auto &kv : b->stats) {
        if (k < capacity) {
            if (names) names[k] = b->stat_names[k].c_str();
            if (launches) launches[k] = kv.second.first;
            if (total_ms) total_ms[k] = kv.second.second;
        }
        ++k;
    }
    return k;
}

// out[0..7]: live scenes, scenes with steps left, used slots, used cells,
// slot capacity, cell capacity, graph captures, graph launches.
extern "C" int sph_batch_info(sph_batch *b, unsigned long long *out) {
    if (!b || !out) return fail("null argument");
    unsigned live = 0, running = 0;
    for (unsigned s = 0; s < b->scene_cap; ++s)
        if (b->seg_of[s] != kNoScene) {
            live += 1;
            if (b->steps_left[s] > 0) running += 1;
        }
    out[0] = live;
    out[1] = running;
    out[2] = b->used_slots;
    out[3] = b->used_cells;
    out[4] = b->slot_cap;
    out[5] = b->cell_cap;
    out[6] = b->graph_captures;
    out[7] = b->graph_launches;
    return 0;
}
