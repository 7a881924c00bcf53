// HIP runtime of libsph_hip.so: device selection, the per-psdata device
// handle, host<->device sync, the binning sequence, named-kernel dispatch and
// the multi-step fast path.
//
// Replaces src/opencl/particle_system_host.c.  Differences in design:
//   * every field has its own device allocation (256-byte aligned) instead of
//     one mirrored blob, and per-particle fields touched by the kernels are
//     kept in cell-sorted slot order, double-buffered so a reorder is one
//     gather; psdata (host) order is restored at every download;
//   * a reorder only moves the fields that are read before they are
//     rewritten: each such field carries the permutation epoch its contents
//     are in and is brought up to date on first use;
//   * launches are asynchronous on one stream; a sync or an explicit
//     synchronize_hip waits;
//   * errors are returned as status codes with a message, never asserted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "host_common.h"
#include "kernels.h"
#include "runtime_internal.h"

using sphhip::FieldKind;
using sphhip::KernelSet;
using sphhip::fail;

namespace {

thread_local std::string t_error;

#define HIP_TRY(expr) SPH_HIP_TRY(expr)

bool g_ready = false;
int g_device = 0;

enum class Access { Read, ReadWrite, WriteAll, WriteSome };

struct DevField {
    std::string name;
    int index = -1;
    FieldKind kind = FieldKind::Other;
    size_t bytes = 0;
    unsigned words = 0;          // 32-bit words per particle (Particle kind)
    bool hot = false;            // kept in slot order, double-buffered
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;
    unsigned long long epoch = 0;
    void *ptr() const { return buf[cur]; }
};

struct KernelStat {
    unsigned long long launches = 0;
    double total_ms = 0.0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};

}  // namespace

struct psdata_hip {
    psdata host{};
    hipStream_t stream = nullptr;
    std::vector<DevField> fields;
    std::map<std::string, int> by_name;
    unsigned pnum = 0, n = 0;
    bool has_grid = false;
    KernelSet kset;
    std::vector<char> shadow;               // device values of the non-particle fields
    unsigned *slot_id[2] = {nullptr, nullptr};
    int slot_cur = 0;
    unsigned *perm = nullptr;               // perm[new slot] = slot one epoch earlier
    unsigned long long epoch = 1;
    sphk::BinScratch bins{};
    void *staging = nullptr;
    size_t staging_bytes = 0;
    double *sums = nullptr;                 // partials + 7 outputs
    sphk::PairCache pairs{};                // neighbour lists + pressures (fluid programs)
    int compensated = 0, polar_fp64 = 1;
    unsigned long long bins_epoch = 1;      // slot-order epoch the grid arrays index
    bool profiling = false;
    std::map<std::string, KernelStat> stats;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> stat_names;    // storage behind get_kernel_stats_hip names

    DevField *field(const char *name) {
        auto it = by_name.find(name);
        return it == by_name.end() ? nullptr : &fields[it->second];
    }
    unsigned *slot_ids() const { return slot_id[slot_cur]; }
};

namespace {

int build_params(psdata_hip *h, sphk::StepParams *pp) {
    return sphhip::params_from_blob(h->host, h->shadow.data(), h->kset, h->n, h->has_grid, h->compensated,
                                    h->polar_fp64, pp);
}

template <typename T>
T *fptr(psdata_hip *h, const char *name) {
    DevField *f = h->field(name);
    return f ? static_cast<T *>(f->ptr()) : nullptr;
}

sphk::FieldPtrs field_ptrs(psdata_hip *h) {
    sphk::FieldPtrs f{};
    f.position = fptr<float>(h, "position");
    f.velocity = fptr<float>(h, "velocity");
    f.force = fptr<float>(h, "force");
    f.acceleration = fptr<float>(h, "acceleration");
    f.density = fptr<float>(h, "density");
    f.density0 = fptr<float>(h, "density0");
    f.originalpos = fptr<float>(h, "originalpos");
    f.rotation = fptr<float>(h, "rotation");
    f.strain = fptr<float>(h, "strain");
    f.stress = fptr<float>(h, "stress");
    f.deformation = fptr<float>(h, "deformation");
    f.apq_temp = fptr<float>(h, "apq_temp");
    f.intermediate = fptr<float>(h, "intermediate");
    f.force_elastic = fptr<float>(h, "force_elastic");
    f.force_viscous = fptr<float>(h, "force_viscous");
    f.cutoff_force = fptr<float>(h, "cutoff_force");
    f.num_neighbours = fptr<unsigned>(h, "num_neighbours");
    f.body_number = fptr<unsigned>(h, "body_number");
    f.plane_constraints_particles = fptr<unsigned>(h, "plane_constraints_particles");
    f.body_dynamic = fptr<unsigned>(h, "body_dynamic");
    f.gridcell = fptr<unsigned>(h, "gridcell");
    f.celloffset = fptr<unsigned>(h, "celloffset");
    f.gridcount = fptr<unsigned>(h, "gridcount");
    return f;
}

// ---- profiling ---------------------------------------------------------------
hipEvent_t take_event(psdata_hip *h) {
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int drain_timings(psdata_hip *h) {
    if (h->pending.empty()) return 0;
    HIP_TRY(hipEventSynchronize(h->pending.back().stop));
    for (PendingTiming &t : h->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t.start, t.stop));
        KernelStat &s = h->stats[t.name];
        s.launches += 1;
        s.total_ms += ms;
        h->event_pool.push_back(t.start);
        h->event_pool.push_back(t.stop);
    }
    h->pending.clear();
    return 0;
}

template <typename F>
int timed(psdata_hip *h, const char *name, F &&launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (h->profiling) {
        a = take_event(h);
        b = take_event(h);
        if (!a || !b) return fail("hipEventCreate failed");
        HIP_TRY(hipEventRecord(a, h->stream));
    }
    hipError_t e = launch();
    if (e != hipSuccess) return fail("%s launch failed: %s", name, hipGetErrorString(e));
    if (h->profiling) {
        HIP_TRY(hipEventRecord(b, h->stream));
        h->pending.push_back(PendingTiming{name, a, b});
        if (h->pending.size() >= 4096) return drain_timings(h);
    }
    return 0;
}

// ---- permutation bookkeeping ----------------------------------------------
// Bring a hot field into the current slot order.
int ensure_current(psdata_hip *h, DevField *f) {
    if (!f || !f->hot || f->epoch == h->epoch) return 0;
    if (f->epoch + 1 != h->epoch) return fail("internal: field %s is %llu epochs stale", f->name.c_str(),
                                             h->epoch - f->epoch);
    void *dst = f->buf[f->cur ^ 1];
    if (timed(h, "reorder_field", [&] {
            return sphk::launch_gather_words(f->ptr(), dst, h->perm, h->n, h->pnum, f->words, h->stream);
        }))
        return -1;
    f->cur ^= 1;
    f->epoch = h->epoch;
    return 0;
}

// Put every hot field back into particle-index order and make the slot map
// the identity (used when n changes).
int reset_permutation(psdata_hip *h) {
    for (DevField &f : h->fields) {
        if (!f.hot) continue;
        if (ensure_current(h, &f)) return -1;
        void *dst = f.buf[f.cur ^ 1];
        if (timed(h, "unpermute_field", [&] {
                return sphk::launch_scatter_words(f.ptr(), dst, h->slot_ids(), h->n, h->pnum, f.words,
                                                  h->stream);
            }))
            return -1;
        f.cur ^= 1;
        HIP_TRY(hipMemcpyAsync(f.buf[f.cur ^ 1], f.ptr(), f.bytes, hipMemcpyDeviceToDevice, h->stream));
    }
    for (int k = 0; k < 2; ++k) HIP_TRY(sphk::launch_iota_u32(h->slot_id[k], h->pnum, h->stream));
    h->epoch += 1;
    for (DevField &f : h->fields)
        if (f.hot) f.epoch = h->epoch;
    return 0;
}

// ---- binning ---------------------------------------------------------------
int compute_bins(psdata_hip *h) {
    h->pairs.valid = 0;
    if (!h->has_grid) return fail("the psdata has no grid arrays (smoothingradius, gridbounds and pnum need values)");
    DevField *pos = h->field("position");
    if (!pos || !pos->hot) return fail("compute_particle_bins needs a per-particle position field");
    DevField *gcell = h->field("gridcell");
    DevField *gcount = h->field("gridcount");
    DevField *coff = h->field("celloffset");
    DevField *cparts = h->field("cellparticles");
    // Fields one epoch behind would fall two behind: update them first.
    for (DevField &f : h->fields)
        if (f.hot && ensure_current(h, &f)) return -1;
    sphk::StepParams p;
    if (build_params(h, &p)) return -1;
    const unsigned n = h->n;
    unsigned *count = static_cast<unsigned *>(gcount->ptr());
    unsigned *offs = static_cast<unsigned *>(coff->ptr());
    if (timed(h, "bin_key_count", [&] {
            return sphk::launch_key_count(static_cast<const float *>(pos->ptr()), n, p.grid, h->bins, count,
                                          h->stream);
        }))
        return -1;
    if (timed(h, "bin_scan", [&] {
            return sphk::launch_exclusive_scan(count, p.grid.ncells, offs, h->bins.tile_sums, h->stream);
        }))
        return -1;
    if (timed(h, "bin_scatter", [&] {
            return sphk::launch_scatter_unsorted(n, h->bins, offs, h->slot_ids(), h->stream);
        }))
        return -1;
    DevField *vel = h->field("velocity");
    const bool move_vel = vel && vel->hot;
    if (timed(h, "bin_rank_place", [&] {
            return sphk::launch_rank_place(
                n, h->bins, offs, count, h->perm, h->slot_id[h->slot_cur ^ 1],
                static_cast<unsigned *>(gcell->ptr()), static_cast<unsigned *>(cparts->ptr()),
                static_cast<const float *>(pos->ptr()), static_cast<float *>(pos->buf[pos->cur ^ 1]),
                move_vel ? static_cast<const float *>(vel->ptr()) : nullptr,
                move_vel ? static_cast<float *>(vel->buf[vel->cur ^ 1]) : nullptr, h->stream);
        }))
        return -1;
    h->slot_cur ^= 1;
    h->epoch += 1;
    pos->cur ^= 1;
    pos->epoch = h->epoch;
    if (move_vel) {
        vel->cur ^= 1;
        vel->epoch = h->epoch;
    }
    gcell->epoch = h->epoch;   // written in the new order in place
    h->bins_epoch = h->epoch;
    return 0;
}

// ---- named kernels -------------------------------------------------------
struct FieldUse {
    const char *name;
    Access access;
    bool required;
};

struct KernelSpec {
    const char *name;
    bool neighbours;
    std::vector<FieldUse> uses;
};

bool has(psdata_hip *h, const char *name) { return h->field(name) != nullptr; }

// 32-bit words per particle of every field the kernels index by particle.
int check_field_shapes(psdata_hip *h) {
    static const struct {
        const char *name;
        unsigned words;
    } shapes[] = {{"position", 3},      {"velocity", 3},     {"force", 3},          {"acceleration", 3},
                  {"originalpos", 3},   {"force_elastic", 3}, {"force_viscous", 3}, {"cutoff_force", 3},
                  {"density", 1},       {"density0", 1},     {"num_neighbours", 1}, {"body_number", 1},
                  {"plane_constraints_particles", 1},        {"gridcell", 1},       {"rotation", 9},
                  {"apq_temp", 9},      {"intermediate", 9}, {"deformation", 9},    {"strain", 6},
                  {"stress", 6}};
    for (const auto &s : shapes) {
        DevField *f = h->field(s.name);
        if (!f) continue;
        if (f->kind != FieldKind::Particle || f->words != s.words || h->host.entry_sizes[f->index] != 4)
            return fail("field %s must hold %u float/uint values per particle (last dimension pnum)", s.name,
                        s.words);
    }
    return 0;
}

int check_and_prepare(psdata_hip *h, const KernelSpec &k) {
    if (k.neighbours && !h->has_grid) return fail("%s needs the grid arrays", k.name);
    if (k.neighbours && h->bins_epoch != h->epoch)
        return fail("%s: the particle order changed since the last compute_particle_bins (n changed or the "
                    "positions were re-placed); call compute_particle_bins first", k.name);
    for (const FieldUse &u : k.uses) {
        DevField *f = h->field(u.name);
        if (!f) {
            if (u.required) return fail("%s needs field '%s', which the psdata does not define", k.name, u.name);
            continue;
        }
        if (f->kind == FieldKind::Particle && !f->hot)
            return fail("internal: field %s used by %s is not device-ordered", u.name, k.name);
        if (u.access != Access::WriteAll && ensure_current(h, f)) return -1;
    }
    if (k.neighbours && ensure_current(h, h->field("gridcell"))) return -1;
    return 0;
}

void mark_written(psdata_hip *h, const KernelSpec &k) {
    for (const FieldUse &u : k.uses)
        if (u.access == Access::WriteAll)
            if (DevField *f = h->field(u.name)) f->epoch = h->epoch;
}

// Field uses of every named kernel (from the USE_FIELD lists of kernels/*.cl).
KernelSpec spec_for(psdata_hip *h, const std::string &name) {
    const Access partial_if_bodies = has(h, "body_number") ? Access::WriteSome : Access::WriteAll;
    if (name == "compute_density")
        return {"compute_density", true,
                {{"position", Access::Read, true}, {"originalpos", Access::Read, false},
                 {"density", Access::WriteAll, true}, {"num_neighbours", Access::WriteAll, false}}};
    if (name == "compute_forces_fluids")
        return {"compute_forces_fluids", true,
                {{"position", Access::Read, true}, {"velocity", Access::Read, true},
                 {"density", Access::Read, true}, {"force", Access::WriteAll, true}}};
    if (name == "step_forward")
        return {"step_forward", false,
                {{"force", Access::Read, true}, {"velocity", Access::ReadWrite, true},
                 {"position", Access::ReadWrite, true}, {"body_number", Access::Read, false},
                 {"acceleration", partial_if_bodies, false}}};
    if (name == "apply_plane_constraints")
        return {"apply_plane_constraints", false,
                {{"plane_constraints_particles", Access::Read, true}, {"plane_constraints", Access::Read, true},
                 {"position", Access::ReadWrite, true}, {"velocity", Access::ReadWrite, true}}};
    if (name == "init_original_position")
        return {"init_original_position", false,
                {{"position", Access::Read, true}, {"originalpos", Access::WriteAll, true}}};
    if (name == "compute_original_density")
        return {"compute_original_density", true,
                {{"position", Access::Read, true}, {"density0", Access::WriteAll, true}}};
    if (name == "compute_rotations_and_strains")
        return {"compute_rotations_and_strains", true,
                {{"position", Access::Read, true}, {"originalpos", Access::Read, true},
                 {"density0", Access::Read, true}, {"body_number", Access::Read, false},
                 {"rotation", partial_if_bodies, true}, {"strain", partial_if_bodies, true},
                 {"apq_temp", partial_if_bodies, false}, {"intermediate", partial_if_bodies, false}}};
    if (name == "compute_stresses_corotational")
        return {"compute_stresses", false, {{"strain", Access::Read, true}, {"stress", Access::WriteAll, true}}};
    if (name == "compute_deformations")
        return {"compute_deformations", true,
                {{"position", Access::Read, true}, {"originalpos", Access::Read, true},
                 {"deformation", Access::WriteAll, true}}};
    if (name == "compute_stresses_svk")
        return {"compute_stresses", false,
                {{"deformation", Access::Read, true}, {"stress", Access::WriteAll, true}}};
    if (name == "compute_forces_solids_corotational" || name == "compute_forces_solids_svk") {
        KernelSpec k{"compute_forces_solids", true,
                     {{"position", Access::Read, true}, {"velocity", Access::Read, true},
                      {"originalpos", Access::Read, true}, {"stress", Access::Read, true},
                      {"density0", Access::Read, true}, {"density", Access::Read, true},
                      {"body_number", Access::Read, false}, {"force", Access::WriteAll, true},
                      {"force_elastic", Access::WriteAll, false}, {"force_viscous", Access::WriteAll, false},
                      {"cutoff_force", Access::WriteAll, false}}};
        if (name == "compute_forces_solids_corotational") k.uses.push_back({"rotation", Access::Read, true});
        return k;
    }
    if (name == "find_particle_bins")
        return {"find_particle_bins", false,
                {{"position", Access::Read, true}, {"gridcell", Access::WriteAll, true}}};
    return {nullptr, false, {}};
}

// Resolve a reference kernel name against the kernel files of the handle.
int resolve_kernel(psdata_hip *h, const char *name, std::string *internal) {
    const std::string n(name);
    const KernelSet &k = h->kset;
    auto need = [&](bool ok, const char *file) {
        if (!ok) return fail("kernel %s is not in this program (add '%s' to the kernel file list)", name, file);
        *internal = n;
        return 0;
    };
    if (n == "compute_density" || n == "step_forward" || n == "find_particle_bins") return need(k.common, "common");
    if (n == "compute_forces_fluids") return need(k.fluids, "fluids");
    if (n == "init_original_position" || n == "compute_original_density") return need(k.solids, "solids");
    if (n == "compute_rotations_and_strains") return need(k.corot, "solid_corotation");
    if (n == "compute_deformations") return need(k.svk, "solid_svk");
    if (n == "apply_plane_constraints") return need(k.constraints, "constraints");
    if (n == "compute_stresses" || n == "compute_forces_solids") {
        if (!k.corot && !k.svk) return fail("kernel %s needs solid_corotation or solid_svk in the kernel file list", name);
        *internal = n + (k.corot ? "_corotational" : "_svk");
        return 0;
    }
    if (n == "set_body_position" || n == "set_body_rotation" || n == "populate_position_cuboid" ||
        n == "rotate_particles")
        return fail("kernel %s takes arguments: call %s_device_hip", name, name);
    if (n == "zero_gridcount" || n == "count_particles_in_bins" || n == "prefix_sum" ||
        n == "copy_celloffset_to_backup" || n == "insert_particles_in_bin_array")
        return fail("kernel %s is one stage of compute_particle_bins; call that instead", name);
    return fail("unknown kernel '%s'", name);
}

int run_kernel(psdata_hip *h, const std::string &internal) {
    KernelSpec k = spec_for(h, internal);
    if (!k.name) return fail("unknown kernel '%s'", internal.c_str());
    if (check_and_prepare(h, k)) return -1;
    sphk::StepParams p;
    if (build_params(h, &p)) return -1;
    sphk::FieldPtrs f = field_ptrs(h);
    hipStream_t s = h->stream;
    int rc = 0;
    // Neighbour lists: fluid programs list r <= h and cache pressures; solid
    // programs list the union of current and reference neighbourhoods (needs
    // the reference positions).  A list stays valid until positions,
    // reference positions or the slot order change.
    const bool solid_program = !h->kset.fluids && (h->kset.solids || h->kset.corot || h->kset.svk);
    const bool listed = h->pairs.list != nullptr && (h->kset.fluids || (solid_program && h->field("originalpos")));
    if (internal == "compute_density" || internal == "step_forward" || internal == "apply_plane_constraints" ||
        internal == "init_original_position" || internal == "find_particle_bins")
        h->pairs.valid = 0;
    const sphk::PairCache cache = h->pairs;
    if (internal == "compute_density" && listed && h->kset.fluids)
        rc = timed(h, "compute_density",
                   [&] { return sphk::launch_compute_density_listed(p, f, cache, s); });
    else if (internal == "compute_density" && listed)
        rc = timed(h, "compute_density",
                   [&] { return sphk::launch_compute_density_solid_listed(p, f, cache, s); });
    else if (internal == "compute_density")
        rc = timed(h, "compute_density", [&] { return sphk::launch_compute_density(p, f, s); });
    else if (internal == "compute_forces_fluids" && cache.valid == sphk::kFluidList)
        rc = timed(h, "compute_forces_fluids",
                   [&] { return sphk::launch_forces_fluids_listed(p, f, cache, false, s); });
    else if (internal == "compute_forces_fluids")
        rc = timed(h, "compute_forces_fluids", [&] { return sphk::launch_forces_fluids(p, f, s); });
    else if (internal == "step_forward")
        rc = timed(h, "step_forward", [&] { return sphk::launch_step_forward(p, f, s); });
    else if (internal == "apply_plane_constraints")
        rc = timed(h, "apply_plane_constraints", [&] { return sphk::launch_plane_constraints(p, f, s); });
    else if (internal == "init_original_position")
        rc = timed(h, "init_original_position", [&] { return sphk::launch_init_original_position(p, f, s); });
    else if (internal == "compute_original_density")
        rc = timed(h, "compute_original_density", [&] { return sphk::launch_compute_original_density(p, f, s); });
    else if (internal == "compute_rotations_and_strains")
        rc = timed(h, "compute_rotations_and_strains",
                   [&] { return sphk::launch_rotations_and_strains(p, f, cache, s); });
    else if (internal == "compute_stresses_corotational")
        rc = timed(h, "compute_stresses_corotational", [&] { return sphk::launch_stresses_corotational(p, f, s); });
    else if (internal == "compute_deformations")
        rc = timed(h, "compute_deformations", [&] { return sphk::launch_deformations_svk(p, f, cache, s); });
    else if (internal == "compute_stresses_svk")
        rc = timed(h, "compute_stresses_svk", [&] { return sphk::launch_stresses_svk(p, f, s); });
    else if (internal == "compute_forces_solids_corotational")
        rc = timed(h, "compute_forces_solids_corotational",
                   [&] { return sphk::launch_forces_solids(p, f, cache, true, s); });
    else if (internal == "compute_forces_solids_svk")
        rc = timed(h, "compute_forces_solids_svk",
                   [&] { return sphk::launch_forces_solids(p, f, cache, false, s); });
    else if (internal == "find_particle_bins")
        rc = timed(h, "find_particle_bins", [&] {
            return sphk::launch_cell_keys(f.position, p.n, p.grid, f.gridcell, s);
        });
    else
        return fail("unknown kernel '%s'", internal.c_str());
    if (rc) return -1;
    mark_written(h, k);
    if (internal == "compute_density" && listed) h->pairs.valid = h->kset.fluids ? sphk::kFluidList : sphk::kSolidList;
    return 0;
}

// compute_forces_fluids then step_forward as one kernel, over a valid fluid
// pair cache: the force pass reads every neighbour from the density pass's
// planes, so it can move each particle in place once its force is known.
// Field bookkeeping as for the two kernels run one after the other.
int run_forces_and_step(psdata_hip *h) {
    const KernelSpec kf = spec_for(h, "compute_forces_fluids"), ks = spec_for(h, "step_forward");
    if (check_and_prepare(h, kf)) return -1;
    mark_written(h, kf);   // force: written in full by the launch below
    if (check_and_prepare(h, ks)) return -1;
    sphk::StepParams p;
    if (build_params(h, &p)) return -1;
    const sphk::FieldPtrs f = field_ptrs(h);
    const sphk::PairCache cache = h->pairs;
    h->pairs.valid = 0;   // the particles move
    if (timed(h, "compute_forces_fluids+step_forward",
              [&] { return sphk::launch_forces_fluids_listed(p, f, cache, true, h->stream); }))
        return -1;
    mark_written(h, ks);
    return 0;
}

// ---- device allocation -------------------------------------------------------
bool hot_field(const std::string &name, const KernelSet &k) {
    static const char *const common[] = {"position", "gridcell", "density", "num_neighbours", "force",
                                         "velocity", "acceleration", "body_number"};
    static const char *const solids[] = {"originalpos", "density0"};
    static const char *const corot[] = {"rotation", "strain", "stress", "apq_temp", "intermediate",
                                        "force_elastic", "force_viscous", "cutoff_force"};
    static const char *const svk[] = {"deformation", "stress", "force_elastic", "force_viscous", "cutoff_force"};
    static const char *const cons[] = {"plane_constraints_particles"};
    static const char *const rigid[] = {"rotation"};
    for (const char *c : common) if (name == c) return true;
    if (k.solids || k.corot || k.svk) for (const char *c : solids) if (name == c) return true;
    if (k.corot) for (const char *c : corot) if (name == c) return true;
    if (k.svk) for (const char *c : svk) if (name == c) return true;
    if (k.constraints) for (const char *c : cons) if (name == c) return true;
    if (k.rigidbody) for (const char *c : rigid) if (name == c) return true;
    return false;
}

void release(psdata_hip *h) {
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (DevField &f : h->fields) {
        if (f.buf[0]) hipFree(f.buf[0]);
        if (f.buf[1]) hipFree(f.buf[1]);
    }
    unsigned *scratch[] = {h->slot_id[0], h->slot_id[1], h->perm, h->bins.key, h->bins.local, h->bins.tmp_id,
                           h->bins.tmp_src, h->bins.tmp_key, h->bins.tile_sums, h->pairs.list, h->pairs.count,
                           reinterpret_cast<unsigned *>(h->pairs.rec)};
    for (unsigned *p : scratch)
        if (p) hipFree(p);
    if (h->staging) hipFree(h->staging);
    if (h->sums) hipFree(h->sums);
    for (PendingTiming &t : h->pending) {
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    for (hipEvent_t e : h->event_pool) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

// Upload one field from the host blob.
int upload_field(psdata_hip *h, const psdata &data, DevField *f) {
    const char *src = static_cast<const char *>(data.data) + data.data_offsets[f->index];
    if (f->bytes == 0) return 0;
    if (f->kind == FieldKind::Other) {
        std::memcpy(h->shadow.data() + data.data_offsets[f->index], src, f->bytes);
    }
    if (!f->hot) {
        HIP_TRY(hipMemcpyAsync(f->ptr(), src, f->bytes, hipMemcpyHostToDevice, h->stream));
        return 0;
    }
    HIP_TRY(hipMemcpyAsync(h->staging, src, f->bytes, hipMemcpyHostToDevice, h->stream));
    if (timed(h, "upload_permute", [&] {
            return sphk::launch_gather_words(h->staging, f->ptr(), h->slot_ids(), h->n, h->pnum, f->words,
                                             h->stream);
        }))
        return -1;
    f->epoch = h->epoch;
    if (h->n < h->pnum) {
        const size_t off = static_cast<size_t>(h->n) * f->words * 4;
        HIP_TRY(hipMemcpyAsync(static_cast<char *>(f->buf[f->cur ^ 1]) + off, static_cast<char *>(f->ptr()) + off,
                               f->bytes - off, hipMemcpyDeviceToDevice, h->stream));
    }
    return 0;
}

int download_field(psdata_hip *h, const psdata &data, DevField *f) {
    char *dst = static_cast<char *>(data.data) + data.data_offsets[f->index];
    if (f->bytes == 0) return 0;
    if (!f->hot) {
        HIP_TRY(hipMemcpyAsync(dst, f->ptr(), f->bytes, hipMemcpyDeviceToHost, h->stream));
        return 0;
    }
    if (ensure_current(h, f)) return -1;
    if (timed(h, "download_unpermute", [&] {
            return sphk::launch_scatter_words(f->ptr(), h->staging, h->slot_ids(), h->n, h->pnum, f->words,
                                              h->stream);
        }))
        return -1;
    HIP_TRY(hipMemcpyAsync(dst, h->staging, f->bytes, hipMemcpyDeviceToHost, h->stream));
    return 0;
}

// Upload a set of fields; a change of n first restores particle order.
int upload_fields(psdata_hip *h, const psdata &data, const std::vector<DevField *> &list) {
    h->pairs.valid = 0;
    for (DevField *f : list) {
        if (f->name != "n") continue;
        unsigned new_n;
        std::memcpy(&new_n, static_cast<const char *>(data.data) + data.data_offsets[f->index], 4);
        if (new_n > h->pnum) return fail("n = %u exceeds pnum = %u", new_n, h->pnum);
        if (new_n != h->n) {
            if (reset_permutation(h)) return -1;
            h->n = new_n;
        }
    }
    for (DevField *f : list)
        if (upload_field(h, data, f)) return -1;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

bool same_psdata(const psdata &a, const psdata &b) { return a.data == b.data && a.num_fields == b.num_fields; }

}  // namespace

// ============================================================================
// Launch tuning
// ============================================================================
namespace sphk {
namespace {
unsigned g_neighbour_block = kMaxNeighbourBlock;
unsigned g_fluid_pair_group = 8;
}
unsigned neighbour_block() { return g_neighbour_block; }
bool set_neighbour_block(unsigned threads) {
    if (threads != 64 && threads != 128 && threads != 256) return false;
    g_neighbour_block = threads;
    return true;
}
unsigned fluid_pair_group() { return g_fluid_pair_group; }
bool set_fluid_pair_group(unsigned entries) {
    if (entries != 8 && entries != 12 && entries != 16) return false;
    g_fluid_pair_group = entries;
    return true;
}
}  // namespace sphk

extern "C" int sph_hip_set_tuning(unsigned neighbour_block) {
    if (!sphk::set_neighbour_block(neighbour_block))
        return fail("neighbour block size must be 64, 128 or 256 (got %u)", neighbour_block);
    return 0;
}

extern "C" int sph_hip_device_name(char *buf, unsigned len) {
    if (!buf || len == 0) return fail("null argument");
    int dev = 0;
    if (sphhip::current_device(&dev)) return -1;
    hipDeviceProp_t prop;
    SPH_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    std::string arch(prop.gcnArchName);
    arch = arch.substr(0, arch.find(':'));   // drop the target features
    std::snprintf(buf, len, "%s %dCU", arch.c_str(), prop.multiProcessorCount);
    return 0;
}

extern "C" int sph_hip_get_tuning(unsigned *neighbour_block) {
    if (!neighbour_block) return fail("null argument");
    *neighbour_block = sphk::neighbour_block();
    return 0;
}

extern "C" int sph_hip_set_fluid_pair_group(unsigned entries) {
    if (!sphk::set_fluid_pair_group(entries))
        return fail("fluid pair group must be 8, 12 or 16 (got %u)", entries);
    return 0;
}

extern "C" int sph_hip_get_fluid_pair_group(unsigned *entries) {
    if (!entries) return fail("null argument");
    *entries = sphk::fluid_pair_group();
    return 0;
}

// ============================================================================
// Helpers shared with the scene batch
// ============================================================================
namespace sphhip {

int fail(const char *fmt, ...) {
    char buf[1024];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, args);
    va_end(args);
    t_error = buf;
    note(2, "sph_hip: %s\n", buf);
    return -1;
}

KernelSet parse_file_list(const char *list) {
    KernelSet k;
    if (!list) return k;
    std::string s(list);
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ',' || s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
        size_t j = i;
        while (j < s.size() && !(s[j] == ',' || s[j] == ' ' || s[j] == '\n' || s[j] == '\r' || s[j] == '\t')) ++j;
        std::string w = s.substr(i, j - i);
        if (w == "common") k.common = true;
        else if (w == "fluids") k.fluids = true;
        else if (w == "solids") k.solids = true;
        else if (w == "solid_corotation") k.corot = true;
        else if (w == "solid_svk") k.svk = true;
        else if (w == "constraints") k.constraints = true;
        else if (w == "rigidbody") k.rigidbody = true;
        else if (!w.empty()) note(1, "sph_hip: ignoring unknown kernel file '%s'\n", w.c_str());
        i = j;
    }
    return k;
}

int current_device(int *device) {
    if (!g_ready && init_hip()) return -1;
    *device = g_device;
    return 0;
}

namespace {

// Field values as the blob holds them.
struct BlobReader {
    const psdata &meta;
    const char *blob;
    int index(const char *name) const { return get_field_psdata(meta, name); }
    bool read_float(int f, unsigned k, float *out) const {
        if (f < 0) return false;
        const unsigned es = meta.entry_sizes[f];
        if (static_cast<size_t>(k + 1) * es > meta.data_sizes[f]) return false;
        const char *p = blob + meta.data_offsets[f] + static_cast<size_t>(k) * es;
        if (es == 8) {
            double d;
            std::memcpy(&d, p, 8);
            *out = static_cast<float>(d);
        } else {
            std::memcpy(out, p, 4);
        }
        return true;
    }
    bool read_u32(int f, unsigned k, unsigned *out) const {
        if (f < 0 || static_cast<size_t>(k + 1) * 4 > meta.data_sizes[f]) return false;
        std::memcpy(out, blob + meta.data_offsets[f] + static_cast<size_t>(k) * 4, 4);
        return true;
    }
    float scalar_or(const char *name, float dflt) const {
        float v;
        return read_float(index(name), 0, &v) ? v : dflt;
    }
    unsigned last_dim(int f) const {
        const unsigned nd = meta.num_dimensions[f];
        if (nd == 0) return 0;
        return meta.dimensions[meta.dimensions_offsets[f] + nd - 1];
    }
};

}  // namespace

int params_from_blob(const psdata &meta, const char *blob, const KernelSet &ks, unsigned n, bool has_grid,
                     int compensated, int polar_fp64, sphk::StepParams *pp) {
    const BlobReader r{meta, blob};
    sphk::StepParams p{};
    p.n = n;
    if (has_grid) {
        const int gr = r.index("gridres");
        unsigned res[3];
        for (unsigned a = 0; a < 3; ++a)
            if (!r.read_u32(gr, a, &res[a])) return fail("gridres is missing");
        p.grid.rx = res[0];
        p.grid.ry = res[1];
        p.grid.rz = res[2];
        p.grid.layer = res[0] * res[1];
        p.grid.ncells = res[0] * res[1] * res[2];
        const int gb = r.index("gridbounds");
        for (unsigned a = 0; a < 3; ++a) {
            if (!r.read_float(gb, 2 * a, &p.grid.lo[a]) || !r.read_float(gb, 2 * a + 1, &p.grid.hi[a]))
                return fail("gridbounds needs 6 values");
        }
    }
    const float hh = r.scalar_or("smoothingradius", 0.0f);
    p.h = hh;
    p.h2 = hh * hh;
    // 315/(64 pi h^9), -45/(pi h^6), 45/(pi h^6): double, rounded once.
    const double hd = static_cast<double>(hh);
    const double h6 = hd * hd * hd * hd * hd * hd, h9 = h6 * hd * hd * hd;
    const double pi = 3.14159265358979323846;
    p.poly6 = static_cast<float>(315.0 / (64.0 * pi * h9));
    p.spiky = static_cast<float>(-45.0 / (pi * h6));
    p.lap = static_cast<float>(45.0 / (pi * h6));
    // Smallest float s with sqrtf(s) > h (sqrtf is correctly rounded on the
    // host and on the device, and monotonic): 0 for h < 0, the smallest
    // subnormal for h = 0; an infinite or NaN h leaves every finite distance
    // in range.
    float s2 = hh * hh;
    if (std::isfinite(hh)) {
        while (s2 > 0.0f && std::sqrt(s2) > hh) s2 = std::nextafter(s2, 0.0f);
        while (!(std::sqrt(s2) > hh)) s2 = std::nextafter(s2, INFINITY);
    } else {
        s2 = INFINITY;
    }
    p.r2_out = s2;
    p.mass = r.scalar_or("mass", 0.0f);
    p.dt = r.scalar_or("timestep", 0.0f);
    p.restdens = r.scalar_or("restdens", 0.0f);
    p.stiffness = r.scalar_or("stiffness", 0.0f);
    p.viscosity = r.scalar_or("viscosity", 0.0f);
    p.youngs = r.scalar_or("youngs_modulus", 0.0f);
    p.poisson = r.scalar_or("poisson", 0.0f);
    const int g = r.index("gravity");
    for (unsigned a = 0; a < 3; ++a)
        if (!r.read_float(g, a, &p.gravity[a])) p.gravity[a] = 0.0f;

    const int sp = r.index("symmetry_planes");
    if (sp >= 0) {
        const unsigned np = r.last_dim(sp);
        if (np > static_cast<unsigned>(sphk::kMaxSymmetryPlanes))
            return fail("at most %d symmetry planes are supported (conf has %u)", sphk::kMaxSymmetryPlanes, np);
        if (static_cast<size_t>(np) * 6 * meta.entry_sizes[sp] > meta.data_sizes[sp])
            return fail("symmetry_planes is smaller than 6 x planes");
        p.nplanes = np;
        for (unsigned s = 0; s < np; ++s) {
            for (unsigned k = 0; k < 6; ++k) r.read_float(sp, 6 * s + k, &p.planes[s][k]);
            if (p.planes[s][3] != 0.0f || p.planes[s][4] != 0.0f || p.planes[s][5] != 0.0f)
                p.active_planes |= 1u << s;
        }
    }
    const int pc = r.index("plane_constraints");
    if (pc >= 0) {
        unsigned nc = r.last_dim(pc);
        if (nc > static_cast<unsigned>(sphk::kMaxPlaneConstraints)) nc = sphk::kMaxPlaneConstraints;
        if (static_cast<size_t>(nc) * 6 * meta.entry_sizes[pc] > meta.data_sizes[pc])
            return fail("plane_constraints is smaller than 6 x planes");
        p.nconstraints = nc;
        for (unsigned c = 0; c < nc; ++c)
            for (unsigned k = 0; k < 6; ++k) r.read_float(pc, 6 * c + k, &p.constraints[c][k]);
    }
    const int bd = r.index("body_dynamic");
    if (bd >= 0) p.body_count = meta.data_sizes[bd] / 4;
    unsigned flag = 0;
    p.eos_tait = r.read_u32(r.index("eos_tait"), 0, &flag) && flag != 0u;
    flag = 0;
    p.mass_consistent = r.read_u32(r.index("mass_consistent"), 0, &flag) && flag != 0u;
    p.svk_szz_zero = r.read_u32(r.index("svk_sigma_zz_zero"), 0, &flag) && flag != 0u;
    p.system = ks.corot ? sphk::kCorotational : ks.svk ? sphk::kSvk : sphk::kFluid;
    p.compensated = compensated;
    p.polar_fp64 = polar_fp64;
    *pp = p;
    return 0;
}

}  // namespace sphhip

// ============================================================================
// C ABI
// ============================================================================
extern "C" const char *sph_hip_last_error(void) { return t_error.c_str(); }

extern "C" int init_hip(void) {
    if (g_ready) return 0;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail("no HIP device available (%s)", hipGetErrorString(e));
    int dev = 0;
    if (const char *env = std::getenv("SPH_HIP_DEVICE")) dev = std::atoi(env);
    if (dev < 0 || dev >= count) return fail("SPH_HIP_DEVICE=%d out of range (%d devices)", dev, count);
    HIP_TRY(hipSetDevice(dev));
    if (sphhip::noalloc_bind_current_device(nullptr)) return -1;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    note(1, "Chosen device: %s (%s, %d CUs)\n", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    g_device = dev;
    g_ready = true;
    return 0;
}

extern "C" int terminate_hip(void) {
    g_ready = false;
    return 0;
}

extern "C" int create_psdata_hip(psdata *data, const char *file_list, psdata_hip **out) {
    if (!out) return fail("create_psdata_hip: out is NULL");
    *out = nullptr;
    if (!g_ready && init_hip()) return -1;
    if (!data || data->num_fields == 0) return fail("create_psdata_hip: empty psdata");
    HIP_TRY(hipSetDevice(g_device));
    psdata_hip *h = new psdata_hip();
    h->host = *data;
    h->kset = sphhip::parse_file_list(file_list);
    if (!h->kset.common) {
        release(h);
        return fail("the kernel file list must include 'common'");
    }
    if (h->kset.corot && h->kset.svk) {
        release(h);
        return fail("solid_corotation and solid_svk both define compute_stresses/compute_forces_solids");
    }
    const int pn = get_field_psdata(*data, "pnum"), nn = get_field_psdata(*data, "n");
    if (pn < 0 || nn < 0) {
        release(h);
        return fail("the psdata needs pnum and n fields");
    }
    std::memcpy(&h->pnum, static_cast<char *>(data->data) + data->data_offsets[pn], 4);
    std::memcpy(&h->n, static_cast<char *>(data->data) + data->data_offsets[nn], 4);
    if (h->n > h->pnum) {
        release(h);
        return fail("n = %u exceeds pnum = %u", h->n, h->pnum);
    }
    if (h->pnum > sphk::kMaxSlots) {
        release(h);
        return fail("pnum = %u exceeds the engine's limit of %u particles", h->pnum, sphk::kMaxSlots);
    }
    std::vector<FieldKind> kinds = sphhip::field_kinds_for(data->data);
    const unsigned total = psdata_data_size(*data);
    h->shadow.assign(static_cast<const char *>(data->data), static_cast<const char *>(data->data) + total);
    size_t max_particle_bytes = 4;
    for (unsigned i = 0; i < data->num_fields; ++i) {
        DevField f;
        f.name = data->names + data->names_offsets[i];
        f.index = static_cast<int>(i);
        f.bytes = data->data_sizes[i];
        if (!kinds.empty()) {
            f.kind = kinds[i];
        } else {   // psdata built elsewhere: infer from the dimensions
            const unsigned nd = data->num_dimensions[i];
            const unsigned last = nd ? data->dimensions[data->dimensions_offsets[i] + nd - 1] : 0;
            if (f.name == "cellparticles") f.kind = FieldKind::Slot;
            else if (f.name == "gridcount" || f.name == "celloffset") f.kind = FieldKind::Cell;
            else if (nd >= 1 && last == h->pnum && h->pnum > 0) f.kind = FieldKind::Particle;
        }
        if (f.kind == FieldKind::Particle) {
            if (h->pnum == 0 || f.bytes % (static_cast<size_t>(h->pnum) * 4) != 0) {
                f.kind = FieldKind::Other;   // not a whole number of 32-bit words per particle
            } else {
                f.words = static_cast<unsigned>(f.bytes / (static_cast<size_t>(h->pnum) * 4));
                f.hot = hot_field(f.name, h->kset);
                max_particle_bytes = std::max(max_particle_bytes, f.bytes);
            }
        }
        // per-slot fields carry kScanPad slots of padding (read, never used,
        // by the unrolled candidate scans)
        const size_t alloc = std::max<size_t>(
            f.bytes + (f.kind == FieldKind::Particle ? static_cast<size_t>(sphk::kScanPad) * f.words * 4 : 0), 4);
        if (hipMalloc(&f.buf[0], alloc) != hipSuccess ||
            (f.hot && hipMalloc(&f.buf[1], alloc) != hipSuccess)) {
            release(h);
            return fail("out of device memory allocating field %s (%zu bytes)", f.name.c_str(), alloc);
        }
        f.epoch = h->epoch;
        h->by_name[f.name] = static_cast<int>(h->fields.size());
        h->fields.push_back(f);
    }
    if (check_field_shapes(h)) {
        release(h);
        return -1;
    }
    h->has_grid = h->field("gridres") && h->field("gridcell") && h->field("gridcount") && h->field("celloffset") &&
                  h->field("cellparticles") && h->field("gridbounds");
    if (h->has_grid) {
        sphk::StepParams p;
        if (build_params(h, &p)) {
            release(h);
            return -1;
        }
        if (p.grid.rx == 0 || p.grid.ry == 0 || p.grid.rz == 0) {
            release(h);
            return fail("grid resolution %ux%ux%u has an empty axis", p.grid.rx, p.grid.ry, p.grid.rz);
        }
    }
    const size_t pn4 = static_cast<size_t>(std::max(h->pnum, 1u)) * 4;
    unsigned ncells = 1;
    if (h->has_grid) {
        DevField *gc = h->field("gridcount");
        ncells = static_cast<unsigned>(std::max<size_t>(gc->bytes / 4, 1));
    }
    const size_t tiles = (ncells + sphk::kScanTile - 1) / sphk::kScanTile + 1;
    void **scratch[] = {reinterpret_cast<void **>(&h->slot_id[0]), reinterpret_cast<void **>(&h->slot_id[1]),
                        reinterpret_cast<void **>(&h->perm),       reinterpret_cast<void **>(&h->bins.key),
                        reinterpret_cast<void **>(&h->bins.local), reinterpret_cast<void **>(&h->bins.tmp_id),
                        reinterpret_cast<void **>(&h->bins.tmp_src), reinterpret_cast<void **>(&h->bins.tmp_key)};
    for (void **p : scratch)
        if (hipMalloc(p, pn4) != hipSuccess) {
            release(h);
            return fail("out of device memory (scratch)");
        }
    if (hipMalloc(reinterpret_cast<void **>(&h->bins.tile_sums), tiles * 4) != hipSuccess ||
        hipMalloc(&h->staging, max_particle_bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&h->sums), (7 * sphk::kReduceBlocks + 8) * sizeof(double)) != hipSuccess ||
        ((h->kset.fluids || ((h->kset.solids || h->kset.corot || h->kset.svk) && h->pnum <= sphk::kListSlot)) &&
         h->has_grid &&
         (hipMalloc(reinterpret_cast<void **>(&h->pairs.list), pn4 * sphk::kListCap) != hipSuccess ||
          hipMalloc(reinterpret_cast<void **>(&h->pairs.count), pn4) != hipSuccess ||
          (h->kset.fluids &&
           hipMalloc(reinterpret_cast<void **>(&h->pairs.rec), sphk::kFluidPlanes * 4 * pn4) != hipSuccess)))) {
        release(h);
        return fail("out of device memory (scratch)");
    }
    h->staging_bytes = max_particle_bytes;
    h->pairs.stride = h->pnum;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        release(h);
        return fail("hipStreamCreate failed");
    }
    for (int k = 0; k < 2; ++k)
        if (sphk::launch_iota_u32(h->slot_id[k], h->pnum, h->stream) != hipSuccess) {
            release(h);
            return fail("iota launch failed");
        }
    std::vector<DevField *> all;
    for (DevField &f : h->fields) all.push_back(&f);
    // Initial contents = the host blob (CL_MEM_COPY_HOST_PTR in the reference).
    for (DevField *f : all) {
        if (upload_field(h, *data, f)) {
            release(h);
            return -1;
        }
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        release(h);
        return fail("initial upload failed");
    }
    *out = h;
    return 0;
}

extern "C" int free_psdata_hip(psdata_hip *pso) {
    release(pso);
    return 0;
}

extern "C" int sync_psdata_host_to_device_hip(psdata data, psdata_hip *h, int full) {
    (void)full;   // metadata is immutable on the device side
    if (!h) return fail("null handle");
    if (!same_psdata(data, h->host)) return fail("psdata does not belong to this handle");
    std::vector<DevField *> all;
    for (DevField &f : h->fields) all.push_back(&f);
    return upload_fields(h, data, all);
}

extern "C" int sync_psdata_device_to_host_hip(psdata data, psdata_hip *h) {
    if (!h) return fail("null handle");
    if (!same_psdata(data, h->host)) return fail("psdata does not belong to this handle");
    for (DevField &f : h->fields)
        if (download_field(h, data, &f)) return -1;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int sync_psdata_fields_host_to_device_hip(psdata data, psdata_hip *h, size_t num_fields,
                                                     const char *const *names) {
    if (!h) return fail("null handle");
    if (!same_psdata(data, h->host)) return fail("psdata does not belong to this handle");
    std::vector<DevField *> list;
    for (size_t i = 0; i < num_fields; ++i) {
        if (!names[i]) continue;
        DevField *f = h->field(names[i]);
        if (!f) {
            note(1, "Field %s not found.\n", names[i]);
            continue;
        }
        list.push_back(f);
    }
    return upload_fields(h, data, list);
}

extern "C" int sync_psdata_fields_device_to_host_hip(psdata data, psdata_hip *h, size_t num_fields,
                                                     const char *const *names) {
    if (!h) return fail("null handle");
    if (!same_psdata(data, h->host)) return fail("psdata does not belong to this handle");
    for (size_t i = 0; i < num_fields; ++i) {
        if (!names[i]) continue;
        DevField *f = h->field(names[i]);
        if (!f) {
            note(1, "Field %s not found.\n", names[i]);
            continue;
        }
        if (download_field(h, data, f)) return -1;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int call_for_all_particles_device_hip(psdata_hip *h, const char *kernel_name) {
    if (!h || !kernel_name) return fail("null argument");
    std::string internal;
    if (resolve_kernel(h, kernel_name, &internal)) return -1;
    return run_kernel(h, internal);
}

extern "C" int compute_particle_bins_device_hip(psdata_hip *h) {
    if (!h) return fail("null handle");
    return compute_bins(h);
}

extern "C" int compute_density_device_hip(psdata_hip *h) {
    return call_for_all_particles_device_hip(h, "compute_density");
}

extern "C" int compute_forces_device_hip(psdata_hip *h) {
    if (!h) return fail("null handle");
    if (h->kset.fluids) return call_for_all_particles_device_hip(h, "compute_forces_fluids");
    return call_for_all_particles_device_hip(h, "compute_forces_solids");
}

extern "C" int step_forward_device_hip(psdata_hip *h) { return call_for_all_particles_device_hip(h, "step_forward"); }

extern "C" int step_device_hip(psdata_hip *h, const char *system, unsigned int num_steps, int apply_constraints) {
    if (!h || !system) return fail("null argument");
    const std::string sys(system);
    std::vector<const char *> seq;
    if (sys == "fluid") {
        seq = {"compute_density", "compute_forces_fluids", "step_forward"};
    } else if (sys == "corotational") {
        if (!h->kset.corot) return fail("system 'corotational' needs solid_corotation in the kernel file list");
        seq = {"compute_density", "compute_rotations_and_strains", "compute_stresses", "compute_forces_solids",
               "step_forward"};
    } else if (sys == "svk") {
        if (!h->kset.svk) return fail("system 'svk' needs solid_svk in the kernel file list");
        seq = {"compute_density", "compute_deformations", "compute_stresses", "compute_forces_solids",
               "step_forward"};
    } else {
        return fail("unknown system '%s' (fluid, corotational, svk)", system);
    }
    if (apply_constraints) seq.push_back("apply_plane_constraints");
    std::vector<std::string> internal(seq.size());
    for (size_t k = 0; k < seq.size(); ++k)
        if (resolve_kernel(h, seq[k], &internal[k])) return -1;
    for (unsigned s = 0; s < num_steps; ++s) {
        if (compute_bins(h)) return -1;
        for (size_t k = 0; k < internal.size(); ++k) {
            if (internal[k] == "compute_forces_fluids" && k + 1 < internal.size() &&
                internal[k + 1] == "step_forward" && h->pairs.valid == sphk::kFluidList) {
                if (run_forces_and_step(h)) return -1;
                ++k;
                continue;
            }
            if (run_kernel(h, internal[k])) return -1;
        }
    }
    return 0;
}

extern "C" int synchronize_hip(psdata_hip *h) {
    if (!h) return fail("null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int set_precision_hip(psdata_hip *h, int compensated_sums, int polar_fp64) {
    if (!h) return fail("null handle");
    h->compensated = compensated_sums & (sphk::kKahanDensity | sphk::kKahanForces);
    h->polar_fp64 = polar_fp64 ? 1 : 0;
    return 0;
}

extern "C" int set_profiling_hip(psdata_hip *h, int enabled) {
    if (!h) return fail("null handle");
    if (!enabled && drain_timings(h)) return -1;
    h->profiling = enabled != 0;
    return 0;
}

extern "C" int reset_kernel_stats_hip(psdata_hip *h) {
    if (!h) return fail("null handle");
    if (drain_timings(h)) return -1;
    h->stats.clear();
    return 0;
}

extern "C" int get_kernel_stats_hip(psdata_hip *h, int capacity, const char **names, unsigned long long *launches,
                                    double *total_ms) {
    if (!h) return fail("null handle");
    if (drain_timings(h)) return -1;
    h->stat_names.clear();
    for (auto &kv : h->stats) h->stat_names.push_back(kv.first);
    int k = 0;
    for (auto &kv : h->stats) {
        if (k < capacity) {
            if (names) names[k] = h->stat_names[k].c_str();
            if (launches) launches[k] = kv.second.launches;
            if (total_ms) total_ms[k] = kv.second.total_ms;
        }
        ++k;
    }
    return k;
}

// ============================================================================
// Scripted rigid bodies and initial-condition helpers
// (src/opencl/particle_system_host.c:530-614)
// ============================================================================
namespace {

// Products in the reference's order (multiplyMatrices, kernels/common.cl:208-220):
// o[r][c] = (a[r][0] b[0][c] + a[r][1] b[1][c]) + a[r][2] b[2][c].
sphk::Mat3 mat_mul(const sphk::Mat3 &a, const sphk::Mat3 &b) {
    sphk::Mat3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            o.m[3 * r + c] = a.m[3 * r] * b.m[c] + a.m[3 * r + 1] * b.m[3 + c] + a.m[3 * r + 2] * b.m[6 + c];
    return o;
}

// The reference's axis factors (kernels/common.cl:575-587, kernels/rigidbody.cl:47-57);
// its y factor carries -sin in row 0 and +sin in row 2.
sphk::Mat3 axis_x(float a) {
    const float c = std::cos(a), s = std::sin(a);
    return sphk::Mat3{{1.0f, 0.0f, 0.0f, 0.0f, c, -s, 0.0f, s, c}};
}
sphk::Mat3 axis_y(float a) {
    const float c = std::cos(a), s = std::sin(a);
    return sphk::Mat3{{c, 0.0f, -s, 0.0f, 1.0f, 0.0f, s, 0.0f, c}};
}
sphk::Mat3 axis_z(float a) {
    const float c = std::cos(a), s = std::sin(a);
    return sphk::Mat3{{c, -s, 0.0f, s, c, 0.0f, 0.0f, 0.0f, 1.0f}};
}

// Body matrix of Euler angles (x, y, x2) = Rx(x2) (Ry(y) Rx(x)) (get_matrix_from_euler).
sphk::Mat3 euler_matrix(const float e[3]) { return mat_mul(axis_x(e[2]), mat_mul(axis_y(e[1]), axis_x(e[0]))); }

sphk::Mat3 transposed(const sphk::Mat3 &a) {
    sphk::Mat3 t;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t.m[3 * r + c] = a.m[3 * c + r];
    return t;
}

// A per-particle field kept in slot order with `words` 32-bit values per particle.
DevField *particle_field(psdata_hip *h, const char *name, unsigned words, const char *who) {
    DevField *f = h->field(name);
    if (!f || f->kind != FieldKind::Particle || !f->hot || f->words != words) {
        fail("%s needs a per-particle field '%s' with %u values per particle", who, name, words);
        return nullptr;
    }
    return f;
}

// A per-body float[3] field holding at least body + 1 entries.
DevField *body_field(psdata_hip *h, const char *name, unsigned body, const char *who) {
    DevField *f = h->field(name);
    if (!f || h->host.entry_sizes[f->index] != 4) {
        fail("%s needs a float field '%s' (3 values per body)", who, name);
        return nullptr;
    }
    if (static_cast<size_t>(body + 1) * 12 > f->bytes) {
        fail("%s: body %u is outside '%s' (%zu bodies)", who, body, name, f->bytes / 12);
        return nullptr;
    }
    return f;
}

int body_count_of(psdata_hip *h, const char *who, unsigned *out) {
    DevField *f = h->field("body_count");
    if (!f || f->bytes < 4 || h->host.entry_sizes[f->index] != 4)
        return fail("%s needs an unsigned field 'body_count'", who);
    std::memcpy(out, h->shadow.data() + h->host.data_offsets[f->index], 4);
    return 0;
}

}  // namespace

extern "C" int populate_position_cuboid_device_hip(psdata_hip *h, float x1, float y1, float z1, float x2, float y2,
                                                   float z2, unsigned xsize, unsigned ysize, unsigned zsize) {
    static const char *who = "populate_position_cuboid";
    if (!h) return fail("null handle");
    DevField *pos = particle_field(h, "position", 3, who);
    if (!pos) return -1;
    DevField *nf = h->field("n");
    const unsigned long long points = static_cast<unsigned long long>(xsize) * ysize * zsize;
    if (points > 0xffffffffull)
        return fail("%s: %llu lattice points; particle indices are 32-bit", who, points);
    if (points == 0 || h->pnum == 0) return 0;   // no lane reaches a particle
    // Lanes at or past pnum write nothing; the lane of particle pnum - 1 sets
    // n = pnum, else the last lattice point sets n = its count.
    const unsigned count = static_cast<unsigned>(std::min<unsigned long long>(points, h->pnum));
    const unsigned new_n = points >= h->pnum ? h->pnum : static_cast<unsigned>(points);
    // The lattice addresses particles by index: return to particle order.
    if (reset_permutation(h)) return -1;
    h->pairs.valid = 0;
    if (new_n != h->n) {
        h->n = new_n;
        char *shadow_n = h->shadow.data() + h->host.data_offsets[nf->index];
        std::memcpy(shadow_n, &new_n, 4);
        HIP_TRY(hipMemcpyAsync(nf->ptr(), shadow_n, 4, hipMemcpyHostToDevice, h->stream));
    }
    const sphk::Box box{{x1, y1, z1}, {x2, y2, z2}};
    if (timed(h, who, [&] {
            return sphk::launch_populate_cuboid(static_cast<float *>(pos->ptr()), count, xsize, ysize, zsize, box,
                                                h->stream);
        }))
        return -1;
    return 0;
}

extern "C" int rotate_particles_device_hip(psdata_hip *h, float angle_x, float angle_y, float angle_z) {
    static const char *who = "rotate_particles";
    if (!h) return fail("null handle");
    DevField *pos = particle_field(h, "position", 3, who);
    if (!pos || ensure_current(h, pos)) return -1;
    h->pairs.valid = 0;
    // Rx (Ry Rz), kernels/common.cl:589-592
    const sphk::Mat3 m = mat_mul(axis_x(angle_x), mat_mul(axis_y(angle_y), axis_z(angle_z)));
    if (timed(h, who, [&] {
            return sphk::launch_rotate(static_cast<float *>(pos->ptr()), h->n, m, h->stream);
        }))
        return -1;
    return 0;
}

extern "C" int set_body_position_device_hip(psdata_hip *h, unsigned body, const float *position) {
    static const char *who = "set_body_position";
    if (!h || !position) return fail("null argument");
    if (!h->kset.rigidbody) return fail("%s is in kernel file 'rigidbody': add it to the kernel file list", who);
    DevField *pos = particle_field(h, "position", 3, who);
    DevField *vel = pos ? particle_field(h, "velocity", 3, who) : nullptr;
    DevField *bn = vel ? particle_field(h, "body_number", 1, who) : nullptr;
    DevField *bp = bn ? body_field(h, "body_position", body, who) : nullptr;
    unsigned body_count = 0;
    if (!bp || body_count_of(h, who, &body_count)) return -1;
    sphk::StepParams p;
    if (build_params(h, &p)) return -1;
    if (ensure_current(h, pos) || ensure_current(h, vel) || ensure_current(h, bn)) return -1;
    h->pairs.valid = 0;
    if (timed(h, who, [&] {
            return sphk::launch_set_body_position(static_cast<float *>(pos->ptr()), static_cast<float *>(vel->ptr()),
                                                  static_cast<const unsigned *>(bn->ptr()), h->n, body,
                                                  body < body_count, static_cast<float *>(bp->ptr()), position, p.dt,
                                                  h->stream);
        }))
        return -1;
    std::memcpy(h->shadow.data() + h->host.data_offsets[bp->index] + 12 * static_cast<size_t>(body), position, 12);
    return 0;
}

extern "C" int set_body_rotation_device_hip(psdata_hip *h, unsigned body, const float *euler_angles) {
    static const char *who = "set_body_rotation";
    if (!h || !euler_angles) return fail("null argument");
    if (!h->kset.rigidbody) return fail("%s is in kernel file 'rigidbody': add it to the kernel file list", who);
    DevField *pos = particle_field(h, "position", 3, who);
    DevField *rot = pos ? particle_field(h, "rotation", 9, who) : nullptr;
    DevField *bn = rot ? particle_field(h, "body_number", 1, who) : nullptr;
    DevField *be = bn ? body_field(h, "body_euler_rotation", body, who) : nullptr;
    unsigned body_count = 0;
    if (!be || body_count_of(h, who, &body_count)) return -1;
    if (ensure_current(h, pos) || ensure_current(h, rot) || ensure_current(h, bn)) return -1;
    // The body matrix is the one of the stored angles, read back as the device holds them.
    float stored[3];
    HIP_TRY(hipMemcpyAsync(stored, static_cast<char *>(be->ptr()) + 12 * static_cast<size_t>(body), 12,
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const sphk::Mat3 bt = transposed(euler_matrix(stored));
    h->pairs.valid = 0;
    if (timed(h, who, [&] {
            return sphk::launch_set_body_rotation(static_cast<float *>(pos->ptr()), static_cast<float *>(rot->ptr()),
                                                  static_cast<const unsigned *>(bn->ptr()), h->n, body,
                                                  body < body_count, bt, static_cast<float *>(be->ptr()),
                                                  euler_angles, h->stream);
        }))
        return -1;
    std::memcpy(h->shadow.data() + h->host.data_offsets[be->index] + 12 * static_cast<size_t>(body), euler_angles,
                12);
    return 0;
}

extern "C" int global_sums_hip(psdata_hip *h, double *out) {
    if (!h || !out) return fail("null argument");
    DevField *pos = h->field("position");
    DevField *vel = h->field("velocity");
    if (!pos || !pos->hot) return fail("global_sums_hip needs position");
    if (ensure_current(h, pos) || ensure_current(h, vel)) return -1;
    sphk::StepParams p;
    if (build_params(h, &p)) return -1;
    sphk::FieldPtrs f = field_ptrs(h);
    double *partials = h->sums;
    double *res = h->sums + 7 * sphk::kReduceBlocks;
    if (timed(h, "global_sums", [&] { return sphk::launch_global_sums(p, f, partials, res, h->stream); })) return -1;
    HIP_TRY(hipMemcpyAsync(out, res, 7 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}
