// The reference's device API under its own names
// (src/opencl/particle_system_host.h:7-80), over the engine's *_hip calls.
// The by-value psdata_opencl carries the engine's handle in ps_prog; errors
// abort after printing the message, as the reference's ASSERTs do.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/sph_hip.h"

namespace {

// Names call_for_all_particles accepts (the kernel file list decides which a
// handle runs).
const char *const kKernelNames[] = {"compute_density",
                                    "step_forward",
                                    "find_particle_bins",
                                    "compute_forces_fluids",
                                    "init_original_position",
                                    "compute_original_density",
                                    "compute_rotations_and_strains",
                                    "compute_deformations",
                                    "compute_stresses",
                                    "compute_forces_solids",
                                    "apply_plane_constraints"};

void check(int rc, const char *call) {
    if (rc == 0) return;
    std::fprintf(stderr, "%s failed: %s\n", call, sph_hip_last_error());
    std::abort();
}

psdata_hip *handle(const psdata_opencl &pso, const char *call) {
    if (!pso.ps_prog) {
        std::fprintf(stderr, "%s: the psdata_opencl holds no device handle\n", call);
        std::abort();
    }
    return static_cast<psdata_hip *>(pso.ps_prog);
}

unsigned grid_cells(const psdata &data) {
    const int f = get_field_psdata(data, "gridres");
    if (f < 0 || data.data_sizes[f] < 12) return 0;
    unsigned res[3];
    std::memcpy(res, static_cast<const char *>(data.data) + data.data_offsets[f], 12);
    return res[0] * res[1] * res[2];
}

}  // namespace

extern "C" {

void init_opencl(void) { check(init_hip(), "init_opencl"); }

void terminate_opencl(void) { check(terminate_hip(), "terminate_opencl"); }

psdata_opencl create_psdata_opencl(psdata *data, const char *file_list) {
    psdata_hip *h = nullptr;
    check(create_psdata_hip(data, file_list, &h), "create_psdata_opencl");
    psdata_opencl pso;
    std::memset(&pso, 0, sizeof(pso));
    pso.host_psdata = *data;
    pso.num_fields = data->num_fields;
    pso.num_kernels = sizeof(kKernelNames) / sizeof(kKernelNames[0]);
    pso.kernel_names = kKernelNames;
    pso.ps_prog = h;
    pso.num_grid_cells = grid_cells(*data);
    pso.po2_workgroup_size = 256;
    pso.num_blocks = (pso.num_grid_cells + 511u) / 512u;
    return pso;
}

void assign_pso_kernel_args(psdata_opencl) {}

void set_kernel_args_to_pso(psdata_opencl, void *) {}

void free_psdata_opencl(psdata_opencl *pso) {
    if (!pso || !pso->ps_prog) return;
    check(free_psdata_hip(static_cast<psdata_hip *>(pso->ps_prog)), "free_psdata_opencl");
    pso->ps_prog = nullptr;
}

void sync_psdata_device_to_host(psdata data, psdata_opencl pso) {
    check(sync_psdata_device_to_host_hip(data, handle(pso, "sync_psdata_device_to_host")),
          "sync_psdata_device_to_host");
}

void sync_psdata_fields_device_to_host(psdata data, psdata_opencl pso, size_t num_fields,
                                       const char *const *field_names) {
    check(sync_psdata_fields_device_to_host_hip(data, handle(pso, "sync_psdata_fields_device_to_host"), num_fields,
                                                field_names),
          "sync_psdata_fields_device_to_host");
}

void sync_psdata_host_to_device(psdata data, psdata_opencl pso, int full) {
    check(sync_psdata_host_to_device_hip(data, handle(pso, "sync_psdata_host_to_device"), full),
          "sync_psdata_host_to_device");
}

void sync_psdata_fields_host_to_device(psdata data, psdata_opencl pso, size_t num_fields,
                                       const char *const *field_names) {
    check(sync_psdata_fields_host_to_device_hip(data, handle(pso, "sync_psdata_fields_host_to_device"), num_fields,
                                                field_names),
          "sync_psdata_fields_host_to_device");
}

void populate_position_cuboid_device_opencl(psdata_opencl pso, float x1, float y1, float z1, float x2, float y2,
                                            float z2, unsigned int xsize, unsigned int ysize, unsigned int zsize) {
    check(populate_position_cuboid_device_hip(handle(pso, "populate_position_cuboid_device_opencl"), x1, y1, z1, x2,
                                              y2, z2, xsize, ysize, zsize),
          "populate_position_cuboid_device_opencl");
}

void rotate_particles_device_opencl(psdata_opencl pso, float angle_x, float angle_y, float angle_z) {
    check(rotate_particles_device_hip(handle(pso, "rotate_particles_device_opencl"), angle_x, angle_y, angle_z),
          "rotate_particles_device_opencl");
}

void call_for_all_particles_device_opencl(psdata_opencl pso, const char *kernel_name) {
    check(call_for_all_particles_device_hip(handle(pso, "call_for_all_particles_device_opencl"), kernel_name),
          "call_for_all_particles_device_opencl");
}

void compute_particle_bins_device_opencl(psdata_opencl pso) {
    check(compute_particle_bins_device_hip(handle(pso, "compute_particle_bins_device_opencl")),
          "compute_particle_bins_device_opencl");
}

void compute_density_device_opencl(psdata_opencl pso) {
    check(compute_density_device_hip(handle(pso, "compute_density_device_opencl")), "compute_density_device_opencl");
}

void compute_forces_device_opencl(psdata_opencl pso) {
    check(compute_forces_device_hip(handle(pso, "compute_forces_device_opencl")), "compute_forces_device_opencl");
}

void step_forward_device_opencl(psdata_opencl pso) {
    check(step_forward_device_hip(handle(pso, "step_forward_device_opencl")), "step_forward_device_opencl");
}

void set_body_position_device_opencl(psdata_opencl pso, unsigned int body, float *position) {
    check(set_body_position_device_hip(handle(pso, "set_body_position_device_opencl"), body, position),
          "set_body_position_device_opencl");
}

void set_body_rotation_device_opencl(psdata_opencl pso, unsigned int body, float *euler_angles) {
    check(set_body_rotation_device_hip(handle(pso, "set_body_rotation_device_opencl"), body, euler_angles),
          "set_body_rotation_device_opencl");
}

}  // extern "C"
