/*
 * sph_hip.h - C ABI of libsph_hip.so, the MI355X (gfx950) SPH step engine.
 *
 * The library replaces libSPH.so of tauroid/opencl-sph for the simulation
 * step.  Host-side functions (config, psdata) keep the reference's names,
 * signatures and struct layout so that ctypes bindings written for libSPH.so
 * bind to them unchanged.  Device-side functions replace the OpenCL runtime
 * (src/opencl/particle_system_host.h) with a HIP runtime: the by-value
 * `psdata_opencl` struct becomes an opaque `psdata_hip *` handle, and every
 * device call returns an int status (0 = success) instead of aborting; the
 * text of the last error is available from sph_hip_last_error().
 *
 * Citations are file:line in the reference repository.
 */
#ifndef SPH_HIP_H_
#define SPH_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * psdata container.  Layout identical to src/datatypes.h:4-25 and
 * src/particle_system.h:18-48 (the Python layer reads these members directly).
 * ------------------------------------------------------------------------- */
typedef enum {
    T_CHAR,
    T_UINT,
    T_INT,
    T_FLOAT32,
    T_FLOAT64,
    TYPE_COUNT
} FieldType;

typedef struct {
    FieldType type;
    union {
        char     *c;
        uint32_t *u;
        int32_t  *i;
        float    *f4;
        double   *f8;
        void     *ptr;
    };
} DataPtr;

typedef struct {
    unsigned int   num_fields;
    const char    *names;
    unsigned int  *names_offsets;
    unsigned int  *dimensions;
    unsigned int  *num_dimensions;
    unsigned int  *dimensions_offsets;
    unsigned int  *entry_sizes;

    void          *data;
    unsigned int  *data_sizes;
    unsigned int  *data_offsets;

    DataPtr       *data_ptrs;

    unsigned int   num_host_fields;
    char         **host_names;
    void         **host_data;
    unsigned int  *host_data_size;
} psdata;

/* ---------------------------------------------------------------------------
 * Logging.  Replaces src/note.h (set_log_level, note).
 * ------------------------------------------------------------------------- */
void set_log_level(unsigned int level);                       /* src/note.c:9-11  */
void note(unsigned int level, const char *format, ...);        /* src/note.c:13-33 */

/* ---------------------------------------------------------------------------
 * Config loader.  Replaces src/config.h:4-6.  load_config returns 1 when the
 * file cannot be read (src/config.c:49-52), 0 otherwise.
 * ------------------------------------------------------------------------- */
int         load_config(const char *relative_path);           /* src/config.c:19-91   */
const char *get_config_section(const char *heading);          /* src/config.c:92-98   */
void        unload_config(void);                               /* src/config.c:99-106  */

/* ---------------------------------------------------------------------------
 * psdata builder and accessors.  Replaces src/build_psdata.h:6-7 and
 * src/particle_system.h:56-77.  The blob layout (field order, type sort,
 * alignment, auto-added grid arrays) is byte-identical to the reference.
 * ------------------------------------------------------------------------- */
void build_psdata_from_string(psdata *data, const char *string);   /* src/build_psdata.c:585-613 */
void build_psdata(psdata *data, const char *path);                   /* src/build_psdata.c:614-636 */
int  get_field_psdata(psdata data, const char *name);                /* src/particle_system.c:84-91 */
void set_field_psdata(psdata *data, const char *name, void *field,
                      unsigned int size, unsigned int offset);       /* src/particle_system.c:93-97 */
unsigned int psdata_names_size(psdata data);                         /* src/particle_system.c:99-103 */
unsigned int psdata_dimensions_size(psdata data);                    /* src/particle_system.c:105-108 */
unsigned int psdata_data_size(psdata data);                          /* src/particle_system.c:110-112 */
int  create_host_field_psdata(psdata *data, const char *name,
                              void *field, unsigned int size);       /* src/particle_system.c:116-155 */
int  get_host_field_psdata(psdata *data, const char *name);          /* src/particle_system.c:157-164 */
void display_psdata(psdata data, const char *const *mask);           /* src/particle_system.c:47-82 */
void free_psdata(psdata *data);                                      /* src/particle_system.c:218-255 */

/* Status of the last build_psdata* call on this thread: 0 = ok.  The
 * reference returns void and asserts; this lets callers detect failure. */
int  psdata_last_build_status(void);

/* ---------------------------------------------------------------------------
 * Device runtime (HIP).  Replaces src/opencl/particle_system_host.h:55-80.
 * ------------------------------------------------------------------------- */
typedef struct psdata_hip psdata_hip;   /* opaque; replaces psdata_opencl (:7-47) */

const char *sph_hip_last_error(void);

/* init_opencl (:55): select the HIP device (the first one, or $SPH_HIP_DEVICE)
 * and create the stream every handle launches on. */
int init_hip(void);
/* terminate_opencl (:60) */
int terminate_hip(void);

/* create_psdata_opencl (:56): mirror `data` on the device.  `file_list` is the
 * $opencl_kernel_files section ("common, fluids", "common, solids,
 * solid_corotation, constraints", ...); it selects which named kernels the
 * handle accepts, exactly as the list of .cl files does in the reference. */
int create_psdata_hip(psdata *data, const char *file_list, psdata_hip **out);
/* free_psdata_opencl (:59) */
int free_psdata_hip(psdata_hip *pso);

/* sync_psdata_* (:62-65).  Device state is kept in cell-sorted order; these
 * calls present and accept it in psdata (particle-index) order. */
int sync_psdata_host_to_device_hip(psdata data, psdata_hip *pso, int full);
int sync_psdata_device_to_host_hip(psdata data, psdata_hip *pso);
int sync_psdata_fields_host_to_device_hip(psdata data, psdata_hip *pso,
                                          size_t num_fields,
                                          const char *const *field_names);
int sync_psdata_fields_device_to_host_hip(psdata data, psdata_hip *pso,
                                          size_t num_fields,
                                          const char *const *field_names);

/* call_for_all_particles_device_opencl (:74): run the named kernel over all
 * particles.  Accepted names: compute_density, compute_forces_fluids,
 * step_forward, init_original_position, compute_original_density,
 * compute_rotations_and_strains, compute_deformations, compute_stresses,
 * compute_forces_solids, apply_plane_constraints, find_particle_bins. */
int call_for_all_particles_device_hip(psdata_hip *pso, const char *kernel_name);

/* compute_particle_bins_device_opencl (:76): key, count, scan, stable insert
 * (ascending particle index within a cell) and cell-sorted reorder. */
int compute_particle_bins_device_hip(psdata_hip *pso);

/* compute_density/forces/step_forward_device_opencl (:78-80) */
int compute_density_device_hip(psdata_hip *pso);
int compute_forces_device_hip(psdata_hip *pso);
int step_forward_device_hip(psdata_hip *pso);

/* No counterpart in the reference: run `num_steps` whole simulation steps of
 * `system` ("fluid", "corotational" or "svk"), the sequences of
 * src/mex/simstep_fluid.c:18-23 and scripts/sph/systems.py:148-189, without
 * host round trips.  When `apply_constraints` is non-zero,
 * apply_plane_constraints runs after step_forward as in
 * scripts/straintest.py:224. */
int step_device_hip(psdata_hip *pso, const char *system, unsigned int num_steps,
                    int apply_constraints);

/* Initial-condition helpers (particle_system_host.c:530-557,590-608; kernels
 * populate_position_cuboid and rotate_particles, kernels/common.cl:533-604).
 * populate: particle i + xsize*j + xsize*ysize*k (below pnum) is placed at
 * corner1 + (index / (size - 1)) * (corner2 - corner1) per axis; n becomes
 * pnum when the lattice reaches it, else the lattice size.  rotate: every
 * active position is multiplied by Rx(angle_x) Ry(angle_y) Rz(angle_z) with
 * the reference's axis factors (its Ry carries -sin in row 0). */
int populate_position_cuboid_device_hip(psdata_hip *pso, float x1, float y1, float z1,
                                        float x2, float y2, float z2, unsigned int xsize,
                                        unsigned int ysize, unsigned int zsize);
int rotate_particles_device_hip(psdata_hip *pso, float angle_x, float angle_y, float angle_z);

/* Scripted rigid bodies (particle_system_host.c:562-588; kernels/rigidbody.cl,
 * which must be in the kernel file list as 'rigidbody').  Particles with
 * body_number == body + 1 (body below body_count) follow the body:
 * set_body_position moves them by position - body_position[body] and gives
 * them that displacement / timestep as velocity; set_body_rotation applies
 * B^T P (B: matrix of the stored Euler angles, P: the particle's rotation
 * field) to position and rotation.  Then the body's stored position / angles
 * become the arguments (float[3]).  A body outside body_position /
 * body_euler_rotation is an error (the reference writes past the array). */
int set_body_position_device_hip(psdata_hip *pso, unsigned int body, const float *position);
int set_body_rotation_device_hip(psdata_hip *pso, unsigned int body, const float *euler_angles);

/* Block until all work queued on the handle's stream has finished. */
int synchronize_hip(psdata_hip *pso);

/* Precision of neighbour sums, a bit mask: 0 = plain FP32 in reference order
 * (default); bit 0 (1) = Kahan-compensated density sums; bit 1 (2) =
 * Kahan-compensated force / deformation sums.  polar_fp64 != 0 computes the corotational
 * polar decomposition and the SVK inverse in FP64 (default 1). */
int set_precision_hip(psdata_hip *pso, int compensated_sums, int polar_fp64);

/* Per-kernel timing with HIP events on the handle's stream (the reference
 * creates a profiling queue at particle_system_host.c:92-93 but never reads
 * it).  get_kernel_stats_hip fills up to `capacity` entries and returns the
 * number of distinct kernels timed (or -1 on error). */
int set_profiling_hip(psdata_hip *pso, int enabled);
int reset_kernel_stats_hip(psdata_hip *pso);
int get_kernel_stats_hip(psdata_hip *pso, int capacity, const char **names,
                         unsigned long long *launches, double *total_ms);

/* Device-side reductions in FP64 over the active particles: total momentum
 * (3), total kinetic energy, and the centre of mass (3).  out has 7 doubles. */
int global_sums_hip(psdata_hip *pso, double *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the reductions of global_sums_hip over caller-owned
 * device arrays, allocating nothing.  position: float[n][3], velocity:
 * float[n][3] or NULL (zero velocities), mass: the particle mass, out7: 7
 * doubles of device memory (momentum, kinetic energy, centre of mass),
 * stream: a hipStream_t or NULL (asynchronous on that stream).  The partial
 * sums live in one library-resident array of 7 x 256 doubles (14 KiB) per
 * GPU, held until the process ends; every entry is rewritten by each call,
 * and calls that overlap on one GPU race for it.  n is at most 1431655765
 * (larger n is an error).  Equal inputs give bitwise the output of
 * global_sums_hip over a handle whose slots hold them in the same order. */
int global_sums_noalloc_hip(const float *position, const float *velocity, unsigned n, float mass, double *out7,
                            void *stream);

/* No counterpart in the reference (its work-group size is fixed at 256,
 * src/opencl/particle_system_host.c:361-528): threads per block (64, 128 or
 * 256) of the neighbour passes that accept any of them, process-wide;
 * chosen by the autotuner (sph.autotune).  Launches captured into a batch's
 * graph keep the size they were captured with. */
int sph_hip_set_tuning(unsigned neighbour_block);
int sph_hip_get_tuning(unsigned *neighbour_block);
/* No counterpart in the reference: list entries whose neighbour loads the
 * fluid force pass issues together (8, 12 or 16), process-wide; more loads
 * in flight against fewer waves per SIMD.  Chosen by the autotuner.  No
 * value changes a result bit (the summation order is the list order). */
int sph_hip_set_fluid_pair_group(unsigned entries);
int sph_hip_get_fluid_pair_group(unsigned *entries);
/* Name, architecture and compute-unit count of the selected device (the key
 * of the autotuner's table). */
int sph_hip_device_name(char *buf, unsigned len);
/* No counterpart in the reference: a self-check of the device code on the
 * selected device.  The pair passes divide by a float through a double
 * reciprocal formed from the hardware estimate and two Newton steps; this
 * compares that reciprocal with the correctly rounded 1.0 / double(b) for
 * every one of the 2^32 float bit patterns b and stores the number of
 * patterns whose results differ in bits (NaN against NaN counts as equal)
 * in *mismatches.  Takes a few milliseconds. */
int sph_hip_check_recip(unsigned long long *mismatches);
/* No counterpart in the reference: the same sweep for the square root of the
 * fluid density pass (the hardware estimate corrected to the rounded root by
 * fused multiply-add residuals), against the compiler's correctly rounded
 * sqrtf, over every float bit pattern. */
int sph_hip_check_sqrt(unsigned long long *mismatches);

/* ---- the reference's device API under its own names ----------------------
 * src/opencl/particle_system_host.h:7-80 (and set_body_position/rotation,
 * defined at src/opencl/particle_system_host.c:562-588): the same names,
 * signatures and by-value device object, so that a caller written for the
 * reference (scripts/sph/bindings.py:911-1033, scripts/sph/__init__.py,
 * scripts/sph/systems.py) runs on this engine with only the library's path
 * changed.  psdata_opencl has the reference's layout (the cl_mem, cl_kernel
 * and cl_program members are pointers); the engine's handle travels in
 * ps_prog, every OpenCL buffer member is NULL, kernel_names lists the kernel
 * names call_for_all_particles accepts and num_grid_cells is the grid's cell
 * count.  Like the reference, these return nothing and abort on an error
 * (after printing sph_hip_last_error to stderr); the *_hip functions above
 * return a status instead.  sync_psdata_host_to_device and
 * sync_psdata_device_to_host move the whole blob, the *_fields_* variants
 * the named fields (a missing name is skipped with a note), as the
 * reference does. */
typedef struct {
    psdata host_psdata;              /* the psdata header as create_psdata_opencl received it */
    unsigned int num_fields;
    void *names, *names_offsets, *dimensions, *num_dimensions, *dimensions_offsets, *field_types, *entry_sizes;
    void *data, *data_sizes, *data_offsets;
    void *block_totals, *backup_prefix_sum;
    size_t num_kernels;
    const char *const *kernel_names;
    const void *kernels;
    void *ps_prog;                   /* psdata_hip * */
    unsigned int num_grid_cells;
    size_t po2_workgroup_size;
    unsigned int num_blocks;
} psdata_opencl;

void init_opencl(void);                                                              /* :55 */
psdata_opencl create_psdata_opencl(psdata *data, const char *file_list);            /* :56 */
void assign_pso_kernel_args(psdata_opencl pso);       /* :57, no-op: arguments are bound per launch */
void set_kernel_args_to_pso(psdata_opencl pso, void *kernel);                        /* :58, no-op */
void free_psdata_opencl(psdata_opencl *pso);                                         /* :59 */
void terminate_opencl(void);                                                         /* :60 */
void sync_psdata_device_to_host(psdata data, psdata_opencl pso);                     /* :62 */
void sync_psdata_fields_device_to_host(psdata data, psdata_opencl pso, size_t num_fields,
                                       const char *const *field_names);              /* :63 */
void sync_psdata_host_to_device(psdata data, psdata_opencl pso, int full);           /* :64 */
void sync_psdata_fields_host_to_device(psdata data, psdata_opencl pso, size_t num_fields,
                                       const char *const *field_names);              /* :65 */
void populate_position_cuboid_device_opencl(psdata_opencl pso, float x1, float y1, float z1, float x2, float y2,
                                            float z2, unsigned int xsize, unsigned int ysize,
                                            unsigned int zsize);                     /* :67-72 */
void rotate_particles_device_opencl(psdata_opencl pso, float angle_x, float angle_y, float angle_z); /* :73 */
void call_for_all_particles_device_opencl(psdata_opencl pso, const char *kernel_name);             /* :74 */
void compute_particle_bins_device_opencl(psdata_opencl pso);                                       /* :76 */
void compute_density_device_opencl(psdata_opencl pso);                                             /* :78 */
void compute_forces_device_opencl(psdata_opencl pso);                                              /* :79 */
void step_forward_device_opencl(psdata_opencl pso);                                                /* :80 */
void set_body_position_device_opencl(psdata_opencl pso, unsigned int body, float *position);      /* host.c:562 */
void set_body_rotation_device_opencl(psdata_opencl pso, unsigned int body, float *euler_angles);  /* host.c:576 */

/* ---- scene batch --------------------------------------------------------
 * No counterpart in the reference, which runs one psdata per OpenCL program
 * and steps it with one launch per kernel (src/opencl/particle_system_host.c:
 * 361-528, scripts/sph/systems.py:148-189).  A batch keeps many independent
 * scenes resident in one pool of particle slots and grid cells and steps them
 * all with one launch per pass; each scene advances one step per batch step
 * until its own step count is used up.  A scene's results are the same as
 * those of the single-scene path at the same precision settings.  Scenes are
 * added from a psdata built from their conf (fields as set from Python) and
 * the kernel file list that create_psdata_hip would take; the system is
 * fluid (fluids), corotational (solid_corotation) or svk (solid_svk).  All
 * calls return 0 on success and -1 on error (message: sph_hip_last_error),
 * except where stated. */
typedef struct sph_batch sph_batch;

/* Device bytes of a pool of `slots` particle slots, `cells` grid cells and
 * `scenes` scene entries (what sph_batch_create allocates). */
size_t sph_batch_device_bytes(unsigned slots, unsigned cells, unsigned scenes);
/* compensated_sums: bit mask as in set_precision_hip, for every scene. */
int sph_batch_create(unsigned slots, unsigned cells, unsigned scenes, int compensated_sums, int polar_fp64,
                     sph_batch **out);
int sph_batch_free(sph_batch *batch);
/* Admits a scene that will run `steps` steps.  Returns 1 (not an error) when
 * the pool has no room for it now; *scene receives its handle.  Solid scenes
 * compute their reference configuration (init_original_position,
 * compute_original_density) before their first step. */
int sph_batch_add(sph_batch *batch, psdata *data, const char *kernel_file_list, unsigned steps, int *scene);
/* 1 when a scene of `slots` particles and `cells` grid cells (plus one
 * padding cell) would be admitted now, else 0 (no allocation). */
int sph_batch_can_admit(sph_batch *batch, unsigned slots, unsigned cells);
/* Releases a scene's slots and cells (finished or not). */
int sph_batch_remove(sph_batch *batch, int scene);
/* Enqueues num_steps batch steps (asynchronous; sph_batch_synchronize or a
 * read waits). */
int sph_batch_step(sph_batch *batch, unsigned num_steps);
/* Capture runs of `steps_per_graph` batch steps (even) into a hipGraph and
 * replay it while the pool layout is unchanged; 0 launches every pass. */
int sph_batch_set_graph_steps(sph_batch *batch, unsigned steps_per_graph);
int sph_batch_synchronize(sph_batch *batch);
int sph_batch_steps_left(sph_batch *batch, int scene, unsigned *steps_left);
/* Copies fields of a scene, in particle order, into the matching fields of
 * `data` (position, velocity, force, density, density0, originalpos,
 * rotation, strain, stress, deformation, body_number). */
int sph_batch_read(sph_batch *batch, int scene, psdata data, size_t num_fields, const char *const *names);
/* Per-pass timing with HIP events on the batch's stream, as
 * set_profiling_hip / get_kernel_stats_hip; steps run without graphs while
 * profiling is on. */
int sph_batch_set_profiling(sph_batch *batch, int enabled);
int sph_batch_reset_kernel_stats(sph_batch *batch);
int sph_batch_get_kernel_stats(sph_batch *batch, int capacity, const char **names, unsigned long long *launches,
                               double *total_ms);
/* out[0..7]: resident scenes, scenes with steps left, used slots, used
 * cells, slot capacity, cell capacity, graph captures, graph launches. */
int sph_batch_info(sph_batch *batch, unsigned long long *out);

#ifdef __cplusplus
}
#endif

#endif /* SPH_HIP_H_ */
