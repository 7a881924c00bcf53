"""ctypes declarations for libsph_hip.so (include/sph_hip.h).

The psdata struct and the host-side functions keep the layout and names of
the reference's generated bindings (scripts/sph/bindings.py:625-760); the
device functions are the HIP replacements of the OpenCL ones
(scripts/sph/bindings.py:961-1062).  Device calls return a status; `check`
turns a non-zero status into SphError carrying the library's message.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, Union, c_char, c_char_p, c_double, c_float, c_int, c_size_t,
                    c_uint, c_ulonglong, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.environ.get("SPH_HIP_LIBRARY", os.path.join(HERE, "libsph_hip.so"))


class SphError(RuntimeError):
    pass


if not os.path.exists(LIBRARY):
    raise ImportError(
        f"{LIBRARY} is missing: build it with `python opencl-sph_amd/build.py` "
        "(there is no CPU fallback for the simulation step)")

_lib = ctypes.CDLL(LIBRARY)

# FieldType (src/datatypes.h:4-11)
T_CHAR, T_UINT, T_INT, T_FLOAT32, T_FLOAT64, TYPE_COUNT = range(6)


class _DataPtrUnion(Union):
    _fields_ = [("c", POINTER(c_char)), ("u", POINTER(c_uint)), ("i", POINTER(ctypes.c_int32)),
                ("f4", POINTER(c_float)), ("f8", POINTER(c_double)), ("ptr", c_void_p)]


class DataPtr(Structure):
    _anonymous_ = ("_u",)
    _fields_ = [("type", c_int), ("_u", _DataPtrUnion)]


class psdata(Structure):
    _fields_ = [
        ("num_fields", c_uint),
        ("names", c_char_p),
        ("names_offsets", POINTER(c_uint)),
        ("dimensions", POINTER(c_uint)),
        ("num_dimensions", POINTER(c_uint)),
        ("dimensions_offsets", POINTER(c_uint)),
        ("entry_sizes", POINTER(c_uint)),
        ("data", c_void_p),
        ("data_sizes", POINTER(c_uint)),
        ("data_offsets", POINTER(c_uint)),
        ("data_ptrs", POINTER(DataPtr)),
        ("num_host_fields", c_uint),
        ("host_names", POINTER(POINTER(c_char))),
        ("host_data", POINTER(c_void_p)),
        ("host_data_size", POINTER(c_uint)),
    ]


def _fn(name, argtypes, restype):
    f = getattr(_lib, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


# ---- logging, config, psdata (same names and signatures as libSPH.so) -----
set_log_level = _fn("set_log_level", [c_uint], None)
load_config = _fn("load_config", [c_char_p], c_int)
_get_config_section = _fn("get_config_section", [c_char_p], ctypes.c_char_p)
unload_config = _fn("unload_config", [], None)
build_psdata_from_string = _fn("build_psdata_from_string", [POINTER(psdata), c_char_p], None)
build_psdata = _fn("build_psdata", [POINTER(psdata), c_char_p], None)
psdata_last_build_status = _fn("psdata_last_build_status", [], c_int)
get_field_psdata = _fn("get_field_psdata", [psdata, c_char_p], c_int)
set_field_psdata = _fn("set_field_psdata", [POINTER(psdata), c_char_p, c_void_p, c_uint, c_uint], None)
psdata_names_size = _fn("psdata_names_size", [psdata], c_uint)
psdata_dimensions_size = _fn("psdata_dimensions_size", [psdata], c_uint)
psdata_data_size = _fn("psdata_data_size", [psdata], c_uint)
create_host_field_psdata = _fn("create_host_field_psdata", [POINTER(psdata), c_char_p, c_void_p, c_uint], c_int)
get_host_field_psdata = _fn("get_host_field_psdata", [POINTER(psdata), c_char_p], c_int)
free_psdata = _fn("free_psdata", [POINTER(psdata)], None)


def get_config_section(heading):
    """Section body as text (None when absent)."""
    if isinstance(heading, str):
        heading = heading.encode()
    raw = _get_config_section(heading)
    return None if raw is None else raw.decode()


# ---- device runtime ---------------------------------------------------------
_last_error = _fn("sph_hip_last_error", [], c_char_p)


def check(status):
    if status != 0:
        msg = _last_error()
        raise SphError(msg.decode() if msg else f"libsph_hip call failed with status {status}")
    return status


_init_hip = _fn("init_hip", [], c_int)
_terminate_hip = _fn("terminate_hip", [], c_int)
_create_psdata_hip = _fn("create_psdata_hip", [POINTER(psdata), c_char_p, POINTER(c_void_p)], c_int)
_free_psdata_hip = _fn("free_psdata_hip", [c_void_p], c_int)
_sync_h2d = _fn("sync_psdata_host_to_device_hip", [psdata, c_void_p, c_int], c_int)
_sync_d2h = _fn("sync_psdata_device_to_host_hip", [psdata, c_void_p], c_int)
_sync_fields_h2d = _fn("sync_psdata_fields_host_to_device_hip", [psdata, c_void_p, c_size_t, POINTER(c_char_p)], c_int)
_sync_fields_d2h = _fn("sync_psdata_fields_device_to_host_hip", [psdata, c_void_p, c_size_t, POINTER(c_char_p)], c_int)
_call_for_all = _fn("call_for_all_particles_device_hip", [c_void_p, c_char_p], c_int)
_bins = _fn("compute_particle_bins_device_hip", [c_void_p], c_int)
_density = _fn("compute_density_device_hip", [c_void_p], c_int)
_forces = _fn("compute_forces_device_hip", [c_void_p], c_int)
_step_forward = _fn("step_forward_device_hip", [c_void_p], c_int)
_step = _fn("step_device_hip", [c_void_p, c_char_p, c_uint, c_int], c_int)
_synchronize = _fn("synchronize_hip", [c_void_p], c_int)
_set_precision = _fn("set_precision_hip", [c_void_p, c_int, c_int], c_int)
_set_profiling = _fn("set_profiling_hip", [c_void_p, c_int], c_int)
_reset_stats = _fn("reset_kernel_stats_hip", [c_void_p], c_int)
_get_stats = _fn("get_kernel_stats_hip", [c_void_p, c_int, POINTER(c_char_p), POINTER(c_ulonglong),
                                          POINTER(c_double)], c_int)
_global_sums = _fn("global_sums_hip", [c_void_p, POINTER(c_double)], c_int)
_populate_cuboid = _fn("populate_position_cuboid_device_hip",
                       [c_void_p, c_float, c_float, c_float, c_float, c_float, c_float, c_uint, c_uint, c_uint], c_int)
_rotate_particles = _fn("rotate_particles_device_hip", [c_void_p, c_float, c_float, c_float], c_int)
_set_body_position = _fn("set_body_position_device_hip", [c_void_p, c_uint, POINTER(c_float)], c_int)
_set_body_rotation = _fn("set_body_rotation_device_hip", [c_void_p, c_uint, POINTER(c_float)], c_int)


def _names_array(names):
    arr = (c_char_p * len(names))()
    arr[:] = [n.encode() if isinstance(n, str) else n for n in names]
    return arr


_set_tuning = _fn("sph_hip_set_tuning", [c_uint], c_int)
_get_tuning = _fn("sph_hip_get_tuning", [POINTER(c_uint)], c_int)
_set_pair_group = _fn("sph_hip_set_fluid_pair_group", [c_uint], c_int)
_get_pair_group = _fn("sph_hip_get_fluid_pair_group", [POINTER(c_uint)], c_int)
_device_name = _fn("sph_hip_device_name", [ctypes.c_char_p, c_uint], c_int)
_check_recip = _fn("sph_hip_check_recip", [POINTER(c_ulonglong)], c_int)
_check_sqrt = _fn("sph_hip_check_sqrt", [POINTER(c_ulonglong)], c_int)


def init_hip():
    check(_init_hip())


def set_tuning(neighbour_block, pair_group=None):
    """Threads per block of the neighbour passes; list entries per load group
    of the fluid force pass (unchanged when None)."""
    check(_set_tuning(int(neighbour_block)))
    if pair_group is not None:
        check(_set_pair_group(int(pair_group)))


def get_tuning():
    out, pg = c_uint(), c_uint()
    check(_get_tuning(ctypes.byref(out)))
    check(_get_pair_group(ctypes.byref(pg)))
    return {"neighbour_block": out.value, "pair_group": pg.value}


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(_device_name(buf, 256))
    return buf.value.decode()


def check_recip():
    """Number of float bit patterns whose device reciprocal (hardware
    estimate + two Newton steps) differs from the correctly rounded
    1.0 / double(b): 0 means the pair passes' divisions are exact."""
    bad = c_ulonglong(0)
    check(_check_recip(ctypes.byref(bad)))
    return int(bad.value)


def check_sqrt():
    """Number of float bit patterns whose device square root in the fluid
    density pass (hardware estimate + residual correction) differs from the
    correctly rounded sqrtf: 0 means the pass's distances are exact."""
    bad = c_ulonglong(0)
    check(_check_sqrt(ctypes.byref(bad)))
    return int(bad.value)


def terminate_hip():
    check(_terminate_hip())


class PsdataHip:
    """Device handle for one psdata (the HIP counterpart of psdata_opencl)."""

    def __init__(self, data, file_list):
        self._data = data
        handle = c_void_p()
        if isinstance(file_list, str):
            file_list = file_list.encode()
        check(_create_psdata_hip(ctypes.byref(data), file_list, ctypes.byref(handle)))
        self.handle = handle

    def __bool__(self):
        return bool(self.handle)

    def free(self):
        if self.handle:
            check(_free_psdata_hip(self.handle))
            self.handle = c_void_p()

    def sync_host_to_device(self, data, full=0):
        check(_sync_h2d(data, self.handle, int(full)))

    def sync_device_to_host(self, data):
        check(_sync_d2h(data, self.handle))

    def sync_fields_host_to_device(self, data, names):
        arr = _names_array(names)
        check(_sync_fields_h2d(data, self.handle, len(names), arr))

    def sync_fields_device_to_host(self, data, names):
        arr = _names_array(names)
        check(_sync_fields_d2h(data, self.handle, len(names), arr))

    def call_for_all_particles(self, kernel):
        check(_call_for_all(self.handle, kernel.encode() if isinstance(kernel, str) else kernel))

    def compute_particle_bins(self):
        check(_bins(self.handle))

    def compute_density(self):
        check(_density(self.handle))

    def compute_forces(self):
        check(_forces(self.handle))

    def step_forward(self):
        check(_step_forward(self.handle))

    def step(self, system, num_steps=1, apply_constraints=False):
        check(_step(self.handle, system.encode(), int(num_steps), int(bool(apply_constraints))))

    def synchronize(self):
        check(_synchronize(self.handle))

    KAHAN_DENSITY, KAHAN_FORCES = 1, 2

    def set_precision(self, compensated_sums=0, polar_fp64=True):
        """compensated_sums: bit mask of KAHAN_DENSITY / KAHAN_FORCES (True = both)."""
        flags = 3 if compensated_sums is True else int(compensated_sums)
        check(_set_precision(self.handle, flags, int(bool(polar_fp64))))

    def set_profiling(self, enabled=True):
        check(_set_profiling(self.handle, int(bool(enabled))))

    def reset_kernel_stats(self):
        check(_reset_stats(self.handle))

    def kernel_stats(self):
        """{kernel name: (launches, total milliseconds)} from HIP events."""
        count = _get_stats(self.handle, 0, None, None, None)
        if count < 0:
            check(-1)
        names = (c_char_p * max(count, 1))()
        launches = (c_ulonglong * max(count, 1))()
        total = (c_double * max(count, 1))()
        count = _get_stats(self.handle, count, names, launches, total)
        if count < 0:
            check(-1)
        return {names[k].decode(): (int(launches[k]), float(total[k])) for k in range(count)}

    def global_sums(self):
        """(momentum[3], kinetic energy, centre of mass[3]) reduced in FP64."""
        out = (c_double * 7)()
        check(_global_sums(self.handle, out))
        return list(out[0:3]), out[3], list(out[4:7])

    # initial-condition helpers and scripted rigid bodies
    # (populate_position_cuboid_device_opencl, rotate_particles_device_opencl,
    # set_body_position/rotation_device_opencl; scripts/sph/bindings.py)
    def populate_position_cuboid(self, corner1, corner2, size):
        c1 = [float(v) for v in corner1]
        c2 = [float(v) for v in corner2]
        check(_populate_cuboid(self.handle, *c1, *c2, *(int(v) for v in size)))

    def rotate_particles(self, angle_x, angle_y, angle_z):
        check(_rotate_particles(self.handle, float(angle_x), float(angle_y), float(angle_z)))

    def set_body_position(self, body, position):
        v = (c_float * 3)(*[float(x) for x in position])
        check(_set_body_position(self.handle, int(body), v))

    def set_body_rotation(self, body, euler_angles):
        v = (c_float * 3)(*[float(x) for x in euler_angles])
        check(_set_body_rotation(self.handle, int(body), v))


# ---- the reference's device names (scripts/sph/bindings.py:911-1062) -------
# The by-value device object and the void, abort-on-error calls of
# src/opencl/particle_system_host.h:7-80, so code written against the
# reference's generated bindings runs unchanged.  The OpenCL members are
# pointers (NULL here); ps_prog carries the engine's psdata_hip handle.
class psdata_opencl(Structure):
    _fields_ = [
        ("host_psdata", psdata),
        ("num_fields", c_uint),
        ("names", c_void_p),
        ("names_offsets", c_void_p),
        ("dimensions", c_void_p),
        ("num_dimensions", c_void_p),
        ("dimensions_offsets", c_void_p),
        ("field_types", c_void_p),
        ("entry_sizes", c_void_p),
        ("data", c_void_p),
        ("data_sizes", c_void_p),
        ("data_offsets", c_void_p),
        ("block_totals", c_void_p),
        ("backup_prefix_sum", c_void_p),
        ("num_kernels", c_size_t),
        ("kernel_names", POINTER(POINTER(c_char))),
        ("kernels", c_void_p),
        ("ps_prog", c_void_p),
        ("num_grid_cells", c_uint),
        ("po2_workgroup_size", c_size_t),
        ("num_blocks", c_uint),
    ]


init_opencl = _fn("init_opencl", [], None)
create_psdata_opencl = _fn("create_psdata_opencl", [POINTER(psdata), c_char_p], psdata_opencl)
assign_pso_kernel_args = _fn("assign_pso_kernel_args", [psdata_opencl], None)
set_kernel_args_to_pso = _fn("set_kernel_args_to_pso", [psdata_opencl, c_void_p], None)
free_psdata_opencl = _fn("free_psdata_opencl", [POINTER(psdata_opencl)], None)
terminate_opencl = _fn("terminate_opencl", [], None)
sync_psdata_device_to_host = _fn("sync_psdata_device_to_host", [psdata, psdata_opencl], None)
sync_psdata_fields_device_to_host = _fn("sync_psdata_fields_device_to_host",
                                        [psdata, psdata_opencl, c_size_t, POINTER(c_char_p)], None)
sync_psdata_host_to_device = _fn("sync_psdata_host_to_device", [psdata, psdata_opencl, c_int], None)
sync_psdata_fields_host_to_device = _fn("sync_psdata_fields_host_to_device",
                                        [psdata, psdata_opencl, c_size_t, POINTER(c_char_p)], None)
populate_position_cuboid_device_opencl = _fn(
    "populate_position_cuboid_device_opencl",
    [psdata_opencl, c_float, c_float, c_float, c_float, c_float, c_float, c_uint, c_uint, c_uint], None)
rotate_particles_device_opencl = _fn("rotate_particles_device_opencl", [psdata_opencl, c_float, c_float, c_float],
                                     None)
call_for_all_particles_device_opencl = _fn("call_for_all_particles_device_opencl", [psdata_opencl, c_char_p], None)
compute_particle_bins_device_opencl = _fn("compute_particle_bins_device_opencl", [psdata_opencl], None)
compute_density_device_opencl = _fn("compute_density_device_opencl", [psdata_opencl], None)
compute_forces_device_opencl = _fn("compute_forces_device_opencl", [psdata_opencl], None)
step_forward_device_opencl = _fn("step_forward_device_opencl", [psdata_opencl], None)
set_body_position_device_opencl = _fn("set_body_position_device_opencl", [psdata_opencl, c_uint, POINTER(c_float)],
                                      None)
set_body_rotation_device_opencl = _fn("set_body_rotation_device_opencl", [psdata_opencl, c_uint, POINTER(c_float)],
                                      None)


# ---- scene batch ----------------------------------------------------------
_batch_bytes = _fn("sph_batch_device_bytes", [c_uint, c_uint, c_uint], c_size_t)
_batch_create = _fn("sph_batch_create", [c_uint, c_uint, c_uint, c_int, c_int, POINTER(c_void_p)], c_int)
_batch_free = _fn("sph_batch_free", [c_void_p], c_int)
_batch_add = _fn("sph_batch_add", [c_void_p, POINTER(psdata), c_char_p, c_uint, POINTER(c_int)], c_int)
_batch_remove = _fn("sph_batch_remove", [c_void_p, c_int], c_int)
_batch_can_admit = _fn("sph_batch_can_admit", [c_void_p, c_uint, c_uint], c_int)
_batch_step = _fn("sph_batch_step", [c_void_p, c_uint], c_int)
_batch_graph = _fn("sph_batch_set_graph_steps", [c_void_p, c_uint], c_int)
_batch_sync = _fn("sph_batch_synchronize", [c_void_p], c_int)
_batch_left = _fn("sph_batch_steps_left", [c_void_p, c_int, POINTER(c_uint)], c_int)
_batch_read = _fn("sph_batch_read", [c_void_p, c_int, psdata, c_size_t, POINTER(c_char_p)], c_int)
_batch_info = _fn("sph_batch_info", [c_void_p, POINTER(c_ulonglong)], c_int)
_batch_prof = _fn("sph_batch_set_profiling", [c_void_p, c_int], c_int)
_batch_reset_stats = _fn("sph_batch_reset_kernel_stats", [c_void_p], c_int)
_batch_stats = _fn("sph_batch_get_kernel_stats", [c_void_p, c_int, POINTER(c_char_p), POINTER(c_ulonglong),
                                                  POINTER(c_double)], c_int)


def _stats(fn, handle):
    count = fn(handle, 0, None, None, None)
    if count < 0:
        check(-1)
    names = (c_char_p * max(count, 1))()
    launches = (c_ulonglong * max(count, 1))()
    total = (c_double * max(count, 1))()
    count = fn(handle, count, names, launches, total)
    if count < 0:
        check(-1)
    return {names[k].decode(): (int(launches[k]), float(total[k])) for k in range(count)}


def batch_device_bytes(slots, cells, scenes):
    return int(_batch_bytes(slots, cells, scenes))


class SphBatch:
    """Device pool of resident scenes stepped together (include/sph_hip.h)."""

    def __init__(self, slots, cells, scenes, compensated_sums=0, polar_fp64=True):
        flags = 3 if compensated_sums is True else int(compensated_sums)
        handle = c_void_p()
        check(_batch_create(slots, cells, scenes, flags, int(bool(polar_fp64)), ctypes.byref(handle)))
        self.handle = handle

    def free(self):
        if self.handle:
            check(_batch_free(self.handle))
            self.handle = c_void_p()

    def add(self, data, file_list, steps):
        """Scene handle, or None when the pool has no room now."""
        if isinstance(file_list, str):
            file_list = file_list.encode()
        scene = c_int(-1)
        rc = _batch_add(self.handle, ctypes.byref(data), file_list, int(steps), ctypes.byref(scene))
        if rc == 1:
            return None
        check(rc)
        return scene.value

    def can_admit(self, slots, cells):
        """Whether a scene of `slots` particles and `cells` cells (padding
        cell included) fits now."""
        rc = _batch_can_admit(self.handle, int(slots), int(cells))
        if rc < 0:
            check(rc)
        return rc == 1

    def remove(self, scene):
        check(_batch_remove(self.handle, int(scene)))

    def step(self, num_steps=1):
        check(_batch_step(self.handle, int(num_steps)))

    def set_graph_steps(self, steps_per_graph):
        check(_batch_graph(self.handle, int(steps_per_graph)))

    def synchronize(self):
        check(_batch_sync(self.handle))

    def steps_left(self, scene):
        out = c_uint()
        check(_batch_left(self.handle, int(scene), ctypes.byref(out)))
        return out.value

    def read(self, scene, data, names):
        arr = _names_array(names)
        check(_batch_read(self.handle, int(scene), data, len(names), arr))

    def set_profiling(self, enabled=True):
        check(_batch_prof(self.handle, int(bool(enabled))))

    def reset_kernel_stats(self):
        check(_batch_reset_stats(self.handle))

    def kernel_stats(self):
        """{pass name: (launches, total milliseconds)} from HIP events."""
        return _stats(_batch_stats, self.handle)

    def info(self):
        out = (c_ulonglong * 8)()
        check(_batch_info(self.handle, out))
        keys = ("scenes", "running", "used_slots", "used_cells", "slot_capacity", "cell_capacity",
                "graph_captures", "graph_launches")
        return dict(zip(keys, (int(v) for v in out)))
