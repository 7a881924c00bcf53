"""The reference's device names (src/opencl/particle_system_host.h:7-80):
psdata_opencl is laid out as the reference's generated ctypes bindings
declare it (scripts/sph/bindings.py:911-959), the library exports every
name, and the abort-on-error behaviour holds.  No device calls: a zeroed
psdata_opencl carries no handle, so these run without a GPU."""
import ctypes
import os
import shutil
import subprocess
import sys
from ctypes import POINTER, Structure, Union, c_char, c_char_p, c_double, c_float, c_int, c_size_t, c_uint, c_void_p

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
INCLUDE = os.path.join(REPO, "include")


# ---- restatement of the reference's ctypes declarations ----------------------
# DataPtr (scripts/sph/bindings.py:644-680), psdata (:683-719),
# psdata_opencl (:911-959).  cl_mem, cl_kernel and cl_program are pointers to
# opaque structs there; any pointer has the same size and alignment.
class _RefUnion(Union):
    _fields_ = [("c", c_char_p), ("u", POINTER(ctypes.c_uint32)), ("i", POINTER(ctypes.c_int32)),
                ("f4", POINTER(c_float)), ("f8", POINTER(c_double)), ("ptr", c_void_p)]


class _RefDataPtr(Structure):
    _fields_ = [("type", c_int), ("unnamed_1", _RefUnion)]


class _RefPsdata(Structure):
    _fields_ = [("num_fields", c_uint), ("names", c_char_p), ("names_offsets", POINTER(c_uint)),
                ("dimensions", POINTER(c_uint)), ("num_dimensions", POINTER(c_uint)),
                ("dimensions_offsets", POINTER(c_uint)), ("entry_sizes", POINTER(c_uint)), ("data", c_void_p),
                ("data_sizes", POINTER(c_uint)), ("data_offsets", POINTER(c_uint)),
                ("data_ptrs", POINTER(_RefDataPtr)), ("num_host_fields", c_uint),
                ("host_names", POINTER(POINTER(c_char))), ("host_data", POINTER(c_void_p)),
                ("host_data_size", POINTER(c_uint))]


cl_mem = c_void_p
cl_kernel = c_void_p
cl_program = c_void_p


class _RefPsdataOpencl(Structure):
    _fields_ = [
        ("host_psdata", _RefPsdata),
        ("num_fields", c_uint),
        ("names", cl_mem),
        ("names_offsets", cl_mem),
        ("dimensions", cl_mem),
        ("num_dimensions", cl_mem),
        ("dimensions_offsets", cl_mem),
        ("field_types", cl_mem),
        ("entry_sizes", cl_mem),
        ("data", cl_mem),
        ("data_sizes", cl_mem),
        ("data_offsets", cl_mem),
        ("block_totals", cl_mem),
        ("backup_prefix_sum", cl_mem),
        ("num_kernels", c_size_t),
        ("kernel_names", POINTER(POINTER(c_char))),
        ("kernels", POINTER(cl_kernel)),
        ("ps_prog", cl_program),
        ("num_grid_cells", c_uint),
        ("po2_workgroup_size", c_size_t),
        ("num_blocks", c_uint),
    ]


REFERENCE_NAMES = (
    "init_opencl", "create_psdata_opencl", "assign_pso_kernel_args", "set_kernel_args_to_pso",
    "free_psdata_opencl", "terminate_opencl", "sync_psdata_device_to_host", "sync_psdata_fields_device_to_host",
    "sync_psdata_host_to_device", "sync_psdata_fields_host_to_device", "populate_position_cuboid_device_opencl",
    "rotate_particles_device_opencl", "call_for_all_particles_device_opencl", "compute_particle_bins_device_opencl",
    "compute_density_device_opencl", "compute_forces_device_opencl", "step_forward_device_opencl",
    "set_body_position_device_opencl", "set_body_rotation_device_opencl")


def layout(struct):
    return {"sizeof": ctypes.sizeof(struct),
            **{name: getattr(struct, name).offset for name, _ in struct._fields_}}


def c_layout(tmp_path, type_name, fields):
    """sizeof and offsetof as the C compiler lays out include/sph_hip.h."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = [f'printf("sizeof %zu\\n", sizeof({type_name}));']
    lines += [f'printf("{f} %zu\\n", offsetof({type_name}, {f}));' for f in fields]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "sph_hip.h"\nint main(void) {\n'
                   + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", INCLUDE, str(src), "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_psdata_opencl_matches_the_reference_ctypes_layout(tmp_path):
    want = layout(_RefPsdataOpencl)
    assert c_layout(tmp_path, "psdata_opencl", [f for f, _ in _RefPsdataOpencl._fields_]) == want
    from sph import bindings as bnd
    assert layout(bnd.psdata_opencl) == want
    # the embedded psdata is the reference's too
    assert c_layout(tmp_path, "psdata", [f for f, _ in _RefPsdata._fields_]) == layout(_RefPsdata)
    assert layout(bnd.psdata) == layout(_RefPsdata)


def test_library_exports_the_reference_names():
    from sph import bindings as bnd
    lib = ctypes.CDLL(bnd.LIBRARY)
    missing = [n for n in REFERENCE_NAMES if not hasattr(lib, n)]
    assert not missing, missing
    for n in REFERENCE_NAMES:
        assert callable(getattr(bnd, n))


def test_free_of_an_empty_object_is_a_no_op():
    from sph import bindings as bnd
    pso = bnd.psdata_opencl()
    bnd.free_psdata_opencl(ctypes.byref(pso))
    assert not pso.ps_prog
    bnd.assign_pso_kernel_args(pso)
    bnd.set_kernel_args_to_pso(pso, None)


@pytest.mark.parametrize("call", ["call_for_all_particles_device_opencl(pso, b'compute_density')",
                                  "compute_particle_bins_device_opencl(pso)",
                                  "step_forward_device_opencl(pso)"])
def test_a_call_without_a_handle_aborts(call):
    """Errors abort after printing the cause, like the reference's ASSERTs
    (src/opencl/particle_system_host.c, HANDLE_CL_ERROR)."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from sph.bindings import *\n"
            "pso = psdata_opencl()\n" % os.path.join(REPO, "opencl-sph_amd")) + call + "\nprint('returned')\n"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "holds no device handle" in r.stderr
