"""A scene driven through the reference's device names
(src/opencl/particle_system_host.h:55-80, called as the reference's
scripts/sph/systems.py does) gives bit for bit what the *_hip path and the
oracle give, and the by-value psdata_opencl carries what the reference's
does (field count, grid cell count, kernel names)."""
import ctypes
import dataclasses

import numpy as np
import pytest

from gpu_util import compare, download, engine, oracle_state, say
from sph import bindings as bnd
from sph import init_data_cpu, free_data_cpu, scenes
from sph.systems import field_view

pytestmark = pytest.mark.gpu

FLUID_SEQ = ("compute_density", "compute_forces_fluids", "step_forward")
SOLID_SEQ = ("compute_density", "compute_rotations_and_strains", "compute_stresses", "compute_forces_solids",
             "step_forward")


def names(*fields):
    arr = (ctypes.c_char_p * len(fields))()
    arr[:] = [f.encode() for f in fields]
    return arr


def load_by_reference_names(scene):
    """psdata from the scene's spec, initial fields written into the host
    blob, the whole blob uploaded: scripts/sph/__init__.py:11-22 and
    scripts/sph/systems.py:62-99."""
    ps = init_data_cpu(scene.spec)
    for name, value in scene.fields.items():
        view = field_view(ps, bnd.get_field_psdata(ps, name.encode()))
        value = np.asarray(value)
        if value.size > 1:
            np.copyto(view[tuple(slice(None, s) for s in value.shape)], value.astype(view.dtype))
        else:
            view[...] = value
    bnd.init_opencl()
    pso = bnd.create_psdata_opencl(ctypes.byref(ps), scene.kernel_files.encode())
    bnd.sync_psdata_host_to_device(ps, pso, 1)
    return ps, pso


def read(ps, pso, *fields):
    bnd.sync_psdata_fields_device_to_host(ps, pso, len(fields), names(*fields))
    return {f: np.array(field_view(ps, bnd.get_field_psdata(ps, f.encode())), copy=True) for f in fields}


def check_object(ps, pso, scene):
    assert pso.ps_prog
    assert pso.num_fields == ps.num_fields
    assert pso.host_psdata.num_fields == ps.num_fields and pso.host_psdata.data == ps.data
    gridres = field_view(ps, bnd.get_field_psdata(ps, b"gridres"))
    assert pso.num_grid_cells == int(np.prod(gridres.astype(np.int64)))
    listed = [ctypes.string_at(pso.kernel_names[k]).decode() for k in range(pso.num_kernels)]
    for k in FLUID_SEQ + SOLID_SEQ + ("find_particle_bins",):
        assert k in listed
    assert not pso.data and not pso.kernels


def test_fluid_through_reference_names(hip_device):
    scene = scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, eos_tait=True, mass_consistent=True)
    steps = 20
    say(f"fluid {steps} steps through the reference names, n={scene.n}")
    ps, pso = load_by_reference_names(scene)
    check_object(ps, pso, scene)
    for _ in range(steps):
        bnd.compute_particle_bins_device_opencl(pso)
        bnd.compute_density_device_opencl(pso)
        bnd.call_for_all_particles_device_opencl(pso, b"compute_forces_fluids")
        bnd.step_forward_device_opencl(pso)
    fields = ("density", "num_neighbours", "force", "velocity", "position")
    got = read(ps, pso, *fields)
    bnd.free_psdata_opencl(ctypes.byref(pso))
    assert not pso.ps_prog
    free_data_cpu(ps)

    s = engine(scene)
    s.run(steps)
    want = download(s, *fields)
    s.free()
    st = oracle_state(scene)
    st.step("fluid", steps)
    for f in ("density", "force", "velocity", "position"):
        compare(f, got[f], want[f])
        compare(f + " (oracle)", got[f], st.f[f])
    assert np.array_equal(got["num_neighbours"], want["num_neighbours"])


def test_corotational_solid_through_reference_names(hip_device):
    scene = scenes.cantilever(nx=16, ny=6, nz=6)
    steps = 10
    say(f"cantilever {steps} steps through the reference names, n={scene.n}")
    ps, pso = load_by_reference_names(scene)
    check_object(ps, pso, scene)
    # reference configuration, scripts/sph/systems.py:139-146
    bnd.call_for_all_particles_device_opencl(pso, b"init_original_position")
    bnd.compute_particle_bins_device_opencl(pso)
    bnd.call_for_all_particles_device_opencl(pso, b"compute_original_density")
    bnd.sync_psdata_device_to_host(ps, pso)
    for _ in range(steps):
        bnd.compute_particle_bins_device_opencl(pso)
        for k in SOLID_SEQ:
            bnd.call_for_all_particles_device_opencl(pso, k.encode())
    fields = ("density", "rotation", "strain", "stress", "force", "velocity", "position")
    got = read(ps, pso, *fields)
    bnd.free_psdata_opencl(ctypes.byref(pso))
    free_data_cpu(ps)

    s = engine(scene)
    s.run(steps)
    want = download(s, *fields)
    s.free()
    for f in fields:
        compare(f, got[f], want[f])


def test_body_calls_through_reference_names(hip_device):
    """set_body_position / set_body_rotation (src/opencl/particle_system_host.c:562-588)
    and rotate_particles through the reference names equal the *_hip calls."""
    scene = scenes.cantilever(nx=12, ny=4, nz=4)
    scene = dataclasses.replace(scene, kernel_files=scene.kernel_files + ", rigidbody")
    ps, pso = load_by_reference_names(scene)
    bnd.call_for_all_particles_device_opencl(pso, b"init_original_position")
    pos3 = (ctypes.c_float * 3)(0.01, -0.02, 0.005)
    ang3 = (ctypes.c_float * 3)(0.1, 0.0, -0.2)
    bnd.set_body_position_device_opencl(pso, 0, pos3)
    bnd.set_body_rotation_device_opencl(pso, 0, ang3)
    bnd.rotate_particles_device_opencl(pso, 0.05, 0.1, 0.0)
    got = read(ps, pso, "position", "originalpos")
    bnd.free_psdata_opencl(ctypes.byref(pso))
    free_data_cpu(ps)

    s = engine(scene, setup=False)
    s.call_for_all_particles("init_original_position")
    s.set_body_position(0, (0.01, -0.02, 0.005))
    s.set_body_rotation(0, (0.1, 0.0, -0.2))
    s.psdata_opencl.rotate_particles(0.05, 0.1, 0.0)
    want = download(s, "position", "originalpos")
    s.free()
    for f in ("position", "originalpos"):
        compare(f, got[f], want[f])


def test_populate_cuboid_through_reference_names(hip_device):
    scene = scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, seed=5)
    ps, pso = load_by_reference_names(scene)
    bnd.populate_position_cuboid_device_opencl(pso, -1, -1, -1, 1, 1, 1, 6, 6, 6)
    got = read(ps, pso, "position", "n")
    bnd.free_psdata_opencl(ctypes.byref(pso))
    free_data_cpu(ps)

    s = engine(scene)
    s.psdata_opencl.populate_position_cuboid((-1, -1, -1), (1, 1, 1), (6, 6, 6))
    want = download(s, "position", "n")
    s.free()
    assert int(got["n"].reshape(-1)[0]) == int(want["n"].reshape(-1)[0]) == 216
    compare("position", got["position"], want["position"])
