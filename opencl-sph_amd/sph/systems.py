"""Particle systems: the step sequences over the device handle.

Mirrors scripts/sph/systems.py:8-189 (ParticleSystem, StrainSolid, SVKSolid)
and adds Fluid, whose step order is the reference's fluid sequence
(src/mex/simstep_fluid.c:18-23: bins, compute_density,
compute_forces_fluids, step_forward).

Fields are zero-copy numpy views of the host psdata blob with the conf's
dimensions reversed (a `3, pnum` field is an (pnum, 3) array), exactly as in
the reference.  sync_fields_* move only the named fields (the reference's
Python passes extra arguments to the whole-blob sync calls,
scripts/sph/systems.py:89-99, so it always moves everything).
"""
import ctypes
import os

import numpy as np

from . import bindings as bnd
from . import free_data, init_data

CONF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "conf")

_NUMPY_TYPES = {
    bnd.T_CHAR: np.dtype("u1"),
    bnd.T_UINT: np.dtype("u4"),
    bnd.T_INT: np.dtype("i4"),
    bnd.T_FLOAT32: np.dtype("f4"),
    bnd.T_FLOAT64: np.dtype("f8"),
}


def field_view(psdata, index):
    """numpy view of field `index` of a psdata (dims reversed)."""
    nd = psdata.num_dimensions[index]
    off = psdata.dimensions_offsets[index]
    dims = [psdata.dimensions[off + d] for d in range(nd)]
    dtype = _NUMPY_TYPES.get(psdata.data_ptrs[index].type, np.dtype("u1"))
    count = int(np.prod(dims)) if dims else 0
    addr = psdata.data + psdata.data_offsets[index]
    if count == 0:
        return np.zeros(tuple(reversed(dims)), dtype=dtype)
    raw = (ctypes.c_char * (count * dtype.itemsize)).from_address(addr)
    return np.frombuffer(raw, dtype=dtype, count=count).reshape(tuple(reversed(dims)))


class ParticleSystem(object):
    config = None

    def __init__(self):
        result = init_data(self.config)
        if result is None:
            raise bnd.SphError(f"could not read config file {self.config}")
        self.psdata, self.psdata_opencl = result
        self.fields = {}

    @property
    def psdata_hip(self):
        return self.psdata_opencl

    def get_fields(self, *args):
        out = []
        for name in args:
            if name not in self.fields:
                index = bnd.get_field_psdata(self.psdata, name.encode())
                if index == -1:
                    raise Exception("{} is not a field.".format(name))
                self.fields[name] = field_view(self.psdata, index)
            out.append(self.fields[name])
        return tuple(out)

    def set_fields(self, *args, **kwargs):
        """Copy values into the host fields, then upload them unless the first
        positional argument is true (scripts/sph/systems.py:62-81).  Arrays
        fill the leading slice of the field; scalars broadcast."""
        host_only = bool(args and args[0])
        names = list(kwargs.keys())
        fields = self.get_fields(*names)
        for field, name in zip(fields, names):
            value = kwargs[name]
            if isinstance(value, np.ndarray) and value.size > 1:
                np.copyto(field[tuple(slice(None, s) for s in value.shape)], value)
            else:
                field[...] = value          # scalars broadcast, cast to the field's type
        if not host_only:
            self.sync_fields_host_to_device(*names)

    def sync_host_to_device(self):
        self.psdata_opencl.sync_host_to_device(self.psdata, 0)

    def sync_device_to_host(self):
        self.psdata_opencl.sync_device_to_host(self.psdata)

    def sync_fields_host_to_device(self, *args):
        if not args:
            return self.sync_host_to_device()
        self.psdata_opencl.sync_fields_host_to_device(self.psdata, list(args))

    def sync_fields_device_to_host(self, *args):
        if not args:
            return self.sync_device_to_host()
        self.psdata_opencl.sync_fields_device_to_host(self.psdata, list(args))

    def set_body_position(self, body, position):
        self.psdata_opencl.set_body_position(body, np.asarray(position).astype(np.dtype('f4')))

    def set_body_rotation(self, body, euler_angles):
        self.psdata_opencl.set_body_rotation(body, np.asarray(euler_angles).astype(np.dtype('f4')))

    def call_for_all_particles(self, kernel):
        self.psdata_opencl.call_for_all_particles(kernel)

    def compute_particle_bins(self):
        self.psdata_opencl.compute_particle_bins()

    def init(self, **kwargs):
        self.set_fields(**kwargs)

    def setup(self):
        pass

    def update(self):
        if not self.psdata or not self.psdata_opencl:
            self.setup()

    def free(self):
        free_data(self.psdata, self.psdata_opencl)
        self.__init__()


class Fluid(ParticleSystem):
    """Weakly compressible fluid (kernels/fluids.cl, src/mex/simstep_fluid.c:18-23)."""
    config = os.path.join(CONF_DIR, "fluid.conf")

    def update(self):
        super(Fluid, self).update()
        self.psdata_opencl.compute_particle_bins()
        self.call_for_all_particles("compute_density")
        self.call_for_all_particles("compute_forces_fluids")
        self.call_for_all_particles("step_forward")

    def run(self, num_steps):
        """num_steps whole steps without host round trips."""
        self.psdata_opencl.step("fluid", num_steps)


class StrainSolid(ParticleSystem):
    """Corotational linear solid (scripts/sph/systems.py:115-168)."""
    config = os.path.join(CONF_DIR, "strainsolid.conf")
    init_values = None
    system = "corotational"

    def init(self, **kwargs):
        super(StrainSolid, self).init(**kwargs)
        if self.init_values is None:
            self.init_values = {k: (v if np.isscalar(v) else np.copy(v)) for k, v in kwargs.items()}

    def setup(self):
        """Reference configuration (scripts/sph/systems.py:139-146)."""
        self.call_for_all_particles("init_original_position")
        self.psdata_opencl.compute_particle_bins()
        self.call_for_all_particles("compute_original_density")
        self.sync_device_to_host()

    def update(self):
        super(StrainSolid, self).update()
        self.psdata_opencl.compute_particle_bins()
        self.call_for_all_particles("compute_density")
        self.call_for_all_particles("compute_rotations_and_strains")
        self.call_for_all_particles("compute_stresses")
        self.call_for_all_particles("compute_forces_solids")
        self.call_for_all_particles("step_forward")

    def run(self, num_steps, apply_constraints=False):
        self.psdata_opencl.step(self.system, num_steps, apply_constraints)

    def reset(self):
        free_data(self.psdata, self.psdata_opencl)
        self.fields = {}
        result = init_data(self.config)
        if result is None:
            raise bnd.SphError(f"could not read config file {self.config}")
        self.psdata, self.psdata_opencl = result
        self.init(**self.init_values)

    def free(self):
        super(StrainSolid, self).free()
        self.init_values = None


class SceneSystem(ParticleSystem):
    """A ParticleSystem built from a scenes.Scene (spec text instead of a
    .conf file), with the scene's initial fields uploaded and, for solids,
    the reference configuration computed."""

    def __init__(self, scene, setup=True):
        from . import init_data_from_string
        self.scene = scene
        self.system = scene.system
        self.psdata, self.psdata_opencl = init_data_from_string(scene.spec, scene.kernel_files)
        self.fields = {}
        self.set_fields(**scene.fields)
        if setup and scene.system in ("corotational", "svk"):
            self.setup()

    def setup(self):
        self.call_for_all_particles("init_original_position")
        self.psdata_opencl.compute_particle_bins()
        self.call_for_all_particles("compute_original_density")

    def update(self):
        self.run(1)

    def run(self, num_steps, apply_constraints=False):
        self.psdata_opencl.step(self.system, num_steps, apply_constraints)

    def free(self):
        free_data(self.psdata, self.psdata_opencl)


class SVKSolid(StrainSolid):
    """St. Venant-Kirchhoff solid (scripts/sph/systems.py:170-189)."""
    config = os.path.join(CONF_DIR, "svksolid.conf")
    init_values = None
    system = "svk"

    def update(self):
        super(StrainSolid, self).update()
        self.psdata_opencl.compute_particle_bins()
        self.call_for_all_particles("compute_density")
        self.call_for_all_particles("compute_deformations")
        self.call_for_all_particles("compute_stresses")
        self.call_for_all_particles("compute_forces_solids")
        self.call_for_all_particles("step_forward")
