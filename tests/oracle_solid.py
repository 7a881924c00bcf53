"""A stand-in for sph.systems.StrainSolid / SVKSolid whose passes run on the
CPU oracle, with its fields in a host psdata built from the same conf, so a
harness (sph.straintest) can drive the engine and the oracle alike and their
results can be compared.  Test infrastructure."""
import os

import numpy as np

import oracle
import sph
from sph import bindings as bnd
from sph.systems import field_view

CONF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opencl-sph_amd", "conf")


class OracleSolid:
    def __init__(self, system="corotational", precision="f32"):
        self.system = system
        self.precision = precision
        self.conf = os.path.join(CONF_DIR, "strainsolid.conf" if system == "corotational" else "svksolid.conf")
        self.init_values = None
        self._build()

    def _build(self):
        if bnd.load_config(self.conf.encode()) > 0:
            raise RuntimeError(f"cannot read {self.conf}")
        try:
            self.psdata = sph.init_data_cpu(bnd.get_config_section("psdata_specification"))
        finally:
            bnd.unload_config()
        self.fields = {}

    def get_fields(self, *names):
        out = []
        for name in names:
            if name not in self.fields:
                index = bnd.get_field_psdata(self.psdata, name.encode())
                if index == -1:
                    raise Exception(f"{name} is not a field.")
                self.fields[name] = field_view(self.psdata, index)
            out.append(self.fields[name])
        return tuple(out)

    def set_fields(self, *args, **kwargs):
        for name, value in kwargs.items():
            (field,) = self.get_fields(name)
            if isinstance(value, np.ndarray) and value.size > 1:
                np.copyto(field[tuple(slice(None, s) for s in value.shape)], value)
            else:
                field[...] = value          # scalars broadcast, cast to the field's type

    def init(self, **kwargs):
        self.set_fields(**kwargs)
        if self.init_values is None:
            self.init_values = {k: (v if np.isscalar(v) else np.copy(v)) for k, v in kwargs.items()}

    # device syncs are no-ops: the oracle works on the host fields directly
    def sync_host_to_device(self):
        pass

    def sync_device_to_host(self):
        pass

    def sync_fields_host_to_device(self, *names):
        pass

    def sync_fields_device_to_host(self, *names):
        pass

    def _state(self):
        g = lambda name: self.get_fields(name)[0]
        n = int(g("n")[0])
        params = {"n": n, "gridbounds": g("gridbounds").reshape(-1).tolist(), "h": float(g("smoothingradius")[0]),
                  "mass": float(g("mass")[0]), "dt": float(g("timestep")[0]), "viscosity": float(g("viscosity")[0]),
                  "youngs": float(g("youngs_modulus")[0]), "poisson": float(g("poisson")[0]),
                  "gravity": g("gravity").reshape(-1).tolist(), "polar_fp64": 1,
                  "symmetry_planes": g("symmetry_planes").reshape(-1, 6),
                  "plane_constraints": g("plane_constraints").reshape(-1, 6)}
        if bnd.get_field_psdata(self.psdata, b"svk_sigma_zz_zero") >= 0:
            params["svk_sigma_zz_zero"] = int(g("svk_sigma_zz_zero")[0])
        names = ["position", "velocity", "originalpos", "density", "density0", "force", "stress"]
        names += ["rotation", "strain"] if self.system == "corotational" else ["deformation"]
        names += [k for k in ("force_elastic", "force_viscous", "cutoff_force")
                  if bnd.get_field_psdata(self.psdata, k.encode()) >= 0]
        fields = {k: g(k)[:n].reshape(n, -1) for k in names}
        fields["body_number"] = g("body_number")[:n]
        fields["body_dynamic"] = g("body_dynamic")
        fields["plane_constraints_particles"] = g("plane_constraints_particles")[:n]
        st = oracle.OracleState(self.precision, params, **fields)
        return st, names, n

    def _store(self, st, names, n):
        for k in names:
            view = self.get_fields(k)[0]
            view[:n] = np.asarray(st.f[k], dtype=view.dtype).reshape(view[:n].shape)

    def setup(self):
        st, names, n = self._state()
        st.setup_solid()
        self._store(st, names, n)

    def update(self):
        st, names, n = self._state()
        st.step(self.system, 1)
        self._store(st, names, n)

    def set_body_position(self, body, position):
        """set_body_position_device_opencl on the host fields
        (src/opencl/particle_system_host.c:562-575, kernels/rigidbody.cl)."""
        g = lambda name: self.get_fields(name)[0]
        n = int(g("n")[0])
        pos, vel, bpos = oracle.set_body_position(
            g("position")[:n], g("velocity")[:n], g("body_number")[:n], n, int(g("body_count")[0]),
            float(g("timestep")[0]), int(body), g("body_position"), np.asarray(position, np.float32))
        g("position")[:n] = pos
        g("velocity")[:n] = vel
        g("body_position")[...] = bpos.reshape(g("body_position").shape)

    def set_body_rotation(self, body, euler_angles):
        """set_body_rotation_device_opencl on the host fields
        (src/opencl/particle_system_host.c:576-588, kernels/rigidbody.cl)."""
        g = lambda name: self.get_fields(name)[0]
        n = int(g("n")[0])
        pos, rot, beul = oracle.set_body_rotation(
            g("position")[:n], g("rotation")[:n], g("body_number")[:n], n, int(g("body_count")[0]), int(body),
            g("body_euler_rotation"), np.asarray(euler_angles, np.float32))
        g("position")[:n] = pos
        g("rotation")[:n] = rot.reshape(g("rotation")[:n].shape)
        g("body_euler_rotation")[...] = beul.reshape(g("body_euler_rotation").shape)

    def call_for_all_particles(self, kernel):
        if kernel != "apply_plane_constraints":
            raise NotImplementedError(kernel)
        st, names, n = self._state()
        st.plane_constraints()
        self._store(st, names, n)

    def reset(self):
        sph.free_data_cpu(self.psdata)
        self._build()
        self.init(**self.init_values)

    def free(self):
        sph.free_data_cpu(self.psdata)
