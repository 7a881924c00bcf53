"""ctypes wrapper of the CPU oracle (oracle/sph_oracle.c).

Test infrastructure: imported only by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker (or the CPU baseline),
never as part of the engine.

OracleState holds a scene in particle-index order (the psdata order) and runs
the reference's passes on it:

    s = OracleState("f32", params, position=..., velocity=..., ...)
    s.step("fluid", 100)
"""
import ctypes
import os
from ctypes import POINTER, c_float, c_double, c_int, c_uint, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
MAXP, MAXC = 5, 32


def _params_type(real):
    class Params(ctypes.Structure):
        _fields_ = [
            ("n", c_uint), ("rx", c_uint), ("ry", c_uint), ("rz", c_uint),
            ("lo", real * 3), ("hi", real * 3),
            ("h", real), ("mass", real), ("dt", real), ("restdens", real), ("stiffness", real),
            ("viscosity", real), ("youngs", real), ("poisson", real),
            ("gravity", real * 3),
            ("nplanes", c_uint), ("planes", (real * 6) * MAXP),
            ("nconstraints", c_uint), ("constraints", (real * 6) * MAXC),
            ("body_count", c_uint), ("polar_fp64", c_int),
            ("eos_tait", c_int), ("mass_consistent", c_int), ("svk_szz_zero", c_int),
            ("compensated", c_int),
        ]
    return Params


_LIBS = {}


def load(precision):
    if precision not in _LIBS:
        path = os.path.join(LIB_DIR, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        real = c_float if precision == "f32" else c_double
        _LIBS[precision] = (lib, real, _params_type(real))
    return _LIBS[precision]


def grid_resolution(gridbounds, h):
    """(unsigned)(|hi - lo| / h) with the float difference taken in float and
    the division in double (src/build_psdata.c:91-99)."""
    gb = np.asarray(gridbounds, dtype=np.float32)
    hf = np.float32(h)
    return [int(abs(float(np.float32(gb[2 * a + 1] - gb[2 * a]))) / float(hf)) for a in range(3)]


def _ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


class OracleState:
    """A scene in particle-index order plus the oracle passes over it."""

    PER_PARTICLE = {"position": 3, "velocity": 3, "force": 3, "acceleration": 3, "originalpos": 3,
                    "density": 1, "density0": 1, "rotation": 9, "strain": 6, "stress": 6, "deformation": 9,
                    "apq_temp": 9, "force_elastic": 3, "force_viscous": 3, "cutoff_force": 3}

    def __init__(self, precision, params, **fields):
        self.lib, self.real_c, Params = load(precision)
        self.dtype = np.float32 if precision == "f32" else np.float64
        p = Params()
        self.n = int(params["n"])
        res = params.get("gridres") or grid_resolution(params["gridbounds"], params["h"])
        p.n, p.rx, p.ry, p.rz = self.n, res[0], res[1], res[2]
        gb = params["gridbounds"]
        for a in range(3):
            p.lo[a], p.hi[a] = gb[2 * a], gb[2 * a + 1]
        for k in ("h", "mass", "dt", "restdens", "stiffness", "viscosity", "youngs", "poisson"):
            setattr(p, k, float(params.get(k, 0.0)))
        for a in range(3):
            p.gravity[a] = params.get("gravity", (0, 0, 0))[a]
        planes = np.asarray(params.get("symmetry_planes", np.zeros((0, 6))), dtype=np.float64).reshape(-1, 6)
        p.nplanes = len(planes)
        for s, row in enumerate(planes):
            for k in range(6):
                p.planes[s][k] = row[k]
        cons = np.asarray(params.get("plane_constraints", np.zeros((0, 6))), dtype=np.float64).reshape(-1, 6)
        p.nconstraints = len(cons)
        for c, row in enumerate(cons):
            for k in range(6):
                p.constraints[c][k] = row[k]
        p.polar_fp64 = int(params.get("polar_fp64", 1))
        p.eos_tait = int(params.get("eos_tait", 0))
        p.mass_consistent = int(params.get("mass_consistent", 0))
        p.svk_szz_zero = int(params.get("svk_sigma_zz_zero", 0))
        p.compensated = int(params.get("compensated", 0))
        self.params = p
        self.ncells = res[0] * res[1] * res[2]
        n = self.n
        self.f = {}
        for name, w in self.PER_PARTICLE.items():
            arr = fields.get(name)
            self.f[name] = (np.zeros((n, w), self.dtype) if arr is None
                            else np.array(arr, dtype=self.dtype, copy=True).reshape(n, w))
        self.f["density"] = self.f["density"].reshape(n)
        self.f["density0"] = self.f["density0"].reshape(n)
        self.num_neighbours = np.zeros(n, np.uint32)
        self.body_number = (None if fields.get("body_number") is None
                            else np.ascontiguousarray(np.asarray(fields["body_number"], np.uint32)))
        self.body_dynamic = (None if fields.get("body_dynamic") is None
                             else np.ascontiguousarray(np.asarray(fields["body_dynamic"], np.uint32)))
        p.body_count = 0 if self.body_dynamic is None else len(self.body_dynamic)
        self.plane_mask = (None if fields.get("plane_constraints_particles") is None
                           else np.ascontiguousarray(np.asarray(fields["plane_constraints_particles"], np.uint32)))
        self.gridcell = np.zeros(n, np.uint32)
        self.gridcount = np.zeros(self.ncells, np.uint32)
        self.celloffset = np.zeros(self.ncells, np.uint32)
        self.cellparticles = np.zeros(n, np.uint32)

    # -- passes ---------------------------------------------------------------
    def _grid(self):
        return (_ptr(self.gridcell), _ptr(self.gridcount), _ptr(self.celloffset), _ptr(self.cellparticles))

    def bins(self):
        self.lib.orc_bins(ctypes.byref(self.params), _ptr(self.f["position"]), *self._grid())

    def density(self):
        self.lib.orc_density(ctypes.byref(self.params), _ptr(self.f["position"]), *self._grid(),
                             _ptr(self.f["density"]), _ptr(self.num_neighbours))

    def original_density(self):
        self.lib.orc_original_density(ctypes.byref(self.params), _ptr(self.f["position"]), *self._grid(),
                                      _ptr(self.f["density0"]))

    def init_original_position(self):
        self.f["originalpos"][:] = self.f["position"]

    def forces_fluids(self):
        self.lib.orc_forces_fluids(ctypes.byref(self.params), _ptr(self.f["position"]), _ptr(self.f["velocity"]),
                                   _ptr(self.f["density"]), *self._grid(), _ptr(self.f["force"]))

    def step_forward(self):
        self.lib.orc_step_forward(ctypes.byref(self.params), _ptr(self.f["force"]), _ptr(self.f["velocity"]),
                                  _ptr(self.f["position"]), _ptr(self.f["acceleration"]),
                                  _ptr(self.body_number), _ptr(self.body_dynamic))

    def plane_constraints(self):
        if self.plane_mask is None:
            return
        self.lib.orc_plane_constraints(ctypes.byref(self.params), _ptr(self.plane_mask), _ptr(self.f["position"]),
                                       _ptr(self.f["velocity"]))

    def rotations_strains(self):
        self.lib.orc_rotations_strains(ctypes.byref(self.params), _ptr(self.f["position"]),
                                       _ptr(self.f["originalpos"]), _ptr(self.f["density0"]),
                                       _ptr(self.body_number), *self._grid(), _ptr(self.f["rotation"]),
                                       _ptr(self.f["strain"]), _ptr(self.f["apq_temp"]))

    def stresses_corot(self):
        self.lib.orc_stresses_corot(ctypes.byref(self.params), _ptr(self.f["strain"]), _ptr(self.f["stress"]))

    def deformations_svk(self):
        self.lib.orc_deformations_svk(ctypes.byref(self.params), _ptr(self.f["position"]),
                                      _ptr(self.f["originalpos"]), *self._grid(), _ptr(self.f["deformation"]))

    def stresses_svk(self):
        self.lib.orc_stresses_svk(ctypes.byref(self.params), _ptr(self.f["deformation"]), _ptr(self.f["stress"]))

    def forces_solids(self, corotational):
        rot = self.f["rotation"] if corotational else None
        self.lib.orc_forces_solids(ctypes.byref(self.params), _ptr(self.f["position"]), _ptr(self.f["velocity"]),
                                   _ptr(self.f["originalpos"]), _ptr(self.f["stress"]), _ptr(rot),
                                   _ptr(self.f["density0"]), _ptr(self.f["density"]), _ptr(self.body_number),
                                   *self._grid(), _ptr(self.f["force"]), _ptr(self.f["force_elastic"]),
                                   _ptr(self.f["force_viscous"]), _ptr(self.f["cutoff_force"]))

    # -- sequences --------------------------------------------------------------
    def setup_solid(self):
        """StrainSolid.setup (scripts/sph/systems.py:139-146)."""
        self.init_original_position()
        self.bins()
        self.original_density()

    def step(self, system, num_steps=1, apply_constraints=False):
        for _ in range(num_steps):
            self.bins()
            self.density()
            if system == "fluid":
                self.forces_fluids()
            elif system == "corotational":
                self.rotations_strains()
                self.stresses_corot()
                self.forces_solids(True)
            elif system == "svk":
                self.deformations_svk()
                self.stresses_svk()
                self.forces_solids(False)
            else:
                raise ValueError(system)
            self.step_forward()
            if apply_constraints:
                self.plane_constraints()


def _bind_signatures():
    for precision in ("f32", "f64"):
        try:
            lib, real, Params = load(precision)
        except ImportError:
            continue
        lib.orc_polar.argtypes = [c_void_p, c_void_p, c_int]
        lib.orc_jacobi_inverse_apply.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
        lib.orc_real_size.restype = c_uint


_bind_signatures()


def polar(apq, precision="f64", fp64=True):
    """R of the oracle's polar decomposition of one 3x3 (row-major)."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(np.asarray(apq, dt).reshape(9))
    r = np.zeros(9, dt)
    lib.orc_polar(_ptr(a), _ptr(r), int(fp64))
    return r.reshape(3, 3)


def apq_times_inverse(apq, aqq_sym, precision="f64", fp64=True):
    """Apq * inverse(Aqq) with Aqq in [00, 11, 22, 01, 02, 12] storage."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(np.asarray(apq, dt).reshape(9))
    q = np.ascontiguousarray(np.asarray(aqq_sym, dt).reshape(6))
    out = np.zeros(9, dt)
    lib.orc_jacobi_inverse_apply(_ptr(a), _ptr(q), _ptr(out), int(fp64))
    return out.reshape(3, 3)


# -- scripted rigid bodies and initial-condition helpers ----------------------
# (kernels/common.cl:533-604, kernels/rigidbody.cl:1-120).  Arrays are in
# particle-index order; every function returns new arrays.
def _arr(a, dt, shape=None):
    out = np.array(a, dtype=dt, copy=True)
    return np.ascontiguousarray(out if shape is None else out.reshape(shape))


def _real_args(lib, real, name, spec):
    fn = getattr(lib, name)
    fn.argtypes = [real if t == "r" else (c_uint if t == "u" else c_void_p) for t in spec]
    return fn


def populate_position_cuboid(position, pnum, n, corner1, corner2, size, precision="f32"):
    """(position, n) after populate_position_cuboid; position has pnum rows."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    pos = _arr(position, dt, (int(pnum), 3))
    nn = np.array([n], np.uint32)
    c1, c2 = _arr(corner1, dt, 3), _arr(corner2, dt, 3)
    fn = _real_args(lib, real, "orc_populate_cuboid", "pupppuuu")
    fn(_ptr(pos), int(pnum), _ptr(nn), _ptr(c1), _ptr(c2), *(int(v) for v in size))
    return pos, int(nn[0])


def rotate_particles(position, n, angle_x, angle_y, angle_z, precision="f32"):
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    pos = _arr(position, dt)
    pos = pos.reshape(-1, 3)
    fn = _real_args(lib, real, "orc_rotate_particles", "purrr")
    fn(_ptr(pos), int(n), float(angle_x), float(angle_y), float(angle_z))
    return pos


def euler_matrix(angles, precision="f64"):
    """Body matrix of Euler angles (x, y, x2) as get_matrix_from_euler forms it."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    e = _arr(angles, dt, 3)
    out = np.zeros(9, dt)
    lib.orc_euler_matrix.argtypes = [c_void_p, c_void_p]
    lib.orc_euler_matrix(_ptr(e), _ptr(out))
    return out.reshape(3, 3)


def set_body_position(position, velocity, body_number, n, body_count, dt_step, body, body_position, target,
                      precision="f32"):
    """(position, velocity, body_position) after set_body_position."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    pos, vel = _arr(position, dt, (-1, 3)), _arr(velocity, dt, (-1, 3))
    bn = _arr(body_number, np.uint32)
    bp = _arr(body_position, dt, (-1, 3))
    t = _arr(target, dt, 3)
    fn = _real_args(lib, real, "orc_set_body_position", "pppuurupp")
    fn(_ptr(pos), _ptr(vel), _ptr(bn), int(n), int(body_count), float(dt_step), int(body), _ptr(bp), _ptr(t))
    return pos, vel, bp


def set_body_rotation(position, rotation, body_number, n, body_count, body, body_euler, angles, precision="f32"):
    """(position, rotation, body_euler_rotation) after set_body_rotation."""
    lib, real, _ = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    pos, rot = _arr(position, dt, (-1, 3)), _arr(rotation, dt, (-1, 9))
    bn = _arr(body_number, np.uint32)
    be = _arr(body_euler, dt, (-1, 3))
    a = _arr(angles, dt, 3)
    fn = _real_args(lib, real, "orc_set_body_rotation", "pppuuupp")
    fn(_ptr(pos), _ptr(rot), _ptr(bn), int(n), int(body_count), int(body), _ptr(be), _ptr(a))
    return pos, rot, be
