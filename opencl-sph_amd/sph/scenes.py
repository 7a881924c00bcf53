"""Synthetic scenes: psdata specifications plus seeded initial conditions.

A Scene carries the $psdata_specification text, the kernel file list, the
initial field values (psdata order) and the scalar parameters, so that the
same scene can be loaded into the engine (`build`) and into any other
consumer of plain arrays.  The field lists follow the reference's confs
(conf/fluid.conf, conf/strainsolid.conf, conf/svksolid.conf) so harness code
written for those layouts finds the same fields.

The workloads are:

    dam_break(...)       weakly compressible fluid block
    cantilever(...)      corotational beam with its x = 0 face clamped
    svk_cube(...)        SVK cube with a smooth random initial velocity
    svk_block(...)       the same for a rectangular block
    request_scene(...)   a block of any of the three for the served stream
    baseline_config(k)   the single-scene benchmark workloads (k = 0..3)

Two optional conf fields extend the fluid model: `eos_tait` (Tait equation
of state p = k((rho/rho0)^7 - 1)) and `mass_consistent` (multiply the fluid
force by the particle mass; the reference's force sums only the neighbour
mass, kernels/fluids.cl:43-56 with kernels/common.cl:708).  Both default to
the reference's behaviour when absent.
"""
import dataclasses
from typing import Dict, List, Optional

import numpy as np

FLUID_KERNELS = "common, fluids"
COROTATIONAL_KERNELS = "common, solids, solid_corotation, constraints"
SVK_KERNELS = "common, solids, solid_svk, constraints"


@dataclasses.dataclass
class Scene:
    name: str
    system: str                     # "fluid", "corotational" or "svk"
    spec: str                       # $psdata_specification text
    kernel_files: str
    fields: Dict[str, np.ndarray]   # initial values in psdata order
    params: Dict[str, object]       # scalars for array-level consumers
    steps: int = 1

    @property
    def n(self):
        return int(self.params["n"])


def _fmt(v):
    if isinstance(v, (list, tuple, np.ndarray)):
        return ", ".join(_fmt(x) for x in v)
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(np.float32(v)))


def spec_text(lines):
    """psdata specification from (name, dims, type, value-or-None) tuples."""
    out = []
    for name, dims, typ, value in lines:
        row = f"{name:<28}| {dims:<14}| {typ}"
        if value is not None:
            row += f" | {_fmt(value)}"
        out.append(row)
    return "\n".join(out) + "\n"


def fluid_spec(pnum, mass, dt, h, gridbounds, restdens, stiffness, viscosity, gravity=(0, 0, -9.81),
               eos_tait=False, mass_consistent=False, body_count=0):
    lines = [
        ("pnum", "1", "unsigned int", pnum),
        ("n", "1", "unsigned int", None),
        ("mass", "1", "float", mass),
        ("timestep", "1", "float", dt),
        ("smoothingradius", "1", "float", h),
        ("gravity", "3", "float", list(gravity)),
        ("position", "3, pnum", "float", None),
        ("posnext", "3, pnum", "float", None),
        ("velocity", "3, pnum", "float", None),
        ("veleval", "3, pnum", "float", None),
        ("acceleration", "3, pnum", "float", None),
        ("force", "3, pnum", "float", None),
        ("density", "pnum", "float", None),
        ("volume", "pnum", "float", None),
        ("num_neighbours", "pnum", "unsigned int", None),
        ("gridbounds", "2, 3", "float", list(gridbounds)),
        ("restdens", "1", "float", restdens),
        ("stiffness", "1", "float", stiffness),
        ("viscosity", "1", "float", viscosity),
        ("eos_tait", "1", "unsigned int", int(bool(eos_tait))),
        ("mass_consistent", "1", "unsigned int", int(bool(mass_consistent))),
    ]
    if body_count:
        # the body fields of the solid confs: step_forward freezes particles
        # of static bodies in every system (kernels/common.cl:699-704)
        lines += [("body_number", "pnum", "unsigned int", None),
                  ("body_count", "1", "unsigned int", body_count),
                  ("body_dynamic", "body_count", "unsigned int", None)]
    return spec_text(lines)


def solid_spec(pnum, mass, dt, h, gridbounds, viscosity, youngs, poisson, gravity=(0, 0, 0), svk=False,
               body_count=2):
    lines = [
        ("pnum", "1", "unsigned int", pnum),
        ("n", "1", "unsigned int", None),
        ("mass", "1", "float", mass),
        ("timestep", "1", "float", dt),
        ("smoothingradius", "1", "float", h),
        ("gravity", "3", "float", list(gravity)),
        ("position", "3, pnum", "float", None),
        ("originalpos", "3, pnum", "float", None),
        ("posnext", "3, pnum", "float", None),
        ("velocity", "3, pnum", "float", None),
        ("veleval", "3, pnum", "float", None),
        ("acceleration", "3, pnum", "float", None),
        ("force", "3, pnum", "float", None),
        ("stress", "6, pnum", "float", None),
    ]
    if svk:
        lines += [("deformation", "3, 3, pnum", "float", None)]
    else:
        lines += [("rotation", "3, 3, pnum", "float", None),
                  ("apq_temp", "3, 3, pnum", "float", None),
                  ("intermediate", "3, 3, pnum", "float", None),
                  ("strain", "6, pnum", "float", None)]
    lines += [
        ("density", "pnum", "float", None),
        ("density0", "pnum", "float", None),
        ("gridbounds", "2, 3", "float", list(gridbounds)),
        ("viscosity", "1", "float", viscosity),
        ("youngs_modulus", "1", "float", youngs),
        ("poisson", "1", "float", poisson),
        ("plane_constraints", "3, 2, 4", "float", None),
        ("plane_constraints_particles", "pnum", "unsigned int", None),
        ("symmetry_planes", "3, 2, 3", "float", None),
        ("cutoff_force", "3, pnum", "float", None),
        ("force_elastic", "3, pnum", "float", None),
        ("force_viscous", "3, pnum", "float", None),
        ("padding", "pnum", "unsigned int", None),
        ("num_neighbours", "pnum", "unsigned int", None),
        ("rigid_mask", "pnum", "unsigned int", None),
        ("body_number", "pnum", "unsigned int", None),
        ("body_count", "1", "unsigned int", body_count),
        ("body_position", "3, body_count", "float", None),
        ("body_euler_rotation", "3, body_count", "float", None),
        ("body_dynamic", "body_count", "unsigned int", None),
    ]
    return spec_text(lines)


def lattice(nx, ny, nz, dx, origin=(0.0, 0.0, 0.0)):
    """(nx*ny*nz, 3) lattice points, x slowest, z fastest."""
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return np.asarray(origin, np.float64) + g * dx, g


def dam_break(nx, ny, nz, dx, h, jitter, seed=0, gridbounds=(0, 1, 0, 1, 0, 1), restdens=1000.0,
              stiffness=2000.0, viscosity=0.1, dt=1e-4, steps=100, eos_tait=True, mass_consistent=True,
              origin=None, static_floor=False, name="dam_break"):
    """static_floor: the block's lowest lattice layer is a static body
    (body_number 1, body_dynamic [0])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if origin is None:
        origin = (gridbounds[0] + h, gridbounds[2] + h, gridbounds[4] + h)
    pts, g = lattice(nx, ny, nz, dx, origin)
    pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    pos = pts.astype(np.float32)
    n = len(pos)
    mass = restdens * dx ** 3
    spec = fluid_spec(n, mass, dt, h, gridbounds, restdens, stiffness, viscosity, eos_tait=eos_tait,
                      mass_consistent=mass_consistent, body_count=1 if static_floor else 0)
    fields = {"position": pos, "n": np.array([n], np.uint32)}
    if static_floor:
        fields["body_number"] = (g[:, 2] == 0).astype(np.uint32)
        fields["body_dynamic"] = np.array([0], np.uint32)
    params = dict(n=n, gridbounds=list(gridbounds), h=h, mass=float(np.float32(mass)), dt=dt, restdens=restdens,
                  stiffness=stiffness, viscosity=viscosity, gravity=(0.0, 0.0, -9.81), eos_tait=int(eos_tait),
                  mass_consistent=int(mass_consistent))
    return Scene(name, "fluid", spec, FLUID_KERNELS, fields, params, steps)


def cantilever(nx=64, ny=16, nz=16, dx=0.01, h=0.02, youngs=1e5, poisson=0.3, rho=1000.0, dt=5e-5,
               viscosity=6e-5, steps=500, name="cantilever"):
    pts, g = lattice(nx, ny, nz, dx)
    pos = pts.astype(np.float32)
    n = len(pos)
    lo, hi = pos.min(0) - 4 * h, pos.max(0) + 4 * h
    gb = [float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]), float(lo[2]), float(hi[2])]
    mass = rho * dx ** 3
    spec = solid_spec(n, mass, dt, h, gb, viscosity, youngs, poisson, gravity=(0, 0, -9.81))
    body = (g[:, 0] == 0).astype(np.uint32)
    rot = np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (n, 1)).reshape(n, 3, 3)
    fields = {"position": pos, "n": np.array([n], np.uint32), "body_number": body,
              "body_dynamic": np.array([0, 1], np.uint32), "rotation": rot}
    params = dict(n=n, gridbounds=gb, h=h, mass=float(np.float32(mass)), dt=dt, viscosity=viscosity,
                  youngs=youngs, poisson=poisson, gravity=(0.0, 0.0, -9.81))
    return Scene(name, "corotational", spec, COROTATIONAL_KERNELS, fields, params, steps)


def smooth_velocity(pos, amplitude, seed, box):
    """Sum of 8 sinusoids with integer wave vectors in {-2..2}^3 \\ 0, unit
    directions and uniform phases, rescaled to max |v| = amplitude."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = np.asarray(box, np.float64)
    v = np.zeros(pos.shape, np.float64)
    for _ in range(8):
        kappa = np.zeros(3)
        while not kappa.any():
            kappa = rng.integers(-2, 3, size=3).astype(np.float64)
        e = rng.normal(size=3)
        e /= np.linalg.norm(e)
        phi = rng.uniform(0, 2 * np.pi)
        a = rng.uniform(0.5, 1.0)
        v += a * e[None, :] * np.sin(2 * np.pi * (pos @ (kappa / L)) + phi)[:, None]
    v *= amplitude / np.max(np.linalg.norm(v, axis=1))
    return v


def svk_cube(nside=64, dx=0.01, h=0.02, youngs=5e5, poisson=0.35, rho=1100.0, dt=2e-5, viscosity=4e-4,
             amplitude=0.05, seed=0, steps=200, name="svk_cube"):
    return svk_block(nside, nside, nside, dx, h, youngs, poisson, rho, dt, viscosity, amplitude, seed, steps, name)


def svk_block(nx, ny, nz, dx=0.01, h=0.02, youngs=5e5, poisson=0.35, rho=1100.0, dt=2e-5, viscosity=4e-4,
              amplitude=0.05, seed=0, steps=200, name="svk_block"):
    pts, _ = lattice(nx, ny, nz, dx)
    pos = pts.astype(np.float32)
    n = len(pos)
    lo, hi = pos.min(0) - 4 * h, pos.max(0) + 4 * h
    gb = [float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]), float(lo[2]), float(hi[2])]
    mass = rho * dx ** 3
    vel = smooth_velocity(pts, amplitude, seed, box=[nx * dx, ny * dx, nz * dx]).astype(np.float32)
    spec = solid_spec(n, mass, dt, h, gb, viscosity, youngs, poisson, gravity=(0, 0, 0), svk=True)
    fields = {"position": pos, "velocity": vel, "n": np.array([n], np.uint32)}
    params = dict(n=n, gridbounds=gb, h=h, mass=float(np.float32(mass)), dt=dt, viscosity=viscosity,
                  youngs=youngs, poisson=poisson, gravity=(0.0, 0.0, 0.0))
    return Scene(name, "svk", spec, SVK_KERNELS, fields, params, steps)


def baseline_config(k, seed=0):
    """The benchmark workloads, by index."""
    if k == 0:
        return dam_break(32, 32, 16, 0.02, 0.04, 1e-4, seed=seed, gridbounds=(0, 1, 0, 1, 0, 1), steps=100,
                         name="dam_break_16k")
    if k == 1:
        return cantilever(name="cantilever_16k")
    if k == 2:
        return svk_cube(seed=seed, name="svk_cube_262k")
    if k == 3:
        return dam_break(256, 256, 64, 0.005, 0.01, 2.5e-5, seed=seed, gridbounds=(0, 2, 0, 2, 0, 1), steps=50,
                         name="dam_break_4m")
    raise ValueError(f"no workload {k}")


def block_dims(n, rng):
    """Rectangular block of about n particles with random aspect: log-aspect
    ratios uniform in [-ln 2, ln 2] per axis, the last axis adjusted so the
    product lands as close to n as integers allow."""
    a = np.exp(rng.uniform(-np.log(2.0), np.log(2.0), size=3))
    a /= np.prod(a) ** (1.0 / 3.0)
    side = n ** (1.0 / 3.0)
    nx = max(2, int(round(side * a[0])))
    ny = max(2, int(round(side * a[1])))
    nz = max(2, int(round(n / (nx * ny))))
    return nx, ny, nz


def request_scene(kind, nx, ny, nz, seed, steps=200, name=None):
    """A served-stream scene: a block of the given lattice size with the
    parameters of the 16k dam break (fluid), the cantilever (corotational,
    x = 0 face clamped) or the SVK cube (svk)."""
    name = name or f"{kind}_{nx}x{ny}x{nz}"
    if kind == "fluid":
        h, dx = 0.04, 0.02
        top = [h + (m - 1) * dx + 2 * h for m in (nx, ny, nz)]
        return dam_break(nx, ny, nz, dx, h, 1e-4, seed=seed, gridbounds=(0, top[0], 0, top[1], 0, top[2]),
                         steps=steps, name=name)
    if kind == "corotational":
        return cantilever(nx, ny, nz, steps=steps, name=name)
    if kind == "svk":
        return svk_block(nx, ny, nz, seed=seed, steps=steps, name=name)
    raise ValueError(f"unknown system {kind}")


def build(scene, setup=True):
    """Load a scene into the engine: returns a ready ParticleSystem."""
    from .systems import SceneSystem
    return SceneSystem(scene, setup=setup)
