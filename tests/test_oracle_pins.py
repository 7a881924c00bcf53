"""Pins of the CPU oracle against mathematics it must satisfy, each chosen so
that a plausible slip (dropped term, wrong sign or index, transposed
operand, wrong power of h) fails:

  cell keys        hand-computed cells for centre, edge, out-of-range and NaN positions
  bins             numpy stable counting sort (bincount / cumsum / stable argsort)
  neighbour sets   O(N^2) brute force with the same float32 distance
  poly6            normalisation: dx^3 sum W over a fine lattice = 1
  density          closed-form lattice sum; mirror planes = explicitly doubled domain
  fluid forces     independent textbook SPH accelerations (Mueller et al. 2003) in numpy;
                   momentum conservation (pairwise antisymmetry)
  Jacobi / polar   R orthonormal, det R = +1, R = U V^T of numpy's SVD
  SVK              F = Apq Aqq^-1 reproduces an affine map exactly; Green strain /
                   St. Venant-Kirchhoff / Cauchy closed forms
  corotational     rest state has zero strain; rigid rotation leaves strain unchanged;
                   Hooke closed form
  integrator       symplectic Euler closed form, static bodies frozen
  constraints      projection onto the plane, normal velocity removed
"""
import math

import numpy as np
import pytest

import oracle
from oracle import OracleState

H = 0.04
PI = math.pi


def lattice_scene(nx, ny, nz, dx, origin, jitter=0.0, seed=0):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.asarray(origin) + g * dx + rng.uniform(-jitter, jitter, g.shape), g


def base_params(n, gb, h=H, **kw):
    p = dict(n=n, gridbounds=list(gb), h=h, mass=1e-3, dt=1e-4, restdens=1000.0, stiffness=2000.0,
             viscosity=0.1, gravity=(0.0, 0.0, -9.81))
    p.update(kw)
    return p


# ---------------------------------------------------------------- grid ---
def test_cell_keys_closed_form():
    gb = [0.0, 1.0, 0.0, 2.0, -1.0, 1.0]
    h = 0.25                                  # res = 4, 8, 8
    pos = np.array([
        [0.125, 0.125, -0.875],               # cell centres
        [0.875, 1.875, 0.875],
        [0.5, 1.0, 0.0],                      # exactly on edges: the upper cell
        [-3.0, 5.0, 0.99],                    # out of range: clamped
        [1.0, 2.0, 1.0],                      # the upper bound maps to the last cell
        [np.nan, 0.3, 0.1],                   # NaN -> 0
    ], np.float32)
    st = OracleState("f32", base_params(len(pos), gb, h=h), position=pos)
    st.bins()
    rx, ry = 4, 8
    want = [(0, 0, 0), (3, 7, 7), (2, 4, 4), (0, 7, 7), (3, 7, 7), (0, 1, 4)]
    assert list(st.gridcell) == [z * rx * ry + y * rx + x for x, y, z in want]


def test_bins_are_numpy_stable_counting_sort():
    rng = np.random.Generator(np.random.PCG64(1))
    pos = rng.uniform(-0.1, 1.1, size=(5000, 3)).astype(np.float32)
    st = OracleState("f32", base_params(len(pos), [0, 1, 0, 1, 0, 1], h=0.1), position=pos)
    st.bins()
    ncell = st.ncells
    assert np.array_equal(st.gridcount, np.bincount(st.gridcell, minlength=ncell))
    assert np.array_equal(st.celloffset, np.concatenate([[0], np.cumsum(st.gridcount)[:-1]]))
    assert np.array_equal(st.cellparticles, np.argsort(st.gridcell, kind="stable"))


def test_neighbour_counts_brute_force():
    rng = np.random.Generator(np.random.PCG64(2))
    pos = rng.uniform(0, 1, size=(1500, 3)).astype(np.float32)
    h = 0.11
    st = OracleState("f32", base_params(len(pos), [0, 1, 0, 1, 0, 1], h=h), position=pos)
    st.bins()
    st.density()
    d = pos[None, :, :] - pos[:, None, :]
    r = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    assert np.array_equal(st.num_neighbours, (r < np.float32(h)).sum(1))


# ------------------------------------------------------------- density ---
def poly6(r, h):
    return np.where(r <= h, 315.0 / (64.0 * PI * h ** 9) * (h * h - r * r) ** 3, 0.0)


def test_poly6_normalisation():
    dx = H / 16
    k = int(H / dx) + 1
    g = np.arange(-k, k + 1) * dx
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    assert abs(dx ** 3 * poly6(r, H).sum() - 1.0) < 2e-3
    # the oracle's kernel at a single pair equals the closed form
    pos = np.array([[0.5, 0.5, 0.5], [0.5 + 0.013, 0.5 - 0.007, 0.5 + 0.002]])
    st = OracleState("f64", base_params(2, [0, 1, 0, 1, 0, 1], mass=1.0), position=pos)
    st.bins()
    st.density()
    r01 = np.linalg.norm(pos[1] - pos[0])
    assert st.f["density"][0] == pytest.approx(float(poly6(0.0, H) + poly6(r01, H)), rel=1e-14)


def test_lattice_density_closed_form():
    dx = H / 2
    pos, g = lattice_scene(9, 9, 9, dx, (0.2, 0.2, 0.2))
    m = 1000 * dx ** 3
    st = OracleState("f64", base_params(len(pos), [0, 1, 0, 1, 0, 1], mass=m), position=pos)
    st.bins()
    st.density()
    k = np.arange(-2, 3)
    v = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) * dx
    want = m * poly6(np.linalg.norm(v, axis=1), H).sum()
    centre = np.where((g == 4).all(1))[0][0]
    assert st.f["density"][centre] == pytest.approx(want, rel=1e-12)


def test_mirror_planes_equal_doubled_domain():
    """A plane at x = 0: particles near it see their mirror images, which is
    the density of the domain reflected explicitly (images farther than 1e-5
    from the plane)."""
    rng = np.random.Generator(np.random.PCG64(3))
    pos = rng.uniform([0.001, 0.3, 0.3], [0.2, 0.6, 0.6], size=(600, 3))
    gb = [-0.3, 0.3, 0.0, 1.0, 0.0, 1.0]
    planes = np.array([[0, 0, 0, 1, 0, 0], [0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0]], float)
    st = OracleState("f64", base_params(len(pos), gb, mass=1.0, symmetry_planes=planes), position=pos)
    st.bins()
    st.density()
    both = np.concatenate([pos, pos * [-1, 1, 1]])
    st2 = OracleState("f64", base_params(len(both), gb, mass=1.0), position=both)
    st2.bins()
    st2.density()
    near = pos[:, 0] <= H
    # densities agree wherever the reflected set is the same set the rule admits
    assert np.allclose(st.f["density"][near], st2.f["density"][:len(pos)][near], rtol=1e-12)
    assert np.allclose(st.f["density"][~near], st2.f["density"][:len(pos)][~near], rtol=1e-12)


# -------------------------------------------------------- fluid forces ---
def textbook_accelerations(pos, vel, rho, m, h, k, rho0, mu):
    """Mueller et al. 2003: a_i = -sum_j m (p_i + p_j)/(2 rho_i rho_j) grad W_spiky
    + mu sum_j m (v_j - v_i)/(rho_i rho_j) lap W_visc, linear EOS."""
    p = k * (rho - rho0)
    d = pos[:, None, :] - pos[None, :, :]          # x_i - x_j
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    inside = r <= h
    rr = np.where(inside, r, 1.0)
    grad = np.where(inside[..., None], -45.0 / (PI * h ** 6) * ((h - rr) ** 2 / rr)[..., None] * d, 0.0)
    lap = np.where(inside, 45.0 / (PI * h ** 6) * (h - rr), 0.0)
    didj = rho[:, None] * rho[None, :]
    a_p = -(m * (p[:, None] + p[None, :]) / (2 * didj))[..., None] * grad
    a_v = mu * (m * lap / didj)[..., None] * (vel[None, :, :] - vel[:, None, :])
    return (a_p + a_v).sum(1)


def test_fluid_force_is_textbook_acceleration():
    rng = np.random.Generator(np.random.PCG64(4))
    pos = rng.uniform(0.3, 0.45, size=(300, 3))
    vel = rng.normal(scale=0.1, size=(300, 3))
    prm = base_params(300, [0, 1, 0, 1, 0, 1], mass=8e-3)
    st = OracleState("f64", prm, position=pos, velocity=vel)
    st.bins()
    st.density()
    st.forces_fluids()
    want = textbook_accelerations(pos, vel, st.f["density"], 8e-3, H, 2000.0, 1000.0, 0.1)
    assert np.allclose(st.f["force"], want, rtol=1e-10, atol=1e-9 * np.abs(want).max())
    # mass_consistent multiplies by the particle's own mass
    st2 = OracleState("f64", dict(prm, mass_consistent=1), position=pos, velocity=vel)
    st2.bins()
    st2.density()
    st2.forces_fluids()
    assert np.allclose(st2.f["force"], 8e-3 * want, rtol=1e-10, atol=1e-9 * 8e-3 * np.abs(want).max())


def test_tait_pressure_and_momentum_conservation():
    rng = np.random.Generator(np.random.PCG64(5))
    pos = rng.uniform(0.3, 0.5, size=(800, 3))
    vel = rng.normal(scale=0.2, size=(800, 3))
    st = OracleState("f64", base_params(800, [0, 1, 0, 1, 0, 1], mass=8e-3, eos_tait=1), position=pos,
                     velocity=vel)
    st.bins()
    st.density()
    st.forces_fluids()
    f = st.f["force"]
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).sum()
    # Tait: 1 % compression gives k ((1.01)^7 - 1)
    one = OracleState("f64", base_params(2, [0, 1, 0, 1, 0, 1], mass=1.0, eos_tait=1),
                      position=np.array([[0.5, 0.5, 0.5], [0.52, 0.5, 0.5]]))
    one.bins()
    one.f["density"][:] = [1010.0, 1000.0]
    one.forces_fluids()
    r = 0.02
    # grad_0 W_spiky(x_0 - x_1) points along +x (x_0 - x_1 = -r x-hat)
    grad0 = 45.0 / (PI * H ** 6) * (H - r) ** 2
    p0 = 2000.0 * (1.01 ** 7 - 1)
    want = -(p0 + 0.0) / (2 * 1010.0 * 1000.0) * grad0     # pushed away from the rarer particle 1
    assert want < 0
    assert one.f["force"][0, 0] == pytest.approx(want, rel=1e-12)


# ------------------------------------------------------- Jacobi / polar ---
def random_rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


@pytest.mark.parametrize("seed", range(6))
def test_polar_is_svd_rotation(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    R0 = random_rotation(rng)
    S = np.eye(3) + 0.05 * (lambda a: a + a.T)(rng.normal(size=(3, 3)))
    A = R0 @ S
    R = oracle.polar(A, "f64", True)
    u, _, vt = np.linalg.svd(A)
    assert np.allclose(R, u @ vt, atol=1e-9)
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-9)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("seed", range(6))
def test_jacobi_inverse(seed):
    rng = np.random.Generator(np.random.PCG64(seed + 10))
    M = rng.normal(size=(3, 3))
    Aqq = M @ M.T + 0.5 * np.eye(3)
    sym = [Aqq[0, 0], Aqq[1, 1], Aqq[2, 2], Aqq[0, 1], Aqq[0, 2], Aqq[1, 2]]
    Apq = rng.normal(size=(3, 3))
    F = oracle.apq_times_inverse(Apq, sym, "f64", True)
    assert np.allclose(F, Apq @ np.linalg.inv(Aqq), rtol=1e-6, atol=1e-8)


# -------------------------------------------------------------- solids ---
def solid_state(kind, deform=None, rotate=None, seed=0, precision="f64"):
    dx, h = 0.01, 0.02
    pos, g = lattice_scene(9, 9, 9, dx, (0.0, 0.0, 0.0))
    gb = [-0.1, 0.2, -0.1, 0.2, -0.1, 0.2]
    prm = dict(n=len(pos), gridbounds=gb, h=h, mass=1e-3, dt=5e-5, viscosity=0.0, youngs=1e5, poisson=0.3,
               gravity=(0.0, 0.0, 0.0), polar_fp64=1)
    st = OracleState(precision, prm, position=pos, rotation=np.tile(np.eye(3), (len(pos), 1, 1)))
    st.setup_solid()
    centre = pos.mean(0)
    x = pos - centre
    if deform is not None:
        x = x @ np.asarray(deform).T
    if rotate is not None:
        x = x @ np.asarray(rotate).T
    st.f["position"][:] = x + centre
    interior = ((g >= 3) & (g <= 5)).all(1)
    return st, interior


def test_svk_deformation_reproduces_affine_map():
    F0 = np.array([[1.02, 0.01, 0.0], [-0.005, 0.99, 0.003], [0.0, 0.004, 1.01]])
    st, interior = solid_state("svk", deform=F0)
    st.bins()
    st.deformations_svk()
    F = st.f["deformation"].reshape(-1, 3, 3)
    assert np.allclose(F[interior], F0, atol=1e-9)


def test_svk_stress_closed_form():
    rng = np.random.Generator(np.random.PCG64(7))
    n = 50
    F = np.eye(3)[None] + 0.03 * rng.normal(size=(n, 3, 3))
    E, nu = 5e5, 0.35
    st = OracleState("f64", dict(n=n, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, youngs=E, poisson=nu),
                     deformation=F.reshape(n, 9))
    st.stresses_svk()
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    G = 0.5 * (np.einsum("nki,nkj->nij", F, F) - np.eye(3))
    S = lam * np.trace(G, axis1=1, axis2=2)[:, None, None] * np.eye(3) + 2 * mu * G
    sig = np.einsum("nij,njk,nlk->nil", F, S, F) / np.linalg.det(F)[:, None, None]
    want = np.stack([sig[:, 0, 0], sig[:, 1, 1], sig[:, 2, 2], sig[:, 0, 1], sig[:, 0, 2], sig[:, 1, 2]], 1)
    assert np.allclose(st.f["stress"], want, rtol=1e-10, atol=1e-8)


def test_corotational_rest_state_and_rotation_invariance():
    st, interior = solid_state("corotational")
    st.bins()
    st.rotations_strains()
    rest = st.f["strain"].copy()
    assert np.abs(rest[interior]).max() < 1e-12
    R = st.f["rotation"].reshape(-1, 3, 3)
    assert np.allclose(R[interior], np.eye(3), atol=1e-12)
    eps = np.array([[1.01, 0.002, 0], [0.002, 0.995, 0], [0, 0, 1.0]])
    rot = random_rotation(np.random.Generator(np.random.PCG64(9)))
    a, _ = solid_state("corotational", deform=eps)
    a.bins()
    a.rotations_strains()
    b, _ = solid_state("corotational", deform=eps, rotate=rot)
    b.bins()
    b.rotations_strains()
    assert np.allclose(a.f["strain"][interior], b.f["strain"][interior], atol=1e-7)
    # A_pq of the rotated body is rot A_pq, so its polar rotation is rot R
    Ra = a.f["rotation"].reshape(-1, 3, 3)[interior]
    Rb = b.f["rotation"].reshape(-1, 3, 3)[interior]
    assert np.allclose(Rb, rot @ Ra, atol=1e-9)
    assert np.allclose(Ra, np.eye(3), atol=1e-4)


def test_hooke_closed_form():
    rng = np.random.Generator(np.random.PCG64(8))
    eps = rng.normal(scale=1e-3, size=(20, 6))
    E, nu = 1e5, 0.3
    st = OracleState("f64", dict(n=20, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, youngs=E, poisson=nu), strain=eps)
    st.stresses_corot()
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    tr = eps[:, :3].sum(1)
    want = np.concatenate([lam * tr[:, None] + 2 * mu * eps[:, :3], 2 * mu * eps[:, 3:]], 1)
    assert np.allclose(st.f["stress"], want, rtol=1e-12)


# ------------------------------------------------- integrator, planes ---
def test_step_forward_closed_form_and_static_bodies():
    rng = np.random.Generator(np.random.PCG64(10))
    n = 40
    pos, vel, f = (rng.normal(size=(n, 3)) for _ in range(3))
    body = np.zeros(n, np.uint32)
    body[:5] = 1
    body[5:10] = 2
    prm = dict(n=n, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, mass=0.5, dt=0.01, gravity=(0.0, -1.0, -9.81))
    st = OracleState("f64", prm, position=pos, velocity=vel, force=f, body_number=body,
                     body_dynamic=np.array([0, 1], np.uint32))
    st.step_forward()
    a = f / 0.5 + np.array([0.0, -1.0, -9.81])
    v = vel + 0.01 * a
    x = pos + 0.01 * v
    assert np.allclose(st.f["position"][5:], x[5:], rtol=1e-15)
    assert np.allclose(st.f["velocity"][5:], v[5:], rtol=1e-15)
    assert np.array_equal(st.f["position"][:5], pos[:5])   # static body 1


def test_plane_constraints_project():
    rng = np.random.Generator(np.random.PCG64(11))
    n = 30
    pos, vel = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    nrm = np.array([1.0, 2.0, -2.0]) / 3.0
    pt = np.array([0.1, -0.2, 0.3])
    cons = np.zeros((4, 6))
    cons[2] = np.concatenate([pt, nrm])
    mask = np.zeros(n, np.uint32)
    mask[::2] = 1 << 2
    st = OracleState("f64", dict(n=n, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, plane_constraints=cons),
                     position=pos, velocity=vel, plane_constraints_particles=mask)
    st.plane_constraints()
    x, v = st.f["position"], st.f["velocity"]
    assert np.allclose((x[::2] - pt) @ nrm, 0, atol=1e-15)
    assert np.allclose(v[::2] @ nrm, 0, atol=1e-15)
    t = vel[::2] - np.outer(vel[::2] @ nrm, nrm)
    assert np.allclose(v[::2], t, atol=1e-15)
    assert np.array_equal(x[1::2], pos[1::2])
