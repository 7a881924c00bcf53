"""Pins of the oracle's corotational passes (oracle/sph_oracle.c
orc_rotations_strains, orc_forces_solids), each chosen so that a plausible
slip fails one of them:

  restatement   an independent FP64 numpy restatement over all pairs
                (tests/corotation_numpy.py, written from
                kernels/solid_corotation.cl:28-195 and :215-403) agrees with the
                FP64 oracle on a jittered, strained, rotated body with static
                body particles: A_pq, R, strain, force, force_elastic,
                force_viscous
  affine map    x = Q (I + eps) X: interior strain ~ eps and R = Q (catches a
                dropped m / rho0, R used for R^T, a dropped 1/2 in the strain)
  linear stress sigma(x0) = sigma0 + linear field: interior elastic force ~
                1/2 V div(sigma) (catches the 1/2 factor, the m^2/(rho0_i
                rho0_j) factor, a sign, sigma_i used for sigma_j)
  equivariance  rotating a strained body rigidly by Q turns force_elastic and
                rotation by Q (catches R vs R^T in the force pass)
  momentum      with no bodies and no planes, sum of force_elastic = 0 for any
                stresses and rotations (catches a dropped f_ij term)
  rest          zero strain and zero velocity give exactly zero force
"""
import numpy as np
import pytest

from oracle import OracleState

import corotation_numpy as ref


def lattice(nx, ny, nz, dx, origin=(0.0, 0.0, 0.0)):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return np.asarray(origin) + g * dx, g


def params(n, gb, h, mass, viscosity=0.0):
    return dict(n=n, gridbounds=list(gb), h=h, mass=mass, dt=5e-5, viscosity=viscosity, youngs=1e5, poisson=0.3,
                gravity=(0.0, 0.0, 0.0), polar_fp64=1)


def random_state(seed, n_side=7, dx=0.01, h=0.021, bodies=True):
    """A jittered lattice body, deformed and rotated, with random symmetric
    stresses, proper rotations, velocities and densities, and a few static
    body particles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x0, g = lattice(n_side, n_side, n_side, dx, (0.05, 0.05, 0.05))
    x0 = x0 + rng.uniform(-0.1 * dx, 0.1 * dx, x0.shape)
    n = len(x0)
    F = ref.rotation_matrix(rng.normal(size=3), 0.4) @ (np.eye(3) + 0.02 * rng.normal(size=(3, 3)))
    c = x0.mean(0)
    x = (x0 - c) @ F.T + c + rng.uniform(-0.05 * dx, 0.05 * dx, x0.shape)
    rot = np.stack([ref.polar_rotation(np.eye(3) + 0.3 * rng.normal(size=(3, 3))) for _ in range(n)])
    stress = 1e3 * rng.normal(size=(n, 6))
    d0 = 1000.0 * (1.0 + 0.05 * rng.uniform(-1, 1, n))
    rho = 1000.0 * (1.0 + 0.05 * rng.uniform(-1, 1, n))
    vel = 0.01 * rng.normal(size=(n, 3))
    body = np.zeros(n, np.uint32)
    if bodies:
        body[g[:, 0] == 0] = 1                      # the x = 0 face: a static body
    gb = [-0.1, 0.3, -0.1, 0.3, -0.1, 0.3]
    return dict(x0=x0, x=x, rot=rot, stress=stress, d0=d0, rho=rho, vel=vel, body=body, gb=gb, h=h,
                mass=1000.0 * dx ** 3, g=g)


def oracle_state(s, viscosity=0.0, **over):
    n = len(s["x"])
    st = OracleState("f64", params(n, s["gb"], s["h"], s["mass"], viscosity), position=s["x"], velocity=s["vel"],
                     originalpos=s["x0"], rotation=s["rot"].reshape(n, 9), stress=s["stress"], density0=s["d0"],
                     density=s["rho"], body_number=s["body"], body_dynamic=np.zeros(1, np.uint32), **over)
    st.bins()
    return st


@pytest.mark.parametrize("seed", [0, 1])
def test_rotations_strains_equal_numpy_restatement(seed):
    s = random_state(seed)
    st = oracle_state(s)
    st.rotations_strains()
    n = len(s["x"])
    cand = ref.candidate_pairs(s["x"], s["gb"], s["h"])
    apq, rot, strain = ref.rotations_strains(s["x"], s["x0"], s["d0"], s["mass"], s["h"], s["body"],
                                             rotation_in=s["rot"], candidates=cand)
    free = s["body"] == 0
    scale = np.abs(apq[free]).max()
    assert np.allclose(st.f["apq_temp"].reshape(n, 3, 3)[free], apq[free], rtol=0, atol=1e-10 * scale)
    # body particles return early: rotation keeps its initial value
    assert np.array_equal(st.f["rotation"].reshape(n, 3, 3)[~free], s["rot"][~free])
    # the oracle's 8-rotation Jacobi against numpy's eigensolver
    assert np.allclose(st.f["rotation"].reshape(n, 3, 3)[free], rot[free], atol=1e-8)
    # sweep 2 on its own, with the oracle's rotation
    _, _, strain2 = ref.rotations_strains(s["x"], s["x0"], s["d0"], s["mass"], s["h"], s["body"],
                                          rotation_used=st.f["rotation"], candidates=cand)
    sc = np.abs(strain2[free]).max()
    assert sc > 1e-3
    assert np.allclose(st.f["strain"][free], strain2[free], rtol=0, atol=1e-10 * sc)
    assert np.allclose(st.f["strain"][free], strain[free], rtol=0, atol=1e-7 * sc)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("bodies", [False, True])
def test_forces_equal_numpy_restatement(seed, bodies):
    s = random_state(seed + 10, bodies=bodies)
    st = oracle_state(s, viscosity=6e-5)
    st.forces_solids(True)
    cand = ref.candidate_pairs(s["x"], s["gb"], s["h"])
    f, fe, fv = ref.forces(s["x"], s["vel"], s["x0"], s["stress"], s["rot"], s["d0"], s["rho"], s["mass"], s["h"],
                           6e-5, s["body"] if bodies else None, candidates=cand)
    for name, want in (("force_elastic", fe), ("force_viscous", fv), ("force", f)):
        got = st.f[name]
        scale = np.abs(want).max()
        assert scale > 0
        assert np.allclose(got, want, rtol=0, atol=1e-10 * scale), name
    # and the SVK variant (no rotations, kernels/solid_svk.cl:182-367)
    st.forces_solids(False)
    _, fe_svk, _ = ref.forces(s["x"], s["vel"], s["x0"], s["stress"], None, s["d0"], s["rho"], s["mass"], s["h"],
                              6e-5, s["body"] if bodies else None, candidates=cand)
    assert np.allclose(st.f["force_elastic"], fe_svk, rtol=0, atol=1e-10 * np.abs(fe_svk).max())


def stretched_body(eps, Q, n_side=13, dx=0.01, h=0.04):
    """A lattice at rest (x = x0), density0 from the oracle's own setup, then
    x = Q (I + eps) (x0 - c) + c."""
    x0, g = lattice(n_side, n_side, n_side, dx)
    n = len(x0)
    c = x0.mean(0)
    gb = [-0.1, 0.25, -0.1, 0.25, -0.1, 0.25]
    mass = 1000.0 * dx ** 3
    st = OracleState("f64", params(n, gb, h, mass), position=x0, rotation=np.tile(np.eye(3).reshape(9), (n, 1)))
    st.setup_solid()
    st.f["position"][:] = (x0 - c) @ (Q @ (np.eye(3) + eps)).T + c
    st.bins()
    r = int(np.ceil(h / dx))
    interior = ((g >= r) & (g < n_side - r)).all(1)
    return st, interior


def test_affine_map_gives_its_strain_and_rotation():
    """kernels/solid_corotation.cl:128-195: for x = Q (I + eps) x0 with eps
    symmetric, R = Q and R^T p - q = eps q, so the strain is eps times the
    lattice's sum of V q (x) gradW(-q), which the spiky kernel's
    normalisation makes ~ I (h = 4 dx here)."""
    eps = np.array([[2e-3, 5e-4, -3e-4], [5e-4, -1e-3, 8e-4], [-3e-4, 8e-4, 1.5e-3]])
    Q = ref.rotation_matrix([1.0, -2.0, 0.5], 0.7)
    st, interior = stretched_body(eps, Q)
    st.rotations_strains()
    n = len(interior)
    R = st.f["rotation"].reshape(n, 3, 3)[interior]
    assert np.allclose(R, Q, atol=1e-6)
    e = ref.sym_full(st.f["strain"][interior])
    err = np.abs(e - eps).max() / np.abs(eps).max()
    assert err < 0.06, err
    # the shear components are halved sums (computeCauchyStrain): compare
    # their ratio to the diagonal, which a missing 1/2 would double
    assert np.allclose(e[:, 0, 1] / eps[0, 1], e[:, 0, 0] / eps[0, 0], rtol=0.12)


def test_linear_stress_field_force_is_half_volume_divergence():
    """kernels/solid_corotation.cl:262-300 with R = I and no bodies:
    f_e,i = 1/2 sum_j m^2/(rho0_i rho0_j) (sigma_i + sigma_j) gradW(x0_i - x0_j),
    whose continuum limit is 1/2 V_i div(sigma)."""
    dx, h, n_side = 0.01, 0.04, 15
    x0, g = lattice(n_side, n_side, n_side, dx)
    n = len(x0)
    mass = 1000.0 * dx ** 3
    gb = [-0.1, 0.25, -0.1, 0.25, -0.1, 0.25]
    st = OracleState("f64", params(n, gb, h, mass), position=x0, rotation=np.tile(np.eye(3).reshape(9), (n, 1)))
    st.setup_solid()
    # a uniform rest density, so that sum_j V_j gradW(x0_i - x0_j) = 0 for
    # every interior particle of the lattice
    st.f["density0"][:] = 1000.0
    # sigma = sigma0 + A x0 componentwise, a divergence that is constant
    rng = np.random.Generator(np.random.PCG64(3))
    sigma0 = 1e3 * rng.normal(size=6)
    A = 1e5 * rng.normal(size=(6, 3))
    st.f["stress"][:] = sigma0 + x0 @ A.T
    st.f["density"][:] = 1000.0
    st.bins()
    st.forces_solids(True)
    # div(sigma)_a = sum_b d sigma_ab / d x_b
    D = np.zeros((3, 3, 3))            # D[a, b, c] = d sigma_ab / d x_c
    for (a, b), k in {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4,
                      (1, 2): 5, (2, 1): 5}.items():
        D[a, b] = A[k]
    div = np.einsum("abb->a", D)
    r = int(np.ceil(h / dx))
    interior = ((g >= r) & (g < n_side - r)).all(1)
    V = mass / st.f["density0"][interior]
    want = 0.5 * V[:, None] * div[None, :]
    got = st.f["force_elastic"][interior]
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 0.06, err


def test_rigid_rotation_turns_elastic_force():
    """force_elastic and rotation of a strained body turn with the body
    (kernels/solid_corotation.cl:262-265: fe = -R_i f_ji + R_j f_ij with
    f_ji, f_ij in the reference configuration)."""
    eps = np.array([[3e-3, 1e-3, 0.0], [1e-3, -2e-3, 5e-4], [0.0, 5e-4, 1e-3]])
    Q0 = ref.rotation_matrix([0.3, 1.0, -0.4], 0.5)
    Q = ref.rotation_matrix([-1.0, 0.2, 0.7], 1.1)
    out = []
    for turn in (np.eye(3), Q):
        st, interior = stretched_body(eps, turn @ Q0, n_side=9, h=0.021)
        st.rotations_strains()
        st.stresses_corot()
        st.forces_solids(True)
        out.append((st.f["force_elastic"].copy(), st.f["rotation"].reshape(-1, 3, 3).copy(), st.f["strain"].copy()))
    (fa, Ra, ea), (fb, Rb, eb) = out
    assert np.abs(fa).max() > 0
    assert np.allclose(Rb, Q @ Ra, atol=1e-9)
    assert np.allclose(eb, ea, rtol=0, atol=1e-9 * np.abs(ea).max())
    assert np.allclose(fb, fa @ Q.T, rtol=0, atol=1e-7 * np.abs(fa).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_elastic_forces_conserve_momentum(seed):
    """Without bodies and planes each pair's terms cancel: the pair (i, j)
    adds -R_i f_ji + R_j f_ij to i and exactly its negative to j."""
    s = random_state(seed + 20, bodies=False)
    st = oracle_state(s)
    st.forces_solids(True)
    fe = st.f["force_elastic"]
    total = fe.sum(0)
    assert np.abs(total).max() < 1e-12 * np.abs(fe).sum(0).max()


def test_rest_state_has_zero_force():
    st, _ = stretched_body(np.zeros((3, 3)), np.eye(3), n_side=7, h=0.021)
    st.rotations_strains()
    st.stresses_corot()
    st.f["velocity"][:] = 0.0
    st.f["density"][:] = st.f["density0"]
    st.forces_solids(True)
    assert np.abs(st.f["strain"]).max() < 1e-13
    for name in ("force", "force_elastic", "force_viscous"):
        assert np.abs(st.f[name]).max() < 1e-13, name
