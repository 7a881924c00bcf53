"""Solid passes on the GPU against the FP32 oracle, bit for bit.

The fixed 8-rotation Jacobi runs the same operation sequence on both sides
in both precisions: FP64 (the default) from half-angle identities with
correctly rounded sqrt and division, FP32 with the reference's atan, cos and
sin, defined once as double polynomials rounded to float (the oracle's copy
is pinned against the C library in tests/test_trig_definition.py).  Every
other operation is IEEE float or double in the reference's order, so every
field matches exactly.  Mirror planes, plane constraints and clamped (static
body) particles are all exercised."""
import numpy as np
import pytest

from gpu_util import compare, download, engine, oracle_state, say
from sph import scenes

pytestmark = pytest.mark.gpu


def small_solid(kind, nside=9, stretch=1.01, planes=True, constraints=True, bodies=True, seed=4, h=0.02,
                jitter=0.0):
    """A cube strained by 1% along x, with mirror planes through its low x
    and y faces, a plane constraint on its high-x face and a clamped layer.
    jitter: the reference lattice's points moved by up to that much."""
    dx = 0.01
    pts, g = scenes.lattice(nside, nside, nside, dx)
    rng = np.random.Generator(np.random.PCG64(seed))
    if jitter:
        pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    n = len(pts)
    lo, hi = pts.min(0) - 4 * h, pts.max(0) + 4 * h
    gb = [float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]), float(lo[2]), float(hi[2])]
    svk = kind == "svk"
    spec = scenes.solid_spec(n, 1e-3, 5e-5, h, gb, 6e-5, 1e5, 0.3, gravity=(0, 0, -9.81), svk=svk)
    pos0 = pts.astype(np.float32)
    fields = {"position": pos0, "n": np.array([n], np.uint32)}
    params = dict(n=n, gridbounds=gb, h=h, mass=float(np.float32(1e-3)), dt=5e-5, viscosity=6e-5, youngs=1e5,
                  poisson=0.3, gravity=(0.0, 0.0, -9.81))
    sym = np.zeros((3, 2, 3), np.float32)
    if planes:
        sym[0] = [[0, 0, 0], [1, 0, 0]]          # x = 0 plane
        sym[1] = [[0, 0, 0], [0, 1, 0]]          # y = 0 plane
    fields["symmetry_planes"] = sym
    params["symmetry_planes"] = sym.reshape(3, 6)
    pc = np.zeros((4, 2, 3), np.float32)
    mask = np.zeros(n, np.uint32)
    if constraints:
        xmax = float(pts[:, 0].max()) * stretch
        pc[0] = [[xmax, 0, 0], [1, 0, 0]]
        mask[g[:, 0] == nside - 1] = 1
    fields["plane_constraints"] = pc
    fields["plane_constraints_particles"] = mask
    params["plane_constraints"] = pc.reshape(4, 6)
    if bodies:
        body = (g[:, 2] == 0).astype(np.uint32)
        fields["body_number"] = body
        fields["body_dynamic"] = np.array([0, 1], np.uint32)
    if not svk:
        fields["rotation"] = np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
    strained = (pts * np.array([stretch, 1.0, 1.0]) + rng.normal(scale=2e-4, size=pts.shape)).astype(np.float32)
    vel = rng.normal(scale=1e-3, size=pts.shape).astype(np.float32)
    scene = scenes.Scene(f"{kind}_small", kind, spec, scenes.SVK_KERNELS if svk else scenes.COROTATIONAL_KERNELS,
                         fields, params)
    return scene, strained, vel


def run_both(kind, steps, polar_fp64=True, compensated=0, **kw):
    scene, strained, vel = small_solid(kind, **kw)
    s = engine(scene, polar_fp64=polar_fp64, compensated=compensated)
    st = oracle_state(scene, polar_fp64=polar_fp64, params={"compensated": compensated})
    # reference configuration from the unstrained lattice, then deform
    got0 = download(s, "originalpos", "density0")
    compare("density0", got0["density0"], st.f["density0"])
    compare("originalpos", got0["originalpos"], st.f["originalpos"])
    s.set_fields(position=strained, velocity=vel)
    st.f["position"][:] = strained
    st.f["velocity"][:] = vel
    s.run(steps, apply_constraints=True)
    st.step(kind, steps, apply_constraints=True)
    names = ["position", "velocity", "density", "stress", "force", "force_elastic", "force_viscous",
             "cutoff_force"]
    names += ["rotation", "strain"] if kind == "corotational" else ["deformation"]
    got = download(s, *names)
    s.free()
    return got, st


def test_cantilever_five_steps(hip_device):
    """The 16 k-particle corotational cantilever workload, bitwise (FP64 polar)."""
    scene = scenes.baseline_config(1)
    s = engine(scene)
    st = oracle_state(scene)
    s.run(5)
    st.step("corotational", 5)
    got = download(s, "position", "velocity", "stress", "rotation", "force")
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))
    # the clamped face did not move
    body = scene.fields["body_number"] > 0
    assert np.array_equal(got["position"][body], scene.fields["position"][body])
    s.free()


@pytest.mark.parametrize("kind", ["corotational", "svk"])
@pytest.mark.parametrize("steps", [1, 20])
@pytest.mark.parametrize("polar_fp64", [True, False], ids=["fp64_polar", "fp32_polar"])
def test_solid_passes_bitwise(hip_device, kind, steps, polar_fp64):
    """The engine's Jacobi and the oracle's are the same IEEE operation
    sequence in both precisions (FP64: half-angle identities with correctly
    rounded sqrt and division; FP32: the reference's atan/cos/sin, defined as
    the same double polynomials on both sides), so every solid field is
    bitwise the oracle's: one step with mirror planes, plane constraints and
    a static layer, and 20 steps without the planes."""
    say(f"{kind}: {steps} step(s), {'FP64' if polar_fp64 else 'FP32'} polar, bitwise")
    got, st = run_both(kind, steps, polar_fp64=polar_fp64, planes=steps == 1)
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))


@pytest.mark.parametrize("kind", ["corotational", "svk"])
@pytest.mark.parametrize("compensated", [1, 3], ids=["kahan_density", "kahan_all"])
def test_kahan_sums_bitwise(hip_device, kind, compensated):
    """Kahan-compensated density and rest-density sums (the benchmarks'
    setting), and those plus the force, A_pq / A_qq and displacement-gradient
    sums, against the oracle's compensated sums, bit for bit."""
    say(f"{kind}: 5 steps, compensated={compensated}, bitwise")
    got, st = run_both(kind, 5, compensated=compensated)
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))


def test_svk_cube_262k_bitwise(hip_device):
    """The 262 k-particle SVK cube (seeded smooth velocity field), three full
    steps with Kahan density sums, bitwise against the oracle."""
    scene = scenes.baseline_config(2)
    say(f"{scene.name}: n={scene.n}, 3 steps, Kahan density, bitwise")
    s = engine(scene, compensated=1)
    st = oracle_state(scene, params={"compensated": 1})
    s.run(3)
    st.step("svk", 3)
    got = download(s, "position", "velocity", "density", "deformation", "stress", "force")
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))
    s.free()


@pytest.mark.parametrize("kind", ["corotational", "svk"])
@pytest.mark.parametrize("compensated", [0, 3], ids=["plain", "kahan_all"])
def test_overflowing_lists_bitwise(hip_device, kind, compensated):
    """h = 2.35 dx: most interior particles have more than 48 neighbours
    within h, more than a neighbour list holds, so their lanes scan the
    candidates in every solid pass while the surface lanes read their
    lists.  Three steps, bitwise."""
    say(f"{kind}: dense packing, 3 steps, compensated={compensated}, bitwise")
    got, st = run_both(kind, 3, compensated=compensated, nside=12, h=0.0235, planes=False)
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))
    nn = st.num_neighbours
    assert (nn > 48).mean() > 0.25 and (nn < 47).mean() > 0.1


@pytest.mark.parametrize("kind", ["corotational", "svk"])
def test_lists_past_the_lds_column_bitwise(hip_device, kind):
    """h = 2.24 dx on a reference lattice jittered by 0.1 dx: about half the
    lists (current or reference neighbours) hold 41-48 entries, more than the
    solid density scan keeps in LDS; the rest of each list goes to the pair
    cache during the scan, and every solid pass reads the whole list.  A few
    lists exceed 48 and scan.  Three steps, Kahan density sums, bitwise."""
    kw = dict(nside=12, h=0.0224, planes=False, jitter=1e-3)
    scene, strained, _ = small_solid(kind, **kw)
    pos0 = scene.fields["position"]
    h = np.float32(0.0224)
    near = lambda x: np.linalg.norm(x[:, None].astype(np.float64) - x[None], axis=2) <= h
    union = (near(pos0) | near(strained)).sum(1)
    assert ((union > 40) & (union <= 48)).mean() > 0.3 and (union > 48).any()
    say(f"{kind}: lists past the LDS column, 3 steps, bitwise")
    got, st = run_both(kind, 3, compensated=1, **kw)
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))


@pytest.mark.parametrize("kind", ["corotational", "svk"])
def test_crowded_rows_take_the_scan(hip_device, kind):
    """4913 particles within one smoothing length, over 8 grid cells: each
    neighbour row holds more than 1024 slots, more than a solid list entry
    addresses, so every lane scans the candidates in every solid pass.  Two
    steps, Kahan density sums, bitwise."""
    if kind == "corotational":
        scene = scenes.cantilever(17, 17, 17, dx=0.001, h=0.02, steps=2, name="crowded")
    else:
        scene = scenes.svk_cube(nside=17, dx=0.001, h=0.02, steps=2, name="crowded")
    say(f"{kind}: crowded rows, 2 steps, n={scene.n}")
    s = engine(scene, compensated=1)
    st = oracle_state(scene, params={"compensated": 1})
    s.run(2)
    st.step(kind, 2)
    names = ["position", "velocity", "density", "stress", "force"]
    names += ["rotation"] if kind == "corotational" else ["deformation"]
    got = download(s, *names)
    s.free()
    for k, v in got.items():
        compare(k, v, st.f[k].reshape(v.shape))
