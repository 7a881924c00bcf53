"""Fluid step on the GPU against the FP32 oracle: density, forces and the
integrator after 1 and 100 steps, both with the reference's own model
(linear EOS, force without the particle's mass) and with the stable model
the benchmarks use (Tait EOS, mass-consistent force).

Bitwise: the engine performs the same IEEE operations in the same order as
the oracle (no FMA contraction, correctly rounded division and sqrt), with
plain and with Kahan-compensated sums, so every field must match exactly."""
import numpy as np
import pytest

from gpu_util import compare, download, engine, oracle_state, say
from sph import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("density", "num_neighbours", "force", "acceleration", "velocity", "position")


def run_pair(scene, steps, compensated=0):
    s = engine(scene, compensated=compensated)
    st = oracle_state(scene, params={"compensated": compensated})
    s.run(steps)
    st.step("fluid", steps)
    got = download(s, *FIELDS)
    s.free()
    return got, st


def check(got, st, exact_counts=True):
    """Bitwise: the engine runs the oracle's operation sequence."""
    for k in ("density", "force", "acceleration", "velocity", "position"):
        compare(k, got[k], st.f[k])
    if exact_counts:
        assert np.array_equal(got["num_neighbours"], st.num_neighbours)


@pytest.mark.parametrize("model", ["reference", "tait"])
def test_one_step(hip_device, model):
    tait = model == "tait"
    scene = scenes.dam_break(32, 32, 16, 0.02, 0.04, 1e-4, eos_tait=tait, mass_consistent=tait)
    say(f"fluid 1 step, model={model}, n={scene.n}")
    got, st = run_pair(scene, 1)
    check(got, st)


def test_hundred_steps_tait(hip_device):
    scene = scenes.baseline_config(0)
    say(f"fluid 100 steps on {scene.name}")
    got, st = run_pair(scene, 100)
    check(got, st)


def test_reference_model_ten_steps(hip_device):
    """The reference's own fluid model at these parameters diverges within a
    few steps; the engine must still follow the oracle through it."""
    scene = scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, eos_tait=False, mass_consistent=False)
    got, st = run_pair(scene, 10)
    check(got, st)


def test_deterministic_repeat(hip_device):
    scene = scenes.dam_break(24, 24, 12, 0.02, 0.04, 1e-4, seed=9)
    outs = []
    for _ in range(2):
        s = engine(scene)
        s.run(20)
        outs.append(download(s, "position", "velocity", "density"))
        s.free()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k])


def test_named_kernel_sequence_matches_fast_path(hip_device):
    """Fluid.update (one call per reference kernel) and the fused step give
    the same state."""
    scene = scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, seed=2)
    a = engine(scene)
    for _ in range(5):
        a.compute_particle_bins()
        for k in ("compute_density", "compute_forces_fluids", "step_forward"):
            a.call_for_all_particles(k)
    b = engine(scene)
    b.run(5)
    ga, gb = download(a, *FIELDS), download(b, *FIELDS)
    for k in FIELDS:
        assert np.array_equal(ga[k], gb[k]), k
    a.free()
    b.free()


def test_large_dam_break_properties(hip_device):
    """4.19 M particles: the grid is a stable permutation, and without walls
    or gravity the internal forces cancel (pairwise antisymmetric)."""
    scene = scenes.baseline_config(3)
    say(f"{scene.name}: n={scene.n}")
    s = engine(scene)
    s.run(2)
    got = download(s, "gridcell", "gridcount", "cellparticles", "force", "density", "position")
    n = scene.n
    cp = got["cellparticles"][:n]
    assert np.array_equal(np.sort(cp), np.arange(n, dtype=np.uint32))
    assert int(got["gridcount"].sum()) == n
    key = got["gridcell"][cp].astype(np.int64)
    assert np.all(np.diff(key) >= 0)
    assert np.all(np.diff(cp.astype(np.int64))[np.diff(key) == 0] > 0)
    f = got["force"].astype(np.float64)
    total = np.abs(f.sum(axis=0))
    assert np.all(total <= 1e-4 * np.abs(f).sum(axis=0) + 1e-12), total
    rho = got["density"]
    assert np.all(np.isfinite(rho)) and rho.min() > 0
    s.free()


@pytest.mark.parametrize("steps", [1, 20])
@pytest.mark.parametrize("compensated", [1, 3], ids=["kahan_density", "kahan_all"])
def test_kahan_sums_bitwise(hip_device, steps, compensated):
    """Kahan-compensated density sums (the benchmarks' setting), and density
    plus force sums, against the oracle's compensated sums: the same
    operation sequence, bit for bit."""
    scene = scenes.dam_break(24, 24, 12, 0.02, 0.04, 1e-4, seed=7, eos_tait=True, mass_consistent=True)
    say(f"fluid {steps} step(s), compensated={compensated}, n={scene.n}")
    got, st = run_pair(scene, steps, compensated=compensated)
    check(got, st)


def test_large_dam_break_three_steps_bitwise(hip_device):
    """The 4.19 M-particle benchmark scene (Tait, mass-consistent, Kahan
    density sums as benchmarked): three full steps (binning and reorder of an
    already sorted state included) bitwise against the oracle."""
    scene = scenes.baseline_config(3)
    say(f"{scene.name}: n={scene.n}, 3 steps, Kahan density, bitwise")
    got, st = run_pair(scene, 3, compensated=1)
    check(got, st)


@pytest.mark.parametrize("compensated", [0, 3], ids=["plain", "kahan_all"])
def test_overflowing_lists_bitwise(hip_device, compensated):
    """Packed at dx = h/2.35, most interior particles have more than 48
    neighbours within h: past the 40 entries of the density scan's LDS
    column, their list entries go straight to the pair cache (up to 64), and
    the few lanes with more take the candidate scan in the density and force
    passes, while the surface lanes keep short lists.  Five steps, bitwise."""
    scene = scenes.dam_break(20, 20, 10, 0.04 / 2.35, 0.04, 2e-3, seed=11)
    say(f"fluid dense packing, 5 steps, compensated={compensated}, n={scene.n}")
    got, st = run_pair(scene, 5, compensated=compensated)
    check(got, st)
    nn = st.num_neighbours
    assert (nn > 48).mean() > 0.25 and (nn < 47).mean() > 0.25


def test_lists_past_the_lds_column_bitwise(hip_device):
    """At h = 2.2 dx most interior lists hold 41-48 entries, more than the
    density scan keeps in LDS: the rest of each list is written to the pair
    cache during the scan and read back by the kernel sum.  Five steps,
    Kahan density sums, bitwise."""
    scene = scenes.dam_break(20, 20, 10, 0.04 / 2.2, 0.04, 2e-3, seed=12)
    say(f"fluid lists past the LDS column, 5 steps, n={scene.n}")
    got, st = run_pair(scene, 5, compensated=1)
    check(got, st)
    nn = st.num_neighbours
    assert ((nn > 40) & (nn <= 64)).mean() > 0.25


def test_crowded_row_takes_the_scan(hip_device):
    """4913 particles in one grid cell: each neighbour row holds more slots
    than a 12-bit list entry addresses (4096), so every lane takes the
    candidate scan in the density and force passes.  Two steps, Kahan
    density sums, bitwise."""
    scene = scenes.dam_break(17, 17, 17, 0.04 / 17, 0.04, 1e-4, seed=5, origin=(0.481, 0.481, 0.481),
                             name="crowded")
    say(f"fluid crowded cell, 2 steps, n={scene.n}")
    got, st = run_pair(scene, 2, compensated=1)
    check(got, st)
    assert st.num_neighbours.max() > 4096
