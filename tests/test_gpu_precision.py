"""Precision against the FP64 oracle (stated tolerances).

One step compares the passes' own arithmetic (identical FP32 inputs on both
sides); a few steps include the drift of the FP32 state.  Tolerances are
normwise relative errors ||gpu - ref||_2 / ||ref||_2 with a margin over the
measured values (tests/precision_report.py prints them):

    fluid, 1 step:   density < 3e-7, force < 1e-5; Kahan sums no worse than plain
    solids, 1 step:  FP64 polar decomposition / inverse at least 10x closer to
                     the FP64 oracle than FP32 (rotation, deformation), < 1e-8
    20 steps:        density < 1e-5; displacement < 5e-3 (fluid), < 1e-3
                     (corotational), < 1e-2 (SVK: stress from F^T F - I cancels
                     about four of FP32's digits)
    FP64 reductions: momentum / kinetic energy / centre of mass equal the FP64
                     sums of the downloaded state to 1e-12 relative
"""
import numpy as np
import pytest

from gpu_util import download, engine, oracle_state, say
from precision_report import rel_errors

pytestmark = pytest.mark.gpu

from sph import scenes  # noqa: E402


def small(system):
    if system == "fluid":
        return scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, seed=5, steps=20, name="fluid_small")
    if system == "corotational":
        return scenes.cantilever(16, 4, 4, steps=20, name="corot_small")
    return scenes.svk_cube(nside=8, steps=20, name="svk_small")


def run(scene, steps, compensated, polar, fields):
    s = engine(scene, compensated=compensated, polar_fp64=polar)
    s.run(steps)
    out = download(s, *fields)
    sums = s.psdata_opencl.global_sums()
    s.free()
    return {k: v[:scene.n] for k, v in out.items()}, sums


def test_fluid_one_step_sum_error():
    sc = small("fluid")
    ref = oracle_state(sc, "f64")
    ref.step("fluid", 1)
    err = {}
    for comp in (0, 1):
        got, _ = run(sc, 1, comp, True, ("density", "force"))
        err[comp] = {k: rel_errors(got[k], ref.f[k])["normwise"] for k in got}
        say(f"  compensated={comp}: density {err[comp]['density']:.3e}, force {err[comp]['force']:.3e}")
        assert err[comp]["density"] < 3e-7
        assert err[comp]["force"] < 1e-5
    assert err[1]["density"] <= err[0]["density"] * 1.05


@pytest.mark.parametrize("system,field", [("corotational", "rotation"), ("svk", "deformation")])
def test_solid_fp64_polar_closer_to_oracle(system, field):
    sc = small(system)
    ref = oracle_state(sc, "f64")
    ref.step(system, 1)
    e32 = rel_errors(run(sc, 1, 0, False, (field,))[0][field], ref.f[field])["normwise"]
    e64 = rel_errors(run(sc, 1, 0, True, (field,))[0][field], ref.f[field])["normwise"]
    say(f"  {system} {field}: FP32 {e32:.3e}, FP64 {e64:.3e}")
    assert e64 < 1e-8
    assert e64 * 10 < e32


@pytest.mark.parametrize("system", ["fluid", "corotational", "svk"])
def test_twenty_steps_drift(system):
    sc = small(system)
    ref = oracle_state(sc, "f64")
    ref.step(system, sc.steps)
    got, sums = run(sc, sc.steps, 1, True, ("position", "velocity", "density"))
    pos0 = np.asarray(sc.fields["position"], np.float64)
    e_rho = rel_errors(got["density"], ref.f["density"])["normwise"]
    e_x = rel_errors(got["position"] - pos0, ref.f["position"] - pos0)["normwise"]
    say(f"  {system}: density {e_rho:.3e}, displacement {e_x:.3e}")
    assert e_rho < 1e-5
    assert e_x < {"fluid": 5e-3, "corotational": 1e-3, "svk": 1e-2}[system]
    # FP64 device reductions = FP64 sums of the same FP32 state
    m = float(sc.params["mass"])
    v = got["velocity"].astype(np.float64)
    x = got["position"].astype(np.float64)
    mom = m * v.sum(0)
    ke = 0.5 * m * (v * v).sum()
    com = x.mean(0)
    scale = m * np.abs(v).sum(0).max()          # momentum can cancel: compare on the terms' scale
    assert np.allclose(sums[0], mom, rtol=0, atol=1e-12 * max(scale, 1e-30))
    assert abs(sums[1] - ke) <= 1e-12 * abs(ke)
    assert np.allclose(sums[2], com, rtol=1e-12, atol=0)


@pytest.mark.parametrize("system", ["fluid", "corotational", "svk"])
def test_tuned_block_sizes_change_no_bit(system):
    """Every neighbour block size and fluid force load group the autotuner
    may pick gives the same bits."""
    from sph import bindings
    sc = small(system)
    fields = ("position", "velocity", "density", "force")
    before = bindings.get_tuning()
    settings = [(64, 8), (128, 8), (256, 8), (256, 12), (256, 16), (64, 16)]
    try:
        outs = {}
        for nb, pg in settings:
            bindings.set_tuning(nb, pg)
            outs[(nb, pg)] = run(sc, 5, 1, True, fields)[0]
    finally:
        bindings.set_tuning(before["neighbour_block"], before["pair_group"])
    for key in settings[1:]:
        for k in fields:
            assert np.array_equal(outs[key][k], outs[settings[0]][k]), (system, key, k)
