"""Density and neighbour counts where the cell geometry is least certain,
bit for bit against the oracle's 27-cell scan: particles exactly on cell
faces, particles clamped into border cells from outside the grid, NaN
positions (in range of everything in the reference: applyKernel tests
"dist > h", kernels/common.cl:78-80),
a grid far from the origin, and a density pass on bins older than the
positions.  A neighbour pass after the slot order changed without new bins
is refused."""
import dataclasses

import numpy as np
import pytest

from gpu_util import compare, download, engine, oracle_state, say
from sph import scenes

pytestmark = pytest.mark.gpu


def with_positions(scene, pos):
    fields = dict(scene.fields)
    fields["position"] = np.ascontiguousarray(pos, np.float32)
    return dataclasses.replace(scene, fields=fields)


def face_lattice():
    # spacing h / 2 from lo + h: every other particle sits on a cell face
    return scenes.dam_break(16, 16, 8, 0.02, 0.04, 0.0, seed=1, name="face_lattice")


def clamped_and_nan():
    sc = scenes.dam_break(16, 16, 8, 0.02, 0.04, 2e-3, seed=2, name="clamped_nan")
    pos = np.array(sc.fields["position"], np.float32)
    pos[0] = (-0.05, 0.1, 0.1)        # below the grid in x: binned into x-cell 0
    pos[1] = (1.3, 0.12, 0.1)         # above the grid: binned into the last x-cell
    pos[2] = (0.1, -0.2, 0.3)
    pos[3] = (np.nan, 0.1, 0.1)       # NaN: cell 0 in x
    pos[4] = (0.06, 0.06, np.nan)
    return with_positions(sc, pos)


def far_grid():
    return scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-3, seed=3, gridbounds=(100, 101, -51, -50, 7, 8),
                            name="far_grid")


def density_pair(scene, stale_steps=0):
    s = engine(scene)
    st = oracle_state(scene)
    try:
        s.psdata_opencl.compute_particle_bins()
        st.bins()
        for _ in range(stale_steps):      # move the particles without rebinning
            s.call_for_all_particles("compute_density")
            s.call_for_all_particles("compute_forces_fluids")
            s.call_for_all_particles("step_forward")
            st.density()
            st.forces_fluids()
            st.step_forward()
        s.call_for_all_particles("compute_density")
        st.density()
        got = download(s, "density", "num_neighbours")
    finally:
        s.free()
    return got, st


@pytest.mark.parametrize("make", [face_lattice, clamped_and_nan, far_grid], ids=lambda f: f.__name__)
def test_trimmed_scan_is_bitwise_the_full_scan(make):
    scene = make()
    say(f"  {scene.name}: n={scene.n}")
    got, st = density_pair(scene)
    assert compare("density", got["density"][:scene.n], st.f["density"]) == 0
    assert np.array_equal(got["num_neighbours"][:scene.n], st.num_neighbours)


def test_density_on_stale_bins_matches_oracle():
    scene = scenes.dam_break(16, 16, 8, 0.02, 0.04, 2e-3, seed=4, dt=2e-3, name="stale_bins")
    got, st = density_pair(scene, stale_steps=3)
    assert compare("density (bins three steps old)", got["density"][:scene.n], st.f["density"]) == 0
    assert np.array_equal(got["num_neighbours"][:scene.n], st.num_neighbours)


def test_neighbour_pass_after_reorder_needs_bins():
    scene = scenes.dam_break(8, 8, 8, 0.02, 0.04, 1e-3, seed=5, name="reorder")
    s = engine(scene)
    try:
        s.run(1)
        s.psdata_opencl.populate_position_cuboid((0.1, 0.1, 0.1), (0.3, 0.3, 0.3), (8, 8, 4))
        from sph import bindings
        with pytest.raises(bindings.SphError, match="compute_particle_bins"):
            s.call_for_all_particles("compute_density")
        s.psdata_opencl.compute_particle_bins()
        s.call_for_all_particles("compute_density")
    finally:
        s.free()
