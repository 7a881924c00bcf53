"""The rigid-body harness (sph.rigidtest, counterpart of
scripts/rigidtest.py:4-110) driving the CPU oracle: required parameters,
body numbering, history layout, and the scripted bodies following their
animations (a static body moves only by set_body_position, so its particles
keep their offsets from the body's position; an unanimated static body does
not move at all)."""
import numpy as np
import pytest

from oracle_solid import OracleSolid
from rigid_scene import sandbox_params
from sph.rigidtest import RigidTest


def test_missing_parameters_are_named():
    with pytest.raises(Exception, match="run_step_count"):
        RigidTest(soft_particles=np.zeros((1, 3)), rigid_bodies=[], total_mass=1.0, strain_solid=None,
                  history_particles="all", history_fields=["position"])


def test_bodies_follow_their_animation_on_the_oracle():
    solid = OracleSolid("corotational")
    try:
        p = sandbox_params(solid, steps=6, rotate=False)
        test = RigidTest(**p)
        n_soft, n_left, n_right = (len(p["soft_particles"]), len(p["rigid_bodies"][0].particles),
                                   len(p["rigid_bodies"][1].particles))
        assert list(test.rigid_body_start_indices) == [n_soft, n_soft + n_left]
        bn = test.body_numbers
        assert (bn[:n_soft] == 0).all() and (bn[n_soft:n_soft + n_left] == 1).all() and (bn[n_soft + n_left:] == 2).all()
        assert test.run() == "success"
        hist = test.history["position"]
        assert hist.shape == (6, n_soft + n_left + n_right, 3)
        start = test.positions.astype(np.float32)
        path = p["rigid_bodies"][0].pos_animation
        for k in range(6):
            moved = hist[k, n_soft:n_soft + n_left] - start[n_soft:n_soft + n_left]
            np.testing.assert_allclose(moved, np.broadcast_to(path[k], moved.shape), rtol=0, atol=2e-7)
            assert np.array_equal(hist[k, n_soft + n_left:], start[n_soft + n_left:])
        assert np.isfinite(hist).all()
    finally:
        solid.free()
