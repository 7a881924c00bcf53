"""The engine through the strain-test harness against the reference
notebook's recorded SVK run "whole_4" (tests/golden/notebook_whole4.json;
see tests/test_known_answers.py for the oracle's side and the provenance):
particle 56 at frame 0, particles 57, 126 and 15 after 3851 steps, with the
reference's stored sigma_zz = 0.  Same tolerances as the oracle's test, for
the FP32 SVK inverse (the reference's) and the FP64 one."""
import pytest

from gpu_util import say
from test_known_answers import FRAME0_TOL, FRAME3850_TOL
from whole4_util import compare_state, golden, run_frames

pytestmark = pytest.mark.gpu


def engine_solid(polar_fp64):
    from sph.systems import SVKSolid
    s = SVKSolid()
    s.psdata_opencl.set_precision(0, polar_fp64)
    return s


def test_frame0_particle_56():
    import numpy as np
    solid = engine_solid(False)
    try:
        state = run_frames(solid, 1)
        want = dict(golden()["states"]["56@0"], deformation=list(np.eye(3).ravel()), stress=[0.0] * 6,
                    force_elastic=[0.0] * 3)
        compare_state(state, 56, want, FRAME0_TOL, say=say)
    finally:
        solid.free()


@pytest.mark.parametrize("polar_fp64", [False, True], ids=["fp32-inverse", "fp64-inverse"])
def test_frame3850_states(polar_fp64):
    solid = engine_solid(polar_fp64)
    try:
        state = run_frames(solid, 3851)
        for pid in (57, 126, 15):
            compare_state(state, pid, golden()["states"][f"{pid}@3850"], FRAME3850_TOL, say=say)
    finally:
        solid.free()
