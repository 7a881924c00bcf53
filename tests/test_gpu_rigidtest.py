"""The rigid-body harness (sph.rigidtest, counterpart of
scripts/rigidtest.py:4-110) on the engine and, through the same harness
code, on the CPU oracle: a soft block between two scripted bodies, the left
one moved and the right one turned every step (set_body_position /
set_body_rotation, kernels/rigidbody.cl), histories equal bit for bit."""
import numpy as np
import pytest

from gpu_util import say
from oracle_solid import OracleSolid
from rigid_scene import sandbox_params

pytestmark = pytest.mark.gpu


def test_rigid_harness_engine_vs_oracle(hip_device):
    from sph.rigidtest import RigidTest
    from sph.systems import StrainSolid
    gpu_solid, cpu_solid = StrainSolid(), OracleSolid("corotational")
    try:
        gpu = RigidTest(**sandbox_params(gpu_solid, steps=12))
        cpu = RigidTest(**sandbox_params(cpu_solid, steps=12))
        assert gpu.run() == "success"
        assert cpu.run() == "success"
        for name in ("position", "velocity"):
            hg, hc = gpu.history[name], cpu.history[name]
            say(f"  {name}: {hg.shape[0]} steps, {hg.shape[1]} particles, max |difference| "
                f"{np.abs(hg - hc).max():.3e}")
            assert np.array_equal(hg, hc), name
        # the turned body rotated about its position, the moved one translated
        n_soft = len(gpu.params.soft_particles)
        assert not np.array_equal(gpu.history["position"][-1, n_soft:], gpu.positions[n_soft:].astype(np.float32))
    finally:
        gpu_solid.free()
        cpu_solid.free()
