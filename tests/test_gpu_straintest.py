"""The strain-test harness (sph.straintest, counterpart of
scripts/straintest.py) on the engine and, through the same harness code, on
the CPU oracle: histories and outputs agree bit for bit, constrained faces
follow the planes, and the history/output memmaps are written in the
reference's layout."""
import os

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def params(solid, **over):
    p = dict(strain_range=np.array([0.01, -0.01]), axis=0, strain_rate=20.0, stabilisation_time=5e-4,
             initial_dimensions=np.array([0.5, 0.3, 0.3]), particle_dimensions=np.array([6, 4, 4]),
             total_mass=45.0, strain_solid=solid, output_particles=np.array([[5, 0, 0], [0, 3, 3]]),
             output_fields=["position", "stress"], history_particles="all",
             history_fields=["position", "velocity"], smoothingradius=0.2)
    p.update(over)
    return p


def engine_solid(system):
    from sph.systems import StrainSolid, SVKSolid
    return StrainSolid() if system == "corotational" else SVKSolid()


@pytest.mark.parametrize("variant,system", [("StrainTest", "corotational"), ("StrainTest", "svk"),
                                            ("StrainTestSym", "corotational"),
                                            ("StrainTestSymPadding", "corotational")])
def test_harness_engine_vs_oracle(variant, system, tmp_path):
    from oracle_solid import OracleSolid
    from sph import straintest
    cls = getattr(straintest, variant)
    gpu_solid = engine_solid(system)
    cpu_solid = OracleSolid(system)
    try:
        gpu = cls(**params(gpu_solid, output_folder=str(tmp_path / "gpu")))
        cpu = cls(**params(cpu_solid))
        assert gpu.run_step_count == cpu.run_step_count > 0
        assert gpu.run() == "success"
        assert cpu.run() == "success"
        assert gpu.history_counter == gpu.run_step_count
        hx_gpu = np.asarray(gpu.history["position"], np.float64)
        hx_cpu = np.asarray(cpu.history["position"], np.float64)
        diff = np.abs(hx_gpu - hx_cpu).max()
        say(f"  {variant}/{system}: {gpu.run_step_count} steps, {len(gpu.positions)} particles, "
            f"max |history position difference| {diff:.3e} m")
        assert np.array_equal(hx_gpu, hx_cpu)            # the same operation sequence, bit for bit
        out_gpu = np.asarray(gpu.output["stress"], np.float64)
        out_cpu = np.asarray(cpu.output["stress"], np.float64)
        assert np.array_equal(out_gpu, out_cpu)
        # the faces held by the moving planes follow them
        axis = 0
        mask = np.asarray(gpu.plane_constraints_particles)[:len(gpu.positions)]
        moving = (mask & (1 << (3 if variant != "StrainTest" else 1))) != 0
        assert moving.any()
        path = np.concatenate([np.concatenate(section) for section in gpu.run_path if len(section)])
        np.testing.assert_allclose(hx_gpu[:, moving, axis], np.repeat(path[:, None], moving.sum(), 1).astype(np.float32),
                                   rtol=0, atol=1e-6)
        # reference file layout
        folder = tmp_path / "gpu"
        hist = np.load(folder / "position_history.npy", mmap_mode="r")
        outp = np.load(folder / "stress_output.npy", mmap_mode="r")
        assert hist.shape == (gpu.run_step_count, len(gpu.positions), 3)
        assert outp.shape == (2, 2, 6)
    finally:
        gpu_solid.free()
        cpu_solid.free()
