"""Run path and bookkeeping of the strain-test harness (sph.straintest), on
the CPU: the harness drives an oracle-backed solid, no device is used."""
import numpy as np
import pytest

from oracle_solid import OracleSolid
from sph.straintest import StrainTest, StrainTestSym, StrainTestSymPadding, column_intersection_indices


def params(solid, **over):
    p = dict(strain_range=np.array([0.01, -0.01]), axis=0, strain_rate=20.0, stabilisation_time=5e-4,
             initial_dimensions=np.array([0.5, 0.3, 0.3]), particle_dimensions=np.array([6, 4, 4]),
             total_mass=45.0, strain_solid=solid, output_particles=np.array([[5, 0, 0], [0, 3, 3]]),
             output_fields=["position", "stress"], history_particles="all",
             history_fields=["position", "velocity"], smoothingradius=0.2)
    p.update(over)
    return p


def test_run_path_layout():
    """Tension targets in the order given, then compression targets in
    reverse order, each a move at plane_frame_step per step (none when the
    plane is already past the target) followed by the hold
    (scripts/straintest.py:115-147)."""
    solid = OracleSolid("corotational")
    try:
        t = StrainTest(**params(solid, strain_range=np.array([0.02, -0.01, 0.01, -0.02])))
    finally:
        solid.free()
    assert list(t.strain_order) == [0, 2, 3, 1]
    half = 0.25
    tension, compression = t.run_path
    assert [s[-1] for s in tension] == pytest.approx([1.02 * half, 1.01 * half])
    assert [s[-1] for s in compression] == pytest.approx([0.98 * half, 0.99 * half])
    # 0.25 -> 0.255 at ~5e-4 per step (the FP32 timestep makes the step a hair
    # under 5e-4, so arange yields 11 positions), then the hold; 0.255 -> 0.2525
    # is downhill: hold only
    assert len(tension[0]) == len(np.arange(0.25, 0.255, t.plane_frame_step)) + t.stabilisation_steps == 16
    assert len(tension[1]) == t.stabilisation_steps
    assert t.plane_frame_step == pytest.approx(20.0 * 1e-4 * 0.25, rel=1e-6)
    assert t.stabilisation_steps == 5
    for strip in tension + compression:
        assert np.all(strip[-t.stabilisation_steps:] == strip[-1])
    np.testing.assert_allclose(np.diff(tension[0][:10]), t.plane_frame_step)
    np.testing.assert_allclose(np.diff(compression[0][:5]), -t.plane_frame_step)
    assert t.run_step_count == sum(len(s) for s in tension + compression)


def test_constraints_and_outputs_layout():
    solid = OracleSolid("corotational")
    try:
        t = StrainTest(**params(solid))
        assert t.positions.shape == (96, 3)
        mask = solid.get_fields("plane_constraints_particles")[0][:96]
        assert np.count_nonzero(mask == 1) == 16 and np.count_nonzero(mask == 2) == 16
        pc = solid.get_fields("plane_constraints")[0]
        assert list(pc[0, :, 0]) == [-0.25, 1] and list(pc[1, :, 0]) == [0.25, 1]
        assert list(t.output_particle_indices) == [15, 80]
        assert t.output["stress"].shape == (2, 2, 6)
        assert t.history["position"].shape == (t.run_step_count, 96, 3)
        mass = solid.get_fields("mass")[0][0]
        assert mass == pytest.approx(45.0 / (5 * 3 * 3))
    finally:
        solid.free()


def test_symmetric_variants_layout():
    solid = OracleSolid("corotational")
    try:
        t = StrainTestSym(**params(solid))
        mask = solid.get_fields("plane_constraints_particles")[0][:len(t.positions)]
        planes = solid.get_fields("symmetry_planes")[0]
        assert t.positions.min() == 0.0
        assert np.all(planes[np.arange(3), 1, np.arange(3)] == 1)
        assert np.count_nonzero(mask & 8) == 16           # moving face at x = 0.25
        assert solid.get_fields("mass")[0][0] == pytest.approx(45.0 / (10 * 6 * 6))
    finally:
        solid.free()
    solid = OracleSolid("corotational")
    try:
        t = StrainTestSymPadding(**params(solid))
        pad = solid.get_fields("padding")[0][:len(t.positions)]
        # spacing 0.05 / 0.05 / 0.05 along the octant, h = 0.2: four layers of padding
        assert t.padding_width == pytest.approx([0.2, 0.2, 0.2])
        assert len(t.positions) == 10 * 8 * 8                # core 6x4x4 + 4 layers beyond each free face
        assert np.count_nonzero(pad) == 10 * 8 * 8 - 6 * 4 * 4
        assert t.get_init_plane_position() == pytest.approx(0.45)
    finally:
        solid.free()


def test_column_intersection_indices():
    x = np.array([[0, 0, 0], [1, 2, 3], [4, 5, 6], [1, 2, 3]])
    assert list(column_intersection_indices(x, np.array([[1, 2, 3]]))) == [1, 3]
    assert list(column_intersection_indices(x, np.array([[4, 5, 6], [0, 0, 0]]))) == [0, 2]
