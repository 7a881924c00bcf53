"""Scripted rigid bodies and initial-condition helpers on the engine against
the oracle (kernels/rigidbody.cl, kernels/common.cl:533-604): bit-exact, with
the engine's slots permuted by earlier steps, and the reference's own flow of
src/testopencl.c:45-50 (6x6x6 lattice on [-1, 1]^3, then a rotation by
(pi/4, 0, pi/6)).  Error paths are checked through the same C ABI."""
import dataclasses
import math

import numpy as np
import pytest

import oracle
from gpu_util import compare, download, engine, say
from sph import bindings, scenes

pytestmark = pytest.mark.gpu


def fluid_small():
    return scenes.dam_break(16, 16, 8, 0.02, 0.04, 1e-4, seed=5, steps=5, name="fluid_small")


def rigid_cantilever():
    sc = scenes.cantilever(16, 4, 4, steps=5, name="corot_rigid")
    return dataclasses.replace(sc, kernel_files=sc.kernel_files + ", rigidbody")


def n_of(system):
    return int(download(system, "n")["n"].reshape(-1)[0])


def test_populate_then_rotate_reference_flow():
    sc = fluid_small()
    pnum = sc.n
    s = engine(sc)
    try:
        s.run(2)                                  # slots now in cell order, not particle order
        state = download(s, "position")["position"][:pnum].reshape(pnum, 3)
        say(f"  populate 6x6x6 on [-1, 1]^3 into pnum {pnum}")
        s.psdata_opencl.populate_position_cuboid((-1, -1, -1), (1, 1, 1), (6, 6, 6))
        want, want_n = oracle.populate_position_cuboid(state, pnum, sc.n, (-1, -1, -1), (1, 1, 1), (6, 6, 6))
        assert n_of(s) == want_n == 216
        got = download(s, "position")["position"][:pnum].reshape(pnum, 3)
        assert compare("position after populate", got, want) == 0
        say("  rotate (pi/4, 0, pi/6)")
        s.psdata_opencl.rotate_particles(math.pi / 4, 0.0, math.pi / 6)
        want = oracle.rotate_particles(want, want_n, np.float32(math.pi / 4), 0.0, np.float32(math.pi / 6))
        got = download(s, "position")["position"][:pnum].reshape(pnum, 3)
        assert compare("position after rotate", got, want) == 0
        # a lattice that reaches pnum sets n = pnum
        s.psdata_opencl.populate_position_cuboid((0.1, 0.1, 0.1), (0.6, 0.5, 0.4), (20, 20, 20))
        want, want_n = oracle.populate_position_cuboid(want, pnum, want_n, (0.1, 0.1, 0.1), (0.6, 0.5, 0.4),
                                                       (20, 20, 20))
        assert n_of(s) == want_n == pnum
        got = download(s, "position")["position"][:pnum].reshape(pnum, 3)
        assert compare("position after populate (n = pnum)", got, want) == 0
        # the step still runs on the new state and its particles
        s.run(1)
        assert np.isfinite(download(s, "position")["position"]).all()
    finally:
        s.free()


def test_rotate_permuted_slots_matches_oracle():
    sc = fluid_small()
    s = engine(sc)
    try:
        s.run(3)
        before = download(s, "position")["position"].reshape(-1, 3)
        s.psdata_opencl.rotate_particles(0.3, -0.7, 1.1)
        got = download(s, "position")["position"].reshape(-1, 3)
        want = oracle.rotate_particles(before, sc.n, np.float32(0.3), np.float32(-0.7), np.float32(1.1))
        assert compare("rotated positions", got, want) == 0
    finally:
        s.free()


def test_set_body_position_and_rotation_match_oracle():
    sc = rigid_cantilever()
    dt = float(np.float32(sc.params["dt"]))
    fields = ("position", "velocity", "rotation", "body_number", "body_position", "body_euler_rotation")
    s = engine(sc)
    try:
        s.run(3)
        st = download(s, *fields)
        n = sc.n
        pos, vel = st["position"].reshape(-1, 3), st["velocity"].reshape(-1, 3)
        rot, bn = st["rotation"].reshape(-1, 9), st["body_number"].reshape(-1)
        bpos, beul = st["body_position"].reshape(-1, 3), st["body_euler_rotation"].reshape(-1, 3)
        body_count = 2
        target = np.array([0.01, -0.005, 0.02], np.float32)
        say(f"  set_body_position(0, {target.tolist()}) on {int((bn == 1).sum())} of {n} particles")
        s.set_body_position(0, target)
        want_p, want_v, want_b = oracle.set_body_position(pos, vel, bn, n, body_count, dt, 0, bpos, target)
        got = download(s, *fields)
        assert compare("position", got["position"].reshape(-1, 3), want_p) == 0
        assert compare("velocity", got["velocity"].reshape(-1, 3), want_v) == 0
        assert compare("body_position", got["body_position"].reshape(-1, 3), want_b) == 0
        # two rotations: the second uses the angles the first stored
        for angles in ([0.2, -0.3, 0.5], [-0.4, 0.1, 0.9]):
            ang = np.array(angles, np.float32)
            s.set_body_rotation(0, ang)
            want_p, want_r, want_e = oracle.set_body_rotation(want_p, rot, bn, n, body_count, 0, beul, ang)
            got = download(s, *fields)
            assert compare(f"position after rotation {angles}", got["position"].reshape(-1, 3), want_p) == 0
            assert compare("rotation field", got["rotation"].reshape(-1, 9), want_r) == 0
            assert compare("body_euler_rotation", got["body_euler_rotation"].reshape(-1, 3), want_e) == 0
            rot, beul = want_r, want_e
        with pytest.raises(bindings.SphError, match="outside"):
            s.set_body_position(2, target)        # body_position holds 2 bodies
        s.run(2)
        assert np.isfinite(download(s, "position")["position"]).all()
    finally:
        s.free()


def test_body_and_helper_errors():
    sc = scenes.cantilever(16, 4, 4, steps=5, name="corot_plain")   # no 'rigidbody' in the kernel files
    s = engine(sc)
    try:
        with pytest.raises(bindings.SphError, match="rigidbody"):
            s.set_body_position(0, np.zeros(3, np.float32))
        with pytest.raises(bindings.SphError, match="rigidbody"):
            s.set_body_rotation(0, np.zeros(3, np.float32))
        with pytest.raises(bindings.SphError, match="set_body_position_device_hip"):
            s.call_for_all_particles("set_body_position")
        with pytest.raises(bindings.SphError, match="32-bit"):
            s.psdata_opencl.populate_position_cuboid((0, 0, 0), (1, 1, 1), (70000, 70000, 2))
        # a lattice with no points changes nothing
        before = download(s, "position", "n")
        s.psdata_opencl.populate_position_cuboid((0, 0, 0), (1, 1, 1), (0, 4, 4))
        after = download(s, "position", "n")
        assert np.array_equal(before["position"], after["position"]) and np.array_equal(before["n"], after["n"])
    finally:
        s.free()
