"""Pins of the oracle's scripted-body and initial-condition functions
(kernels/common.cl:533-604, kernels/rigidbody.cl:1-120) against mathematics
they must satisfy, each chosen so that a plausible slip fails:

  populate_position_cuboid  the reference's own case (src/testopencl.c:45: 6x6x6 on
                            [-1, 1]^3) against numpy's linspace lattice, x fastest;
                            n = pnum / lattice size / unchanged; rows past pnum untouched
  rotate_particles          scipy's intrinsic X-Y-Z rotation with the reference's y sign
                            (src/testopencl.c:50 angles); zero angles = identity bitwise;
                            lengths kept; rows past n untouched
  Euler body matrix         scipy's intrinsic X-Y-X rotation (same y sign)
  set_body_position         rigid translation of exactly the tagged particles (pair
                            distances kept), velocity = displacement / dt, store of the
                            new body position, no motion for a body >= body_count
  set_body_rotation         with identity particle rotations: x -> B^T x, rotation -> B^T;
                            with a non-symmetric rotation field: numpy's (B^T P) x and
                            B^T P P
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle

PI = math.pi


def test_populate_reference_case_is_the_lattice():
    pos, n = oracle.populate_position_cuboid(np.zeros((216, 3)), 216, 0, [-1, -1, -1], [1, 1, 1], [6, 6, 6])
    assert n == 216
    axis = np.linspace(-1.0, 1.0, 6)
    k, j, i = np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij")
    want = np.stack([axis[i], axis[j], axis[k]], -1).reshape(-1, 3)     # particle i + 6 j + 36 k
    np.testing.assert_allclose(pos, want, rtol=0, atol=2e-7)
    assert np.array_equal(pos[0], [-1, -1, -1])
    # the FP64 oracle agrees to its own rounding
    pos64, _ = oracle.populate_position_cuboid(np.zeros((216, 3)), 216, 0, [-1, -1, -1], [1, 1, 1], [6, 6, 6],
                                               precision="f64")
    np.testing.assert_allclose(pos64, want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("pnum,size,n_after", [(100, (4, 5, 6), 100), (200, (4, 5, 6), 120), (120, (4, 5, 6), 120),
                                               (50, (0, 5, 6), 7)])
def test_populate_sets_n_and_leaves_the_rest(pnum, size, n_after):
    start = np.full((pnum, 3), 7.0, np.float32)
    pos, n = oracle.populate_position_cuboid(start, pnum, 7, [0, 0, 0], [3, 4, 5], size)
    assert n == n_after
    points = size[0] * size[1] * size[2]
    placed = min(points, pnum)
    assert np.all(pos[placed:] == 7.0)
    if placed:
        # integer-spaced lattice: exact coordinates
        p = np.arange(placed)
        want = np.stack([p % size[0], (p // size[0]) % size[1], p // (size[0] * size[1])], -1)
        assert np.array_equal(pos[:placed], want.astype(np.float32))


def test_rotate_is_scipys_rotation():
    rng = np.random.Generator(np.random.PCG64(3))
    x = rng.uniform(-1, 1, (50, 3))
    ax, ay, az = PI / 4, 0.3, PI / 6
    want = x @ Rotation.from_euler("XYZ", [ax, -ay, az]).as_matrix().T
    got64 = oracle.rotate_particles(x, 50, ax, ay, az, precision="f64")
    np.testing.assert_allclose(got64, want, rtol=0, atol=1e-14)
    got32 = oracle.rotate_particles(x.astype(np.float32), 50, ax, ay, az)
    np.testing.assert_allclose(got32, want, rtol=0, atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(got64, axis=1), np.linalg.norm(x, axis=1), rtol=1e-14)
    # zero angles: identity, bit for bit; rows past n untouched
    x32 = x.astype(np.float32)
    assert np.array_equal(oracle.rotate_particles(x32, 50, 0, 0, 0), x32)
    part = oracle.rotate_particles(x32, 20, ax, ay, az)
    assert np.array_equal(part[20:], x32[20:])
    assert not np.array_equal(part[:20], x32[:20])


def test_euler_matrix_is_scipys_xyx():
    for e in ([0.3, 0.2, 0.1], [PI / 3, -1.1, 2.5], [0, 0, 0]):
        want = Rotation.from_euler("XYX", [e[2], -e[1], e[0]]).as_matrix()
        np.testing.assert_allclose(oracle.euler_matrix(e), want, rtol=0, atol=1e-15)


def _bodies(n=40, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    bn = rng.integers(0, 3, n).astype(np.uint32)      # 0 = no body, 1 = body 0, 2 = body 1
    return pos, vel, bn


def test_set_body_position_is_a_rigid_translation():
    pos, vel, bn = _bodies()
    stored = np.array([[0.5, 0.5, 0.5], [1.0, -2.0, 0.25]], np.float32)
    target = np.array([1.5, 0.25, -0.75], np.float32)
    dt = np.float32(1e-3)
    p2, v2, s2 = oracle.set_body_position(pos, vel, bn, 40, 2, dt, 1, stored, target)
    moved = bn == 2
    assert moved.any() and (~moved).any()
    assert np.array_equal(p2[~moved], pos[~moved]) and np.array_equal(v2[~moved], vel[~moved])
    d = (target.astype(np.float64) - stored[1])
    np.testing.assert_allclose(p2[moved], pos[moved].astype(np.float64) + d, rtol=0, atol=4e-7)
    np.testing.assert_allclose(v2[moved], np.broadcast_to(d / float(dt), v2[moved].shape), rtol=1e-6)
    a, b = p2[moved].astype(np.float64), pos[moved].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(a[:, None] - a[None], axis=-1),
                               np.linalg.norm(b[:, None] - b[None], axis=-1), atol=2e-6)
    assert np.array_equal(s2[1], target) and np.array_equal(s2[0], stored[0])
    # a body at or past body_count moves nothing, its position is still stored
    p3, v3, s3 = oracle.set_body_position(pos, vel, bn, 40, 1, dt, 1, stored, target)
    assert np.array_equal(p3, pos) and np.array_equal(v3, vel) and np.array_equal(s3[1], target)


def test_set_body_rotation_matches_numpy():
    pos, _, bn = _bodies(seed=2)
    n = len(pos)
    euler = np.array([[0.0, 0.0, 0.0], [0.4, -0.7, 1.3]], np.float64)
    new = np.array([0.1, 0.2, 0.3])
    B = Rotation.from_euler("XYX", [euler[1, 2], -euler[1, 1], euler[1, 0]]).as_matrix()
    tagged = bn == 2
    # identity particle rotations: x -> B^T x, rotation -> B^T
    eye = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    p2, r2, e2 = oracle.set_body_rotation(pos, eye, bn, n, 2, 1, euler, new, precision="f64")
    np.testing.assert_allclose(p2[tagged], pos[tagged] @ B, rtol=0, atol=1e-14)
    np.testing.assert_allclose(r2[tagged], np.tile(B.T.reshape(1, 9), (tagged.sum(), 1)), rtol=0, atol=1e-15)
    assert np.array_equal(p2[~tagged], pos[~tagged]) and np.array_equal(e2[1], new) and np.array_equal(e2[0], euler[0])
    # a non-symmetric rotation field
    rng = np.random.Generator(np.random.PCG64(4))
    P = rng.uniform(-1, 1, (n, 3, 3))
    p3, r3, _ = oracle.set_body_rotation(pos, P.reshape(n, 9), bn, n, 2, 1, euler, new, precision="f64")
    T = np.einsum("ab,nbc->nac", B.T, P)
    np.testing.assert_allclose(p3[tagged], np.einsum("nab,nb->na", T, pos.astype(np.float64))[tagged], atol=1e-13)
    np.testing.assert_allclose(r3[tagged].reshape(-1, 3, 3), np.einsum("nab,nbc->nac", T, P)[tagged], atol=1e-13)
    # body 0 of body_count 1 with zero stored angles: B = I, x unchanged; body 1 >= body_count moves nothing
    p4, _, _ = oracle.set_body_rotation(pos, eye, bn, n, 1, 1, euler, new, precision="f64")
    assert np.array_equal(p4, pos)
