"""An independent FP64 numpy restatement of the corotational solid passes,
O(N^2) over all particle pairs instead of the 27-cell grid walk, used only to
check the CPU oracle (tests/test_oracle_corotation_pins.py).  Test
infrastructure.

Written from kernels/solid_corotation.cl and the helpers it calls:
  applyKernel / applyKernelGradient / applyLapKernel   kernels/common.cl:78-90
  contributeApq                                        kernels/solids.cl:3-9
  addOuterProduct, multiply(Sym)MatrixVector           kernels/common.cl:221-256
  getR (R = Apq (Apq^T Apq)^-1/2)                      kernels/common.cl:306-331
  computeCauchyStrain, computeStress                   kernels/solid_corotation.cl:1-26
  compute_rotations_and_strains                        kernels/solid_corotation.cl:28-195
  compute_forces_solids                                kernels/solid_corotation.cl:215-403
No symmetry planes: the mirror images are pinned separately (doubled domain).
Matrices are 3x3 row-major, index [row, column] = I_rc; symmetric tensors are
stored as [00, 11, 22, 01, 02, 12] (kernels/common.cl:27-35).
"""
import math

import numpy as np

PI = 3.1415926535          # the reference's PI macro (kernels/common.cl:15)


def poly6(r, h):
    w = 315.0 / 64.0 / PI / h ** 9 * (h * h - r * r) ** 3
    return np.where(r > h, 0.0, w)


def grad_spiky(v, h):
    """applyKernelGradient(v): direction * -45/(pi h^6) (h - |v|)^2, zero past h."""
    r = np.linalg.norm(v, axis=-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        g = v / r[..., None] * (-45.0 / PI / h ** 6 * (h - r) ** 2)[..., None]
    return np.where((r > h)[..., None], 0.0, g)


def lap_visc(r, h):
    return np.where(r > h, 0.0, 45.0 / PI / h ** 6 * (h - r))


def sym_full(s):
    """[00, 11, 22, 01, 02, 12] -> full symmetric 3x3 (leading axes kept)."""
    s = np.asarray(s, np.float64)
    out = np.empty(s.shape[:-1] + (3, 3))
    for (r, c), k in {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4,
                      (1, 2): 5, (2, 1): 5}.items():
        out[..., r, c] = s[..., k]
    return out


def polar_rotation(apq):
    """R = Apq (Apq^T Apq)^(-1/2) for each 3x3 (getR)."""
    s = np.einsum("...ki,...kj->...ij", apq, apq)
    w, p = np.linalg.eigh(s)
    inv_sqrt = np.einsum("...ik,...k,...jk->...ij", p, 1.0 / np.sqrt(w), p)
    return apq @ inv_sqrt


def rotations_strains(pos, opos, density0, mass, h, body_number=None, rotation_in=None, rotation_used=None,
                      candidates=None):
    """(apq, rotation, strain) of compute_rotations_and_strains.

    Body particles (body_number > 0) return early: apq 0, their rotation and
    strain untouched (rotation_in kept, strain NaN marks "not written").
    rotation_used, when given, replaces getR's result in the strain sweep, so
    that sweep 2 can be checked on its own against another polar solver.
    candidates: the visited pairs (candidate_pairs); all pairs when None."""
    pos, opos = np.asarray(pos, np.float64), np.asarray(opos, np.float64)
    n = len(pos)
    p = pos[None, :, :] - pos[:, None, :]            # [i, j] = x_j - x_i
    q = opos[None, :, :] - opos[:, None, :]          # x0_j - x0_i
    off = ~np.eye(n, dtype=bool)                     # j != i
    if candidates is not None:
        off &= candidates
    w = poly6(np.linalg.norm(p, axis=-1), h) * off
    apq = np.einsum("ij,ija,ijb->iab", w, p, q)      # sum W(|p|) p (x) q
    rot = polar_rotation(apq)
    rot_s = rot if rotation_used is None else np.asarray(rotation_used, np.float64).reshape(n, 3, 3)
    # u = R^T p - q; sum_{j != i} (u m / rho0_i) (x) gradW(-q)
    u = np.einsum("iba,ijb->ija", rot_s, p) - q
    g = grad_spiky(-q, h) * off[..., None]
    g = np.where(off[..., None], g, 0.0)
    du = np.einsum("ija,ijb->iab", u * (mass / np.asarray(density0, np.float64))[:, None, None], g)
    strain = np.stack([du[:, 0, 0], du[:, 1, 1], du[:, 2, 2], 0.5 * (du[:, 0, 1] + du[:, 1, 0]),
                       0.5 * (du[:, 0, 2] + du[:, 2, 0]), 0.5 * (du[:, 1, 2] + du[:, 2, 1])], 1)
    if body_number is not None:
        body = np.asarray(body_number) > 0
        apq[body] = 0.0
        if rotation_in is not None:
            rot[body] = np.asarray(rotation_in, np.float64).reshape(n, 3, 3)[body]
        strain[body] = np.nan
    return apq, rot, strain


def stresses(strain, youngs, poisson):
    lam = youngs * poisson / (1.0 + poisson) / (1.0 - 2.0 * poisson)
    mu = youngs / 2.0 / (1.0 + poisson)
    e = np.asarray(strain, np.float64)
    tr = e[:, :3].sum(1)
    return np.concatenate([lam * tr[:, None] + 2.0 * mu * e[:, :3], 2.0 * mu * e[:, 3:]], 1)


def forces(pos, vel, opos, stress, rotation, density0, density, mass, h, viscosity, body_number=None,
           candidates=None):
    """(force, force_elastic, force_viscous) of compute_forces_solids with a
    rotation field (corotational) or rotation=None (the SVK variant,
    kernels/solid_svk.cl:182-367, which applies no rotations)."""
    pos, vel, opos = (np.asarray(a, np.float64) for a in (pos, vel, opos))
    n = len(pos)
    d0 = np.asarray(density0, np.float64)
    rho = np.asarray(density, np.float64)
    sig = sym_full(stress)
    off = ~np.eye(n, dtype=bool)
    if candidates is not None:
        off &= candidates
    X = opos[:, None, :] - opos[None, :, :]                   # [i, j] = x0_i - x0_j
    vivj = mass * mass / d0[:, None] / d0[None, :]
    gx = np.where(off[..., None], grad_spiky(X, h), 0.0)       # gradW(x)
    gmx = np.where(off[..., None], grad_spiky(-X, h), 0.0)     # gradW(-x)
    f_ji = -vivj[..., None] * np.einsum("iab,ijb->ija", sig, gx)
    f_ij = -vivj[..., None] * np.einsum("jab,ijb->ija", sig, gmx)
    if body_number is not None:
        bj = (np.asarray(body_number) > 0)[None, :, None]
        f_ij = np.where(bj, -f_ji, f_ij)
    if rotation is None:
        fe = -f_ji + f_ij
    else:
        R = np.asarray(rotation, np.float64).reshape(n, 3, 3)
        fe = -np.einsum("iab,ijb->ija", R, f_ji) + np.einsum("jab,ijb->ija", R, f_ij)
    rcur = np.linalg.norm(pos[None, :, :] - pos[:, None, :], axis=-1)
    fv = (mass / rho)[None, :, None] * (vel[None, :, :] - vel[:, None, :]) * lap_visc(rcur, h)[..., None]
    fv = np.where(off[..., None], fv, 0.0)
    f_e = 0.5 * fe.sum(1)
    f_v = viscosity * fv.sum(1)
    return f_e + f_v, f_e, f_v


def candidate_pairs(pos, gridbounds, h):
    """[i, j]: j lies in one of the 27 grid cells around i's cell, the set the
    reference visits (FOR_PARTICLES_IN_RANGE, kernels/common.cl:49-68, over
    the cells of find_particle_bins, :372-391, with res = |hi - lo| / h)."""
    pos = np.asarray(pos, np.float64)
    gb = np.asarray(gridbounds, np.float64).reshape(3, 2)
    res = np.array([int(abs(float(np.float32(hi) - np.float32(lo))) / float(np.float32(h))) for lo, hi in gb])
    cell = np.floor((pos - gb[:, 0]) / (gb[:, 1] - gb[:, 0]) * res).astype(np.int64)
    cell = np.clip(cell, 0, res - 1)
    return (np.abs(cell[:, None, :] - cell[None, :, :]) <= 1).all(-1)


def rotation_matrix(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * k + (1 - math.cos(angle)) * (k @ k)
