"""Algorithmic traffic and work of each engine kernel, for roofline figures.

Bytes rule: a pass reads every per-particle input field of particle i once
and writes every output field once; neighbour-j reads count zero (they are
served from L2/LDS); a per-cell array counts once per pass that touches it.
Flops rule: 8 flops per candidate pair (difference, squared distance), plus
a per-in-range-pair cost for the pass's pair arithmetic.

`n` is the number of active particles and `cells` the number of grid cells.
"""

MI355X_HBM_GBS = 8000.0        # HBM3E, GB/s
MI355X_FP32_TFLOPS = 157.3     # dense FP32 (vector), TFLOP/s

# bytes per particle, bytes per cell, flops per candidate, flops per in-range pair
_FLUID_KERNELS = {
    "bin_key_count": (20, 8, 0, 0),
    "bin_scan": (0, 8, 0, 0),
    "bin_scatter": (28, 0, 0, 0),
    "bin_rank_place": (84, 0, 0, 0),
    "compute_density": (24, 8, 8, 4),
    "compute_forces_fluids": (44, 8, 8, 31),
    "step_forward": (72, 0, 0, 0),
}
# Passes that run two of the above as one kernel (not counted again in the
# step totals): the force pass with the integrator fused in reads x, v and
# density and writes force, acceleration, x and v.
_FUSED_KERNELS = {
    "compute_forces_fluids+step_forward": (76, 8, 8, 31),
}


# Batch passes: (systems the pass runs for, bytes per active particle, bytes
# per grid cell).  Every pass also reads each slot's scene entry (4 B), not
# counted, as padding slots and scene tables are overhead of the batch.
# The batch writes only `force` from the solid force passes (no exported
# components), and gathers the persistent state of each system in binning.
BATCH_PASSES = {
    "batch_key_count": (("fluid", "corotational", "svk"), 20, 8),
    "batch_scan": ((), 0, 8),
    "batch_scatter": (("fluid", "corotational", "svk"), 28, 0),
    "batch_rank": (("fluid", "corotational", "svk"), 20, 0),
    "compute_density_fluid": (("fluid",), 24, 8),
    "compute_forces_fluids": (("fluid",), 44, 8),
    "compute_density_solid": (("corotational", "svk"), 16, 8),
    "compute_rotations_and_strains": (("corotational",), 88, 8),
    "compute_stresses_corotational": (("corotational",), 48, 0),
    "compute_forces_solids_corotational": (("corotational",), 120, 8),
    "compute_deformations": (("svk",), 60, 8),
    "compute_stresses_svk": (("svk",), 60, 0),
    "compute_forces_solids_svk": (("svk",), 84, 8),
    "step_forward": (("fluid", "corotational", "svk"), 72, 0),
}
# Batch passes: (systems the pass runs for, flops per candidate pair, flops
# per in-range pair (self included), flops per in-range pair when Kahan
# density sums are on).  The neighbour passes are bound by vector issue:
# their fraction is taken against the FP32 vector peak.  Candidate pairs are
# those of the 27 cells around each particle's cell; the pair arithmetic per
# in-range pair: density 4 (+4 compensated), fluid force 31, rotations and
# strains 24 (A_pq) + 52 (displacement gradient) over two candidate sweeps,
# SVK deformations 24 (A_pq) + 15 (A_qq), solid forces 110.  The FP64 polar
# decomposition / A_qq inverse per particle is not counted.
BATCH_FLOPS = {
    "compute_density_fluid": (("fluid",), 8, 4, 8),
    "compute_forces_fluids": (("fluid",), 8, 31, 31),
    "compute_density_solid": (("corotational", "svk"), 8, 4, 8),
    "compute_rotations_and_strains": (("corotational",), 16, 76, 76),
    "compute_deformations": (("svk",), 8, 39, 39),
    "compute_forces_solids_corotational": (("corotational",), 8, 110, 110),
    "compute_forces_solids_svk": (("svk",), 8, 110, 110),
}


def batch_flops(name, pair_steps_by_system, kahan_density):
    """Algorithmic flops of batch pass `name` over a run, from each system's
    candidate and in-range pair-steps {system: (candidates, in_range)}."""
    m = BATCH_FLOPS.get(name)
    if m is None:
        return None
    systems, per_cand, per_pair, per_pair_kahan = m
    total = 0.0
    for k in systems:
        cand, pairs = pair_steps_by_system.get(k, (0, 0))
        total += per_cand * cand + (per_pair_kahan if kahan_density else per_pair) * pairs
    return total


# gathered words per particle in binning: pos, vel, id (+ opos, density0,
# body; + rotation, strain for corotational), read and written, + perm
GATHER_BYTES = {"fluid": 2 * 7 * 4 + 4, "corotational": 2 * 27 * 4 + 4, "svk": 2 * 12 * 4 + 4}


def batch_bytes_per_launch(name, particle_steps_by_system, launches, cell_steps=0):
    """Average algorithmic bytes of one launch of batch pass `name`, from the
    particle-steps each system ran while the pass was timed."""
    if name == "batch_gather":
        total = sum(GATHER_BYTES[k] * v for k, v in particle_steps_by_system.items())
        return total / launches if launches else None
    m = BATCH_PASSES.get(name)
    if m is None or not launches:
        return None
    systems, per_particle, per_cell = m
    total = per_particle * sum(particle_steps_by_system.get(k, 0) for k in systems) + per_cell * cell_steps
    return total / launches


def kernel_model(name):
    return _FLUID_KERNELS.get(name) or _FUSED_KERNELS.get(name)


def bytes_per_launch(name, n, cells):
    m = kernel_model(name)
    if m is None:
        return None
    return m[0] * n + m[1] * cells


def flops_per_launch(name, candidates, in_range):
    m = kernel_model(name)
    if m is None:
        return None
    return m[2] * candidates + m[3] * in_range


def step_bytes(n, cells):
    return sum(bytes_per_launch(k, n, cells) for k in _FLUID_KERNELS)


def t_roof_seconds(n, cells, candidates, in_range):
    """Sum over passes of max(bytes / HBM peak, flops / FP32 peak)."""
    t = 0.0
    for k in _FLUID_KERNELS:
        b = bytes_per_launch(k, n, cells)
        f = flops_per_launch(k, candidates, in_range)
        t += max(b / (MI355X_HBM_GBS * 1e9), f / (MI355X_FP32_TFLOPS * 1e12))
    return t
