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


def kernel_model(name):
    return _FLUID_KERNELS.get(name)


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
