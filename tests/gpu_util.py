"""Shared helpers of the GPU parity tests: the same Scene loaded into the HIP
engine and into the CPU oracle, and comparison reports."""
import sys

import numpy as np

import oracle
from sph import scenes

PARTICLE_FIELDS = ("position", "velocity", "force", "acceleration", "density", "density0", "rotation", "strain",
                   "stress", "deformation", "force_elastic", "force_viscous", "cutoff_force", "originalpos")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine(scene, setup=True, compensated=False, polar_fp64=True):
    s = scenes.build(scene, setup=False)
    s.psdata_opencl.set_precision(compensated, polar_fp64)
    if setup and scene.system in ("corotational", "svk"):
        s.setup()
    return s


def oracle_state(scene, precision="f32", polar_fp64=True, setup=True, **overrides):
    params = dict(scene.params)
    params["polar_fp64"] = int(polar_fp64)
    params.update(overrides.pop("params", {}))
    fields = {k: v for k, v in scene.fields.items() if k in oracle.OracleState.PER_PARTICLE}
    for k in ("body_number", "body_dynamic", "plane_constraints_particles"):
        if k in scene.fields:
            fields[k] = scene.fields[k]
    fields.update(overrides)
    st = oracle.OracleState(precision, params, **fields)
    if setup and scene.system in ("corotational", "svk"):
        st.setup_solid()
    return st


def download(system, *names):
    system.sync_fields_device_to_host(*names)
    out = {}
    for name in names:
        (view,) = system.get_fields(name)
        out[name] = np.array(view, copy=True)
    return out


def compare(name, got, want, rtol=0.0, atol=0.0):
    """Max abs/rel difference and mismatch count; asserts within tolerance.

    Equal means the same bit pattern (so -0.0 differs from +0.0); any NaN
    equals any NaN, whose payload bits are not specified."""
    got = np.asarray(got).reshape(np.asarray(want).shape)
    want = np.asarray(want)
    if got.dtype.kind == "f" and got.dtype == want.dtype:
        bits = np.dtype(f"u{got.dtype.itemsize}")
        same = (np.ascontiguousarray(got).view(bits) == np.ascontiguousarray(want).view(bits))
        same |= np.isnan(got) & np.isnan(want)
    elif got.dtype.kind == "f":
        same = (got == want) | (np.isnan(got) & np.isnan(want))
    else:
        same = got == want
    mism = int((~same).sum())
    diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
    maxabs = float(np.nanmax(diff)) if diff.size else 0.0
    scale = np.abs(want.astype(np.float64))
    tol = atol + rtol * scale
    bad = ~same if rtol == 0.0 and atol == 0.0 else ~same & ~(diff <= tol)
    say(f"  {name}: {mism}/{want.size} differ bitwise, max |diff| {maxabs:.3e}, outside tolerance {int(bad.sum())}")
    assert not bad.any(), f"{name}: {int(bad.sum())} entries outside rtol={rtol} atol={atol} (max |diff| {maxabs})"
    return mism
