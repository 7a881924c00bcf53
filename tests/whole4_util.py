"""The reference notebook's SVK strain test "whole_4" (tests/golden/
notebook_whole4.json): the harness of sph.straintest with the notebook's
parameters on any solid object (engine or oracle), stepped along the run
path for a given number of history frames."""
import json
import os

import numpy as np

from sph.straintest import StrainTestSymPadding

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "notebook_whole4.json")
FIELDS = ("position", "density", "stress", "deformation", "force_elastic", "force_viscous")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def run_frames(solid, frames, reference_szz=True):
    """Set the run's material on `solid`, build the harness and step it
    through history frames 0..frames-1; returns {field: array} of the last."""
    g = golden()["run"]
    mat = g["material"]
    solid.set_fields(timestep=mat["timestep"], viscosity=mat["viscosity"], youngs_modulus=mat["youngs_modulus"],
                     poisson=mat["poisson"], gravity=np.asarray(mat["gravity"]),
                     svk_sigma_zz_zero=int(reference_szz))
    i = g["particle_dimensions"][0]
    x, y, z = np.meshgrid(np.arange(i), np.arange(i), np.arange(i), indexing="ij")
    everyone = np.vstack((x.ravel(), y.ravel(), z.ravel())).T
    sr = g["strain_range"]
    test = StrainTestSymPadding(
        strain_range=np.linspace(sr["start"], sr["stop"], num=sr["num"]), axis=g["axis"],
        strain_rate=g["strain_rate"], stabilisation_time=g["stabilisation_time"],
        initial_dimensions=np.asarray(g["initial_dimensions"]), total_mass=g["total_mass"], strain_solid=solid,
        output_fields=["density"], history_particles=everyone[:1], history_fields=["position"],
        output_particles=everyone[:1], particle_dimensions=np.asarray(g["particle_dimensions"]),
        smoothingradius=g["smoothingradius"])
    plan = [q for section in test.run_path for strip in section for q in strip]
    solid.sync_host_to_device()
    solid.setup()
    for plane in plan[:frames]:
        test.set_plane_constraints(plane)
        solid.sync_fields_host_to_device("plane_constraints")
        solid.update()
        solid.call_for_all_particles("apply_plane_constraints")
    solid.sync_fields_device_to_host(*FIELDS)
    return {k: np.array(solid.get_fields(k)[0], copy=True) for k in FIELDS}


def compare_state(state, pid, want, tol, say=print):
    """Max deviation per field of particle pid from the notebook's values;
    asserts each within tol[field] (absolute; density relative)."""
    got = {
        "position": state["position"][pid].ravel(), "density": float(state["density"][pid]),
        "stress": state["stress"][pid].ravel(), "deformation": state["deformation"][pid].ravel(),
        "force_elastic": state["force_elastic"][pid].ravel(), "force_viscous": state["force_viscous"][pid].ravel(),
    }
    worst = {}
    for k, t in tol.items():
        if k == "density":
            d = abs(got[k] - want[k]) / abs(want[k])
        else:
            d = float(np.max(np.abs(np.asarray(got[k], np.float64) - np.asarray(want[k], np.float64))))
        worst[k] = d
        assert d <= t, f"particle {pid} {k}: got {got[k]}, notebook {want[k]} (deviation {d:.3e} > {t})"
    say(f"  particle {pid}: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    return worst
