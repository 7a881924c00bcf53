"""Error of every engine output against the FP64 oracle.

    python tests/precision_report.py [--scenes dam_break_16k,cantilever_16k,svk_cube_262k] [--out report.json]

For each scene, the same inputs run through the FP64 oracle (the reference's
algorithm with every operation in double) and through the GPU engine under
each precision setting:

    fp32            plain FP32 neighbour sums, FP32 polar decomposition / inverse
    fp32+polar64    plain FP32 sums, FP64 polar decomposition / inverse
    kahan+polar64   Kahan-compensated density and force sums, FP64 polar

after one step (the passes' own arithmetic: both sides see identical FP32
inputs) and after the scene's steps (including the drift of the FP32
state), the relative error of each output field is reported: normwise
||gpu - ref||_2 / ||ref||_2 and max_i |gpu_i - ref_i| / max_i |ref_i|.  For
position the error is taken on the displacement from the initial position
(the quantity the step changes).  Global quantities: total momentum and
kinetic energy and centre of mass (FP64 reductions on the device), the
cantilever's tip deflection (mean z of the free end) and the SVK strain
energy sum_i V_i 1/2 S:E from the deformation gradients.

Test infrastructure: it runs the oracle as the checker, never the engine.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "opencl-sph_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, HERE)

SETTINGS = {"fp32": (0, False), "fp32+polar64": (0, True), "kahan+polar64": (3, True)}
FIELDS = {"fluid": ("position", "velocity", "density", "force"),
          "corotational": ("position", "velocity", "density", "force", "stress", "rotation", "strain"),
          "svk": ("position", "velocity", "density", "force", "stress", "deformation")}


def rel_errors(got, ref):
    got = np.asarray(got, np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, np.float64).reshape(len(ref), -1)
    d = got - ref
    nref = np.linalg.norm(ref)
    mref = np.max(np.abs(ref)) if ref.size else 0.0
    return {"normwise": float(np.linalg.norm(d) / nref) if nref > 0 else float(np.linalg.norm(d)),
            "max": float(np.max(np.abs(d)) / mref) if mref > 0 else float(np.max(np.abs(d)))}


def strain_energy(deformation, youngs, poisson, volume):
    """sum_i V_i 1/2 S:E with E = (F^T F - I)/2, S = lambda tr(E) I + 2 mu E, in FP64."""
    F = np.asarray(deformation, np.float64).reshape(-1, 3, 3)
    E = 0.5 * (np.einsum("nki,nkj->nij", F, F) - np.eye(3))
    lam = youngs * poisson / ((1 + poisson) * (1 - 2 * poisson))
    mu = youngs / (2 * (1 + poisson))
    trE = np.trace(E, axis1=1, axis2=2)
    S = lam * trE[:, None, None] * np.eye(3) + 2 * mu * E
    return float(np.sum(volume * 0.5 * np.einsum("nij,nij->n", S, E)))


def global_sums(pos, vel, mass):
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    m = float(mass)
    return {"momentum": (m * vel.sum(0)).tolist(), "kinetic_energy": float(0.5 * m * (vel * vel).sum()),
            "centre_of_mass": pos.mean(0).tolist()}


def tip_deflection(pos, pos0):
    x0 = np.asarray(pos0, np.float64)
    tip = x0[:, 0] >= x0[:, 0].max() - 1e-6
    return float(np.asarray(pos, np.float64)[tip, 2].mean() - x0[tip, 2].mean())


def run_scene(scene, steps, log):
    import gpu_util
    out = {"scene": scene.name, "system": scene.system, "particles": scene.n, "steps": steps, "settings": {}}
    t = time.perf_counter()
    ref = gpu_util.oracle_state(scene, "f64", polar_fp64=True)
    ref.step(scene.system, steps)
    log(f"{scene.name}: FP64 oracle {steps} steps in {time.perf_counter() - t:.1f} s")
    pos0 = np.asarray(scene.fields["position"], np.float64)
    refs = {k: ref.f[k] for k in FIELDS[scene.system]}
    refs["position"] = refs["position"] - pos0
    out["reference_globals"] = global_sums(ref.f["position"], ref.f["velocity"], scene.params["mass"])
    if scene.system == "corotational":
        out["reference_tip_deflection"] = tip_deflection(ref.f["position"], pos0)
    if scene.system == "svk":
        vol = scene.params["mass"] / 1100.0
        out["reference_strain_energy"] = strain_energy(ref.f["deformation"], scene.params["youngs"],
                                                       scene.params["poisson"], vol)
    for name, (comp, polar) in SETTINGS.items():
        if scene.system == "fluid" and name == "fp32+polar64":
            continue                        # no polar decomposition in a fluid
        s = gpu_util.engine(scene, compensated=comp, polar_fp64=polar)
        s.run(steps)
        got = gpu_util.download(s, *FIELDS[scene.system])
        sums = s.psdata_opencl.global_sums()
        s.free()
        got = {k: v[:scene.n] for k, v in got.items()}
        got["position"] = got["position"].astype(np.float64) - pos0
        entry = {k: rel_errors(got[k], refs[k]) for k in FIELDS[scene.system]}
        g_ref = out["reference_globals"]
        entry["kinetic_energy"] = abs(sums[1] - g_ref["kinetic_energy"]) / max(abs(g_ref["kinetic_energy"]), 1e-300)
        entry["momentum_abs"] = float(np.linalg.norm(np.array(sums[0]) - np.array(g_ref["momentum"])))
        entry["centre_of_mass_abs"] = float(np.linalg.norm(np.array(sums[2]) - np.array(g_ref["centre_of_mass"])))
        if scene.system == "corotational":
            tip = tip_deflection(got["position"] + pos0, pos0)
            entry["tip_deflection"] = {"value": tip,
                                       "rel_error": abs(tip - out["reference_tip_deflection"])
                                       / abs(out["reference_tip_deflection"])}
        if scene.system == "svk":
            e = strain_energy(got["deformation"], scene.params["youngs"], scene.params["poisson"],
                              scene.params["mass"] / 1100.0)
            entry["strain_energy"] = {"value": e, "rel_error": abs(e - out["reference_strain_energy"])
                                      / abs(out["reference_strain_energy"])}
        out["settings"][name] = entry
        log(f"  {name}: " + ", ".join(f"{k} {v['normwise']:.2e}" for k, v in entry.items()
                                      if isinstance(v, dict) and "normwise" in v))
    return out


def run_one_step(scene, log):
    """Errors of the first step's passes (density, forces, stresses), which
    see identical FP32 inputs on both sides: the arithmetic error of the
    passes themselves, without the drift of the FP32 state."""
    import gpu_util
    ref = gpu_util.oracle_state(scene, "f64", polar_fp64=True)
    ref.step(scene.system, 1)
    fields = [k for k in ("density", "force", "stress", "rotation", "deformation") if k in FIELDS[scene.system]]
    out = {}
    for name, (comp, polar) in SETTINGS.items():
        if scene.system == "fluid" and name == "fp32+polar64":
            continue
        s = gpu_util.engine(scene, compensated=comp, polar_fp64=polar)
        s.run(1)
        got = gpu_util.download(s, *fields)
        s.free()
        out[name] = {k: rel_errors(got[k][:scene.n], ref.f[k]) for k in fields}
        if scene.system != "fluid":
            # A solid at rest has stress and elastic force ~0 in FP64, so
            # these two are measured on the material's scale instead:
            # stress / E (strain units) and force / (E V^(2/3)).
            E = float(scene.params["youngs"])
            vol = float(scene.params["mass"]) / (1000.0 if scene.system == "corotational" else 1100.0)
            for k, scale in (("stress", E), ("force", E * vol ** (2.0 / 3.0))):
                d = np.asarray(got[k][:scene.n], np.float64).reshape(scene.n, -1) - ref.f[k].reshape(scene.n, -1)
                out[name][k] = {"normwise": float(np.sqrt(np.mean(d * d)) / scale),
                                "max": float(np.max(np.abs(d)) / scale), "scale": k + " / " +
                                ("E" if k == "stress" else "(E V^(2/3))")}
        log(f"  1 step, {name}: " + ", ".join(f"{k} {v['normwise']:.3e}" for k, v in out[name].items()))
    return out


def main():
    from sph import bindings, scenes
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", default="dam_break_16k,cantilever_16k,svk_cube_262k")
    ap.add_argument("--steps", default="", help="comma list of step counts overriding each scene's own")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    bindings.set_log_level(3)
    import sph
    sph.init_hip()
    index = {"dam_break_16k": 0, "cantilever_16k": 1, "svk_cube_262k": 2, "dam_break_4m": 3}
    names = args.scenes.split(",")
    steps = [int(s) for s in args.steps.split(",")] if args.steps else [None] * len(names)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    report = []
    for name, n in zip(names, steps):
        scene = scenes.baseline_config(index[name])
        entry = run_scene(scene, n or scene.steps, log)
        entry["one_step"] = run_one_step(scene, log)
        report.append(entry)
    text = json.dumps(report, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
