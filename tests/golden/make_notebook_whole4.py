"""Write tests/golden/notebook_whole4.json: the particle states that the
reference's own notebook printed for its SVK strain-test run "whole_4"
(scripts/View generated data.ipynb, cell 1 output: particle 56 at history
frame 0, particles 57, 126 and 15 at frame 3850), with the run's parameters
(scripts/Strain solid run through (whole domain)-SVK.ipynb, cells 5 and 7,
loop index i = 4; material from conf/svksolid.conf:4-6,21-23).  The numbers
are copied from the notebook's stored output text, which this script reads.

    python tests/golden/make_notebook_whole4.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
NOTEBOOK = "/root/reference/scripts/View generated data.ipynb"


def parse(text):
    out = {}
    for block in re.split(r"(?=Particle \d+ at frame \d+:)", text):
        m = re.match(r"Particle (\d+) at frame (\d+):", block)
        if not m:
            continue
        rec = {}
        for name in ("position", "stress", "deformation", "force_elastic", "density", "force_viscous"):
            f = re.search(name + r":\s*(.*?)(?=\n\w+:|\Z)", block, re.S)
            nums = re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", f.group(1))
            vals = [float(v) for v in nums]
            rec[name] = vals[0] if name == "density" else vals
        out[f"{m.group(1)}@{m.group(2)}"] = rec
    return out


def main():
    nb = json.load(open(NOTEBOOK))
    texts = ["".join(o.get("text", "")) for o in nb["cells"][1].get("outputs", [])]
    states = parse("".join(t for t in texts if "Particle" in t))
    data = {
        "source": "scripts/View generated data.ipynb cell 1 output (run results/whole_4)",
        "run": {"harness": "StrainTestSymPadding", "particle_dimensions": [4, 4, 4],
                "initial_dimensions": [0.6, 0.6, 0.6], "total_mass": 6.0, "smoothingradius": 0.2,
                "strain_range": {"start": -0.5, "stop": 0.5, "num": 15}, "axis": 0, "strain_rate": 0.07,
                "stabilisation_time": 10.0,
                "material": {"timestep": 0.01, "viscosity": 0.0004, "youngs_modulus": 10.0, "poisson": 0.2,
                             "gravity": [0.0, 0.0, 0.0]}},
        "states": states,
    }
    path = os.path.join(HERE, "notebook_whole4.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print(path, sorted(states))


if __name__ == "__main__":
    main()
