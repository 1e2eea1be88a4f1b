"""Copies the reference's homogeneous uniaxial run (one Hex8, J2 with linear +
Voce hardening, finite logarithmic kinematics, 13 steps of 20 increments:
pulled to 65 % stretch, then pushed back to zero) into homogeneous_uniaxial.json.

Source (data written by the reference's own sparlab_form with "%.17g",
src/io/FormingWriter.cpp:15):
  benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab/deck.json
  benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab/out/step_*_s*_{nodes,elements}.csv
The numbers are kept as the files hold them (parsed with float(), which
round-trips %.17g exactly). Run from the repository root, where
/root/reference is mounted: python tests/golden/copy_homogeneous_uniaxial.py
"""
import csv
import glob
import json
import os
import re

SRC = "/root/reference/benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "homogeneous_uniaxial.json")


def rows(path):
    with open(path) as f:
        r = csv.reader(f)
        head = next(r)
        return head, [[float(v) for v in line] for line in r]


def main():
    deck = json.load(open(os.path.join(SRC, "deck.json")))
    steps = []
    for path in glob.glob(os.path.join(SRC, "out", "step_*_elements.csv")):
        k = int(re.search(r"step_(\d+)_", os.path.basename(path)).group(1))
        eh, erows = rows(path)
        nh, nrows = rows(path.replace("_elements.csv", "_nodes.csv"))
        assert eh == ["element", "eq_plastic_strain", "von_mises_Pa"], eh
        assert nh == ["node", "X", "Y", "Z", "ux", "uy", "uz"], nh
        steps.append(dict(step=k, source=os.path.basename(path).replace("_elements.csv", ""),
                          nodes=nrows, elements=erows))
    steps.sort(key=lambda s: s["step"])
    out = dict(description="the reference's homogeneous uniaxial run: one Hex8, J2 with linear + Voce hardening, "
                           "finite logarithmic kinematics, 13 steps of 20 increments",
               source="benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab",
               material=deck["material"], kinematics=deck["forming"]["kinematics"], mesh=deck["mesh"],
               node_columns=["node", "X", "Y", "Z", "ux", "uy", "uz"],
               element_columns=["element", "eq_plastic_strain", "von_mises_Pa"], steps=steps)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {len(steps)} steps")


if __name__ == "__main__":
    main()
