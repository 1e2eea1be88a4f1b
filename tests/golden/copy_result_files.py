"""Takes the data of the reference's own result directory of the kin_h0833
deck (benchmarks/physics_audit/challenge/cases/kin_h0833/output, written by
sparlab_form: src/io/FormingWriter.cpp, src/io/CsvWriter.cpp) into
tests/golden/results/:

  kin_h0833.npz   the values of its tables, as read back from the text:
                  u_<stem> (4802, 3) displacements of the node table,
                  elements_<stem> (2304, 2) eq_plastic_strain, von_mises_Pa,
                  for the steps form and release; tool_forces (395, 11) the
                  numeric columns of tool_forces.csv (step, increment, t, cx,
                  cy, cz, fx, fy, fz, active_nodes, max_penetration_m)
  kin_h0833.json  each table's SHA-256 and size in bytes, the tool column of
                  tool_forces.csv, and summary.json's steps (name, type,
                  completed, files_stem)

The reference coordinates of the node tables are not copied: they are the
deck's mesh (tests/golden/decks/kin_h0833.json).

Run from the repository root where /root/reference is mounted:
python tests/golden/copy_result_files.py
"""
import csv
import hashlib
import json
import os

import numpy as np

SRC = "/root/reference/benchmarks/physics_audit/challenge/cases/kin_h0833/output"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "results")
STEMS = ("step_1_form", "step_3_release")


def rows(path):
    with open(path, newline="") as f:
        r = list(csv.reader(f))
    return r[0], r[1:]


def main():
    os.makedirs(OUT, exist_ok=True)
    arrays, digests = {}, {}
    for stem in STEMS:
        for kind in ("nodes", "elements"):
            name = f"{stem}_{kind}.csv"
            path = os.path.join(SRC, name)
            data = open(path, "rb").read()
            digests[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
            header, body = rows(path)
            if kind == "nodes":
                assert header == ["node", "X", "Y", "Z", "ux", "uy", "uz"]
                assert [int(r[0]) for r in body] == list(range(len(body)))
                arrays[f"u_{stem}"] = np.array([[float(x) for x in r[4:7]] for r in body])
            else:
                assert header == ["element", "eq_plastic_strain", "von_mises_Pa"]
                assert [int(r[0]) for r in body] == list(range(len(body)))
                arrays[f"elements_{stem}"] = np.array([[float(x) for x in r[1:3]] for r in body])
    path = os.path.join(SRC, "tool_forces.csv")
    data = open(path, "rb").read()
    digests["tool_forces.csv"] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    header, body = rows(path)
    assert header == ["step", "increment", "t", "tool", "cx", "cy", "cz", "fx", "fy", "fz", "active_nodes",
                      "max_penetration_m"]
    arrays["tool_forces"] = np.array([[float(x) for i, x in enumerate(r) if i != 3] for r in body])
    tools = sorted({r[3] for r in body})
    with open(os.path.join(SRC, "summary.json")) as f:
        summary = json.load(f)
    steps = [{k: s[k] for k in ("name", "type", "completed", "files_stem")} for s in summary["steps"]]
    np.savez_compressed(os.path.join(OUT, "kin_h0833.npz"), **arrays)
    with open(os.path.join(OUT, "kin_h0833.json"), "w") as f:
        json.dump({"digests": digests, "tools": tools, "steps": steps}, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
