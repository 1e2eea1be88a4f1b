"""Result files of a forming run in sparlab_form's output contract.

The precomp package reads a sparlab_form run through its result directory
(python/precomp/fea/results.py:320-365, load_result): summary.json's `steps`
(name, completed, files_stem), and per completed step `<files_stem>_nodes.csv`
(node, X, Y, Z, ux, uy, uz: reference coordinates and displacement) and
`<files_stem>_elements.csv` (element, eq_plastic_strain, von_mises_Pa), with
tool_forces.csv beside them. The writers here produce the same files, byte for
byte, from the device path's outputs, so that a run of the drop-in or of the
batched form (one directory per problem) is loaded as a sparlab_form run is.

Reference: src/io/FormingWriter.cpp:13-17 (every number with 17 significant
digits, `%.17g`), :29-39 (the node table), :56-65 (forming_step_stem), :67-99
(a step's tables and its increment snapshots), :102-121 (tool_forces.csv),
:123-196 (the summary's fields); src/io/CsvWriter.cpp:59-98 (the header, the
integer index column, rows ended by a newline); src/fem/Forming.cpp:1255-1281
(the element table's values: the von Mises stress of elastoplastic_stress, and
its largest equivalent plastic strain for an elastoplastic element, 0 for an
elastic one). VTK files (write_vtk, FormingWriter.cpp:41-54) are not written.
"""
import json
import os

import numpy as np

NODE_COLUMNS = ("node", "X", "Y", "Z", "ux", "uy", "uz")
ELEMENT_COLUMNS = ("element", "eq_plastic_strain", "von_mises_Pa")
FORCE_COLUMNS = ("step", "increment", "t", "tool", "cx", "cy", "cz", "fx", "fy", "fz", "active_nodes",
                 "max_penetration_m")
_STEM_BYTES = frozenset(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789-_")


def number(v):
    """A number as the reference writes it: 17 significant digits, `%g`
    notation (FormingWriter.cpp:13-17; CsvWriter's stream precision 17)."""
    return "%.17g" % float(v)


def forming_step_stem(index, name):
    """`step_<index + 1>_<s>`, s the step name with every byte but ASCII
    letters, digits, `-` and `_` replaced by `_`, "step" if empty
    (FormingWriter.cpp:56-65)."""
    clean = "".join(chr(b) if b in _STEM_BYTES else "_" for b in str(name).encode("utf-8"))
    return f"step_{int(index) + 1}_{clean or 'step'}"


def _table(header, columns, index=None):
    """CSV text: the header, then per row the integer index (0, 1, ... unless
    given) and the columns' numbers (CsvWriter::row)."""
    cols = [np.asarray(c, dtype=np.float64).reshape(-1) for c in columns]
    n = cols[0].shape[0] if cols else 0
    if any(c.shape[0] != n for c in cols):
        raise ValueError("columns of different lengths")
    idx = range(n) if index is None else [int(i) for i in index]
    fmt = ",".join(["%d"] + ["%.17g"] * len(cols))
    lines = [",".join(header)]
    lines.extend(fmt % row for row in zip(idx, *(c.tolist() for c in cols)))
    return "\n".join(lines) + "\n"


def node_table(coords, u):
    """The node table of a step (write_nodes, FormingWriter.cpp:29-39): the
    reference coordinates (n, 3) and the displacement (n, 3 or 3 n)."""
    x = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    d = np.asarray(u, dtype=np.float64).reshape(-1, 3)
    if d.shape != x.shape:
        raise ValueError(f"displacement of {d.shape[0]} nodes for {x.shape[0]} nodes")
    return _table(NODE_COLUMNS, [x[:, 0], x[:, 1], x[:, 2], d[:, 0], d[:, 1], d[:, 2]])


def element_table(plastic_strain, von_mises):
    """The element table of a step (FormingWriter.cpp:76-84)."""
    return _table(ELEMENT_COLUMNS, [plastic_strain, von_mises])


def element_values(stress, elastoplastic):
    """(eq_plastic_strain, von_mises_Pa) per element from the stress recovery
    of every element (NonlinearSystem.stress: von_mises and
    max_equivalent_plastic_strain), as the forming driver fills them
    (Forming.cpp:1263-1279): the plastic strain of an elastic element is 0."""
    vm = np.asarray(stress["von_mises"], dtype=np.float64).reshape(-1)
    if elastoplastic:
        eps = np.asarray(stress["max_equivalent_plastic_strain"], dtype=np.float64).reshape(-1)
    else:
        eps = np.zeros_like(vm)
    return eps, vm


def tool_force_table(records):
    """tool_forces.csv (FormingWriter.cpp:102-121): per record (step from 1,
    increment, t, tool name, centre (3), force on the tool (3), active nodes,
    largest penetration)."""
    lines = [",".join(FORCE_COLUMNS)]
    for r in records:
        step, inc, t, tool, c, f, active, pen = r
        c = [float(v) for v in c]
        f = [float(v) for v in f]
        lines.append(",".join([str(int(step)), str(int(inc)), number(t), str(tool), *map(number, c),
                               *map(number, f), str(int(active)), number(pen)]))
    return "\n".join(lines) + "\n"


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)


class ResultDirectory:
    """A forming run's result directory in the reference's output contract.

    coords (n, 3) and conn (e, 8): the mesh. add_step writes a step's node
    table (and its element table when given the element values) at once;
    add_tool_record collects tool_forces.csv's rows; close writes
    tool_forces.csv and summary.json (case, version, completed, termination,
    steps with name, type, completed, files_stem, max_displacement_m and the
    caller's fields, mesh, warnings, files). Only the files written here are
    listed; what the reference adds beyond the loader's contract (timing,
    analysis, tools) is taken from `extra` when the caller has it.
    """

    def __init__(self, directory, coords, conn, case="", version=None, rule_points=None, mesh_source=""):
        self.directory = str(directory)
        os.makedirs(self.directory, exist_ok=True)
        self.coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
        self.num_elements = int(np.asarray(conn).reshape(-1, 8).shape[0])
        self.case = case
        self.version = version
        self.rule_points = rule_points
        self.mesh_source = mesh_source
        self.steps = []
        self.files = []
        self.forces = []

    def path(self, name):
        return os.path.join(self.directory, name)

    def add_step(self, name, u, type="form", completed=True, plastic_strain=None, von_mises=None,
                 snapshots=(), **fields):
        """Writes step len(steps)'s tables (none for a step that did not
        complete, as the reference) and records its summary entry. snapshots:
        (increment, displacement) pairs, written as
        `<stem>_inc_<increment>_nodes.csv`."""
        k = len(self.steps)
        stem = forming_step_stem(k, name)
        d = np.asarray(u, dtype=np.float64).reshape(-1, 3)
        entry = {"name": str(name), "type": str(type), "completed": bool(completed)}
        entry.update(fields)
        entry["files_stem"] = stem
        entry["max_displacement_m"] = float(np.sqrt((d * d).sum(axis=1)).max()) if d.size else 0.0
        self.steps.append(entry)
        if not completed:
            return []
        written = [stem + "_nodes.csv"]
        _write(self.path(written[0]), node_table(self.coords, d))
        if von_mises is not None:
            eps = np.zeros(self.num_elements) if plastic_strain is None else plastic_strain
            if np.asarray(von_mises).size != self.num_elements or np.asarray(eps).size != self.num_elements:
                raise ValueError("element values of another mesh")
            written.append(stem + "_elements.csv")
            _write(self.path(written[-1]), element_table(eps, von_mises))
        for inc, us in snapshots:
            written.append(f"{stem}_inc_{int(inc)}_nodes.csv")
            _write(self.path(written[-1]), node_table(self.coords, us))
        self.files.extend(written)
        return written

    def add_tool_record(self, step, increment, t, tool, centre, force, active_nodes, max_penetration):
        """One row of tool_forces.csv: step counted from 1, the force the
        sheet exerts on the tool [N]."""
        self.forces.append((step, increment, t, tool, centre, force, active_nodes, max_penetration))

    def close(self, completed=None, termination="", warnings=(), **extra):
        """Writes tool_forces.csv and summary.json; returns the summary."""
        _write(self.path("tool_forces.csv"), tool_force_table(self.forces))
        files = ["summary.json"] + self.files + ["tool_forces.csv"]
        x = self.coords
        summary = {"case": self.case, "sparlab_version": self.version or "",
                   "completed": all(s["completed"] for s in self.steps) if completed is None else bool(completed),
                   "termination": str(termination)}
        summary.update(extra)
        summary["steps"] = self.steps
        summary["mesh"] = {"source": str(self.mesh_source), "element_type": "Hex8",
                           "element_formulation": "incompatible_modes"}
        if self.rule_points is not None:
            summary["mesh"]["stiffness_rule_points"] = int(self.rule_points)
        summary["mesh"].update({"dim": 3, "nodes": int(x.shape[0]), "elements": self.num_elements,
                                "dofs": 3 * int(x.shape[0]),
                                "bounding_box_min_m": x.min(axis=0).tolist() if x.size else [0.0] * 3,
                                "bounding_box_max_m": x.max(axis=0).tolist() if x.size else [0.0] * 3})
        summary["warnings"] = list(warnings)
        summary["files"] = files
        with open(self.path("summary.json"), "w") as f:
            json.dump(summary, f, indent=1)
        return summary
