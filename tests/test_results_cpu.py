"""Result files in the reference's output contract (sparlab_gpu.results).

Pinned by the reference's own result directory of the kin_h0833 deck
(benchmarks/physics_audit/challenge/cases/kin_h0833/output, written by
src/io/FormingWriter.cpp and src/io/CsvWriter.cpp; fixtures made by
tests/golden/copy_result_files.py): from the deck's mesh (the golden deck,
through precomp.deck_model) and the values of each table as read back from
the reference's text, the writers reproduce every table, byte for byte: the
node tables of the form and release steps (4802 nodes), their element tables
(2304 elements) and tool_forces.csv (395 rows). The step stems equal the
summary's files_stem. A written directory meets the loader's contract
(python/precomp/fea/results.py:303-365): summary steps with name, completed
and files_stem, a node table for every completed step with the same nodes and
reference coordinates, no table for a stopped step.
"""
import csv
import hashlib
import json
import os
import re

import numpy as np
import pytest

from sparlab_gpu import precomp as P
from sparlab_gpu import results as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture():
    with open(os.path.join(GOLDEN, "results", "kin_h0833.json")) as f:
        meta = json.load(f)
    arrays = np.load(os.path.join(GOLDEN, "results", "kin_h0833.npz"), allow_pickle=False)
    return meta, arrays


def digest(text):
    data = text.encode("ascii")
    return {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}


def test_tables_reproduce_the_reference_files():
    meta, arrays = fixture()
    model = P.deck_model(P.read_deck(os.path.join(GOLDEN, "decks", "kin_h0833.json")))
    for stem in ("step_1_form", "step_3_release"):
        u = arrays[f"u_{stem}"]
        assert u.shape == model.coords.shape
        assert digest(R.node_table(model.coords, u)) == meta["digests"][f"{stem}_nodes.csv"], stem
        el = arrays[f"elements_{stem}"]
        assert el.shape == (model.conn.shape[0], 2)
        assert digest(R.element_table(el[:, 0], el[:, 1])) == meta["digests"][f"{stem}_elements.csv"], stem
    tf = arrays["tool_forces"]
    (tool,) = meta["tools"]
    records = [(r[0], r[1], r[2], tool, r[3:6], r[6:9], r[9], r[10]) for r in tf]
    assert digest(R.tool_force_table(records)) == meta["digests"]["tool_forces.csv"]


def test_a_changed_value_changes_the_table():
    """The digest check sees one value off by one ulp."""
    meta, arrays = fixture()
    model = P.deck_model(P.read_deck(os.path.join(GOLDEN, "decks", "kin_h0833.json")))
    u = arrays["u_step_3_release"].copy()
    u[1234, 2] = np.nextafter(u[1234, 2], np.inf)
    assert digest(R.node_table(model.coords, u)) != meta["digests"]["step_3_release_nodes.csv"]


def test_step_stems():
    meta, _ = fixture()
    for i, s in enumerate(meta["steps"]):
        assert R.forming_step_stem(i, s["name"]) == s["files_stem"]
    # every byte but ASCII letters, digits, '-' and '_' becomes '_' (two for a
    # two-byte UTF-8 letter); an empty name is "step"
    assert R.forming_step_stem(2, "Drück 2") == "step_3_Dr__ck_2"
    assert R.forming_step_stem(0, "a/b.c-d_e") == "step_1_a_b_c-d_e"
    assert R.forming_step_stem(1, "") == "step_2_step"


def test_numbers_read_back_exactly():
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-300, 300, 2000),
                        [0.0, -0.0, 5e-324, -2.2250738585072014e-308, 1.7976931348623157e308, 0.1, 1e-5, 123456789.0]])
    for x in v:
        s = R.number(x)
        assert np.float64(float(s)).tobytes() == np.float64(x).tobytes(), (x, s)
    assert R.number(-0.0) == "-0" and R.number(0.02) == "0.02" and R.number(1e-5) == "1.0000000000000001e-05"


def read_table(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def test_result_directory_meets_the_loader_contract(tmp_path):
    rng = np.random.default_rng(5)
    coords = rng.standard_normal((27, 3))
    conn = rng.integers(0, 27, size=(4, 8))
    d = R.ResultDirectory(tmp_path / "run", coords, conn, case="c", version="v", rule_points=8)
    u1 = rng.standard_normal((27, 3)) * 1e-3
    u2 = rng.standard_normal(81) * 1e-3
    vm = rng.random(4) * 1e8
    eps = rng.random(4) * 0.1
    d.add_step("form", u1, plastic_strain=eps, von_mises=vm, snapshots=[(5, u1 * 0.5)], increments=10)
    d.add_step("unload", u2, type="release", completed=False)
    d.add_step("release 1", u2, type="release")
    d.add_tool_record(1, 1, 0.25, "tool", (0.0, 0.0, 1e-3), (0.0, 0.0, 12.5), 3, 1e-6)
    summary = d.close(termination="stopped in step 2")
    with open(tmp_path / "run" / "summary.json") as f:
        assert json.load(f) == summary
    assert summary["completed"] is False
    stem_re = re.compile(r"^step_(\d+)_[A-Za-z0-9_-]+$")
    ref0 = None
    for i, s in enumerate(summary["steps"]):
        assert {"name", "completed", "files_stem"} <= set(s)
        assert stem_re.match(s["files_stem"]) and s["files_stem"] == R.forming_step_stem(i, s["name"])
        path = tmp_path / "run" / f"{s['files_stem']}_nodes.csv"
        if not s["completed"]:
            assert not path.exists()
            continue
        header, rows = read_table(path)
        assert tuple(header) == R.NODE_COLUMNS
        vals = np.array([[float(x) for x in r] for r in rows])
        assert np.array_equal(vals[:, 0], np.arange(27))
        if ref0 is None:
            ref0 = vals[:, 1:4]
        assert np.array_equal(vals[:, 1:4], ref0) and np.array_equal(ref0, coords)
        u = u1 if s["name"] == "form" else u2.reshape(-1, 3)
        assert np.array_equal(vals[:, 4:7], u)
        assert s["max_displacement_m"] == np.sqrt((u * u).sum(axis=1)).max()
    header, rows = read_table(tmp_path / "run" / "step_1_form_elements.csv")
    assert tuple(header) == R.ELEMENT_COLUMNS
    vals = np.array([[float(x) for x in r] for r in rows])
    assert np.array_equal(vals[:, 1], eps) and np.array_equal(vals[:, 2], vm)
    assert (tmp_path / "run" / "step_1_form_inc_5_nodes.csv").exists()
    assert not (tmp_path / "run" / "step_3_release_1_elements.csv").exists()
    header, rows = read_table(tmp_path / "run" / "tool_forces.csv")
    assert tuple(header) == R.FORCE_COLUMNS and rows == [["1", "1", "0.25", "tool", "0", "0", "0.001", "0", "0",
                                                          "12.5", "3", "9.9999999999999995e-07"]]
    assert summary["files"][0] == "summary.json" and "step_3_release_1_nodes.csv" in summary["files"]
    with pytest.raises(ValueError):
        d.add_step("bad", u1, von_mises=np.zeros(3))


def test_element_values_follow_the_driver():
    stress = {"von_mises": np.array([1.0, 2.0]), "max_equivalent_plastic_strain": np.array([0.1, 0.0])}
    eps, vm = R.element_values(stress, elastoplastic=True)
    assert np.array_equal(eps, [0.1, 0.0]) and np.array_equal(vm, [1.0, 2.0])
    eps, vm = R.element_values(stress, elastoplastic=False)
    assert np.array_equal(eps, [0.0, 0.0]) and np.array_equal(vm, [1.0, 2.0])
