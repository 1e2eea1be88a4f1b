"""A device run written in the reference's output contract and read back
(sparlab_gpu.results; the contract: python/precomp/fea/results.py:303-365,
written by src/io/FormingWriter.cpp).

A 12 x 10 x 2 sheet (top surface at z = 0) pressed into a dome, J2 + Voce:
the system's evaluation, commit and end-of-step stress recovery on the
device, as the forming driver runs them at the end of a step
(src/fem/Forming.cpp:1255-1281), then a form step with its element table and
a release step. Read back: the node tables hold the reference coordinates and
the device displacement exactly (17 significant digits read back to the same
doubles); the element table holds the device stress recovery's von Mises
stress and largest equivalent plastic strain exactly, and those agree with
the oracle's incompatible_elastoplastic_stress on the system's committed
history within 1e-12 of the largest; the height map of the released top
surface from the files equals the one from the device arrays bit for bit and
the oracle's ray-cast bit for bit.
"""
import csv
import json

import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material
from sparlab_gpu import results as R
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def read_rows(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    return rows[0], np.array([[float(x) for x in r] for r in rows[1:]])


def top_surface(coords, u):
    """(surface (ny, nx, 3), grid) of the deformed top nodes (reference z = 0)
    on the reference top-node grid."""
    top = np.flatnonzero(coords[:, 2] == 0.0)
    xs, ys = np.unique(coords[top, 0]), np.unique(coords[top, 1])
    order = np.lexsort((coords[top, 0], coords[top, 1]))
    cur = (coords + u.reshape(-1, 3))[top][order]
    grid = (xs[0], ys[0], xs.size, ys.size, xs[1] - xs[0])
    return np.ascontiguousarray(cur.reshape(ys.size, xs.size, 3)), grid


def test_device_run_written_and_read_back(device, tmp_path):
    import torch
    m = G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0)
    coords, conn = W.structured_hex_mesh(12, 10, 2, 0.012, 0.010, 0.001, -0.006, -0.005, -0.001)
    rng = np.random.default_rng(21)
    r2 = (coords[:, 0] ** 2 + coords[:, 1] ** 2) / 0.003 ** 2
    u = np.zeros_like(coords)
    u[:, 2] = -1.5e-4 * np.exp(-r2)
    u[:, :2] = 2e-5 * coords[:, :2] / 0.006 * np.exp(-r2)[:, None]
    u += rng.normal(0.0, 1e-7, u.shape)
    rule = G.Rule(2, 0, 2)
    sysm = G.NonlinearSystem(coords, conn, m, rule)
    sysm.evaluate(u)
    sysm.commit()
    stress = sysm.stress(u)
    eps, vm = R.element_values(stress, m.plastic)
    assert m.plastic and (eps > 0).any() and (eps == 0).any()
    # the oracle's stress recovery on the system's committed history
    states, alpha = sysm.history()
    x0 = coords[conn]
    ue = u[conn].reshape(conn.shape[0], 24)
    ref = O.element_stress_batch(x0, ue, oracle_material(m), states, alpha, rule=(2, 0, 2), kinematics=2,
                                 threads=8)
    assert (ref["status"] == 0).all()
    assert np.abs(vm - ref["stress"][:, 24]).max() <= 1e-12 * np.abs(ref["stress"][:, 24]).max()
    assert np.abs(eps - ref["stress"][:, 25]).max() <= 1e-12 * np.abs(ref["stress"][:, 25]).max()

    u_rel = 0.8 * u
    d = R.ResultDirectory(tmp_path / "run", coords, conn, case="dome", version=G.lib().sparlab_version().decode(),
                          rule_points=rule.points, mesh_source="structured_hex 12 x 10 x 2")
    d.add_step("form", u, plastic_strain=eps, von_mises=vm, increments=1)
    d.add_step("release", u_rel, type="release")
    summary = d.close(termination="reached the end of its window")
    assert summary["completed"] and [s["files_stem"] for s in summary["steps"]] == ["step_1_form", "step_2_release"]
    with open(tmp_path / "run" / "summary.json") as f:
        assert json.load(f)["mesh"]["nodes"] == coords.shape[0]

    header, nodes = read_rows(tmp_path / "run" / "step_1_form_nodes.csv")
    assert tuple(header) == R.NODE_COLUMNS
    assert np.array_equal(nodes[:, 0], np.arange(coords.shape[0]))
    assert np.array_equal(nodes[:, 1:4], coords) and np.array_equal(nodes[:, 4:7], u)
    header, elements = read_rows(tmp_path / "run" / "step_1_form_elements.csv")
    assert tuple(header) == R.ELEMENT_COLUMNS
    assert np.array_equal(elements[:, 1], eps) and np.array_equal(elements[:, 2], vm)

    _, rel = read_rows(tmp_path / "run" / "step_2_release_nodes.csv")
    surf_file, grid = top_surface(rel[:, 1:4], rel[:, 4:7])
    surf_mem, grid_mem = top_surface(coords, u_rel)
    assert grid == grid_mem and np.array_equal(surf_file, surf_mem)
    z_f, hit_f = G.surface_raycast_device(torch.as_tensor(surf_file, device=device), grid)
    z_m, hit_m = G.surface_raycast_device(torch.as_tensor(surf_mem, device=device), grid)
    torch.cuda.synchronize()
    z_o, hit_o, folded = O.surface_raycast(surf_file, grid)
    assert folded == 0
    z_f, hit_f = z_f.cpu().numpy(), hit_f.cpu().numpy().astype(bool)
    assert np.array_equal(z_f, z_m.cpu().numpy()) and np.array_equal(hit_f, hit_m.cpu().numpy().astype(bool))
    assert np.array_equal(hit_f, np.asarray(hit_o, dtype=bool)) and hit_f.sum() > 0.5 * hit_f.size
    assert np.array_equal(z_f[hit_f], np.asarray(z_o)[hit_f])
