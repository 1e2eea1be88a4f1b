"""The precomp side on the device (sparlab_gpu.precomp).

DeckBatch: three decks in one process (the reference's fine-mesh deck
kin_h0833, a coarser two-layer variant of it, and the homogeneous uniaxial
deck with the incompatible-mode element), two batched systems (the first two
share material, rule and kinematics); every deck's f_int, tangent values and
scalars equal those of its own NonlinearSystem, bit for bit (energy and
gross within 1e-14: their sums run in another tree), over two increments
with a commit between; then every deck's result directory written from the
batch (write_results: node table, element table from the end-of-step stress
recovery, summary) equals the one written from its own system, byte for
byte.

shape_sensitivity: dJ/dX at an equilibrium state (J = g . u on the free
DOFs; a plate clamped at x = 0 with the tip's deflection prescribed, under
the load that holds a bending field in equilibrium) against central
differences of J over equilibria re-solved by Newton with the device block
solve at X +- h d, h = 1e-4 of the element size: elastic (a symmetric
tangent: conjugate gradients for K^T) and plastic (Hill48 + Chaboche:
BiCGStab on the transposed non-symmetric block), the plastic case also with
the 2 x 2 x 5 box rule of the reference's fine-mesh decks.
"""
import copy
import json
import os

import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import precomp as P
from sparlab_gpu import results as R
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def deck(name):
    with open(os.path.join(DECKS, name)) as f:
        return json.load(f)


def decks():
    fine = deck("kin_h0833.json")
    coarse = copy.deepcopy(fine)
    coarse["mesh"].update(nx=24, ny=24, nz=2)
    homo = deck("homogeneous_uniaxial.json")
    homo["model"]["element_formulation"] = "incompatible_modes"
    homo["forming"]["kinematics"] = "finite_logarithmic"
    return [fine, coarse, homo]


def displacements(models, scale):
    us = []
    for m in models:
        x, y = m.coords[:, 0], m.coords[:, 1]
        u = np.zeros_like(m.coords)
        if m.conn.shape[0] > 1:   # a dome pressed into the sheet
            u[:, 2] = scale * W._family_depth("dome", x, y, 0.015, 0.003, 0.5)
        else:                     # uniaxial stretch of the cube
            u[:, 0] = scale * 0.5 * m.coords[:, 0]
            u[:, 1] = -scale * 0.2 * m.coords[:, 1]
            u[:, 2] = -scale * 0.2 * m.coords[:, 2]
        us.append(u.reshape(-1))
    return us


def test_deck_batch_equals_each_deck_alone(device, tmp_path):
    import torch
    batch = P.DeckBatch(decks())
    assert len(batch.groups) == 2 and sorted(len(g[0]) for g in batch.groups) == [1, 2]
    plain = [m.system() for m in batch.models]
    for step, scale in enumerate((0.03, 0.05)):
        us = displacements(batch.models, scale)
        print(f"deck batch: increment {step}", flush=True)
        out, scalars = batch.evaluate(us)
        torch.cuda.synchronize()
        for i, (m, s) in enumerate(zip(batch.models, plain)):
            ev = s.evaluate(us[i], want_tangent=True, column_major=True)
            fint, values = out[i]
            assert np.array_equal(fint.cpu().numpy(), ev.internal), (step, i)
            assert np.array_equal(values.cpu().numpy(), ev.tangent_values), (step, i)
            sc = scalars[i]
            assert sc["failed"] == 0 and sc["yielding"] == ev.yielding_points
            assert abs(sc["energy"] - ev.energy) <= 1e-14 * abs(ev.energy)
            assert abs(sc["gross"] - ev.gross) <= 1e-14 * abs(ev.gross)
            s.commit()
        batch.commit()
    assert scalars[0]["yielding"] > 0
    # every deck's result directory from the batch equals the one written
    # from its own system, file for file and byte for byte
    dirs = [tmp_path / f"batch_{i}" for i in range(len(plain))]
    summaries = batch.write_results(dirs, us, termination="reached the end of its window", increments=2)
    version = G.lib().sparlab_version().decode()
    for i, (m, s) in enumerate(zip(batch.models, plain)):
        eps, vm = R.element_values(s.stress(us[i]), m.material.plastic)
        d = R.ResultDirectory(tmp_path / f"plain_{i}", m.coords, m.conn, case=m.name, version=version,
                              rule_points=m.rule.points, mesh_source=m.mesh_source)
        d.add_step("form", us[i], plastic_strain=eps, von_mises=vm, increments=2)
        assert d.close(termination="reached the end of its window") == summaries[i]
        names = sorted(os.listdir(dirs[i]))
        assert names == sorted(os.listdir(tmp_path / f"plain_{i}"))
        assert names == ["step_1_form_elements.csv", "step_1_form_nodes.csv", "summary.json", "tool_forces.csv"]
        for name in names:
            assert (dirs[i] / name).read_bytes() == (tmp_path / f"plain_{i}" / name).read_bytes(), (i, name)
    assert summaries[0]["mesh"]["source"] == "structured_hex 48 x 48 x 1"
    rows = (dirs[0] / "step_1_form_elements.csv").read_text().split()[1:]
    assert len(rows) == batch.models[0].conn.shape[0] and any(float(r.split(",")[1]) > 0.0 for r in rows)


def plate():
    coords, conn = W.structured_hex_mesh(6, 4, 2, 0.012, 0.008, 0.002)
    nn = coords.shape[0]
    clamped = np.isclose(coords[:, 0], 0.0)
    tip = np.isclose(coords[:, 0], 0.012)
    fixed = np.zeros(3 * nn, bool)
    fixed[np.repeat(clamped, 3)] = True
    fixed[3 * np.nonzero(tip)[0] + 2] = True
    return coords, conn, tip, np.nonzero(~fixed)[0].astype(np.int32)


def solve(s, coords, free, u0, load, device, symmetric):
    """Newton for f_int(u, X)_free = load with u's prescribed DOFs fixed at
    u0's, from u0; returns u (device) and the tangent's column-major values
    evaluated there."""
    import torch
    s.set_coords(coords)
    u = u0.clone()
    d_free = torch.as_tensor(free, device=device).long()
    f = torch.zeros_like(u)
    values = torch.empty(s.nnz, dtype=torch.float64, device=device)
    s.partition_device(d_free.int())
    for it in range(30):
        s.evaluate_device(u, f, values, want_tangent=True, column_major=False)
        ev = s.result()
        r = f[d_free] - load
        norm = float(r.norm())
        if norm <= 1e-13 * ev.gross:
            break
        block = s.free_block_device(values)
        du, _, _ = s.solve_free_block_device(block, -r, tolerance=1e-11, symmetric=symmetric)
        u[d_free] += du
    else:
        raise AssertionError(f"no equilibrium: residual {norm:.3e} of {ev.gross:.3e}")
    s.evaluate_device(u, f, values, want_tangent=True, column_major=True)
    s.result()
    return u, values


@pytest.mark.parametrize("case", ["elastic", "plastic", "plastic_2x2x5"])
def test_shape_sensitivity_matches_central_differences(device, case):
    """The base state u* is a bending field (every node on a parabola of the
    tip deflection, the clamped end and the tip's z prescribed); the load is
    its own internal force on the free DOFs, b = f_int(u*, X)_free, so u* is
    an equilibrium at X; X +- h d are re-solved from u* by Newton."""
    import torch
    coords, conn, tip, free = plate()
    if case == "elastic":
        m = G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0)
        deflection = 2e-6
    else:
        m = W.microbench_material()
        deflection = 1.5e-4
    rule = G.Rule(2, 5, 2) if case.endswith("2x2x5") else G.Rule(2, 0, 2)
    s = G.NonlinearSystem(coords, conn, m, rule)
    sym = m.symmetric_tangent
    x = coords[:, 0] / 0.012
    ustar = np.zeros_like(coords)
    ustar[:, 2] = -deflection * x * x
    ustar[:, 0] = deflection * 2.0 * x * (coords[:, 2] + 0.001) / 0.012  # plane sections stay plane
    u0 = torch.as_tensor(ustar.reshape(-1), device=device)
    d_free = torch.as_tensor(free, device=device).long()
    f = torch.zeros_like(u0)
    s.evaluate_device(u0, f, None, want_tangent=False)
    ev = s.result()
    assert (ev.yielding_points > 0) == case.startswith("plastic")
    load = f[d_free].clone()
    u, values = solve(s, coords, free, u0, load, device, sym)
    assert torch.equal(u, u0)   # already in equilibrium
    rng = np.random.default_rng(4)
    g = rng.normal(size=free.size)
    grad, its = P.shape_sensitivity(s, u, values, torch.as_tensor(free, device=device), torch.as_tensor(g, device=device))
    grad = grad.cpu().numpy()
    direction = rng.normal(size=coords.shape) * 0.002
    h = 1e-4
    j = []
    for sign in (1.0, -1.0):
        up, _ = solve(s, coords + sign * h * direction, free, u0, load, device, sym)
        j.append(float(g @ up[d_free].cpu().numpy()))
    s.set_coords(coords)
    fd = (j[0] - j[1]) / (2 * h)
    an = float((grad * direction).sum())
    print(f"shape sensitivity ({case}): adjoint {an:.10e}, central differences {fd:.10e}, "
          f"{its} iterations of the transposed solve", flush=True)
    assert abs(fd - an) <= 1e-5 * abs(an)
