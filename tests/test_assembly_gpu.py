"""Parity of the device assembly, force gather, nonlinear system and contact
with the oracle (src/fem/Assembler.cpp:86-215, src/fem/NonlinearSystem.hpp:
192-268, src/fem/RigidTool.cpp:398-555, src/fem/Forming.cpp:280-297).

Bars: the sparsity pattern bit-exact; assembled values and f_int bit-exact
(the gather adds the reference's numbers in the reference's order); the gross
norm within 1e-14 (its sum order is a tree); contact per-node values within
1e-13 relative, resultants within 1e-13.
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def csc_to_csr_of_k(outer, inner, csc_values):
    """The reference's column-major values of K -> row-major values of K on the
    same (structurally symmetric) pattern."""
    n = outer.size - 1
    cols = np.repeat(np.arange(n), np.diff(outer))
    # entry (inner[q], cols[q]) of K; in the row-major arrays row r = inner[q]
    order = np.lexsort((cols, inner))
    return csc_values[order]


def test_assembly_workload_bit_exact(device):
    """100 x 100 x 4 plate, random non-symmetric element matrices."""
    import torch
    w = W.assembly_workload()
    coords, conn, fe, ke = w["coords"], w["conn"], w["fe"], w["ke"]
    nn = coords.shape[0]
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33), G.Rule(), G.FINITE_LOGARITHMIC)
    row_ptr, col_idx = sysm.pattern()
    o_outer, o_inner, _ = O.pattern(nn, conn)
    assert np.array_equal(row_ptr, o_outer) and np.array_equal(col_idx, o_inner)
    assert sysm.nnz == 10_600_317
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    d_fe, d_ke = t(fe), t(ke)
    values = torch.empty(sysm.nnz, dtype=torch.float64, device=device)
    fint = torch.empty(3 * nn, dtype=torch.float64, device=device)
    gross = torch.empty(1, dtype=torch.float64, device=device)
    ref_csc = O.scatter(nn, conn, ke.reshape(-1, 576))
    ref_f, ref_gross = O.gather_force(nn, conn, fe)
    # column-major: the reference's own value order
    sysm.assemble_device(fe=d_fe, ke=d_ke, fint=fint, values=values, gross=gross, column_major=True)
    torch.cuda.synchronize()
    assert np.array_equal(values.cpu().numpy(), ref_csc)
    assert np.array_equal(fint.cpu().numpy(), ref_f)
    assert abs(gross.item() - ref_gross) <= 1e-14 * ref_gross
    # row-major values of K
    sysm.assemble_device(ke=d_ke, values=values, column_major=False)
    torch.cuda.synchronize()
    ref_csr = csc_to_csr_of_k(o_outer, o_inner, ref_csc)
    assert np.array_equal(values.cpu().numpy(), ref_csr)
    # element matrices stored column-major (Eigen's storage of the reference's
    # Ke): the same values bit for bit in either order
    d_ke_cm = t(ke.reshape(-1, 24, 24).transpose(0, 2, 1))
    del d_ke
    for column_major, ref in ((True, ref_csc), (False, ref_csr)):
        values.fill_(float("nan"))
        sysm.assemble_device(ke=d_ke_cm, values=values, column_major=column_major, ke_column_major=True)
        torch.cuda.synchronize()
        assert np.array_equal(values.cpu().numpy(), ref), column_major
    # an unknown layout flag is refused before any work is queued
    import ctypes
    rc = G.lib().sparlab_system_assemble_device(sysm.h, None, ctypes.c_void_p(d_ke_cm.data_ptr()), 4, None,
                                                ctypes.c_void_p(values.data_ptr()), None, None)
    assert rc == -1


def jittered_mesh(nx, ny, nz, lx, ly, lz, seed):
    coords, conn = W.structured_hex_mesh(nx, ny, nz, lx, ly, lz)
    rng = np.random.default_rng(seed)
    h = min(lx / nx, ly / ny, lz / nz)
    coords = coords + rng.uniform(-0.15, 0.15, coords.shape) * h
    return coords, conn


@pytest.mark.parametrize("rule", [(2, 0, 2), (2, 5, 2), (2, 7, 2)], ids=["2x2x2", "2x2x5", "2x2x7"])
def test_system_evaluate_commit_and_history(device, rule):
    """NonlinearSystem::evaluate / commit on a small plastic mesh vs the oracle
    element by element, the oracle's scatter and gather; two steps so the
    second starts from a committed plastic history. The 2 x 2 x 2 rule and
    the box rules through the thickness of the reference's fine-mesh decks
    (2 x 2 x 5, e.g. benchmarks/physics_audit/challenge/cases/kin_h0833/deck/
    deck.json, and 2 x 2 x 7)."""
    coords, conn = jittered_mesh(5, 4, 2, 0.01, 0.008, 0.002, seed=3)
    m = G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0,
                   hill=dict(r0=0.7, r45=0.6, r90=0.8), backstresses=((6000e6, 60.0), (1500e6, 10.0)))
    om = oracle_material(m)
    sysm = G.NonlinearSystem(coords, conn, m, G.Rule(*rule))
    nn, ne = coords.shape[0], conn.shape[0]
    rng = np.random.default_rng(8)
    states = np.zeros((ne, G.Rule(*rule).points, 39))
    alpha = np.zeros((ne, 9))
    for step in range(2):
        u = rng.normal(0, 2e-5, (nn, 3)) + (step + 1) * 1e-4 * coords[:, [2, 0, 1]]
        ev = sysm.evaluate(u, want_tangent=True, column_major=True)
        x0 = coords[conn]
        ue = u[conn].reshape(ne, 24)
        ref = O.element_batch(x0, ue, om, states, alpha=alpha, rule=rule, kinematics=2, threads=4)
        assert (ref["status"] == 0).all()
        f_ref, gross_ref = O.gather_force(nn, conn, ref["force"])
        k_ref = O.scatter(nn, conn, ref["tangent"].reshape(ne, 576))
        assert np.abs(ev.internal - f_ref).max() <= 1e-12 * np.abs(f_ref).max()
        assert np.abs(ev.tangent_values - k_ref).max() <= 1e-12 * np.abs(k_ref).max()
        assert abs(ev.gross - gross_ref) <= 1e-12 * gross_ref
        assert abs(ev.energy - ref["scalars"][:, 0].sum()) <= 1e-12 * abs(ref["scalars"][:, 0].sum())
        assert ev.yielding_points == int(ref["scalars"][:, 1].sum())
        fe, ke = sysm.element_outputs()
        assert np.abs(fe - ref["force"]).max() <= 1e-12 * np.abs(ref["force"]).max()
        trial_states, trial_alpha = sysm.history(trial=True)
        assert np.abs(trial_alpha - ref["alpha"]).max() <= 1e-10 * max(np.abs(ref["alpha"]).max(), 1e-30)
        sysm.commit()
        committed, committed_alpha = sysm.history()
        assert np.array_equal(committed_alpha, trial_alpha)
        states, alpha = ref["states"], ref["alpha"]
    assert ev.yielding_points > 0
    # restart: set the history of a fresh system to the committed one
    other = G.NonlinearSystem(coords, conn, m, G.Rule(*rule))
    other.set_history(committed, committed_alpha)
    s2, a2 = other.history()
    assert np.array_equal(a2, committed_alpha)


def test_system_reports_element_failure(device):
    coords, conn = jittered_mesh(2, 2, 1, 1.0, 1.0, 0.5, seed=1)
    sysm = G.NonlinearSystem(coords, conn, G.Material(200e9, 0.3, yield_stress=250e6, hardening_modulus=1e9))
    u = np.zeros_like(coords)
    ev = sysm.evaluate(u)
    assert ev.yielding_points == 0
    u[:, 2] = -3.0 * coords[:, 2]          # turns every element inside out
    with pytest.raises(G.SolverError):
        sysm.evaluate(u)


@pytest.mark.parametrize("shape", ["sphere", "plane", "cylinder"])
@pytest.mark.parametrize("friction", [0.0, 0.1])
@pytest.mark.parametrize("symmetric", [False, True])
def test_contact_matches_oracle(device, shape, friction, symmetric):
    """Slave nodes of the plate's top against a tool, with and without a
    committed friction history, vs the oracle's ToolContact::evaluate."""
    import torch
    # a 10 x 10 x 2 mm plate, 0.25 mm cells, every top node a slave
    coords, _ = W.structured_hex_mesh(40, 40, 2, 0.01, 0.01, 0.002, -0.005, -0.005, -0.002)
    nodes = np.nonzero(np.isclose(coords[:, 2], 0.0))[0].astype(np.int32)
    nn = coords.shape[0]
    rng = np.random.default_rng(21)
    u = rng.normal(0, 2e-6, (nn, 3))
    start = np.zeros((nn, 3))
    radius = 0.005
    if shape == "sphere":
        centre, kw = (0.003, -0.002, radius - 2e-4), dict(radius=radius)
    elif shape == "plane":
        centre, kw = (0.0, 0.0, -5e-5), dict(normal=(0.0, 0.1, -1.0))
    else:
        centre, kw = (0.0, 0.001, radius - 2e-4), dict(radius=radius, axis=(1.0, 0.2, 0.0))
    centre0 = (centre[0] - 3e-4, centre[1] + 1e-4, centre[2])
    kappa_a = rng.uniform(0.5e7, 1.5e7, nodes.size)
    hist = (rng.integers(0, 2, nodes.size).astype(np.int32), rng.normal(0, 5.0, (nodes.size, 3)),
            rng.normal(0, 1e-3, (nodes.size, 3)))
    tool_ref = O.tool_record({"plane": 0, "sphere": 1, "cylinder": 2}[shape], centre, centre0,
                             radius=kw.get("radius", 0.0), normal=kw.get("normal", (0, 0, 1)),
                             axis=kw.get("axis", (0, 0, 1)), friction=friction)
    ref = O.contact(tool_ref, coords, u.ravel(), start.ravel(), nodes, kappa_a, history=hist,
                    friction_symmetric=symmetric)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    tool = G.make_tool(shape, centre, centre0, friction=friction, **kw)
    out = G.contact_evaluate_device(tool, t(coords.ravel()), t(u.ravel()), t(start.ravel()), t(nodes, torch.int32),
                                    t(kappa_a), history=(t(hist[0], torch.int32), t(hist[1].ravel()),
                                                         t(hist[2].ravel())), friction_symmetric=symmetric)
    assert ref["n_active"] > 20
    assert out["n_active"] == ref["n_active"]
    assert np.array_equal(out["active"].cpu().numpy(), ref["active"])
    res = out["residual"].cpu().numpy()
    assert np.abs(res - ref["residual"]).max() <= 1e-13 * np.abs(ref["residual"]).max()
    act = ref["active"] == 1
    blk = out["blocks"].cpu().numpy()[act]
    assert np.abs(blk - ref["blocks"][act]).max() <= 1e-13 * np.abs(ref["blocks"][act]).max()
    if friction > 0:
        assert np.array_equal(out["trial_present"].cpu().numpy(), ref["t_present"])
        tf = out["trial_force"].cpu().numpy()[act]
        assert np.abs(tf - ref["t_force"][act]).max() <= 1e-13 * max(np.abs(ref["t_force"][act]).max(), 1e-300)
        assert out["n_slipping"] == ref["n_slipping"]
    for k in ("gross", "round_off"):
        assert abs(out[k] - ref[k]) <= 1e-13 * ref[k]
    assert np.abs(out["resultant"] - ref["resultant"]).max() <= 1e-13 * np.abs(ref["resultant"]).max()
    assert out["max_penetration"] == ref["max_penetration"]
    assert out["symmetric"] == ref["symmetric"]


def test_contact_blocks_into_the_assembled_matrix(device):
    """add_node_block on the device equals the oracle's on the column-major values."""
    import torch
    w = W.assembly_workload(nx=12, ny=10, nz=2, contact_nodes=60)
    coords, conn = w["coords"], w["conn"]
    nn, ne = coords.shape[0], conn.shape[0]
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33))
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    values = torch.empty(sysm.nnz, dtype=torch.float64, device=device)
    sysm.assemble_device(ke=t(w["ke"]), values=values, column_major=True)
    nodes = w["contact_nodes"]
    rng = np.random.default_rng(2)
    blocks = rng.normal(size=(nodes.size, 3, 3))
    active = rng.integers(0, 2, nodes.size).astype(np.int32)
    first = sysm.diagonal_blocks()[nodes]
    row_ptr, _ = sysm.pattern()
    row_len = (row_ptr[3 * nodes + 1] - row_ptr[3 * nodes]).astype(np.int32)
    G.add_node_blocks_device(t(blocks), t(active, torch.int32), t(first, torch.int32), t(row_len, torch.int32),
                             values, column_major=True)
    torch.cuda.synchronize()
    ref = O.scatter(nn, conn, w["ke"].reshape(ne, 576))
    ref = O.add_node_blocks(nn, conn, ref, nodes[active == 1], blocks[active == 1].reshape(-1, 9))
    assert np.array_equal(values.cpu().numpy(), ref)


@pytest.mark.parametrize("renumber", [False, True])
@pytest.mark.parametrize("column_major", [True, False])
def test_stiffness_gross_bit_exact(device, column_major, renumber):
    """stiffness_gross (NonlinearSystem.hpp:55-67): |K| |x| equal to the
    oracle's column loop bit for bit, from the reference's column-major values
    or the row-major values of the same (non-symmetric) K; the norm over the
    free DOFs within 1e-14 (its sum is a tree); zero entries of x skipped.
    With the nodes numbered at random, neighbourhoods and their places in
    each other's columns are irregular."""
    import torch
    w = W.assembly_workload(nx=24, ny=20, nz=3, contact_nodes=10)
    coords, conn = w["coords"], w["conn"]
    nn, ne = coords.shape[0], conn.shape[0]
    if renumber:
        perm = np.random.default_rng(3).permutation(nn)  # new number of each old node
        new_coords = np.empty_like(coords)
        new_coords[perm] = coords
        coords, conn = new_coords, perm[conn].astype(conn.dtype)
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33))
    outer, inner, _ = O.pattern(nn, conn)
    csc = O.scatter(nn, conn, w["ke"].reshape(ne, 576))
    rng = np.random.default_rng(31)
    x = rng.normal(size=3 * nn)
    x[rng.choice(3 * nn, nn, replace=False)] = 0.0
    free = np.sort(rng.choice(3 * nn, 2 * nn, replace=False)).astype(np.int32)
    norm, product = O.stiffness_gross(nn, conn, csc, x, free)
    values = csc if column_major else csc_to_csr_of_k(outer, inner, csc)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    d_product = torch.full((3 * nn,), float("nan"), dtype=torch.float64, device=device)
    out, _ = sysm.stiffness_gross_device(t(values), t(x), t(free, torch.int32), column_major=column_major,
                                         product=d_product)
    torch.cuda.synchronize()
    assert np.array_equal(d_product.cpu().numpy(), product)
    assert abs(out.item() - norm) <= 1e-14 * norm
    # a free DOF outside [0, 3 nn) is caught by the kernel's bounds check: a NaN norm
    bad = free.copy()
    bad[-1] = 3 * nn
    out, _ = sysm.stiffness_gross_device(t(values), t(x), t(bad, torch.int32), column_major=column_major)
    torch.cuda.synchronize()
    assert np.isnan(out.item())
