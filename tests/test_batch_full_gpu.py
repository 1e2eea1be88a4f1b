"""The batched form at its full size: the 64 forming problems of the batched
workload (dome, truncated cone, elliptic cone, pyramid; 2 000 - 40 000 Hex8
each, 1.34 M in all; one forming increment from a virgin history, Hill48 +
two Chaboche backstresses + Voce, logarithmic strain), one captured
evaluation of the whole batch, against the oracle.

Bars: every problem's assembled internal force f_int within 1e-12 of the
oracle's (every element of every problem evaluated by the oracle, gathered
to the nodes; max |delta| / max |ref| per problem), and for a seeded sample
of 16 deforming elements per problem the element force and tangent within
1e-12 (the element tests' bars); no failed element. The sample is drawn
among elements whose nodal displacements differ by at least a tenth of the
problem's largest such difference: outside the formed pocket the sheet does
not move, or moves rigidly, and an element force there is zero or round-off.
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material, rel
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def test_full_batched_workload_against_the_oracle(device):
    import torch
    m = W.microbench_material()
    om = oracle_material(m)
    rule = G.Rule(2, 0, 2)
    probs = W.batched_problems()
    total = sum(p["conn"].shape[0] for p in probs)
    print(f"full batched workload: {len(probs)} problems, {total} elements", flush=True)
    assert len(probs) == 64 and total > 1_000_000
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
    u = torch.as_tensor(np.concatenate([p["u"] for p in probs]).reshape(-1), device=device)
    fint = torch.zeros_like(u)
    values = torch.empty(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(u, fint, values)
    sc = b.result()
    assert (sc["failed"] == 0).all()
    assert (sc["yielding"] > 0).sum() >= 32
    del values
    fint = fint.cpu().numpy()
    rng = np.random.default_rng(64)
    worst_f = worst_fe = worst_ke = 0.0
    for p, prob in enumerate(probs):
        conn, coords, up = prob["conn"], prob["coords"], prob["u"]
        ne, nn = conn.shape[0], coords.shape[0]
        x0 = coords[conn]
        ue = up[conn].reshape(ne, 24)
        ref = O.element_batch(x0, ue, om, np.zeros((ne, 8, O.STATE)), kinematics=2, want_tangent=False)
        assert (ref["status"] == 0).all()
        ref_f, _ = O.gather_force(nn, conn, ref["force"])
        d0, d1 = 3 * b.node_offset[p], 3 * b.node_offset[p + 1]
        worst_f = max(worst_f, np.abs(fint[d0:d1] - ref_f).max() / np.abs(ref_f).max())
        spread = np.abs(ue.reshape(ne, 8, 3) - ue.reshape(ne, 8, 3)[:, :1, :]).max(axis=(1, 2))
        pool = np.nonzero(spread >= 0.1 * spread.max())[0]
        sample = np.sort(rng.choice(pool, size=min(16, pool.size), replace=False))
        outs = [b.system.element_outputs(first=b.element_offset[p] + e, count=1) for e in sample]
        fe = np.concatenate([o[0] for o in outs])
        ke = np.concatenate([o[1] for o in outs])
        rk = O.element_batch(x0[sample], ue[sample], om, np.zeros((sample.size, 8, O.STATE)), kinematics=2)
        worst_fe = max(worst_fe, rel(fe, rk["force"]).max())
        worst_ke = max(worst_ke, rel(ke, rk["tangent"]).max())
    print(f"  f_int {worst_f:.1e}, sampled element force {worst_fe:.1e}, tangent {worst_ke:.1e}", flush=True)
    assert worst_f <= 1e-12
    assert worst_fe <= 1e-12 and worst_ke <= 1e-12


def test_full_batched_newton_solve_equals_plain_multigrid_solves(device):
    """One Newton correction of all 64 problems (rims clamped; BiCGStab to
    1e-8, each problem preconditioned by its own multigrid hierarchy): the
    segmented solve over the union's block against each problem solved alone
    with the plain multigrid solve. Bar: per problem, the correction, the
    iteration count and the relative residual equal bit for bit."""
    import torch
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    probs = W.batched_problems()

    def rim_free(coords):
        x, y = coords[:, 0], coords[:, 1]
        rim = np.isclose(x, x.min()) | np.isclose(x, x.max()) | np.isclose(y, y.min()) | np.isclose(y, y.max())
        return np.nonzero(np.repeat(~rim, 3))[0].astype(np.int32)

    frees = [rim_free(p["coords"]) for p in probs]
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
    b.system.set_block_preconditioner("multigrid")
    u = torch.as_tensor(np.concatenate([p["u"] for p in probs]).reshape(-1), device=device)
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(u, fint, values, column_major=False)
    assert (b.result()["failed"] == 0).all()
    free = np.concatenate([f + 3 * b.node_offset[i] for i, f in enumerate(frees)]).astype(np.int32)
    d_free = torch.as_tensor(free, device=device)
    b.system.partition_device(d_free)
    block = b.system.free_block_device(values)
    rows = b.free_segments(free)
    print(f"full batched Newton solve: {len(probs)} problems, {free.size} free DOFs", flush=True)
    x, its, rel_b, status = b.system.solve_free_block_segments_device(block, -fint[d_free.long()], rows,
                                                                      tolerance=1e-8, symmetric=False)
    assert (status == 0).all()
    x = x.cpu().numpy()
    del b, block, values
    torch.cuda.empty_cache()
    for p, (prob, fr) in enumerate(zip(probs, frees)):
        s = G.NonlinearSystem(prob["coords"], prob["conn"], m, rule)
        s.set_block_preconditioner("multigrid")
        up = torch.as_tensor(prob["u"].reshape(-1), device=device)
        f = torch.zeros_like(up)
        v = torch.zeros(s.nnz, dtype=torch.float64, device=device)
        s.evaluate_device(up, f, v, column_major=False)
        s.result()
        dp = torch.as_tensor(fr, device=device)
        s.partition_device(dp)
        blk = s.free_block_device(v)
        xp, itp, relp = s.solve_free_block_device(blk, -f[dp.long()], tolerance=1e-8, symmetric=False)
        assert np.array_equal(x[rows[p]:rows[p + 1]], xp.cpu().numpy()), f"problem {p}: x"
        assert its[p] == itp and rel_b[p] == relp, (p, its[p], itp, rel_b[p], relp)
