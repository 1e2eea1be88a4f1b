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
