"""The batched form (sparlab_batch_*) against each problem's own plain
device system, which the other GPU tests pin to the oracle.

A batch is the block-diagonal union of its problems, so every problem's
f_int, CSR values, trial history and per-problem scalars must equal those of
the problem evaluated alone: bit for bit for f_int, the values and the
history (the same kernels add the same numbers in the same order), within
1e-14 for energy and gross (their sums run in another tree). The captured
HIP graph is replayed several times with eager work and synchronisations in
between, and every replay must equal the direct launches; the autotuned
elements per wavefront must not change a bit.
"""
import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def small_problems(seed=3):
    """Four small uneven sheets with forming increments large enough to yield."""
    rng = np.random.default_rng(seed)
    out = []
    for (nx, ny, nz), fam in zip(((3, 3, 1), (4, 2, 2), (5, 5, 1), (2, 3, 2)), W.FAMILIES):
        coords, conn = W.structured_hex_mesh(nx, ny, nz, 0.01, 0.01, 0.001, -0.005, -0.005, -0.001)
        w = W._family_depth(fam, coords[:, 0], coords[:, 1], 0.005, 0.003, 0.5)
        u = np.zeros_like(coords)
        u[:, 2] = 0.05 * w
        u += rng.normal(0.0, 2e-6, size=u.shape)
        out.append(dict(coords=coords, conn=conn, u=u))
    return out


def union(problems):
    return np.concatenate([p["u"] for p in problems]).reshape(-1)


def plain_results(problems, material, rule, commit_first=False):
    """Each problem alone: (fint, values, scalars, trial history)."""
    res = []
    for p in problems:
        s = G.NonlinearSystem(p["coords"], p["conn"], material, rule)
        if commit_first:
            s.evaluate(p["u"] * 0.5, want_tangent=True)
            s.commit()
        ev = s.evaluate(p["u"], want_tangent=True)
        st, al = s.history(trial=True)
        res.append(dict(fint=ev.internal, values=ev.tangent_values, energy=ev.energy, gross=ev.gross,
                        yielding=ev.yielding_points, states=st, alpha=al))
    return res


def batch_results(b, problems, device, graphs=True, column_major=False):
    import torch
    b.set_graphs(graphs)
    u = torch.as_tensor(union(problems), device=device)
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(u, fint, values, want_tangent=True, column_major=column_major)
    sc = b.result()
    return fint.cpu().numpy(), values.cpu().numpy(), sc


def check_against_plain(b, problems, plain, fint, values, sc):
    st, al = b.system.history(trial=True)
    for p, ref in enumerate(plain):
        d0, d1 = 3 * b.node_offset[p], 3 * b.node_offset[p + 1]
        v0, v1 = b.value_offset[p], b.value_offset[p + 1]
        e0, e1 = b.element_offset[p], b.element_offset[p + 1]
        assert np.array_equal(fint[d0:d1], ref["fint"]), f"problem {p}: f_int"
        assert np.array_equal(values[v0:v1], ref["values"]), f"problem {p}: values"
        assert np.array_equal(st[e0:e1], ref["states"]), f"problem {p}: trial history"
        assert np.array_equal(al[e0:e1], ref["alpha"]), f"problem {p}: modes"
        assert sc["failed"][p] == 0
        assert sc["yielding"][p] == ref["yielding"]
        assert abs(sc["energy"][p] - ref["energy"]) <= 1e-14 * abs(ref["energy"])
        assert abs(sc["gross"][p] - ref["gross"]) <= 1e-14 * abs(ref["gross"])


def test_batch_equals_each_problem_alone(device):
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    problems = small_problems()
    plain = plain_results(problems, m, rule)
    assert sum(r["yielding"] for r in plain) > 0, "the increments should yield somewhere"
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in problems], m, rule)
    fint, values, sc = batch_results(b, problems, device, graphs=False)
    check_against_plain(b, problems, plain, fint, values, sc)


def test_graph_replays_equal_direct_launches(device):
    import torch
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    problems = small_problems()
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in problems], m, rule)
    ref_f, ref_v, _ = batch_results(b, problems, device, graphs=False)
    b.set_graphs(True)
    u = torch.as_tensor(union(problems), device=device)
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    for replay in range(4):
        fint.fill_(float("nan"))
        values.fill_(float("nan"))
        b.evaluate_device(u, fint, values, want_tangent=True)
        torch.cuda.synchronize()
        assert np.array_equal(fint.cpu().numpy(), ref_f), f"replay {replay}: f_int"
        assert np.array_equal(values.cpu().numpy(), ref_v), f"replay {replay}: values"
        # the program's own eager work between replays
        scratch = torch.randn(4096, dtype=torch.float64, device=device)
        (scratch * 2.0).sum().item()


def test_commit_and_next_increment(device):
    """Evaluate at half the increment, commit, evaluate the full increment:
    the captured graphs follow the swapped history buffers."""
    import torch
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    problems = small_problems()
    plain = plain_results(problems, m, rule, commit_first=True)
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in problems], m, rule)
    u = torch.as_tensor(union(problems), device=device)
    half = u * 0.5
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(half, fint, values, want_tangent=True)
    b.result()
    b.commit()
    b.evaluate_device(u, fint, values, want_tangent=True)
    sc = b.result()
    check_against_plain(b, problems, plain, fint.cpu().numpy(), values.cpu().numpy(), sc)


def test_autotune_keeps_results(device):
    import torch
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    problems = small_problems()
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in problems], m, rule)
    ref_f, ref_v, _ = batch_results(b, problems, device, graphs=False)
    u = torch.as_tensor(union(problems), device=device)
    chunk = b.autotune(u)
    assert chunk in (8, 16, 32, 64, 128, 256)
    f2, v2, _ = batch_results(b, problems, device, graphs=True)
    assert np.array_equal(f2, ref_f) and np.array_equal(v2, ref_v)


def test_failure_is_reported_per_problem(device):
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    problems = small_problems()
    problems[2]["u"] = problems[2]["u"].copy()
    conn = problems[2]["conn"]
    problems[2]["u"][conn[0, 4:]] += np.array([0.0, 0.0, -0.0025])  # invert element 0 of problem 2
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in problems], m, rule)
    _, _, sc = batch_results(b, problems, device, graphs=True)
    assert sc["failed"][2] >= 1 and sc["first_failed"][2] == 0
    assert all(sc["failed"][p] == 0 and sc["first_failed"][p] == -1 for p in (0, 1, 3))


def test_forming_batch_builds_and_evaluates(device):
    """Eight problems of the batched workload generator (uneven meshes)."""
    import torch
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    probs = W.batched_problems(count=8, min_elements=200, max_elements=2000)
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
    u = torch.as_tensor(np.concatenate([p["u"] for p in probs]).reshape(-1), device=device)
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(u, fint, values)
    sc = b.result()
    assert (sc["failed"] == 0).all()
    assert (sc["yielding"] > 0).any()
    # one problem alone, bit for bit
    p = 5
    s = G.NonlinearSystem(probs[p]["coords"], probs[p]["conn"], m, rule)
    ev = s.evaluate(probs[p]["u"])
    d0, d1 = 3 * b.node_offset[p], 3 * b.node_offset[p + 1]
    v0, v1 = b.value_offset[p], b.value_offset[p + 1]
    assert np.array_equal(fint.cpu().numpy()[d0:d1], ev.internal)
    assert np.array_equal(values.cpu().numpy()[v0:v1], ev.tangent_values)
