"""The batched form's block solve (sparlab_system_solve_free_block_segments_device):
the union's free-free block is block-diagonal over the problems, and every
problem's segment must be solved exactly as the plain block solve solves that
problem alone, which the block-solve tests pin by residuals.

Bars: per problem, x, the iteration count, the relative residual and how the
solve ended equal the plain solve's bit for bit (the same block cut, partial
sums, scalars, reads and restarts) for conjugate gradients (J2 + Voce: a
symmetric tangent) and BiCGStab (Hill48 + two Chaboche backstresses), with
either preconditioner: the diagonal, or the multigrid V-cycle (each
problem's own hierarchy, the problems' set-ups and V-cycles in shared
launches; the problems here coarsen to one, two and three levels); a
problem whose block is not positive definite fails alone (the others still
equal their plain solves); an iteration limit leaves every problem with the
plain solve's last iterate and residual; segments that cut a problem's
coupled rows apart are refused.
"""
import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

SIZES = ((6, 6, 2), (10, 8, 2), (4, 4, 1), (12, 12, 2), (7, 9, 2), (28, 24, 2))


def problems():
    out = []
    for i, (nx, ny, nz) in enumerate(SIZES):
        fam = W.FAMILIES[i % len(W.FAMILIES)]
        coords, conn = W.structured_hex_mesh(nx, ny, nz, 0.012, 0.012, 0.001, -0.006, -0.006, -0.001)
        w = W._family_depth(fam, coords[:, 0], coords[:, 1], 0.006, 0.003, 0.5)
        u = np.zeros_like(coords)
        u[:, 2] = 0.04 * w
        out.append(dict(coords=coords, conn=conn, u=u))
    return out


def rim_free(coords):
    """Every DOF except those of the rim's nodes (the clamped edge)."""
    x, y = coords[:, 0], coords[:, 1]
    rim = np.isclose(x, x.min()) | np.isclose(x, x.max()) | np.isclose(y, y.min()) | np.isclose(y, y.max())
    return np.nonzero(np.repeat(~rim, 3))[0].astype(np.int32)


def law(name):
    if name == "J2":
        return G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0)
    return W.microbench_material()


def setup(device, name):
    import torch
    probs = problems()
    m = law(name)
    rule = G.Rule(2, 0, 2)
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
    u = torch.as_tensor(np.concatenate([p["u"] for p in probs]).reshape(-1), device=device)
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=device)
    b.evaluate_device(u, fint, values, want_tangent=True, column_major=False)
    sc = b.result()
    assert (sc["failed"] == 0).all()
    frees = [rim_free(p["coords"]) for p in probs]
    free = np.concatenate([f + 3 * b.node_offset[i] for i, f in enumerate(frees)]).astype(np.int32)
    b.system.partition_device(torch.as_tensor(free, device=device))
    block = b.system.free_block_device(values).clone()
    rows = b.free_segments(free)
    rhs = -fint[torch.as_tensor(free, device=device).long()]
    return probs, m, rule, b, frees, rows, block, rhs, m.symmetric_tangent


def plain_solve(device, p, m, rule, free, negate=False, precond="diagonal", levels=None, **kw):
    """The problem alone: (x, iterations, relative residual, status); the
    multigrid hierarchy's level count appended to levels."""
    import torch
    s = G.NonlinearSystem(p["coords"], p["conn"], m, rule)
    s.set_block_preconditioner(precond)
    u = torch.as_tensor(p["u"].reshape(-1), device=device)
    f = torch.zeros_like(u)
    values = torch.zeros(s.nnz, dtype=torch.float64, device=device)
    s.evaluate_device(u, f, values, want_tangent=True, column_major=False)
    s.result()
    d_free = torch.as_tensor(free, device=device)
    s.partition_device(d_free)
    block = s.free_block_device(values).clone()
    if negate:
        block.neg_()
    rhs = -f[d_free.long()]
    x = torch.zeros_like(rhs)
    try:
        _, its, rel = s.solve_free_block_device(block, rhs, x=x, **kw)
        if levels is not None and precond == "multigrid":
            levels.append(len(s.block_preconditioner_levels()[0]))
        return x.cpu().numpy(), its, rel, 0
    except G.SolverError:
        return x.cpu().numpy(), None, None, 1


@pytest.mark.parametrize("precond", ["diagonal", "multigrid"])
@pytest.mark.parametrize("name", ["J2", "Hill48_Chaboche"])
def test_segments_equal_plain_solves(device, name, precond):
    probs, m, rule, b, frees, rows, block, rhs, sym = setup(device, name)
    assert sym == (name == "J2")
    b.system.set_block_preconditioner(precond)
    print(f"segmented block solve: {name}, {precond}, {len(probs)} problems, {rhs.numel()} rows", flush=True)
    x, its, rel, status = b.system.solve_free_block_segments_device(block, rhs, rows, tolerance=1e-10,
                                                                    symmetric=sym)
    x = x.cpu().numpy()
    assert (status == 0).all(), status
    assert len(set(its.tolist())) > 1   # the problems stop at different iterations
    levels = []
    for p, prob in enumerate(probs):
        xp, itp, relp, stp = plain_solve(device, prob, m, rule, frees[p], precond=precond, levels=levels,
                                         tolerance=1e-10, symmetric=sym)
        assert stp == 0
        assert np.array_equal(x[rows[p]:rows[p + 1]], xp), f"problem {p}: x"
        assert its[p] == itp and rel[p] == relp, (p, its[p], itp, rel[p], relp)
    if precond == "multigrid":
        assert {1, 2, 3} <= set(levels), levels
        # again: the segments' hierarchies are kept and set up anew
        x2, its2, rel2, _ = b.system.solve_free_block_segments_device(block, rhs, rows, tolerance=1e-10,
                                                                      symmetric=sym)
        assert np.array_equal(x2.cpu().numpy(), x) and np.array_equal(its2, its) and np.array_equal(rel2, rel)


def test_a_failing_problem_fails_alone(device):
    """Problem 1's block negated: conjugate gradients refuse it at the start
    (a negative diagonal); the others converge as they do alone."""
    probs, m, rule, b, frees, rows, block, rhs, sym = setup(device, "J2")
    ptr, _ = b.system.partition_pattern_device(device)
    ptr = ptr.cpu().numpy()
    block[int(ptr[rows[1]]):int(ptr[rows[2]])].neg_()
    x, its, rel, status = b.system.solve_free_block_segments_device(block, rhs, rows, tolerance=1e-10,
                                                                    symmetric=True)
    x = x.cpu().numpy()
    assert status[1] != 0 and (np.delete(status, 1) == 0).all(), status
    for p, prob in enumerate(probs):
        xp, itp, relp, stp = plain_solve(device, prob, m, rule, frees[p], negate=(p == 1), tolerance=1e-10,
                                         symmetric=True)
        assert stp == (1 if p == 1 else 0)
        assert np.array_equal(x[rows[p]:rows[p + 1]], xp), f"problem {p}: x"
        if p != 1:
            assert its[p] == itp and rel[p] == relp


@pytest.mark.parametrize("precond,limit", [("diagonal", 40), ("multigrid", 4)])
@pytest.mark.parametrize("name", ["J2", "Hill48_Chaboche"])
def test_iteration_limit(device, name, precond, limit):
    probs, m, rule, b, frees, rows, block, rhs, sym = setup(device, name)
    b.system.set_block_preconditioner(precond)
    x, its, rel, status = b.system.solve_free_block_segments_device(block, rhs, rows, tolerance=1e-12,
                                                                    max_iterations=limit, symmetric=sym)
    x = x.cpu().numpy()
    assert (status != 0).sum() >= 3
    for p, prob in enumerate(probs):
        xp, itp, relp, stp = plain_solve(device, prob, m, rule, frees[p], precond=precond, tolerance=1e-12,
                                         max_iterations=limit, symmetric=sym)
        assert stp == (1 if status[p] != 0 else 0), p
        assert np.array_equal(x[rows[p]:rows[p + 1]], xp), f"problem {p}: x"
        if stp == 0:
            assert its[p] == itp and rel[p] == relp
        else:
            assert its[p] == limit


def test_segments_must_not_cut_coupled_rows(device):
    probs, m, rule, b, frees, rows, block, rhs, sym = setup(device, "J2")
    cut = rows.copy()
    cut[2] += 3   # problem 2's first node moved into problem 1
    with pytest.raises(RuntimeError, match="couples two segments"):
        b.system.solve_free_block_segments_device(block, rhs, cut, symmetric=True)
    with pytest.raises(RuntimeError, match="cover"):
        b.system.solve_free_block_segments_device(block, rhs, rows[:-1], symmetric=True)
