"""The device solve of the free-free block (sparlab_system_solve_free_block_device)
against a direct solve of the same system: where the reference factorises
and solves a symmetric tangent (PartitionedFactor::factorize and solve,
src/fem/Forming.cpp:104-172), the device runs preconditioned conjugate
gradients (BiCGStab for a non-symmetric block), the preconditioner the
block's diagonal or a smoothed-aggregation multigrid V-cycle, so the check is
the mathematics of K_ff x = b, not a copy of the factorisation's round-off.
Every check runs with both preconditioners.

Bars: the reported and the recomputed true residual ||b - K x|| / ||b|| both
within the tolerance asked (1e-12); the solution within 1e-7 of scipy's
sparse LU solution (the block's condition number times the tolerance bounds
the difference); two solves bit-identical (fixed-order sums, the multigrid
set-up included); b = 0 gives x = 0 after no iteration; a block that is not
positive definite and an iteration limit that is too small raise
SolverError; b and x that overlap are refused. The multigrid hierarchy of a
40 x 30 x 4 plate has at least three levels, takes at most a tenth of the
diagonal's iterations, and is rebuilt alike after new coordinates and a new
partition.
"""
import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


PRECONDITIONERS = ["diagonal", "multigrid"]


def plate_system(device, material=None, nx=12, ny=10, nz=2, precond="diagonal"):
    import torch
    coords, conn = W.structured_hex_mesh(nx, ny, nz, 0.012, 0.010, 0.002)
    m = material or G.Material(70e9, 0.33)
    sysm = G.NonlinearSystem(coords, conn, m)
    sysm.set_block_preconditioner(precond)
    nn = coords.shape[0]
    u = torch.zeros(3 * nn, dtype=torch.float64, device=device)
    fint = torch.empty(3 * nn, dtype=torch.float64, device=device)
    values = torch.empty(sysm.nnz, dtype=torch.float64, device=device)
    sysm.evaluate_device(u, fint, values, column_major=True)
    ends = np.isclose(coords[:, 0], coords[:, 0].min()) | np.isclose(coords[:, 0], coords[:, 0].max())
    free = np.nonzero(np.repeat(~ends, 3))[0].astype(np.int32)
    sysm.partition_device(torch.as_tensor(free, device=device))
    block = sysm.free_block_device(values)
    ptr, index = sysm.partition_pattern_device(device)
    torch.cuda.synchronize()
    return sysm, block, ptr, index, free


def as_scipy(block, ptr, index):
    import scipy.sparse as sp
    n = ptr.shape[0] - 1
    return sp.csr_matrix((block.cpu().numpy(), index.cpu().numpy(), ptr.cpu().numpy()), shape=(n, n))


@pytest.mark.parametrize("precond", PRECONDITIONERS)
def test_block_solve_matches_direct_solve(device, precond):
    import torch
    from scipy.sparse.linalg import spsolve
    sysm, block, ptr, index, free = plate_system(device, precond=precond)
    k = as_scipy(block, ptr, index)
    assert abs(k - k.T).max() <= 1e-12 * abs(k).max()  # the elastic tangent is symmetric
    rng = np.random.default_rng(3)
    b = rng.normal(size=free.size) * 1e3
    d_b = torch.as_tensor(b, device=device)
    x, its, rel = sysm.solve_free_block_device(block, d_b, tolerance=TOL)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    true_rel = np.linalg.norm(b - k @ xs) / np.linalg.norm(b)
    assert rel <= TOL and true_rel <= 1.5 * TOL and abs(true_rel - rel) <= 1e-3 * TOL + 0.05 * rel
    ref = spsolve(k.tocsc(), b)
    assert np.linalg.norm(xs - ref) <= 1e-7 * np.linalg.norm(ref), (its, rel)
    assert its > 0
    # the same solve again, bit for bit
    x2, its2, rel2 = sysm.solve_free_block_device(block, d_b, tolerance=TOL)
    torch.cuda.synchronize()
    assert its2 == its and rel2 == rel and torch.equal(x2, x)


@pytest.mark.parametrize("precond", PRECONDITIONERS)
def test_block_solve_of_a_zero_right_hand_side(device, precond):
    import torch
    sysm, block, ptr, index, free = plate_system(device, nx=4, ny=3, precond=precond)
    b = torch.zeros(free.size, dtype=torch.float64, device=device)
    x, its, rel = sysm.solve_free_block_device(block, b, tolerance=TOL)
    torch.cuda.synchronize()
    assert its == 0 and rel == 0.0 and (x == 0).all()


@pytest.mark.parametrize("precond", PRECONDITIONERS)
def test_block_solve_refuses_what_it_cannot_solve(device, precond):
    import torch
    sysm, block, ptr, index, free = plate_system(device, nx=6, ny=5, precond=precond)
    b = torch.as_tensor(np.random.default_rng(8).normal(size=free.size), device=device)
    with pytest.raises(G.SolverError, match="not positive definite"):
        sysm.solve_free_block_device(-block, b, tolerance=TOL)
    with pytest.raises(G.SolverError, match="did not converge"):
        sysm.solve_free_block_device(block, b, tolerance=TOL, max_iterations=1)
    with pytest.raises(RuntimeError, match="overlap"):
        sysm.solve_free_block_device(block, b, x=b, tolerance=TOL)
    # still solvable afterwards
    _, _, rel = sysm.solve_free_block_device(block, b, tolerance=TOL)
    assert rel <= TOL


@pytest.mark.parametrize("precond", PRECONDITIONERS)
def test_nonsymmetric_block_solve_matches_direct_solve(device, precond):
    """BiCGStab (symmetric = 0) on a block that is not symmetric: the plate's
    elastic block plus, on every node's 3 x 3 diagonal block, c (I + S) with S
    a seeded skew matrix (the shape frictional contact adds: a penalty and a
    non-symmetric tangential part), read by rows; against scipy's sparse LU
    solve of the same matrix."""
    import torch
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve
    sysm, block, ptr, index, free = plate_system(device, precond=precond)
    k = as_scipy(block, ptr, index)
    rng = np.random.default_rng(21)
    node = free // 3
    c = 0.5 * k.diagonal().mean()
    rows, cols, vals = [], [], []
    for start in np.nonzero(np.r_[True, node[1:] != node[:-1]])[0]:
        m = int(np.sum(node[start:start + 3] == node[start]))
        a = rng.uniform(-1, 1, (m, m))
        blk = c * (np.eye(m) + (a - a.T))
        for i in range(m):
            for j in range(m):
                rows.append(start + i)
                cols.append(start + j)
                vals.append(blk[i, j])
    data = k.data.copy()  # in place: the block's own entries (explicit zeros kept)
    for r, col, v in zip(rows, cols, vals):
        lo, hi = k.indptr[r], k.indptr[r + 1]
        data[lo + np.searchsorted(k.indices[lo:hi], col)] += v
    kn = sp.csr_matrix((data, k.indices, k.indptr), shape=k.shape)
    assert abs(kn - kn.T).max() > 1e-3 * abs(kn).max()
    d_block = torch.as_tensor(kn.data, device=device)
    b = rng.normal(size=free.size) * 1e3
    d_b = torch.as_tensor(b, device=device)
    x, its, rel = sysm.solve_free_block_device(d_block, d_b, tolerance=TOL, symmetric=False)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    true_rel = np.linalg.norm(b - kn @ xs) / np.linalg.norm(b)
    assert rel <= TOL and true_rel <= 1.5 * TOL
    ref = spsolve(kn.tocsc(), b)
    assert np.linalg.norm(xs - ref) <= 1e-7 * np.linalg.norm(ref), (its, rel)
    x2, its2, rel2 = sysm.solve_free_block_device(d_block, d_b, tolerance=TOL, symmetric=False)
    torch.cuda.synchronize()
    assert its2 == its and rel2 == rel and torch.equal(x2, x)
    # a symmetric block solves the same way by either method
    xs_sym, _, rel_sym = sysm.solve_free_block_device(block, d_b, tolerance=TOL, symmetric=False)
    torch.cuda.synchronize()
    ref_sym = spsolve(as_scipy(block, ptr, index).tocsc(), b)
    assert rel_sym <= TOL and np.linalg.norm(xs_sym.cpu().numpy() - ref_sym) <= 1e-7 * np.linalg.norm(ref_sym)
    # negative diagonal entries (the non-symmetric block negated, which the
    # reference's LU solves): BiCGStab preconditions with 1 / |K_ii|
    xn, _, rel_n = sysm.solve_free_block_device(-d_block, d_b, tolerance=TOL, symmetric=False)
    torch.cuda.synchronize()
    assert rel_n <= TOL and np.linalg.norm(xn.cpu().numpy() + ref) <= 1e-7 * np.linalg.norm(ref)


def test_multigrid_hierarchy_and_iterations(device):
    import torch
    from scipy.sparse.linalg import spsolve
    coords, conn = W.structured_hex_mesh(40, 30, 4, 0.012, 0.010, 0.002)  # plate_system's
    sysm, block, ptr, index, free = plate_system(device, nx=40, ny=30, nz=4)
    k = as_scipy(block, ptr, index)
    b = torch.as_tensor(np.random.default_rng(5).normal(size=free.size) * 1e3, device=device)
    x_d, its_d, rel_d = sysm.solve_free_block_device(block, b, tolerance=1e-10)
    sysm.set_block_preconditioner("multigrid")
    x_m, its_m, rel_m = sysm.solve_free_block_device(block, b, tolerance=1e-10)
    torch.cuda.synchronize()
    rows, nnz = sysm.block_preconditioner_levels()
    print(f"multigrid levels {rows} (nonzeros {nnz}); iterations {its_m} against the diagonal's {its_d}", flush=True)
    assert len(rows) >= 3 and rows[0] == free.size and rows[-1] <= 256
    assert all(rows[i + 1] < rows[i] for i in range(len(rows) - 1))
    assert its_m * 10 <= its_d
    ref = spsolve(k.tocsc(), b.cpu().numpy())
    for x in (x_d, x_m):
        assert np.linalg.norm(x.cpu().numpy() - ref) <= 1e-6 * np.linalg.norm(ref)
    # new coordinates (the same ones) and a new partition rebuild the hierarchy alike
    sysm.set_coords(coords)
    assert sysm.block_preconditioner_levels() == ([], [])
    x2, its2, rel2 = sysm.solve_free_block_device(block, b, tolerance=1e-10)
    sysm.partition_device(torch.as_tensor(free, device=device))
    x3, its3, rel3 = sysm.solve_free_block_device(block, b, tolerance=1e-10)
    torch.cuda.synchronize()
    assert its2 == its_m == its3 and rel2 == rel_m == rel3
    assert torch.equal(x2, x_m) and torch.equal(x3, x_m)
