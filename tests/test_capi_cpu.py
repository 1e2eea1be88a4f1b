"""The C ABI without a device: the library loads and exports every symbol
include/sparlab_gpu.h declares; the host-side parts (material derivation,
integration rules, state layout, the sparsity pattern) against the oracle.

The pattern is the reference's (src/fem/Assembler.cpp:86-173); its known
answers: the brute-force pattern of all element DOF pairs (the triplet path,
Assembler.cpp:228-259), the closed-form nnz of a structured grid
(9 prod_d (3 n_d - 2)), structural symmetry, and the oracle's restatement bit
for bit.
"""
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from sparlab_gpu import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sparlab_gpu.h")).read()
    return sorted(set(re.findall(r"\b(sparlab_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = G.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"{name} is declared in include/sparlab_gpu.h but not exported"
    assert lib.sparlab_version().decode().startswith("sparlab-gpu")


def test_material_derivation_matches_oracle():
    for hill in (dict(r0=0.7, r45=0.6, r90=0.8), dict(r0=1.9, r45=1.5, r90=2.3, rolling_direction=(1.0, 0.5, 0.0)),
                 dict(r0=1.9, r45=1.5, r90=2.3, rolling_direction=(1.0, 0.4, 0.3), sheet_normal=(-0.3, 0.1, 1.0))):
        m = G.Material(70e9, 0.33, yield_stress=95e6, hill=hill)
        p, _ = O.hill_yield_matrix(hill["r0"], hill["r45"], hill["r90"], rd=hill.get("rolling_direction", (1, 0, 0)),
                                   nd=hill.get("sheet_normal", (0, 0, 1)))
        assert np.abs(m.yield_matrix - p).max() <= 1e-15 * np.abs(p).max()
    vm = G.Material(200e9, 0.3, yield_stress=250e6)
    assert np.array_equal(vm.yield_matrix, O.von_mises_yield_matrix())
    with pytest.raises(RuntimeError):
        G.Material(200e9, 0.3, hardening_modulus=1e9)          # hardening without a yield stress
    with pytest.raises(RuntimeError):
        G.Material(200e9, 0.6)                                  # nu outside (-1, 0.5)
    with pytest.raises(RuntimeError):
        G.Material(200e9, 0.3, yield_stress=1e8, hill=dict(r0=-1.0))


def test_rules_and_state_layout():
    for rule in ((2, 0, 2), (2, 2, 2), (1, 0, 2), (2, 1, 2), (2, 1, 0), (2, 5, 2), (2, 7, 2), (2, 7, 0), (3, 0, 2)):
        r = G.Rule(*rule)
        assert r.points == O.rule_points(*rule)
    with pytest.raises(RuntimeError):
        G.Rule(4, 0, 2).points            # 64 points: more than the kernels take
    with pytest.raises(RuntimeError):
        G.Rule(5, 0, 2).points
    m = W.microbench_material()
    assert m.state_fields == 8 + 6 * 2
    ref = W.microbench(16, seed=1)["states"]
    back = G.states_to_reference(G.states_from_reference(ref, m), m, 16, 8)
    assert np.array_equal(back[:, :, 0:6], ref[:, :, 0:6])
    assert np.array_equal(back[:, :, 12:24], ref[:, :, 12:24])
    assert np.array_equal(back[:, :, 36], ref[:, :, 36])


def brute_force_pattern(nn, conn):
    rows = set()
    for e in range(conn.shape[0]):
        dofs = [3 * n + k for n in conn[e] for k in range(3)]
        for i in dofs:
            for j in dofs:
                rows.add((i, j))
    rows = sorted(rows)
    row_ptr = np.zeros(3 * nn + 1, dtype=np.int64)
    for i, _ in rows:
        row_ptr[i + 1] += 1
    return np.cumsum(row_ptr), np.array([j for _, j in rows])


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 2), (5, 4, 3)])
def test_pattern_brute_force_and_closed_form(dims):
    nx, ny, nz = dims
    coords, conn = W.structured_hex_mesh(nx, ny, nz, 1.0, 1.0, 1.0)
    nn = coords.shape[0]
    row_ptr, col_idx, bo = G.pattern(nn, conn)
    rp, ci = brute_force_pattern(nn, conn)
    assert np.array_equal(row_ptr, rp) and np.array_equal(col_idx, ci)
    # 9 prod_d (3 n_d - 2) with n_d = d + 1 nodes per direction
    assert col_idx.size == 9 * np.prod([3 * (n + 1) - 2 for n in dims])
    o_outer, o_inner, o_bo = O.pattern(nn, conn)
    assert np.array_equal(row_ptr, o_outer) and np.array_equal(col_idx, o_inner) and np.array_equal(bo, o_bo)


def test_pattern_of_the_assembly_workload():
    """100 x 100 x 4: 51 005 nodes, nnz = 9 * 301 * 301 * 13 = 10 600 317."""
    coords, conn = W.structured_hex_mesh(100, 100, 4, 0.1, 0.1, 0.004)
    assert coords.shape[0] == 51005
    row_ptr, col_idx, _ = G.pattern(coords.shape[0], conn)
    assert col_idx.size == 10_600_317
    o_outer, o_inner, _ = O.pattern(coords.shape[0], conn)
    assert np.array_equal(row_ptr, o_outer) and np.array_equal(col_idx, o_inner)


def test_pattern_of_an_unstructured_mesh():
    """A perturbed numbering: elements in random order, nodes renumbered."""
    coords, conn = W.structured_hex_mesh(4, 3, 2, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(4)
    perm = rng.permutation(coords.shape[0]).astype(np.int32)
    conn = perm[conn][rng.permutation(conn.shape[0])]
    nn = coords.shape[0]
    row_ptr, col_idx, bo = G.pattern(nn, conn)
    rp, ci = brute_force_pattern(nn, conn)
    assert np.array_equal(row_ptr, rp) and np.array_equal(col_idx, ci)
    o_outer, o_inner, o_bo = O.pattern(nn, conn)
    assert np.array_equal(bo, o_bo)


def test_pattern_refuses_bad_connectivity():
    coords, conn = W.structured_hex_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    bad = conn.copy()
    bad[0, 3] = 99
    with pytest.raises(RuntimeError):
        G.pattern(coords.shape[0], bad)
    bad = conn.copy()
    bad[0, 3] = bad[0, 2]
    with pytest.raises(RuntimeError):
        G.pattern(coords.shape[0], bad)


def test_union_pattern_is_block_diagonal():
    """The batched form's premise: problems that share no nodes give a union
    pattern whose rows of problem p are problem p's pattern shifted by its
    first DOF, and whose values of problem p are one contiguous slice."""
    meshes = [W.structured_hex_mesh(*d, 1.0, 1.0, 1.0) for d in ((2, 2, 1), (3, 1, 2), (1, 1, 1), (4, 3, 1))]
    node_off = np.cumsum([0] + [c.shape[0] for c, _ in meshes])
    uconn = np.concatenate([k + node_off[i] for i, (_, k) in enumerate(meshes)]).astype(np.int32)
    row_ptr, col_idx, _ = G.pattern(int(node_off[-1]), uconn)
    for i, (c, k) in enumerate(meshes):
        rp, ci, _ = G.pattern(c.shape[0], k)
        r0, r1 = 3 * node_off[i], 3 * node_off[i + 1]
        v0, v1 = row_ptr[r0], row_ptr[r1]
        assert np.array_equal(row_ptr[r0:r1 + 1] - v0, rp)
        assert np.array_equal(col_idx[v0:v1] - 3 * node_off[i], ci)


def test_workspace_size_mirrors_the_library():
    """sparlab_gpu.workspace_doubles (used to size device scratch from Python)
    equals the library's sparlab_element_workspace_doubles."""
    import ctypes
    lib = G.lib()
    for points_rule in (G.Rule(2, 0, 2), G.Rule(1, 0, 2), G.Rule(2, 1, 2)):
        for ne in (0, 1, 7, 1000, 1_000_003):
            got = lib.sparlab_element_workspace_doubles(ne, ctypes.byref(points_rule.c))
            assert got == G.workspace_doubles(ne, points_rule.points), (ne, points_rule.points)


def test_element_batch_refuses_a_misaligned_workspace():
    """The condensation copies workspace records in 16-byte chunks: a caller's
    workspace off a 16-byte boundary is refused before any device call (the
    pointers below are never dereferenced)."""
    import ctypes
    lib = G.lib()
    mat = G.Material(200e9, 0.3, yield_stress=250e6)
    rule = G.Rule(2, 0, 2)
    fake = 1 << 20
    b = G.CElementBatch(*([ctypes.c_void_p(fake)] * 3), 8, 1, *([ctypes.c_void_p(fake)] * 11),
                        ctypes.c_void_p(fake + 8), None)
    lib.sparlab_last_error.restype = ctypes.c_char_p
    rc = lib.sparlab_elements_evaluate_device(ctypes.byref(mat.c), ctypes.byref(rule.c), 2, 1, ctypes.byref(b), None)
    assert rc == -1
    assert b"16-byte aligned" in lib.sparlab_last_error()


def test_oracle_stiffness_gross_exact_on_integers():
    """stiffness_gross (NonlinearSystem.hpp:55-67) of the oracle: with small
    integer values and displacements every product and sum is exact in FP64,
    so |K| |x| must equal the exact integer product bit for bit (a transposed
    or misplaced value changes it: K is not symmetric), and the norm over the
    free DOFs equals the exact norm to rounding; with random real values it
    matches a dense |K| |x|."""
    import scipy.sparse as sp
    coords, conn = W.structured_hex_mesh(4, 3, 2, 1.0, 1.0, 1.0)
    nn = coords.shape[0]
    outer, inner, _ = O.pattern(nn, conn)
    nnz = inner.size
    rng = np.random.default_rng(12)
    values = rng.integers(-8, 9, nnz).astype(np.float64)
    x = rng.integers(-5, 6, 3 * nn).astype(np.float64)
    free = np.sort(rng.choice(3 * nn, 2 * nn, replace=False)).astype(np.int32)
    k = sp.csc_matrix((values, inner, outer), shape=(3 * nn, 3 * nn))
    exact = np.abs(k).astype(np.int64) @ np.abs(x).astype(np.int64)
    norm, product = O.stiffness_gross(nn, conn, values, x, free)
    assert np.array_equal(product, exact.astype(np.float64))
    assert not np.array_equal(product, (np.abs(k).T @ np.abs(x)))
    assert abs(norm - np.sqrt(float(np.sum(exact[free].astype(np.float64) ** 2)))) <= 1e-15 * norm
    values = rng.normal(size=nnz)
    x = rng.normal(size=3 * nn)
    k = sp.csc_matrix((values, inner, outer), shape=(3 * nn, 3 * nn))
    norm, product = O.stiffness_gross(nn, conn, values, x, free)
    dense = np.abs(k.toarray()) @ np.abs(x)
    assert np.abs(product - dense).max() <= 1e-14 * np.abs(dense).max()


@pytest.mark.parametrize("which", ["clamped_edges", "random", "all", "none"])
def test_oracle_free_block_is_the_submatrix(which):
    """PartitionedFactor::build + assemble (Forming.cpp:89-98, 208-236) of the
    oracle against scipy's K[free][:, free] of the column-major matrix, with
    values that are not symmetric (a transposed block, a dropped or misplaced
    entry or a renumbering off by one fails): pattern and values equal, and
    position sends exactly the entries whose row and column are both free."""
    import scipy.sparse as sp
    coords, conn = W.structured_hex_mesh(5, 4, 2, 1.0, 1.0, 0.5)
    nn = coords.shape[0]
    n = 3 * nn
    outer, inner, _ = O.pattern(nn, conn)
    rng = np.random.default_rng(31)
    values = rng.normal(size=inner.size)
    if which == "clamped_edges":
        fixed = np.nonzero(np.isclose(coords[:, 0], 0.0) | np.isclose(coords[:, 0], 1.0))[0]
        mask = np.ones(n, bool)
        mask[(3 * fixed[:, None] + np.arange(3)).ravel()] = False
        mask[2::3] &= ~np.isclose(coords[:, 2], 0.0)  # and the bottom's z
        free = np.nonzero(mask)[0]
    elif which == "random":
        free = np.sort(rng.choice(n, n // 2, replace=False))
    elif which == "all":
        free = np.arange(n)
    else:
        free = np.zeros(0, np.int64)
    free = free.astype(np.int32)
    b_outer, b_inner, b_values, position = O.free_block(nn, conn, values, free)
    k = sp.csc_matrix((values, inner, outer), shape=(n, n))
    sub = k[free][:, free].tocsc()
    sub.sort_indices()
    assert np.array_equal(b_outer, sub.indptr) and np.array_equal(b_inner, sub.indices)
    assert np.array_equal(b_values, sub.data)
    cols = np.repeat(np.arange(n), np.diff(outer))
    both = np.isin(inner, free) & np.isin(cols, free)
    assert np.array_equal(position >= 0, both)
    assert np.array_equal(np.sort(position[both]), np.arange(b_inner.size))
    if which == "all":
        assert np.array_equal(b_values, values)


def test_device_wrappers_refuse_tensors_the_library_would_overrun():
    """The wrappers of device entry points that take no sizes check dtype,
    contiguity, device and length before calling the library (a wrong tensor
    would otherwise be read or written out of bounds on the GPU)."""
    import types
    import torch
    fake = types.SimpleNamespace(nnz=10, num_nodes=4, _block=(5, 7), h=None)
    cpu = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(ValueError, match="device"):
        G.NonlinearSystem.free_block_device(fake, cpu)
    with pytest.raises(ValueError, match="float64"):
        G.NonlinearSystem.free_block_device(fake, torch.zeros(10, dtype=torch.float32))
    with pytest.raises(ValueError, match="contiguous"):
        G.NonlinearSystem.free_block_device(fake, torch.zeros(20, dtype=torch.float64)[::2])
    with pytest.raises(ValueError):
        G.NonlinearSystem.solve_free_block_device(fake, torch.zeros(7, dtype=torch.float64), cpu[:5])
    with pytest.raises(ValueError):
        G.NonlinearSystem.stiffness_gross_device(fake, torch.zeros(9, dtype=torch.float64), cpu[:12],
                                                 torch.zeros(3, dtype=torch.int32))
    with pytest.raises(ValueError, match="int32"):
        G.NonlinearSystem.partition_device(fake, torch.zeros(3, dtype=torch.int64))
    with pytest.raises(ValueError, match="device|needed"):
        G._require(torch.zeros(3, dtype=torch.float64), "t", 4, device=None)
