"""Parity of the device free-free block (sparlab_system_partition_device,
sparlab_system_partition_pattern_device, sparlab_system_free_block_device)
with the oracle's PartitionedFactor::build and assemble
(src/fem/Forming.cpp:89-98, 208-236), on the same inputs.

Bar: bit-exact (index work and copies): the block's pointers and indices
equal the oracle's, its values equal the oracle's for the reference's
column-major values, and for row-major values the block's rows (the
transpose of the column-major block of K^T). Edge cases: every DOF free (the
block is the whole matrix), none, a single DOF, clamped edges, and lists
that do not strictly ascend or leave the range (refused, no partition).
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def mesh():
    coords, conn = W.structured_hex_mesh(9, 7, 3, 0.009, 0.007, 0.003)
    return coords, conn


def free_lists(coords, rng):
    n = 3 * coords.shape[0]
    mask = np.ones(n, bool)
    fixed = np.nonzero(np.isclose(coords[:, 0], coords[:, 0].min()) | np.isclose(coords[:, 0], coords[:, 0].max()))[0]
    mask[(3 * fixed[:, None] + np.arange(3)).ravel()] = False
    return {
        "clamped": np.nonzero(mask)[0],
        "random": np.sort(rng.choice(n, n // 3, replace=False)),
        "all": np.arange(n),
        "one": np.array([n // 2]),
        "none": np.zeros(0, np.int64),
    }


@pytest.mark.parametrize("which", ["clamped", "random", "all", "one", "none"])
def test_free_block_matches_oracle(device, which):
    import torch
    coords, conn = mesh()
    nn = coords.shape[0]
    rng = np.random.default_rng(41)
    free = free_lists(coords, rng)[which].astype(np.int32)
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33))
    nnz = sysm.nnz
    values = rng.normal(size=nnz)  # not symmetric: a transposed entry shows
    outer, inner, b_values, _ = O.free_block(nn, conn, values, free)
    d_free = torch.as_tensor(free, device=device)
    got_nnz = sysm.partition_device(d_free)
    assert got_nnz == inner.size
    ptr, index = sysm.partition_pattern_device(device)
    torch.cuda.synchronize()
    assert np.array_equal(ptr.cpu().numpy(), outer) and np.array_equal(index.cpu().numpy(), inner)
    cm = sysm.free_block_device(torch.as_tensor(values, device=device))
    torch.cuda.synchronize()
    assert np.array_equal(cm.cpu().numpy(), b_values)
    # row-major values of the same K (the entry (r, c) at row r) give the
    # block's rows: the column-major block of K^T
    o_full, i_full, _ = O.pattern(nn, conn)
    cols = np.repeat(np.arange(3 * nn), np.diff(o_full))
    k = {(int(r), int(c)): v for r, c, v in zip(i_full, cols, values)}
    row_major = np.array([k[(int(r), int(c))] for r, c in zip(cols, i_full)])
    rm = sysm.free_block_device(torch.as_tensor(row_major, device=device))
    torch.cuda.synchronize()
    _, _, b_transposed, _ = O.free_block(nn, conn, row_major, free)
    assert np.array_equal(rm.cpu().numpy(), b_transposed)
    # and read as rows they are K_ff itself
    import scipy.sparse as sp
    kff = sp.csc_matrix((values, i_full, o_full), shape=(3 * nn, 3 * nn))[free][:, free]
    rows = sp.csr_matrix((rm.cpu().numpy(), index.cpu().numpy(), ptr.cpu().numpy()), shape=kff.shape)
    assert (rows != kff).nnz == 0
    if which == "all":
        assert np.array_equal(cm.cpu().numpy(), values)


def test_free_block_of_the_assembled_tangent(device):
    """The block of a real assembled tangent (evaluate, then partition) is the
    oracle's block of the same values, and a second partition replaces the
    first."""
    import torch
    coords, conn = mesh()
    nn = coords.shape[0]
    m = G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0)
    sysm = G.NonlinearSystem(coords, conn, m)
    u = 2e-3 * coords[:, [2, 0, 1]]
    fint = torch.empty(3 * nn, dtype=torch.float64, device=device)
    values = torch.empty(sysm.nnz, dtype=torch.float64, device=device)
    sysm.evaluate_device(torch.as_tensor(u.reshape(-1), device=device), fint, values, column_major=True)
    torch.cuda.synchronize()
    vals = values.cpu().numpy()
    rng = np.random.default_rng(5)
    for which in ("clamped", "random"):
        free = free_lists(coords, rng)[which].astype(np.int32)
        sysm.partition_device(torch.as_tensor(free, device=device))
        got = sysm.free_block_device(values)
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), O.free_block(nn, conn, vals, free)[2])


@pytest.mark.parametrize("bad", ["descending", "repeated", "negative", "beyond"])
def test_partition_refuses_bad_lists(device, bad):
    import torch
    coords, conn = mesh()
    n = 3 * coords.shape[0]
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33))
    free = {"descending": [5, 4, 9], "repeated": [3, 7, 7, 8], "negative": [-1, 2, 3], "beyond": [0, 1, n]}[bad]
    with pytest.raises(RuntimeError, match="strictly ascending"):
        sysm.partition_device(torch.as_tensor(np.array(free, np.int32), device=device))
    with pytest.raises(RuntimeError, match="no partition"):
        sysm.free_block_device(torch.zeros(sysm.nnz, dtype=torch.float64, device=device))
    with pytest.raises(RuntimeError, match="no partition"):
        sysm.partition_pattern_device(device)
    # a good list afterwards partitions as usual
    assert sysm.partition_device(torch.arange(n, dtype=torch.int32, device=device)) == sysm.nnz
