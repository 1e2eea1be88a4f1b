"""Parity of the device contact increment set-up with the oracle:
ToolContact::begin_increment (src/fem/RigidTool.cpp:345-393,
sparlab_contact_candidates_device) and the semismooth anticipation
ToolContact::anticipate (:557-603, sparlab_contact_anticipate_device), on the
same inputs.

Bars: the candidate lists (slave indices, in slave order) and counts
bit-exact; the anticipation's flags and count exact, its residual and blocks
within 1e-13 of their largest entry, as the contact evaluation's (the gap
g = d - R cancels: one rounding of d, which the device's fused multiply-adds
place differently, is about 1e-14 of g); its blocks reach the assembled
matrix through add_node_blocks as the oracle's add_node_block.
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = {"plane": 0, "sphere": 1, "cylinder": 2}


def tools(shape, radius=0.005):
    """The same tool as a device record and an oracle record: a path from c0
    to c1 over the plate's top, at a slant."""
    c0, c1 = (-0.003, 0.001, radius - 1e-4), (0.002, -0.002, radius - 3e-4)
    kw = dict(radius=radius, normal=(0.1, -0.2, 1.0), axis=(0.3, 1.0, 0.1))
    if shape == "plane":
        c0, c1 = (0.0, 0.0, 1e-4), (0.0, 0.0, -1e-4)
    dev = G.make_tool(shape, c1, c0, **kw)
    ora = O.tool_record(SHAPES[shape], c1, c0, radius=radius, normal=kw["normal"], axis=kw["axis"])
    return dev, ora


def plate():
    coords, _ = W.structured_hex_mesh(60, 60, 2, 0.02, 0.02, 0.002, -0.01, -0.01, -0.002)
    top = np.nonzero(np.isclose(coords[:, 2], 0.0))[0].astype(np.int32)
    rng = np.random.default_rng(17)
    u = rng.normal(0, 4e-5, coords.shape)
    return coords, top, u


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_candidates_match_oracle(device, shape):
    import torch
    coords, top, u0 = plate()
    nodes = np.random.default_rng(4).permutation(top)
    dev_tool, ora_tool = tools(shape)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    for margin in (0.0, 2e-4, 1e-3):
        want = O.contact_candidates(ora_tool, coords, u0, nodes, margin)
        cand, count = G.contact_candidates_device(dev_tool, t(coords.ravel()), t(u0.ravel()), t(nodes, torch.int32),
                                                  margin)
        torch.cuda.synchronize()
        n = int(count.item())
        assert n == want.size and np.array_equal(cand.cpu().numpy()[:n], want), (shape, margin)
    assert 0 < want.size < nodes.size
    # no slaves: an empty list
    cand, count = G.contact_candidates_device(dev_tool, t(coords.ravel()), t(u0.ravel()),
                                              torch.zeros(0, dtype=torch.int32, device=device), 1e-3)
    torch.cuda.synchronize()
    assert int(count.item()) == 0


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_anticipation_matches_oracle(device, shape):
    import torch
    coords, top, u = plate()
    nn = coords.shape[0]
    rng = np.random.default_rng(23)
    du = np.zeros_like(u)
    du[:, 2] = rng.uniform(-6e-4, 2e-4, nn)  # some top nodes pushed into the tool
    dev_tool, ora_tool = tools(shape)
    cands = O.contact_candidates(ora_tool, coords, u, top, 1e-3)
    nodes = top[cands]
    ka = rng.uniform(0.5e7, 2e7, nodes.size)
    r_want, b_want, a_want, n_want = O.contact_anticipate(ora_tool, coords, u, du, nodes, ka)
    assert 0 < n_want < nodes.size
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    residual = torch.zeros(3 * nn, dtype=torch.float64, device=device)
    blocks, active, added = G.contact_anticipate_device(dev_tool, t(coords.ravel()), t(u.ravel()), t(du.ravel()),
                                                        t(nodes, torch.int32), t(ka), residual)
    torch.cuda.synchronize()
    assert np.array_equal(active.cpu().numpy(), a_want) and int(added.item()) == n_want
    r, b = residual.cpu().numpy(), blocks.cpu().numpy().reshape(-1, 9)
    assert np.abs(r - r_want).max() <= 1e-13 * np.abs(r_want).max()
    assert np.abs(b - b_want).max() <= 1e-13 * np.abs(b_want).max()
    # the blocks into the assembled matrix, as add_node_block
    conn = W.structured_hex_mesh(60, 60, 2, 0.02, 0.02, 0.002, -0.01, -0.01, -0.002)[1]
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33))
    values = torch.zeros(sysm.nnz, dtype=torch.float64, device=device)
    row_ptr, _ = sysm.pattern()
    first = t(sysm.diagonal_blocks()[nodes], torch.int32)
    row_len = t((row_ptr[3 * nodes + 1] - row_ptr[3 * nodes]).astype(np.int32), torch.int32)
    G.add_node_blocks_device(blocks, active, first, row_len, values, column_major=True)
    torch.cuda.synchronize()
    ref = O.add_node_blocks(nn, conn, np.zeros(sysm.nnz), nodes[a_want == 1], b_want[a_want == 1])
    assert np.abs(values.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
