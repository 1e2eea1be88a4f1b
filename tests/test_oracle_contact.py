"""The oracle's contact against the reference's known answers
(tests/test_forming.cpp:194-330): the node blocks are the exact derivative of
the contact residual (central differences, h = 1e-9, <= 1e-6 of the largest
block entry) for sphere, plane and cylinder, frictionless, sticking, slipping
and new to contact; slip makes the tangent non-symmetric; the symmetric
friction tangent gives the same forces, symmetric blocks, and for the plane
positive semi-definite ones.
"""
import numpy as np
import pytest

import oracle_lib as O
from sparlab_gpu import workloads as W


def block_model():
    """A 10 x 10 x 1 mm block, 20 x 20 x 1 elements, its top at z = 0."""
    coords, conn = W.structured_hex_mesh(20, 20, 1, 0.01, 0.01, 0.001, 0.0, 0.0, -0.001)
    top = np.nonzero(np.isclose(coords[:, 2], 0.0))[0].astype(np.int32)
    return coords, conn, top


@pytest.mark.parametrize("shape", ["sphere", "plane", "cylinder"])
@pytest.mark.parametrize("regime", ["frictionless", "stick", "slip", "new"])
def test_contact_tangent_is_derivative(shape, regime):
    coords, _, top = block_model()
    nn = coords.shape[0]
    rng = np.random.default_rng(7)
    u = 0.02 * rng.uniform(-1e-4, 1e-4, 3 * nn)
    start = u.copy()
    friction = 0.0 if regime == "frictionless" else (0.1 if regime == "slip" else 0.8)
    code = {"sphere": 1, "plane": 0, "cylinder": 2}[shape]
    centre = (0.005, 0.005, -2e-5) if shape == "plane" else (0.0052, 0.0047, 0.005 - 1e-4)
    tool = O.tool_record(code, centre, centre, radius=0.005, normal=(0.1, -0.2, 1.0), axis=(1.0, 0.3, 0.0),
                         friction=friction)
    # kappa A: the reference's kappa_j A_j = s E sqrt(A_j) with s = 10, E = 70 GPa, A = (0.5 mm)^2
    kappa_a = np.full(top.size, 10.0 * 70e9 * 0.0005)
    hist = None
    ev0 = O.contact(tool, coords, u, start, top, kappa_a, want_tangent=False)
    assert ev0["n_active"] >= 3
    if regime in ("stick", "slip"):
        present = np.zeros(top.size, np.int32)
        force = np.zeros((top.size, 3))
        rel = np.zeros((top.size, 3))
        c = np.array(centre)
        for i, node in enumerate(top):
            x = coords[node] + u[3 * node:3 * node + 3]
            if shape == "plane":
                nrm = np.array([0.1, -0.2, 1.0]) / np.linalg.norm([0.1, -0.2, 1.0])
                g = nrm @ (x - c)
            else:
                r = x - c
                if shape == "cylinder":
                    a = np.array([1.0, 0.3, 0.0]) / np.linalg.norm([1.0, 0.3, 0.0])
                    r = r - (a @ r) * a
                nrm = r / np.linalg.norm(r)
                g = np.linalg.norm(r) - 0.005
            if g >= 0:
                continue
            along = np.array([1.0, 0.4, 0.0])
            tangent = along - (along @ nrm) * nrm
            tangent /= np.linalg.norm(tangent)
            present[i] = 1
            force[i] = (5.0 if regime == "stick" else 50.0) * (tangent + 0.2 * nrm)
            rel[i] = (x - c) - (1e-8 if regime == "stick" else 1e-3) * tangent
        hist = (present, force, rel)
    ev = O.contact(tool, coords, u, start, top, kappa_a, history=hist)
    act = np.nonzero(ev["active"])[0]
    top_entry = np.abs(ev["blocks"][act]).max()
    h = 1e-9
    worst = 0.0
    for i in act:
        node = top[i]
        for c in range(3):
            d = 3 * node + c
            up, dn = u.copy(), u.copy()
            up[d] += h
            dn[d] -= h
            fd = (O.contact(tool, coords, up, start, top, kappa_a, history=hist, want_tangent=False)["residual"] -
                  O.contact(tool, coords, dn, start, top, kappa_a, history=hist, want_tangent=False)["residual"]) / (2 * h)
            fd[3 * node:3 * node + 3] -= ev["blocks"][i][:, c]
            worst = max(worst, np.abs(fd).max() / top_entry)
    assert worst < 1e-6
    if regime == "slip":
        assert ev["n_slipping"] == ev["n_active"]
        assert not ev["symmetric"]
    if regime in ("stick", "new"):
        assert ev["n_slipping"] == 0
    if regime == "frictionless":
        assert ev["symmetric"]
    es = O.contact(tool, coords, u, start, top, kappa_a, history=hist, friction_symmetric=True)
    assert es["symmetric"]
    assert np.abs(es["residual"] - ev["residual"]).max() == 0.0
    lowest, top_ev = 0.0, 0.0
    for i in act:
        b = es["blocks"][i]
        assert np.abs(b - b.T).max() == 0.0
        w = np.linalg.eigvalsh(b)
        lowest, top_ev = min(lowest, w.min()), max(top_ev, w.max())
    if shape == "plane":
        assert lowest >= -1e-12 * top_ev


def test_contact_refuses_a_node_through_the_tool():
    coords, _, top = block_model()
    u = np.zeros(3 * coords.shape[0])
    tool = O.tool_record(1, (0.005, 0.005, 0.0), radius=0.005)
    with pytest.raises(O.OracleError):
        O.contact(tool, coords, u, u, top, np.ones(top.size))


# ---------------------------------------------------------------------------
# The increment set-up (ToolContact::begin_increment, RigidTool.cpp:345-393)
# and the semismooth anticipation (ToolContact::anticipate, :557-603), pinned
# by case analysis independent of the oracle's projection formula.
# ---------------------------------------------------------------------------
def path_distance(px, py, pz, a0, a1, h):
    """Distance from (px, py, pz) to the segment from (a0, 0, h) to (a1, 0, h),
    a0 < a1, by cases on where px falls."""
    ex = np.where(px < a0, px - a0, np.where(px > a1, px - a1, 0.0))
    return np.sqrt(ex * ex + py * py + (pz - h) ** 2)


def candidate_points(seed):
    rng = np.random.default_rng(seed)
    xy = np.stack(np.meshgrid(np.linspace(-0.01, 0.01, 21), np.linspace(-0.01, 0.01, 21)), -1).reshape(-1, 2)
    coords = np.concatenate([xy, rng.uniform(-2e-4, 2e-4, (xy.shape[0], 1))], 1)
    u0 = rng.uniform(-5e-5, 5e-5, coords.shape)
    nodes = rng.permutation(coords.shape[0])[:300].astype(np.int32)
    return coords, u0, nodes


@pytest.mark.parametrize("shape", ["sphere", "cylinder", "plane"])
def test_candidates_match_case_analysis(shape):
    coords, u0, nodes = candidate_points(3)
    a0, a1, h, radius = -0.004, 0.003, 0.0052, 0.005
    margin = 1.2e-4
    x = (coords + u0)[nodes]
    if shape == "sphere":
        tool = O.tool_record(1, (a1, 0.0, h), (a0, 0.0, h), radius=radius)
        want = path_distance(x[:, 0], x[:, 1], x[:, 2], a0, a1, h) <= radius + margin
    elif shape == "cylinder":  # axis y: the distance in the x-z plane
        tool = O.tool_record(2, (a1, 0.0, h), (a0, 0.0, h), radius=radius, axis=(0.0, 1.0, 0.0))
        want = path_distance(x[:, 0], 0.0, x[:, 2], a0, a1, h) <= radius + margin
    else:  # normal +z, from z0 down to z1
        z0, z1 = -1e-4, -2e-4
        tool = O.tool_record(0, (0.0, 0.0, z1), (0.0, 0.0, z0), normal=(0.0, 0.0, 2.0))
        want = np.minimum(x[:, 2] - z0, x[:, 2] - z1) <= margin
    got = O.contact_candidates(tool, coords, u0, nodes, margin)
    assert np.array_equal(got, np.nonzero(want)[0]) and 0 < got.size < nodes.size
    # a larger margin keeps every candidate
    wider = O.contact_candidates(tool, coords, u0, nodes, 2 * margin)
    assert set(got) <= set(wider)


def test_anticipation_cases():
    """Sphere R = 5 mm at the origin, kappa A = 1e7: (a) outside, driven inside:
    added with residual kappa A g n and block kappa A n n^T; (b) outside,
    staying outside; (c) already inside; (d) deep in the tool (skipped);
    (e) outside, driven through the tool (added)."""
    radius, ka = 0.005, 1e7
    tool = O.tool_record(1, (0.0, 0.0, 0.0), radius=radius)
    dirs = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, -0.6, 0.8]])
    dist = np.array([5.2e-3, 5.3e-3, 4.9e-3, 2.0e-3, 5.1e-3])
    step = np.array([-4.0e-4, -1.0e-4, -2.0e-4, -1.0e-4, -4.0e-3])
    coords = dirs * dist[:, None]
    du = dirs * step[:, None]
    nodes = np.arange(5, dtype=np.int32)
    res, blocks, active, added = O.contact_anticipate(tool, coords, np.zeros(15), du.reshape(-1), nodes, np.full(5, ka))
    assert list(active) == [1, 0, 0, 0, 1] and added == 2
    for i in range(5):
        if active[i]:
            g = dist[i] - radius
            assert np.allclose(res[3 * i:3 * i + 3], ka * g * dirs[i], rtol=1e-14, atol=0.0)
            assert np.allclose(blocks[i].reshape(3, 3), ka * np.outer(dirs[i], dirs[i]), rtol=1e-14, atol=1e-7)
        else:
            assert not res[3 * i:3 * i + 3].any() and not blocks[i].any()
