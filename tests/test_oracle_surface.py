"""The oracle's height map (_interpolate_structured + raycast_top,
python/precomp/fea/results.py:262-283, python/precomp/geometry/raycast.py:
27-100, before the nearest fill) against what the mathematics fixes.

The reference has no test of this path with known answers; it is pinned by:
  * a plane: a surface whose heights are z = a x + b y + c of its deformed
    plan positions (an affine plan map) is interpolated exactly (to
    round-off), and exactly the grid nodes inside the deformed plan (a
    parallelogram) are hit;
  * brute force: every triangle against every grid node (no bounding boxes)
    on a curved, non-affinely deformed surface gives the same map bit for bit
    (a missed or extra node of a bounding box, a wrong cell index or corner
    order fails);
  * a surface folded over in plan view is refused and nothing is cast.
"""
import numpy as np
import pytest

import oracle_lib as O


def structured(nxs, nys, hs, warp=None):
    x, y = np.meshgrid(np.arange(nxs) * hs, np.arange(nys) * hs)
    p = np.stack([x, y, np.zeros_like(x)], axis=-1)
    return p if warp is None else warp(p)


def test_plane_is_interpolated_exactly_inside_the_plan():
    A = np.array([[1.03, 0.07], [-0.05, 0.98]])
    shift = np.array([0.0013, -0.0021])
    p = structured(13, 9, 1e-3)
    xy = p[..., :2] @ A.T + shift
    a, b, c = 0.3, -0.2, -1e-3
    surf = np.concatenate([xy, (a * xy[..., 0] + b * xy[..., 1] + c)[..., None]], axis=-1)
    grid = (-0.001, -0.002, 31, 25, 0.5e-3)
    z, hit, folded = O.surface_raycast(surf, grid)
    assert folded == 0
    gx = grid[0] + np.arange(grid[2]) * grid[4]
    gy = grid[1] + np.arange(grid[3]) * grid[4]
    X, Y = np.meshgrid(gx, gy)
    exact = a * X + b * Y + c
    assert np.abs(z[hit] - exact[hit]).max() <= 1e-15
    st = np.linalg.solve(A, np.stack([X - shift[0], Y - shift[1]]).reshape(2, -1)).reshape(2, *X.shape)
    span = (12e-3, 8e-3)
    margin = np.minimum(np.minimum(st[0], span[0] - st[0]), np.minimum(st[1], span[1] - st[1]))
    clear = np.abs(margin) > 1e-8
    assert np.array_equal(hit[clear], (margin > 0)[clear])
    assert hit.sum() > 0.3 * hit.size and (~hit).sum() > 0


def brute_force(surf, grid):
    nys, nxs = surf.shape[:2]
    x0, y0, nx, ny, h = grid
    tris = []
    for cy in range(nys - 1):
        for cx in range(nxs - 1):
            a, b, c, d = surf[cy, cx], surf[cy, cx + 1], surf[cy + 1, cx + 1], surf[cy + 1, cx]
            tris.append((a, b, c))
            tris.append((a, c, d))
    z = np.full((ny, nx), -np.inf)
    for a, b, c in tris:
        v0, v1 = b[:2] - a[:2], c[:2] - a[:2]
        den = v0[0] * v1[1] - v1[0] * v0[1]
        for jj in range(ny):
            for ii in range(nx):
                px = (x0 + ii * h) - a[0]
                py = (y0 + jj * h) - a[1]
                u = (px * v1[1] - v1[0] * py) / den
                v = (v0[0] * py - px * v0[1]) / den
                if u >= -1e-6 and v >= -1e-6 and u + v <= 1.0 + 1e-6:
                    z[jj, ii] = max(z[jj, ii], (a[2] + u * (b[2] - a[2])) + v * (c[2] - a[2]))
    return z, np.isfinite(z)


def test_curved_surface_matches_brute_force():
    rng = np.random.default_rng(3)

    def warp(p):
        q = p.copy()
        q[..., 0] += 1e-4 * np.sin(p[..., 1] * 700.0) + rng.uniform(-5e-5, 5e-5, p.shape[:2])
        q[..., 1] += 1e-4 * np.cos(p[..., 0] * 500.0) + rng.uniform(-5e-5, 5e-5, p.shape[:2])
        q[..., 2] = -2e-3 * np.exp(-((p[..., 0] - 4e-3) ** 2 + (p[..., 1] - 3e-3) ** 2) / 4e-6)
        return q

    surf = structured(9, 7, 1e-3, warp)
    for grid in [(0.0, 0.0, 9, 7, 1e-3), (-0.5e-3, 0.2e-3, 23, 17, 0.4e-3)]:
        z, hit, folded = O.surface_raycast(surf, grid)
        zb, hb = brute_force(surf, grid)
        assert folded == 0 and np.array_equal(hit, hb) and np.array_equal(z[hit], zb[hb])
        assert hit.any() and (np.isinf(z) == ~hit).all()


def test_folded_surface_is_refused():
    surf = structured(6, 5, 1e-3)
    surf[2, 3, 0] += 1.6e-3  # a node pushed past its neighbour: cells fold over
    z, hit, folded = O.surface_raycast(surf, (0.0, 0.0, 6, 5, 1e-3))
    assert folded > 0 and not hit.any()
