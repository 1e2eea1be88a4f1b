"""Parity of the device height map (sparlab_surface_raycast_device) with the
oracle's _interpolate_structured + raycast_top (python/precomp/fea/
results.py:262-283, python/precomp/geometry/raycast.py:27-100, before the
nearest fill), on the same surfaces and grids.

Bar: bit-exact heights and hit flags (the device keeps numpy's order of
operations; the highest hit is an order-free maximum). Cases: a plane under
an affine plan map, a curved non-affinely deformed surface on the surface's
own node grid and on finer offset grids, a 201 x 201 surface on a 400 x 400
grid, a grid beside the surface (nothing hit), and a folded surface
(refused).
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G

pytestmark = pytest.mark.gpu


def structured(nxs, nys, hs):
    x, y = np.meshgrid(np.arange(nxs) * hs, np.arange(nys) * hs)
    return np.stack([x, y, np.zeros_like(x)], axis=-1)


def warped(nxs, nys, hs, seed):
    rng = np.random.default_rng(seed)
    p = structured(nxs, nys, hs)
    q = p.copy()
    span = hs * (nxs - 1)
    q[..., 0] += 0.1 * hs * np.sin(p[..., 1] * 7.0 / span) + rng.uniform(-0.05, 0.05, p.shape[:2]) * hs
    q[..., 1] += 0.1 * hs * np.cos(p[..., 0] * 5.0 / span) + rng.uniform(-0.05, 0.05, p.shape[:2]) * hs
    q[..., 2] = -0.2 * span * np.exp(-((p[..., 0] - 0.5 * span) ** 2 + (p[..., 1] - 0.4 * span) ** 2) / (0.1 * span ** 2))
    return q


def check(device, surf, grid):
    import torch
    z, hit = G.surface_raycast_device(torch.as_tensor(np.ascontiguousarray(surf), device=device), grid)
    torch.cuda.synchronize()
    zo, ho, folded = O.surface_raycast(surf, grid)
    assert folded == 0
    z, hit = z.cpu().numpy(), hit.cpu().numpy()
    assert np.array_equal(hit, ho), grid
    assert np.array_equal(z, zo), grid  # -inf where not hit, on both sides
    return hit


def test_plane_under_an_affine_map(device):
    A = np.array([[1.03, 0.07], [-0.05, 0.98]])
    p = structured(13, 9, 1e-3)
    xy = p[..., :2] @ A.T + np.array([0.0013, -0.0021])
    surf = np.concatenate([xy, (0.3 * xy[..., 0] - 0.2 * xy[..., 1] - 1e-3)[..., None]], axis=-1)
    hit = check(device, surf, (-0.001, -0.002, 31, 25, 0.5e-3))
    assert hit.any() and not hit.all()


@pytest.mark.parametrize("grid", [(0.0, 0.0, 9, 7, 1e-3), (-0.5e-3, 0.2e-3, 23, 17, 0.4e-3),
                                  (0.25e-3, 0.25e-3, 64, 48, 0.125e-3)])
def test_curved_surface(device, grid):
    assert check(device, warped(9, 7, 1e-3, 3), grid).any()


def test_large_surface(device):
    surf = warped(201, 201, 1e-4, 5)
    hit = check(device, surf, (-0.001, -0.001, 400, 400, 0.055e-3))
    assert hit.mean() > 0.5


def test_grid_beside_the_surface(device):
    hit = check(device, warped(9, 7, 1e-3, 3), (0.05, 0.05, 10, 10, 1e-3))
    assert not hit.any()


def test_folded_surface_is_refused(device):
    import torch
    surf = structured(6, 5, 1e-3)
    surf[2, 3, 0] += 1.6e-3
    with pytest.raises(RuntimeError, match="fold over"):
        G.surface_raycast_device(torch.as_tensor(surf, device=device), (0.0, 0.0, 6, 5, 1e-3))
