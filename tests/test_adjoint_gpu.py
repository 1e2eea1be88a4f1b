"""The differentiation form: vector-Jacobian products of the assembled
internal force (sparlab_system_vjp_device) at the converged modes and a fixed
committed history, on a plastically deforming Hill48 + Chaboche sheet with
logarithmic strain (the displacement and coordinate checks also with small
strain and the Green-Lagrange finite strain).

Checks (the reference has no adjoint; these pin it by mathematics):
  * the same adjoint machinery in the displacements reproduces v^T K_t of the
    forward tangent, which the element tests pin to the oracle (1e-10);
  * coordinates: central finite differences of v^T f_int along a random
    direction, h = 1e-6 of the element size (1e-6 relative);
  * coordinates: translation invariance (moving every node by the same vector
    at fixed u leaves f_int unchanged, so the node gradients sum to zero);
  * parameters (r0, r45, r90, C_i, gamma_i): central finite differences with
    a relative step of 1e-6 (1e-6 relative);
  * the torch.autograd.Function wrapper returns the same products;
  * the 2 x 2 x 5 box rule (20 points, 32-lane teams) passes the first three
    checks as the 2 x 2 x 2 rule does.
"""
import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu


def sheet(nx=4, ny=4, nz=2):
    coords, conn = W.structured_hex_mesh(nx, ny, nz, 0.01, 0.01, 0.001, -0.005, -0.005, -0.001)
    w = W._family_depth("dome", coords[:, 0], coords[:, 1], 0.005, 0.003, 0.5)
    rng = np.random.default_rng(11)
    u1 = np.zeros_like(coords)
    u1[:, 2] = 0.06 * w
    u2 = u1 * 1.6 + rng.normal(0.0, 1e-6, size=u1.shape)
    return coords, conn, u1.reshape(-1), u2.reshape(-1)


KINEMATICS = [G.SMALL_STRAIN, G.FINITE, G.FINITE_LOGARITHMIC]


RULES = {"2x2x2": (2, 0, 2), "2x2x5": (2, 5, 2)}


def prepared_system(device, material=None, kinematics=G.FINITE_LOGARITHMIC, rule="2x2x2"):
    """A system with a plastic committed history, evaluated at the next increment."""
    import torch
    coords, conn, u1, u2 = sheet()
    m = material or W.microbench_material()
    s = G.NonlinearSystem(coords, conn, m, G.Rule(*RULES[rule]), kinematics)
    s.evaluate(u1)
    s.commit()
    ev = s.evaluate(u2)
    assert ev.yielding_points > 0
    return s, coords, conn, torch.as_tensor(u2, device=device), ev


def fint_at(s, u_dev, coords=None, material=None):
    import torch
    if coords is not None:
        s.set_coords(coords)
    if material is not None:
        s.set_material(material)
    f = torch.zeros(3 * s.num_nodes, dtype=torch.float64, device=u_dev.device)
    s.evaluate_device(u_dev, f, None, want_tangent=False)
    s.result()
    return f.cpu().numpy()


@pytest.mark.parametrize("rule", list(RULES))
@pytest.mark.parametrize("kinematics", KINEMATICS)
def test_displacement_vjp_is_tangent_transpose(device, kinematics, rule):
    import torch
    s, coords, conn, u, ev = prepared_system(device, kinematics=kinematics, rule=rule)
    row_ptr, col_idx = s.pattern()
    n = 3 * s.num_nodes
    vals = ev.tangent_values
    rng = np.random.default_rng(5)
    v = rng.normal(size=n)
    ktv = np.zeros(n)
    for i in range(n):
        sl = slice(row_ptr[i], row_ptr[i + 1])
        ktv[col_idx[sl]] += vals[sl] * v[i]
    g = s.vjp_device(u, torch.as_tensor(v, device=device), coords=False, displacement=True)
    got = g["displacement"].cpu().numpy()
    assert np.abs(got - ktv).max() <= 1e-10 * np.abs(ktv).max()


@pytest.mark.parametrize("rule", list(RULES))
@pytest.mark.parametrize("kinematics", KINEMATICS)
def test_coordinate_vjp_matches_central_differences(device, kinematics, rule):
    import torch
    s, coords, conn, u, _ = prepared_system(device, kinematics=kinematics, rule=rule)
    rng = np.random.default_rng(6)
    v = rng.normal(size=3 * s.num_nodes)
    g = s.vjp_device(u, torch.as_tensor(v, device=device), coords=True)["coords"].cpu().numpy()
    size = 0.01 / 4
    h = 1e-6
    worst = 0.0
    for trial in range(3):
        d = rng.normal(size=coords.shape) * size
        fp = fint_at(s, u, coords + h * d)
        fm = fint_at(s, u, coords - h * d)
        fd = v @ (fp - fm) / (2.0 * h)
        an = g @ d.reshape(-1)
        worst = max(worst, abs(fd - an) / abs(an))
    s.set_coords(coords)
    assert worst <= 1e-6, worst


@pytest.mark.parametrize("kinematics", KINEMATICS)
def test_coordinate_vjp_translation_invariance(device, kinematics):
    import torch
    s, coords, conn, u, _ = prepared_system(device, kinematics=kinematics)
    v = np.random.default_rng(7).normal(size=3 * s.num_nodes)
    g = s.vjp_device(u, torch.as_tensor(v, device=device), coords=True)["coords"].cpu().numpy().reshape(-1, 3)
    assert np.abs(g.sum(axis=0)).max() <= 1e-9 * np.abs(g).sum(axis=0).max()


@pytest.mark.parametrize("rule", list(RULES))
def test_parameter_vjp_matches_central_differences(device, rule):
    import torch
    m = W.microbench_material()
    s, coords, conn, u, _ = prepared_system(device, m, rule=rule)
    v = np.random.default_rng(8).normal(size=3 * s.num_nodes)
    g = s.vjp_device(u, torch.as_tensor(v, device=device), coords=False, parameters=True)["parameters"].cpu().numpy()
    theta = G.material_parameters(m)
    active = [0, 1, 2, 3, 4, 5, 6]
    assert np.all(g[7:] == 0.0)
    scaled = np.array([g[k] * theta[k] for k in active])
    for k in active:
        h = 1e-6 * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fp = fint_at(s, u, material=G.with_parameters(m, tp))
        fm = fint_at(s, u, material=G.with_parameters(m, tm))
        fd = v @ (fp - fm) / (2.0 * h)
        err = abs(fd - g[k]) * theta[k] / np.abs(scaled).max()
        assert err <= 1e-6, (G.PARAMETER_NAMES[k], fd, g[k], err)
    s.set_material(m)


def test_autograd_function(device):
    import torch
    m = W.microbench_material()
    s, coords, conn, u, _ = prepared_system(device, m)
    X = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    theta = torch.tensor(G.material_parameters(m), dtype=torch.float64, requires_grad=True)
    v = torch.as_tensor(np.random.default_rng(9).normal(size=3 * s.num_nodes), device=device)
    f = G.InternalForce.apply(X, theta, s, u)
    (f * v).sum().backward()
    ref = s.vjp_device(u, v, coords=True, parameters=True)
    assert np.array_equal(X.grad.numpy().reshape(-1), ref["coords"].cpu().numpy())
    assert np.array_equal(theta.grad.numpy(), ref["parameters"].cpu().numpy())


@pytest.mark.parametrize("size", [(20, 25, 4), (200, 250, 4)], ids=["20x25x4", "200x250x4"])
def test_spiral_forming_state_matches_central_differences(device, size):
    """The differentiation workload's state (spiral incremental forming, mid
    path) at 20 x 25 x 4 and at its full 200 x 250 x 4 (200 000 elements):
    coordinate and parameter products against central differences, h = 1e-6
    of the element size and of each parameter (1e-4 for the parameters of the
    full sheet, see below)."""
    import torch
    nx, ny, nz = size
    print(f"spiral forming state {nx} x {ny} x {nz}", flush=True)
    m = W.microbench_material()
    coords, conn, u1, u2 = W.differentiation_workload(nx=nx, ny=ny, nz=nz)
    s = G.NonlinearSystem(coords, conn, m, G.Rule(2, 0, 2))
    u = torch.as_tensor(u1, device=device)
    f = torch.zeros(3 * s.num_nodes, dtype=torch.float64, device=device)
    values = torch.empty(s.nnz, dtype=torch.float64, device=device)
    s.evaluate_device(u, f, values)
    s.result()
    s.commit()
    u = torch.as_tensor(u2, device=device)
    s.evaluate_device(u, f, values)
    ev = s.result()
    assert ev.yielding_points > 0 and ev.failed_elements == 0
    del values
    rng = np.random.default_rng(12)
    v = rng.normal(size=3 * s.num_nodes)
    g = s.vjp_device(u, torch.as_tensor(v, device=device), coords=True, parameters=True)
    gx, gp = g["coords"].cpu().numpy(), g["parameters"].cpu().numpy()
    d = rng.normal(size=coords.shape) * (0.020 / nx)
    h = 1e-6
    fd = v @ (fint_at(s, u, coords + h * d) - fint_at(s, u, coords - h * d)) / (2.0 * h)
    s.set_coords(coords)
    assert abs(fd - gx @ d.reshape(-1)) <= 1e-6 * abs(gx @ d.reshape(-1))
    theta = G.material_parameters(m)
    dt = rng.normal(size=7) * theta[:7]
    # The parameter step: 1e-6, except on the full sheet, where a central
    # difference at 1e-6 is limited by round-off (eps sum |v_i f_i| / h, about
    # 1e-6 of the product there: the oracle's own differences at h = 1e-6,
    # 1e-5, 1e-4 give -3.0035957, -3.0035991, -3.0035987, with no local
    # iteration count changing); at 1e-4 the truncation error is about 1e-8.
    hp = 1e-4 if nx * ny * nz > 100_000 else h
    tp, tm = theta.copy(), theta.copy()
    tp[:7] += hp * dt
    tm[:7] -= hp * dt
    fd = v @ (fint_at(s, u, material=G.with_parameters(m, tp)) - fint_at(s, u, material=G.with_parameters(m, tm))) / (2.0 * hp)
    s.set_material(m)
    an = gp[:7] @ dt
    assert abs(fd - an) <= 1e-6 * abs(an)
