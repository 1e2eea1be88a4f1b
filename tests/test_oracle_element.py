"""The oracle's incompatible-mode Hex8 against the reference's known answers.

Restates tests/test_incompatible_modes.cpp:112-685 of the reference:
quadrature exactness, the linear condensed stiffness (symmetric, six rigid
modes, patch test, exact pure bending, frame and numbering invariance, rule
independence on a box), the consistent tangent against central differences
for every kinematics and law, the local iteration through load reversals,
homogeneous deformation, objectivity and the tangent at rest.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_lib as O

DISTORTED = np.array([[0.05, 1.10, 1.20, -0.05, 0.10, 0.95, 1.05, 0.00],
                      [0.00, 0.10, 1.15, 0.90, 0.05, -0.05, 1.05, 1.10],
                      [0.00, -0.05, 0.10, 0.05, 0.90, 1.05, 1.20, 1.00]]).T.copy()


def box(a=1.0, b=1.0, c=1.0):
    return np.array([[0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0],
                     [0, 0, c], [a, 0, c], [a, b, c], [0, b, c]], dtype=float)


def rotation():
    axis = np.array([1.0, 2.0, 3.0])
    return Rotation.from_rotvec(0.7 * axis / np.linalg.norm(axis)).as_matrix()


def deformation(x, s):
    ue = np.zeros(24)
    for a in range(8):
        p = x[a]
        ue[3 * a:3 * a + 3] = s * np.array([0.8 * p[0] + 0.6 * p[0] * p[2] + 0.2 * p[1],
                                            -0.3 * p[1] + 0.5 * p[1] * p[2] - 0.4 * p[0] * p[1],
                                            -0.4 * p[2] - 0.5 * p[0] * p[0] + 0.3 * p[0] * p[1] * p[2]])
    return ue


def elastic():
    return O.material(200e9, 0.3)


def j2_steel():
    return O.material(200e9, 0.3, sy0=250e6, H=2e9, Hkin=1e9)


def hill_sheet():
    return O.material(200e9, 0.3, sy0=250e6, H=1.5e9, hill=dict(r0=1.9, r45=1.5, r90=2.3, rd=(1.0, 0.5, 0.0)))


def chaboche_steel():
    return O.material(200e9, 0.3, sy0=250e6, Q=100e6, delta=20.0, backstresses=((40e9, 400.0), (5e9, 20.0)))


def tangent_error(k, ue, h, force):
    err = 0.0
    for j in range(24):
        up, um = ue.copy(), ue.copy()
        up[j] += h
        um[j] -= h
        col = (force(up) - force(um)) / (2 * h)
        err = max(err, np.abs(col - k[:, j]).max())
    return err / np.abs(k).max()


# ---------------------------------------------------------------------------
# Quadrature (test_incompatible_modes.cpp:112-194)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", range(1, 8))
def test_extended_gauss_legendre_exact(n):
    xi, _, zeta, w = O.rule(2, n, 2)
    through = zeta[::4]
    wt = w[::4] / 1.0
    assert len(through) == (n if n != 2 else 2)
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs((wt * through ** k).sum() - exact) <= 2e-15
    assert np.all(np.diff(through) > 0)


def test_box_rule_equal_orders_is_cube_rule():
    for n in (1, 2, 3, 4):
        a = O.rule(n, n, 2)
        b = O.rule(n, 0, 2)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    # xi fastest
    xi, eta, zeta, w = O.rule(2, 0, 2)
    assert xi[0] < xi[1] and eta[0] == eta[1] and zeta[0] == zeta[3]
    assert w.sum() == pytest.approx(8.0, rel=1e-15)


# ---------------------------------------------------------------------------
# The linear condensed stiffness (test_incompatible_modes.cpp:212-366)
# ---------------------------------------------------------------------------
def standard_linear_stiffness(x0, E=200e9, nu=0.3, rule=(2, 0, 2)):
    """The compatible Hex8 stiffness: the small-strain elastic standard force is
    linear in u, so its central difference is exact."""
    m = O.material(E, nu)
    k = np.zeros((24, 24))
    for j in range(24):
        e = np.zeros(24)
        e[j] = 1.0
        k[:, j] = O.standard_force(x0, e, m, rule=rule, kinematics=0)
    return k


@pytest.mark.parametrize("name,coords,tp", [("box", box(1.0, 0.7, 0.4), 0),
                                            ("thin box", box(1.0, 1.0, 0.01), 0),
                                            ("distorted", DISTORTED, 0),
                                            ("distorted 2x2x5", DISTORTED, 5)])
def test_linear_stiffness_rigid_modes(name, coords, tp):
    k = O.linear_stiffness(coords, 200e9, 0.3, rule=(2, tp, 2))
    assert np.abs(k - k.T).max() <= 1e-15 * np.abs(k).max()
    lam = np.linalg.eigvalsh(k)
    top = np.abs(lam).max()
    rigid = np.abs(lam[:6]).max()
    assert rigid <= 1e-12 * top
    assert lam[6] > 1e3 * rigid
    # softer than the compatible element (Loewner order)
    ks = standard_linear_stiffness(coords, rule=(2, tp, 2))
    assert np.linalg.eigvalsh(0.5 * ((ks - k) + (ks - k).T)).min() >= -1e-10 * top


def test_patch_linear_field_modes_zero():
    """A linear field leaves the modes at zero and the condensed force is the
    compatible one (test_incompatible_modes.cpp:256-279)."""
    grad = np.array([[3.0e-4, 1.0e-4, -0.5e-4], [2.0e-4, -2.0e-4, 0.7e-4], [-0.5e-4, 0.4e-4, 1.5e-4]])
    ue = np.concatenate([grad @ DISTORTED[a] + np.array([1e-3, 0, 0]) for a in range(8)])
    for n in (2, 3):
        for t in (0, 3, 7):
            if n ** 2 * (t if t else n) > 112:
                continue
            rule = (n, t, 2)
            k = O.linear_stiffness(DISTORTED, 200e9, 0.3, rule=rule)
            ks = standard_linear_stiffness(DISTORTED, rule=rule)
            f = k @ ue
            f0 = ks @ ue
            assert np.abs(f - f0).max() <= 1e-11 * np.abs(f0).max()
            el = O.element(DISTORTED, ue, elastic(), rule=rule, kinematics=0)
            assert np.abs(el["alpha"]).max() <= 1e-13 * np.abs(ue).max()


def pure_bending(x, kappa, nu, z0):
    z = x[2] - z0
    return np.array([kappa * x[0] * z, -nu * kappa * x[1] * z, -0.5 * kappa * (x[0] ** 2 + nu * (z * z - x[1] ** 2))])


@pytest.mark.parametrize("slenderness", [1.0, 10.0, 100.0])
def test_exact_pure_bending(slenderness):
    """K* u = f exactly for pure bending (test_incompatible_modes.cpp:281-319)."""
    young, nu, kappa, h, width = 70e9, 0.3, 1e-3, 0.01, 0.02
    length = slenderness * h
    coords = box(length, width, h)
    ue = np.concatenate([pure_bending(coords[a], kappa, nu, 0.5 * h) for a in range(8)])
    f = np.zeros(24)
    for a in range(8):
        sx = 1.0 if coords[a, 0] > 0 else -1.0
        sz = 1.0 if coords[a, 2] > 0 else -1.0
        f[3 * a] = sx * sz * 0.5 * width * young * kappa * h * h / 12.0
    k = O.linear_stiffness(coords, young, nu)
    gross = (np.abs(k) @ np.abs(ue)).max()
    assert np.abs(k @ ue - f).max() <= 1e-13 * gross


def test_frame_invariance_and_renumbering():
    k = O.linear_stiffness(DISTORTED, 200e9, 0.3)
    r = rotation()
    q = np.kron(np.eye(8), r)
    kr = O.linear_stiffness(DISTORTED @ r.T, 200e9, 0.3)
    assert np.abs(kr - q @ k @ q.T).max() <= 1e-12 * np.abs(k).max()
    perm = [1, 2, 3, 0, 5, 6, 7, 4]
    kp = O.linear_stiffness(DISTORTED[perm], 200e9, 0.3)
    for a in range(8):
        for b in range(8):
            blk = kp[3 * a:3 * a + 3, 3 * b:3 * b + 3] - k[3 * perm[a]:3 * perm[a] + 3, 3 * perm[b]:3 * perm[b] + 3]
            assert np.abs(blk).max() <= 1e-12 * np.abs(k).max()


def test_rules_agree_on_a_box():
    coords = box(1.3, 0.8, 0.2)
    k2 = O.linear_stiffness(coords, 200e9, 0.3, rule=(2, 0, 2))
    for rule in ((3, 0, 2), (4, 0, 2), (2, 6, 2)):
        assert np.abs(O.linear_stiffness(coords, 200e9, 0.3, rule=rule) - k2).max() <= 1e-12 * np.abs(k2).max()


# ---------------------------------------------------------------------------
# The non-linear element (test_incompatible_modes.cpp:433-685)
# ---------------------------------------------------------------------------
CASES = [("small elastic", elastic, 0, 1e-3, 0), ("small J2", j2_steel, 0, 4e-3, 0),
         ("small Hill48", hill_sheet, 0, 4e-3, 3), ("small Chaboche", chaboche_steel, 0, 4e-3, 0),
         ("finite elastic", elastic, 1, 0.05, 0), ("finite J2", j2_steel, 1, 0.02, 0),
         ("finite Hill48", hill_sheet, 1, 0.02, 5), ("finite Chaboche", chaboche_steel, 1, 0.02, 0),
         ("log elastic", elastic, 2, 0.05, 0), ("log J2", j2_steel, 2, 0.1, 0),
         ("log Hill48", hill_sheet, 2, 0.1, 3), ("log Chaboche", chaboche_steel, 2, 0.1, 0)]


@pytest.mark.parametrize("name,make,kin,scale,tp", CASES, ids=[c[0] for c in CASES])
def test_condensed_tangent_is_derivative(name, make, kin, scale, tp):
    m = make()
    rule = (2, tp, 2)
    npts = O.rule_points(*rule)
    plastic = m[2] > 0
    first = O.element(DISTORTED, deformation(DISTORTED, scale), m, rule=rule, kinematics=kin)
    committed = first["states"] if plastic else O.virgin_states(npts)
    alpha = first["alpha"]
    ue = deformation(DISTORTED, 1.3 * scale)
    for a in range(8):
        ue[3 * a + 2] += 0.2 * scale * DISTORTED[a, 0] * DISTORTED[a, 1]
    el = O.element(DISTORTED, ue, m, states=committed, alpha=alpha, rule=rule, kinematics=kin)
    if plastic:
        assert el["yielding"] == npts
    assert el["iterations"] >= 1
    assert el["iterations"] <= (1 if (kin == 0 and not plastic) else 5)
    symmetric = all(m[13 + i] <= 0 for i in range(int(m[8])))
    assert el["symmetric"] == symmetric
    asym = np.abs(el["tangent"] - el["tangent"].T).max()
    if symmetric:
        assert asym == 0.0
    else:
        assert asym > 1e-6 * np.abs(el["tangent"]).max()
    force = lambda v: O.element(DISTORTED, v, m, states=committed, alpha=alpha, rule=rule, kinematics=kin,
                                want_tangent=False)["force"]
    assert tangent_error(el["tangent"], ue, 1e-8, force) <= 1e-6
    if not plastic:
        err = 0.0
        for j in range(24):
            up, um = ue.copy(), ue.copy()
            up[j] += 1e-7
            um[j] -= 1e-7
            wp = O.element(DISTORTED, up, m, states=committed, alpha=alpha, rule=rule, kinematics=kin,
                           want_tangent=False)["energy"]
            wm = O.element(DISTORTED, um, m, states=committed, alpha=alpha, rule=rule, kinematics=kin,
                           want_tangent=False)["energy"]
            err = max(err, abs((wp - wm) / 2e-7 - el["force"][j]))
        assert err <= 1e-6 * np.abs(el["force"]).max()


@pytest.mark.parametrize("make", [j2_steel, chaboche_steel], ids=["J2", "Chaboche"])
@pytest.mark.parametrize("kin", [0, 2], ids=["small", "log"])
def test_local_iteration_through_reversal(make, kin):
    m = make()
    rule = (2, 5, 2)
    scale = 4e-3 if kin == 0 else 0.02
    first = O.element(DISTORTED, deformation(DISTORTED, scale), m, rule=rule, kinematics=kin, want_tangent=False)
    assert first["yielding"] > 0
    ue = deformation(DISTORTED, -6.0 * scale)
    for a in range(8):
        ue[3 * a] += 2.0 * scale * DISTORTED[a, 2] ** 2
    el = O.element(DISTORTED, ue, m, states=first["states"], alpha=first["alpha"], rule=rule, kinematics=kin)
    assert el["yielding"] > 0
    assert el["iterations"] <= 15
    again = O.element(DISTORTED, ue, m, states=first["states"], alpha=el["alpha"], rule=rule, kinematics=kin,
                      want_tangent=False)
    assert again["iterations"] <= 1
    assert np.abs(again["force"] - el["force"]).max() <= 1e-8 * np.abs(el["force"]).max()


@pytest.mark.parametrize("make", [lambda: O.material(200e9, 0.3, sy0=250e6), j2_steel, chaboche_steel, hill_sheet],
                         ids=["perfect", "J2", "Chaboche", "Hill48"])
@pytest.mark.parametrize("tp", [5, 7])
@pytest.mark.parametrize("factor", [-6.0, -10.0, -15.0])
def test_local_iteration_large_log_reversals(make, tp, factor):
    """Logarithmic strains of 2 % reversed to 12-30 % in one increment, K_aa
    indefinite (test_incompatible_modes.cpp:549-595)."""
    m = make()
    rule = (2, tp, 2)
    first = O.element(DISTORTED, deformation(DISTORTED, 0.02), m, rule=rule, kinematics=2, want_tangent=False)
    ue = deformation(DISTORTED, factor * 0.02)
    el = O.element(DISTORTED, ue, m, states=first["states"], alpha=first["alpha"], rule=rule, kinematics=2,
                   want_tangent=False)
    assert el["yielding"] > 0
    assert el["iterations"] <= 20
    again = O.element(DISTORTED, ue, m, states=first["states"], alpha=el["alpha"], rule=rule, kinematics=2,
                      want_tangent=False)
    assert again["iterations"] <= 1
    assert np.abs(again["force"] - el["force"]).max() <= 1e-8 * np.abs(el["force"]).max()


@pytest.mark.parametrize("make", [elastic, j2_steel, hill_sheet], ids=["elastic", "J2", "Hill48"])
@pytest.mark.parametrize("kin", [0, 1, 2])
def test_homogeneous_deformation_leaves_modes_zero(make, kin):
    f = np.array([[1.04, 0.03, -0.02], [0.01, 0.97, 0.05], [-0.03, 0.02, 1.02]])
    m = make()
    rule = (2, 3, 2)
    scale = 0.1 if kin == 0 else 1.0
    ue = np.concatenate([scale * (f - np.eye(3)) @ DISTORTED[a] for a in range(8)])
    el = O.element(DISTORTED, ue, m, rule=rule, kinematics=kin, want_tangent=False)
    ref = O.standard_force(DISTORTED, ue, m, rule=rule, kinematics=kin)
    assert np.abs(el["alpha"]).max() <= 1e-12
    assert np.abs(el["force"] - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("kin", [1, 2])
def test_objectivity_and_rigid_rotation(kin):
    m = j2_steel()
    r = rotation()
    rigid = np.concatenate([(r - np.eye(3)) @ DISTORTED[a] for a in range(8)])
    at_rest = O.element(DISTORTED, rigid, m, kinematics=kin, want_tangent=False)
    assert np.abs(at_rest["force"]).max() <= 1e-3
    assert np.abs(at_rest["alpha"]).max() <= 1e-14
    ue = deformation(DISTORTED, 0.02)
    rotated = np.concatenate([r @ (DISTORTED[a] + ue[3 * a:3 * a + 3]) - DISTORTED[a] for a in range(8)])
    el = O.element(DISTORTED, ue, m, kinematics=kin, want_tangent=False)
    er = O.element(DISTORTED, rotated, m, kinematics=kin, want_tangent=False)
    assert el["yielding"] == 8
    for a in range(8):
        assert np.abs(er["force"][3 * a:3 * a + 3] - r @ el["force"][3 * a:3 * a + 3]).max() <= \
            1e-9 * np.abs(el["force"]).max()
    for k in range(3):
        assert np.abs(er["alpha"][3 * k:3 * k + 3] - r @ el["alpha"][3 * k:3 * k + 3]).max() <= \
            1e-9 * np.abs(el["alpha"]).max()


@pytest.mark.parametrize("kin", [0, 1, 2])
def test_tangent_at_rest_is_linear(kin):
    k = O.linear_stiffness(DISTORTED, 200e9, 0.3)
    el = O.element(DISTORTED, np.zeros(24), elastic(), kinematics=kin)
    assert np.abs(el["tangent"] - k).max() <= 1e-12 * np.abs(k).max()
    assert el["iterations"] == 0
