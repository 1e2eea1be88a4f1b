"""The oracle's stress recovery (incompatible_elastoplastic_stress,
src/fem/IncompatibleModes.cpp:602-664) against closed forms.

The reference has no test of its own for these measures; they are pinned by
mathematics:
  * a homogeneous deformation F of a distorted element (its modes stay zero,
    so every point sees F) with an elastic material has closed forms for every
    kinematics, computed here independently of the oracle (numpy eigh / logm):
      - logarithmic: E_log = 1/2 ln(F^T F), Hencky T = K tr(E_log) I + 2G dev(E_log),
        S = U^-1 T U^-1 (T coaxial with U = (F^T F)^1/2), tau = R T R^T,
        sigma = tau / det F;
      - finite: E = 1/2 (F^T F - I), S = K tr(E) I + 2G dev(E), sigma = F S F^T / det F;
      - small strain: eps = sym(F - I), sigma = K tr(eps) I + 2G dev(eps);
  * objectivity of a plastic state: rotating the deformed element (and its
    modes) rotates the Cauchy and Kirchhoff stresses and leaves the second
    Piola-Kirchhoff stress, the log strain and the scalars unchanged;
  * the committed equivalent plastic strain comes back as the largest over
    the points (the return from the committed state is an elastic check).
Bars: 1e-10 relative (closed forms through a 3x3 eigen-decomposition), 1e-9
relative for objectivity.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_lib as O

DISTORTED = np.array([[0.05, 1.10, 1.20, -0.05, 0.10, 0.95, 1.05, 0.00],
                      [0.00, 0.10, 1.15, 0.90, 0.05, -0.05, 1.05, 1.10],
                      [0.00, -0.05, 0.10, 0.05, 0.90, 1.05, 1.20, 1.00]]).T.copy()
F0 = np.array([[1.04, 0.03, -0.02], [0.01, 0.97, 0.05], [-0.03, 0.02, 1.02]])
E_MOD, NU = 200e9, 0.3


def voigt(t):
    return np.array([t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[1, 2], t[2, 0]])


def tensor(v):
    return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])


def linear_elastic(e):
    k = E_MOD / (3.0 * (1.0 - 2.0 * NU))
    g = E_MOD / (2.0 * (1.0 + NU))
    tr = np.trace(e)
    return k * tr * np.eye(3) + 2.0 * g * (e - tr / 3.0 * np.eye(3))


def closed_form(f, kin):
    """(cauchy, pk2, kirchhoff, log strain, max strain, min jacobian)."""
    if kin == 0:
        eps = 0.5 * ((f - np.eye(3)) + (f - np.eye(3)).T)
        s = linear_elastic(eps)
        return s, s, np.zeros((3, 3)), np.zeros((3, 3)), np.abs(eps).max(), 1.0
    j = np.linalg.det(f)
    c = f.T @ f
    if kin == 1:
        e = 0.5 * (c - np.eye(3))
        s = linear_elastic(e)
        return f @ s @ f.T / j, s, np.zeros((3, 3)), np.zeros((3, 3)), np.abs(e).max(), j
    w, q = np.linalg.eigh(c)
    elog = q @ np.diag(0.5 * np.log(w)) @ q.T
    t = linear_elastic(elog)
    u_inv = q @ np.diag(1.0 / np.sqrt(w)) @ q.T
    s = u_inv @ t @ u_inv
    r = f @ u_inv
    tau = r @ t @ r.T
    return tau / j, s, tau, elog, np.abs(elog).max(), j


def von_mises(s):
    a, b, c = s[0, 0] - s[1, 1], s[1, 1] - s[2, 2], s[2, 2] - s[0, 0]
    return np.sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (s[0, 1] ** 2 + s[1, 2] ** 2 + s[2, 0] ** 2))


def recover(x0, ue, m, states=None, alpha=None, kin=2, rule=(2, 0, 2)):
    npts = O.rule_points(*rule)
    st = O.virgin_states(npts)[None] if states is None else states[None]
    out = O.element_stress_batch(x0[None], ue[None], m, st, None if alpha is None else alpha[None], rule=rule,
                                 kinematics=kin)
    assert out["status"][0] == 0
    return out["stress"][0]


@pytest.mark.parametrize("kin", [0, 1, 2])
def test_homogeneous_elastic_stress_closed_forms(kin):
    m = O.material(E_MOD, NU)
    f = np.eye(3) + (0.1 if kin == 0 else 1.0) * (F0 - np.eye(3))
    ue = np.concatenate([(f - np.eye(3)) @ DISTORTED[a] for a in range(8)])
    got = recover(DISTORTED, ue, m, kin=kin, rule=(2, 3, 2))
    cauchy, pk, tau, elog, max_strain, jac = closed_form(f, kin)
    scale = np.abs(cauchy).max()
    assert np.abs(got[0:6] - voigt(cauchy)).max() <= 1e-10 * scale
    assert np.abs(got[6:12] - voigt(pk)).max() <= 1e-10 * scale
    assert np.abs(got[12:18] - voigt(tau)).max() <= 1e-10 * scale
    assert np.abs(got[18:24] - voigt(elog)).max() <= 1e-10 * max(np.abs(elog).max(), 1e-300)
    assert abs(got[24] - von_mises(cauchy)) <= 1e-10 * scale
    assert got[25] == 0.0
    assert abs(got[26] - max_strain) <= 1e-10 * max_strain
    assert abs(got[27] - jac) <= 1e-12


@pytest.mark.parametrize("kin", [1, 2])
def test_plastic_stress_is_objective(kin):
    m = O.material(200e9, 0.3, sy0=250e6, H=2e9, Hkin=1e9)
    r = Rotation.from_rotvec(0.7 * np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)).as_matrix()
    ue = np.zeros(24)
    for a in range(8):
        p = DISTORTED[a]
        ue[3 * a:3 * a + 3] = 0.02 * np.array([0.8 * p[0] + 0.6 * p[0] * p[2] + 0.2 * p[1],
                                               -0.3 * p[1] + 0.5 * p[1] * p[2] - 0.4 * p[0] * p[1],
                                               -0.4 * p[2] - 0.5 * p[0] * p[0] + 0.3 * p[0] * p[1] * p[2]])
    el = O.element(DISTORTED, ue, m, kinematics=kin, want_tangent=False)
    assert el["yielding"] == 8
    states, alpha = el["states"], el["alpha"]
    got = recover(DISTORTED, ue, m, states, alpha, kin)
    rotated = np.concatenate([r @ (DISTORTED[a] + ue[3 * a:3 * a + 3]) - DISTORTED[a] for a in range(8)])
    alpha_r = np.concatenate([r @ alpha[3 * k:3 * k + 3] for k in range(3)])
    rot = recover(DISTORTED, rotated, m, states, alpha_r, kin)
    scale = np.abs(got[0:6]).max()
    assert np.abs(rot[0:6] - voigt(r @ tensor(got[0:6]) @ r.T)).max() <= 1e-9 * scale
    assert np.abs(rot[6:12] - got[6:12]).max() <= 1e-9 * scale
    assert np.abs(rot[12:18] - voigt(r @ tensor(got[12:18]) @ r.T)).max() <= 1e-9 * scale
    assert np.abs(rot[18:24] - got[18:24]).max() <= 1e-9 * max(np.abs(got[18:24]).max(), 1e-300)
    assert np.abs(rot[24:] - got[24:]).max() <= 1e-9 * np.abs(got[24:]).max()
    # the committed equivalent plastic strain comes back as the largest over the points
    assert got[25] == states[:, 36].max() and got[25] > 0.0
