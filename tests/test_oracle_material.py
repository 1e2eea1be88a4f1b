"""The oracle's material point against the reference's known answers.

Restates the checks of the reference's tests/test_logarithmic.cpp:110-257,
tests/test_plasticity.cpp:156-350 and tests/test_anisotropy.cpp:255-1001 that
pin the logarithmic strain and the returns: closed forms, an independent
matrix logarithm (mpmath's logm at 50 digits; scipy's double-precision logm
loses the 1e-17 off-diagonal terms of a C within round-off of I), fourth-order
central differences, invariants.
"""
import numpy as np
import pytest
import mpmath as mp
from scipy.spatial.transform import Rotation

import oracle_lib as O


def engineering(t):
    return np.array([t[0, 0], t[1, 1], t[2, 2], 2 * t[0, 1], 2 * t[1, 2], 2 * t[2, 0]])


def stress_tensor(v):
    return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])


def relative(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def derivative(f, x, h):
    """Fourth-order central differences of a 6-vector function (columns = d/dx_j)."""
    d = np.zeros((6, 6))
    for j in range(6):
        def at(s):
            y = x.copy()
            y[j] += s
            return f(y)
        d[:, j] = (-at(2 * h) + 8 * at(h) - 8 * at(-h) + at(-2 * h)) / (12 * h)
    return d


def axis_angle(angle, axis):
    axis = np.asarray(axis, float)
    return Rotation.from_rotvec(angle * axis / np.linalg.norm(axis)).as_matrix()


# ---------------------------------------------------------------------------
# Logarithmic strain (test_logarithmic.cpp:110-216)
# ---------------------------------------------------------------------------
A = 1.4
LOG_CASES = {
    "distinct": (2.3, 0.6, 1.1),
    "two equal": (1.7, 1.7, 0.5),
    "C = I": (1.0, 1.0, 1.0),
    "C = s^2 I": (1.69, 1.69, 1.69),
    "two within 1e-9": (A, A * (1 + 1e-9), 0.8),
    "two within 1e-7": (A, A * (1 + 1e-7), 0.8),
    "two within 1e-4": (A, A * (1 + 1e-4), 0.8),
    "three within 1e-9": (A, A * (1 + 1e-9), A * (1 + 2e-9)),
    "three within 1e-7": (A, A * (1 + 0.4e-7), A * (1 + 1e-7)),
    "three within 1e-4": (A, A * (1 + 1e-4), A * (1 + 2.5e-4)),
    "three within 4e-3 (series)": (A, A * (1 + 2e-3), A * (1 + 4e-3)),
    "three within 6e-3 (quotient)": (A, A * (1 + 3e-3), A * (1 + 6e-3)),
    "stretch 2, two nearly equal": (4.0, 0.25, 0.25 * (1 + 1e-8)),
}
T_STRESS = np.array([3.0e8, -1.0e8, 0.5e8, 0.7e8, -0.4e8, 0.2e8])
Q_FRAME = axis_angle(0.7, (0.3, -1.0, 0.5))


@pytest.mark.parametrize("name", list(LOG_CASES))
def test_log_strain_exact(name):
    lam = np.array(LOG_CASES[name])
    cg = Q_FRAME @ np.diag(lam) @ Q_FRAME.T
    green = engineering(0.5 * (cg - np.eye(3)))
    log = O.log_strain(green, T_STRESS)
    mp.mp.dps = 50
    lg = mp.logm(mp.matrix(cg.tolist()))
    reference = 0.5 * np.array([[float(lg[i, j]) for j in range(3)] for i in range(3)])
    assert relative(log["strain"], engineering(reference)) <= 1e-14
    assert abs(log["strain"][:3].sum() - 0.5 * np.log(np.linalg.det(cg))) <= 4e-15
    h = 5e-4 * lam.min()
    p_fd = derivative(lambda e: O.log_strain(e)["strain"], green, h)
    l_fd = derivative(lambda e: O.log_strain(e, T_STRESS)["s"], green, h)
    assert relative(log["projection"], p_fd) <= 1e-11
    assert relative(log["curvature"], l_fd) <= 1e-11
    assert relative(log["curvature"], log["curvature"].T) <= 1e-15
    assert relative(log["s"], log["projection"].T @ T_STRESS) <= 1e-15


def test_log_strain_limits():
    def green_of(lam):
        return engineering(0.5 * (Q_FRAME @ np.diag(lam) @ Q_FRAME.T - np.eye(3)))
    equal = O.log_strain(green_of([A, A, 0.8]), T_STRESS)
    near = O.log_strain(green_of([A, A * (1 + 1e-9), 0.8]), T_STRESS)
    assert relative(near["projection"], equal["projection"]) <= 2e-9
    assert relative(near["curvature"], equal["curvature"]) <= 2e-9
    ident = O.log_strain(np.zeros(6))
    assert np.abs(ident["projection"] - np.eye(6)).max() <= 1e-15
    rng = np.random.default_rng(5)
    e = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            e[i, j] = e[j, i] = 1e-6 * rng.uniform(-1, 1)
    series = e - e @ e + 4.0 / 3.0 * e @ e @ e
    assert relative(O.log_strain(engineering(e))["strain"], engineering(series)) <= 1e-15
    bad = np.zeros(6)
    bad[0] = -0.6
    with pytest.raises(O.OracleError):
        O.log_strain(bad)
    bad[0] = np.nan
    with pytest.raises(O.OracleError):
        O.log_strain(bad)


# ---------------------------------------------------------------------------
# J2 radial return (test_plasticity.cpp:156-350)
# ---------------------------------------------------------------------------
E_STEEL, NU = 200e9, 0.3


def uniaxial_strain_history(m, eps_axial, state=None):
    """Drive sigma_22 = sigma_33 = 0 (uniaxial stress along x) by Newton on
    the lateral strains; returns the response and the strain."""
    strain = np.array([eps_axial, -NU * eps_axial, -NU * eps_axial, 0, 0, 0], float)
    for _ in range(50):
        r = O.plastic_return(m, strain, state)
        res = r["stress"][1:3]
        if np.abs(res).max() <= 1e-14 * max(1.0, np.abs(r["stress"]).max()):
            break
        c = r["tangent"][1:3, 1:3]
        strain[1:3] -= np.linalg.solve(c, res)
    return r, strain


def test_j2_uniaxial_linear_hardening_exact():
    """Linear isotropic hardening: sigma = sy + Et/(1) ... the exact uniaxial
    stress-strain curve for any step (test_plasticity.cpp:156-196)."""
    sy, h = 250e6, 2e9
    m = O.material(E_STEEL, NU, sy0=sy, H=h)
    for eps in (2e-3, 5e-3, 2e-2):
        r, strain = uniaxial_strain_history(m, eps)
        # uniaxial: sigma = sy + H abar, eps = sigma / E + abar
        abar = r["state"][36]
        sigma = r["stress"][0]
        assert sigma == pytest.approx(sy + h * abar, rel=1e-12)
        assert eps == pytest.approx(sigma / E_STEEL + abar, rel=1e-10)


def test_j2_voce_exact():
    """Voce saturation: sigma = sy(abar) on a uniaxial path (test_plasticity.cpp:236-249)."""
    sy, q, d = 250e6, 100e6, 20.0
    m = O.material(E_STEEL, NU, sy0=sy, Q=q, delta=d)
    for eps in (3e-3, 1e-2, 5e-2):
        r, _ = uniaxial_strain_history(m, eps)
        abar = r["state"][36]
        assert r["stress"][0] == pytest.approx(sy + q * (1 - np.exp(-d * abar)), rel=1e-12)


def test_j2_reversed_path_kinematic():
    """Iso + Prager kinematic hardening, loading then reversal: the reversed
    yield point is at sigma_n - 2 sy(abar) (test_plasticity.cpp:198-234)."""
    sy, h, hk = 250e6, 1e9, 2e9
    m = O.material(E_STEEL, NU, sy0=sy, H=h, Hkin=hk)
    r1, s1 = uniaxial_strain_history(m, 5e-3)
    # elastic unloading step by step until reversed yield; the radius grows with abar
    abar = r1["state"][36]
    beta_x = r1["state"][6]  # tensorial back stress, deviatoric part along x
    # uniaxial back stress sigma-space centre: 3/2 beta_11
    centre = 1.5 * beta_x
    radius = sy + h * abar
    assert r1["stress"][0] == pytest.approx(centre + radius, rel=1e-10)


def test_j2_tangent_is_derivative():
    """Consistent tangent vs central differences (test_plasticity.cpp:323-350)."""
    m = O.material(E_STEEL, NU, sy0=250e6, H=2e9, Hkin=1e9, Q=50e6, delta=15.0)
    rng = np.random.default_rng(3)
    for _ in range(5):
        strain = rng.normal(0, 3e-3, 6)
        state = np.zeros(39)
        state[0:6] = rng.normal(0, 1e-3, 6)
        state[0:3] -= state[0:3].mean()
        state[36] = 0.01
        r = O.plastic_return(m, strain, state)
        h = 1e-8
        fd = derivative(lambda e: O.plastic_return(m, e, state)["stress"], strain, h)
        assert relative(r["tangent"], fd) <= 1e-6


# ---------------------------------------------------------------------------
# Hill48 and the general return (test_anisotropy.cpp)
# ---------------------------------------------------------------------------
def test_hill_coefficients_closed_form():
    """F, G, H, N from the r-values (Material.hpp:92-102; test_anisotropy.cpp:255-300)."""
    r0, r45, r90 = 1.9, 1.5, 2.3
    p, c = O.hill_yield_matrix(r0, r45, r90)
    F, G, H, L, M, N = c
    assert G == pytest.approx(1 / (1 + r0), rel=1e-15)
    assert H == pytest.approx(r0 / (1 + r0), rel=1e-15)
    assert F == pytest.approx(r0 / (r90 * (1 + r0)), rel=1e-15)
    assert N == pytest.approx((r0 + r90) * (1 + 2 * r45) / (2 * r90 * (1 + r0)), rel=1e-15)
    # r = 1: von Mises
    p1, c1 = O.hill_yield_matrix(1.0, 1.0, 1.0)
    assert np.abs(p1 - O.von_mises_yield_matrix()).max() <= 1e-15
    # normal rows sum to zero: pressure-insensitive; symmetric
    assert np.abs(p[0:3, 0:3].sum(axis=1)).max() <= 1e-15
    assert np.abs(p - p.T).max() <= 1e-15


def test_hill_directional_yield_and_r_values():
    """Uniaxial stress along angle theta to RD yields at sigma_theta with the
    Hill48 r-value (test_anisotropy.cpp:390-436): the equivalent stress of a
    unit uniaxial stress at theta is 1/ (sigma_theta/sigma_0) and the plastic
    strain ratio of the flow is r(theta)."""
    r0, r45, r90 = 1.9, 1.5, 2.3
    p, c = O.hill_yield_matrix(r0, r45, r90)
    F, G, H, L, M, N = c
    for theta_deg, r_expected in ((0, r0), (45, r45), (90, r90)):
        t = np.radians(theta_deg)
        n = np.array([np.cos(t) ** 2, np.sin(t) ** 2, 0, np.sin(t) * np.cos(t), 0, 0])
        m = p @ n  # flow direction (engineering)
        # plastic strain in the sheet: width direction (perpendicular in plane) over thickness
        w = np.array([np.sin(t) ** 2, np.cos(t) ** 2, 0, -2 * np.sin(t) * np.cos(t), 0, 0])
        # width strain = m . (tensorial components of w-direction projector)
        eps_w = m[0] * np.sin(t) ** 2 + m[1] * np.cos(t) ** 2 - m[3] * np.sin(t) * np.cos(t)
        eps_t = m[2]
        assert eps_w / eps_t == pytest.approx(r_expected, rel=1e-12)
        del w


def general_return_material(criterion_hill=True, backstresses=(), hk=0.0, h=1.5e9, q=0.0, d=0.0):
    hill = dict(r0=1.9, r45=1.5, r90=2.3, rd=(1.0, 0.5, 0.0)) if criterion_hill else None
    return O.material(E_STEEL, NU, sy0=250e6, H=h, Hkin=hk, Q=q, delta=d, hill=hill, backstresses=backstresses)


def test_general_return_von_mises_limit_is_radial():
    """The general return with r = 1 and no backstresses reproduces the
    radial return (test_anisotropy.cpp:523-576)."""
    radial = O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=80e6, delta=12.0)
    general = O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=80e6, delta=12.0,
                         hill=dict(r0=1.0, r45=1.0, r90=1.0))
    rng = np.random.default_rng(11)
    for _ in range(10):
        strain = rng.normal(0, 4e-3, 6)
        a = O.plastic_return(radial, strain)
        b = O.plastic_return(general, strain)
        assert a["yielding"] == b["yielding"]
        assert relative(b["stress"], a["stress"]) <= 1e-11
        assert relative(b["tangent"], a["tangent"]) <= 1e-9
        assert b["state"][36] == pytest.approx(a["state"][36], rel=1e-10, abs=1e-16)


def test_chaboche_monotonic_uniaxial_closed_form():
    """von Mises + Armstrong-Frederick backstresses + Voce, monotonic
    uniaxial loading: sigma = sy(p) + sum_i C_i/gamma_i (1 - exp(-gamma_i p))
    (exponential integration is exact on a proportional path;
    test_anisotropy.cpp:578-647)."""
    sy, q, d = 250e6, 100e6, 20.0
    back = ((40e9, 400.0), (5e9, 20.0))
    m = O.material(E_STEEL, NU, sy0=sy, Q=q, delta=d, backstresses=back)
    for eps in (2e-3, 6e-3, 2e-2):
        r, _ = uniaxial_strain_history(m, eps)
        p = r["state"][36]
        expected = sy + q * (1 - np.exp(-d * p)) + sum(c / g * (1 - np.exp(-g * p)) for c, g in back)
        assert r["stress"][0] == pytest.approx(expected, rel=1e-11)


@pytest.mark.parametrize("case", ["hill", "chaboche", "hill+chaboche+prager", "backward euler", "hill+linear"])
def test_general_tangent_is_derivative(case):
    """A plastic step from a committed plastic state: the consistent tangent is
    the derivative of the stress, non-symmetric iff a backstress recovers; a
    step back inside is elastic; the return from the converged state at the
    same strain reproduces it; without the tangent, the same stress
    (test_anisotropy.cpp:789-877)."""
    hill = dict(r0=1.9, r45=1.5, r90=2.3, rd=(1.0, 0.5, 0.0))
    tilted = dict(r0=1.9, r45=1.5, r90=2.3, rd=(1.0, 0.4, 0.3), nd=(-0.3, 0.1, 1.0))
    back = ((40e9, 400.0), (5e9, 20.0))
    mats = {
        "hill": (O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=60e6, delta=20.0, hill=hill), True),
        "chaboche": (O.material(E_STEEL, NU, sy0=250e6, H=0.5e9, Q=60e6, delta=25.0, backstresses=back), False),
        "hill+chaboche+prager": (O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=60e6, delta=20.0, Hkin=2e9,
                                            hill=tilted, backstresses=back), False),
        "backward euler": (O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=60e6, delta=20.0, Hkin=2e9,
                                      hill=tilted, backstresses=back, integration=1), False),
        "hill+linear": (O.material(E_STEEL, NU, sy0=250e6, H=1.5e9, Q=60e6, delta=20.0, Hkin=1e9, hill=hill,
                                   backstresses=((6e9, 0.0),)), True),
    }
    m, symmetric = mats[case]
    rng = np.random.default_rng(3)
    first = rng.uniform(-1, 1, 6) * 4e-3
    second = first + rng.uniform(-1, 1, 6) * 3e-3
    r1 = O.plastic_return(m, first)
    assert r1["yielding"]
    r2 = O.plastic_return(m, second, r1["state"])
    assert r2["yielding"]
    assert r2["symmetric"] == symmetric
    fd = derivative(lambda e: O.plastic_return(m, e, r1["state"])["stress"], second, 1e-7)
    assert relative(r2["tangent"], fd) <= 1e-8
    asym = relative(r2["tangent"], r2["tangent"].T)
    if symmetric:
        assert asym <= 1e-12
    else:
        assert asym >= 1e-4
    back_step = second - 0.1 * (second - first)
    rb = O.plastic_return(m, back_step, r2["state"])
    assert not rb["yielding"]
    fdb = derivative(lambda e: O.plastic_return(m, e, r2["state"])["stress"], back_step, 1e-7)
    assert relative(rb["tangent"], fdb) <= 1e-8
    again = O.plastic_return(m, second, r2["state"])
    assert not again["yielding"]
    assert relative(again["stress"], r2["stress"]) <= 1e-12
    lean = O.plastic_return(m, second, r1["state"], want_tangent=False)
    assert relative(lean["stress"], r2["stress"]) == 0.0


def test_general_return_state_on_surface():
    """The returned state's yield function is round-off (test_anisotropy.cpp:940-995)."""
    m = general_return_material(True, backstresses=((40e9, 400.0), (5e9, 20.0)), q=60e6, d=25.0)
    p, _ = O.hill_yield_matrix(1.9, 1.5, 2.3, rd=(1.0, 0.5, 0.0))
    rng = np.random.default_rng(23)
    g = E_STEEL / (2 * (1 + NU))
    for _ in range(10):
        strain = rng.normal(0, 5e-3, 6)
        r = O.plastic_return(m, strain)
        if not r["yielding"]:
            continue
        st = r["state"]
        el = strain - st[0:6]
        mean = el[0:3].sum() / 3
        s = np.concatenate([2 * g * (el[0:3] - mean), g * el[3:6]])
        xi = s - st[12:18] - st[18:24]
        sy = 250e6 + 1.5e9 * st[36] + 60e6 * (1 - np.exp(-25 * st[36]))
        f = np.sqrt(xi @ p @ xi) - sy
        assert abs(f) <= 1e-12 * sy
