"""Known answers that pin the oracle where finite differences cannot: the
stored energies, the backward-Euler integration of the backstresses, the
Coulomb slip force and the penalty normal force, and a whole elastoplastic
path that the reference itself computed.

Each check returns its worst relative error; the test asserts it against the
bar and then runs the same check with a mistake planted in the oracle
(oracle/om.hpp, enum Plant: a flipped Voce energy term, 1/2 for 3/4 in the
backstress energy, chi = dl in backward Euler, a slip limit of mu kappa A, a
normal force 1e-6 too large, a Voce rate 1 % high) and asserts that it fails.

Sources of the known answers (the reference's own tests and data):
  tests/test_plasticity.cpp:137-153 (sigma_y' and the isotropic energy are
    the derivative and the integral of the hardening law), :156-196 (uniaxial
    linear hardening: exact curve and stored energy sigma^2/2E + H_tot eps_p^2/2);
  tests/test_anisotropy.cpp:438-475 (Hill48 uniaxial along any direction:
    exact curve and energy sigma^2/2E + H abar^2/2), :578-644 (Chaboche:
    monotonic closed form; backward Euler is first order);
  tests/test_forming.cpp:374-460 (flat punch: F = E A delta / H / (1 + E /
    (kappa H))), :500-600 (a slipping node carries mu p_N along its slip);
  benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab/out (one
    Hex8, J2 + Voce, logarithmic strain, pulled and pushed back: the element's
    equivalent plastic strain and von Mises stress at every step end).
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from sparlab_gpu import workloads as W

E, NU = 200e9, 0.3
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return abs(a - b) / abs(b)


def uniaxial_stress(m, eps_x, state=None, lateral=None):
    """sigma_ii = 0 except sigma_11: Newton on the other five strain
    components (shears included, which an anisotropic law couples)."""
    strain = np.zeros(6) if lateral is None else lateral.copy()
    strain[0] = eps_x
    if lateral is None:
        strain[1] = strain[2] = -NU * eps_x
    for _ in range(60):
        r = O.plastic_return(m, strain, state)
        res = r["stress"][1:]
        if np.abs(res).max() <= 1e-13 * max(1.0, abs(r["stress"][0])):
            break
        strain[1:] -= np.linalg.solve(r["tangent"][1:, 1:], res)
    return r, strain


def check_plant(check, plant, bar):
    with O.planted(plant):
        err = check()
    assert err > bar, f"planted mistake {plant} passed the check ({err:.3e} <= {bar:.1e})"


# ---------------------------------------------------------------------------
# Hardening laws and stored energies
# ---------------------------------------------------------------------------
def hardening_error():
    """test_plasticity.cpp:137-153: d sigma_y / da = slope, d psi_iso / da =
    sigma_y - sigma_y0, psi_iso(0) = 0 (linear + Voce law)."""
    m = O.material(E, NU, sy0=250e6, H=1.5e9, Q=120e6, delta=25.0)
    worst = abs(O.hardening(m, 0.0)[2])
    h = 1e-7
    for a in (0.003, 0.04, 0.2):
        y, slope, _ = O.hardening(m, a)
        fd_y = (O.hardening(m, a + h)[0] - O.hardening(m, a - h)[0]) / (2 * h)
        fd_psi = (O.hardening(m, a + h)[2] - O.hardening(m, a - h)[2]) / (2 * h)
        worst = max(worst, rel(fd_y, slope), rel(fd_psi, y - 250e6))
    return worst


def test_hardening_derivatives_and_energy_integral():
    assert hardening_error() <= 1e-6
    check_plant(hardening_error, O.PLANT_VOCE_ENERGY_SIGN, 1e-6)


def linear_uniaxial_energy_error():
    """test_plasticity.cpp:156-196: linear isotropic + Prager, one step and 40
    steps to eps = 0.01: sigma = E (sy + Ht eps) / (E + Ht), energy =
    sigma^2 / 2E + Ht eps_p^2 / 2 with Ht = H + H_kin."""
    worst = 0.0
    for h, hk in ((0.0, 0.0), (2e9, 0.0), (0.0, 5e9), (1e9, 3e9)):
        m = O.material(E, NU, sy0=250e6, H=h, Hkin=hk)
        ht = h + hk
        target = 0.01
        exact = E * (250e6 + ht * target) / (E + ht)
        eps_p = target - exact / E
        r, _ = uniaxial_stress(m, target)
        state, lat = None, None
        for k in range(1, 41):
            r40, lat = uniaxial_stress(m, target * k / 40.0, state, lat)
            state = r40["state"]
        for res in (r, r40):
            worst = max(worst, rel(res["stress"][0], exact),
                        rel(res["energy"], exact * exact / (2 * E) + 0.5 * ht * eps_p * eps_p))
    return worst


def test_linear_hardening_uniaxial_stress_and_energy():
    assert linear_uniaxial_energy_error() <= 1e-10


def hill_phi(r0, r45, r90, theta):
    """Hill48 sigma_eq^2 / sigma^2 of uniaxial stress at theta to RD, from
    F, G, H, N of the r-values (Material.cpp:140-151 in closed form)."""
    g = 1.0 / (1.0 + r0)
    hh = r0 / (1.0 + r0)
    f = r0 / (r90 * (1.0 + r0))
    n = (r0 + r90) * (1.0 + 2.0 * r45) / (2.0 * r90 * (1.0 + r0))
    c, s = np.cos(theta), np.sin(theta)
    return f * s ** 4 + g * c ** 4 + hh * (c * c - s * s) ** 2 + 2.0 * n * s * s * c * c


def hill_uniaxial_energy_error():
    """test_anisotropy.cpp:438-475: Hill48 (r 1.9 / 1.5 / 2.3) with linear
    isotropic hardening loaded at theta to RD: sigma = E (k sy + k^2 H eps) /
    (E + k^2 H), k = phi^-1/2, and energy sigma^2/2E + H abar^2/2."""
    hmod, sy, target = 2e9, 250e6, 0.012
    worst = 0.0
    for deg in (0.0, 22.5, 45.0, 67.5, 90.0):
        th = np.radians(deg)
        m = O.material(E, NU, sy0=sy, H=hmod, hill=dict(r0=1.9, r45=1.5, r90=2.3,
                                                        rd=(np.cos(-th), np.sin(-th), 0.0)))
        k = 1.0 / np.sqrt(hill_phi(1.9, 1.5, 2.3, th))
        exact = E * (k * sy + k * k * hmod * target) / (E + k * k * hmod)
        state, lat = None, None
        for step in range(1, 41):
            r, lat = uniaxial_stress(m, target * step / 40.0, state, lat)
            state = r["state"]
        abar = state[36]
        worst = max(worst, rel(r["stress"][0], exact), rel(abar, k * (target - exact / E)),
                    rel(r["energy"], exact * exact / (2 * E) + 0.5 * hmod * abar * abar))
    return worst


def test_hill48_uniaxial_stress_and_energy():
    assert hill_uniaxial_energy_error() <= 1e-10


def chaboche_energy_error():
    """von Mises + Voce + two Armstrong-Frederick backstresses, monotonic
    uniaxial stress (exponential integration is exact on this path,
    test_anisotropy.cpp:578-620): X_i = C_i/g_i (1 - exp(-g_i p)),
    sigma = sy(p) + sum X_i, and the stored energy sigma^2/2E + H p^2/2 +
    Q (p - (1 - exp(-d p))/d) + sum X_i^2 / 2 C_i (the quadratic energy of
    each backstress, X_i = C_i times its strain-like variable)."""
    sy, h, q, d = 250e6, 1.5e9, 100e6, 20.0
    back = ((40e9, 400.0), (5e9, 20.0))
    m = O.material(E, NU, sy0=sy, H=h, Q=q, delta=d, backstresses=back)
    worst = 0.0
    for eps in (2e-3, 6e-3, 2e-2):
        r, _ = uniaxial_stress(m, eps)
        p = r["state"][36]
        x = [c / g * -np.expm1(-g * p) for c, g in back]
        sigma = sy + h * p + q * -np.expm1(-d * p) + sum(x)
        energy = (sigma * sigma / (2 * E) + 0.5 * h * p * p + q * (p + np.expm1(-d * p) / d)
                  + sum(xi * xi / (2 * c) for xi, (c, _) in zip(x, back)))
        worst = max(worst, rel(r["stress"][0], sigma), rel(r["energy"], energy))
    return worst


def test_chaboche_uniaxial_stress_and_energy():
    assert chaboche_energy_error() <= 1e-10
    check_plant(chaboche_energy_error, O.PLANT_VOCE_ENERGY_SIGN, 1e-10)
    check_plant(chaboche_energy_error, O.PLANT_KINEMATIC_ENERGY, 1e-10)


# ---------------------------------------------------------------------------
# Backward-Euler integration of the backstresses
# ---------------------------------------------------------------------------
BACK = ((40e9, 400.0), (5e9, 20.0))


def backward_euler_material():
    return O.material(E, NU, sy0=250e6, H=1.5e9, Q=100e6, delta=20.0, backstresses=BACK, integration=1)


def backward_euler_step_error():
    """The backward-Euler update alpha (1 + g dl) = alpha_n + (2/3) C dl m
    (Plasticity.cpp:192-212) on uniaxial stress is X = (X_n + C dp) / (1 +
    g dp) per backstress, X = (3/2) alpha_11: three steps from a virgin state,
    each checked against that recursion and sigma = sy(p) + sum X."""
    m = backward_euler_material()
    sy = lambda p: 250e6 + 1.5e9 * p + 100e6 * -np.expm1(-20.0 * p)
    state, lat, x, p_old = None, None, [0.0, 0.0], 0.0
    worst = 0.0
    for eps in (4e-3, 1e-2, 2.5e-2):
        r, lat = uniaxial_stress(m, eps, state, lat)
        state = r["state"]
        p = state[36]
        dp = p - p_old
        x = [(xi + c * dp) / (1.0 + g * dp) for xi, (c, g) in zip(x, BACK)]
        worst = max(worst, rel(r["stress"][0], sy(p) + sum(x)))
        for i, xi in enumerate(x):  # alpha_i,11 in the state (39-record: alpha_i at 12 + 6 i)
            worst = max(worst, rel(1.5 * state[12 + 6 * i], xi))
        p_old = p
    return worst


def test_backward_euler_step_closed_form():
    assert backward_euler_step_error() <= 1e-10
    check_plant(backward_euler_step_error, O.PLANT_BACKWARD_EULER_CHI, 1e-10)


def backward_euler_convergence():
    """test_anisotropy.cpp:627-644: at eps = 0.02 the backward-Euler stress
    error against the exact (exponential) solution halves as the steps double."""
    m = backward_euler_material()
    target = 0.02

    def exact_sigma(eps):
        f = lambda p: 250e6 + 1.5e9 * p + 100e6 * -np.expm1(-20.0 * p) + sum(
            c / g * -np.expm1(-g * p) for c, g in BACK)
        lo, hi = 0.0, eps
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if mid + f(mid) / E - eps > 0.0:
                hi = mid
            else:
                lo = mid
        return f(0.5 * (lo + hi))

    sigma = exact_sigma(target)
    errors = []
    for steps in (10, 20, 40, 80):
        state, lat = None, None
        for k in range(1, steps + 1):
            r, lat = uniaxial_stress(m, target * k / steps, state, lat)
            state = r["state"]
        errors.append(abs(r["stress"][0] - sigma) / sigma)
    return errors


def test_backward_euler_is_first_order():
    errors = backward_euler_convergence()
    assert errors[0] > 1e-4
    for a, b in zip(errors, errors[1:]):
        assert a / b == pytest.approx(2.0, rel=0.1), errors


# ---------------------------------------------------------------------------
# Contact: Coulomb slip and the penalty normal force
# ---------------------------------------------------------------------------
def slip_error():
    """A slipping node carries |F_T| = mu kappa A |g| tangent to the tool
    (RigidTool.cpp:398-555; test_forming.cpp:500-600); from a history with no
    force, against its slip. Plane, sphere and cylinder; random nodes,
    penetrations, histories and slips; stick nodes carry -kappa_T delta."""
    rng = np.random.default_rng(17)
    mu, ka = 0.2, 3.0e7
    worst = 0.0
    slipped = stuck = 0
    tools = (O.tool_record(shape=0, centre=(0, 0, 0), normal=(0, 0, -1), friction=mu),
             O.tool_record(shape=1, centre=(0, 0, 0.0105), radius=0.01, friction=mu),
             O.tool_record(shape=2, centre=(0, 0, 0.0105), radius=0.01, axis=(0, 1, 0), friction=mu))
    for tool in tools:
        ns = 40
        coords = np.zeros((ns, 3))
        coords[:, 0] = rng.uniform(-2e-3, 2e-3, ns)
        coords[:, 1] = rng.uniform(-2e-3, 2e-3, ns)
        u = np.zeros((ns, 3))
        u[:, 2] = rng.uniform(1e-6, 2e-5, ns)  # upwards, into the tool
        x = coords + u
        slip = rng.normal(size=(ns, 3))
        scale = np.where(rng.uniform(size=ns) < 0.7, 1e-4, 1e-9)  # mostly slipping, some sticking
        hp = np.ones(ns, np.int32)
        hf = np.zeros((ns, 3))
        with_force = rng.uniform(size=ns) < 0.5
        hf[with_force] = rng.normal(size=(int(with_force.sum()), 3)) * 50.0
        hr = (x - tool[1:4]) - slip * scale[:, None]
        nodes = np.arange(ns, dtype=np.int32)
        out = O.contact(tool, coords, u.reshape(-1), np.zeros(3 * ns), nodes, np.full(ns, ka),
                        history=(hp, hf, hr))
        for i in range(ns):
            if not out["active"][i]:
                continue
            ft = out["t_force"][i]
            force = -out["residual"][3 * i:3 * i + 3]
            fn = force - ft  # the normal part p_N n
            pn = np.linalg.norm(fn)
            nvec = fn / pn
            worst = max(worst, abs(ft @ nvec) / max(np.linalg.norm(ft), 1e-300) if np.linalg.norm(ft) > 0 else 0.0)
            w = (x[i] - tool[1:4]) - hr[i]
            delta = w - (w @ nvec) * nvec
            trial_free = -ka * delta  # the trial with no force in the history
            if np.linalg.norm(hf[i]) == 0.0 and np.linalg.norm(trial_free) <= mu * pn:
                stuck += 1
                worst = max(worst, np.linalg.norm(ft - trial_free) / np.linalg.norm(trial_free))
            elif np.linalg.norm(hf[i]) == 0.0 or np.linalg.norm(ft) > 0.999 * mu * pn:
                slipped += 1
                worst = max(worst, rel(np.linalg.norm(ft), mu * pn))
                if np.linalg.norm(hf[i]) == 0.0:
                    worst = max(worst, np.linalg.norm(ft / np.linalg.norm(ft) + delta / np.linalg.norm(delta)))
    assert slipped > 20 and stuck >= 3
    return worst


def test_slipping_nodes_carry_mu_times_the_normal_force():
    assert slip_error() <= 1e-12
    check_plant(slip_error, O.PLANT_SLIP_RADIUS, 1e-12)


def flat_punch_error():
    """test_forming.cpp:374-460: a 4 x 4 x 10 mm column of one Hex8 in
    section and four through its height, on rollers, pressed by a rigid
    plane lowered by delta, small strain, elastic: the punch force is
    E A delta / H / (1 + E / (kappa H)) with kappa = s E / sqrt(A / 4), for
    s = 10, 100, 1000 (the penalty in series with the column), every top
    node active."""
    a_side, height, delta = 0.004, 0.01, 1.0e-5
    area = a_side * a_side
    coords, conn = W.structured_hex_mesh(1, 1, 4, a_side, a_side, height)
    nn = coords.shape[0]
    k = np.zeros((3 * nn, 3 * nn))
    for e in range(conn.shape[0]):
        ke = O.linear_stiffness(coords[conn[e]], E, NU)
        dofs = np.array([3 * n + c for n in conn[e] for c in range(3)])
        k[np.ix_(dofs, dofs)] += ke
    fixed = np.zeros(3 * nn, bool)
    tol = 1e-12
    for n, (x, y, z) in enumerate(coords):
        if abs(z) < tol:
            fixed[3 * n + 2] = True
            if abs(x) < tol and abs(y) < tol:
                fixed[3 * n] = fixed[3 * n + 1] = True
            if abs(x - a_side) < tol and abs(y) < tol:
                fixed[3 * n + 1] = True
    top = np.nonzero(np.abs(coords[:, 2] - height) < tol)[0].astype(np.int32)
    free = ~fixed
    rigid = E * area * delta / height
    worst = 0.0
    for s in (10.0, 100.0, 1000.0):
        kappa = s * E / np.sqrt(area / 4.0)
        ka = np.full(top.size, kappa * area / 4.0)
        tool = O.tool_record(shape=0, centre=(0, 0, height - delta), normal=(0, 0, -1))
        u = np.zeros(3 * nn)
        for _ in range(6):
            c = O.contact(tool, coords, u, np.zeros(3 * nn), top, ka)
            r = k @ u + c["residual"]
            jac = k.copy()
            for i, node in enumerate(top):
                d = slice(3 * node, 3 * node + 3)
                jac[d, d] += c["blocks"][i]
            du = np.zeros(3 * nn)
            du[free] = np.linalg.solve(jac[np.ix_(free, free)], -r[free])
            u += du
            if np.abs(du).max() <= 1e-16:
                break
        c = O.contact(tool, coords, u, np.zeros(3 * nn), top, ka)
        assert c["n_active"] == 4
        exact = rigid / (1.0 + E / (kappa * height))
        worst = max(worst, rel(c["resultant"][2], exact),
                    rel(c["max_penetration"], exact / (area * kappa)))
    return worst


def test_flat_punch_force_with_the_penalty_in_series():
    assert flat_punch_error() <= 1e-9
    check_plant(flat_punch_error, O.PLANT_NORMAL_FORCE, 1e-9)


# ---------------------------------------------------------------------------
# The reference's own homogeneous uniaxial run
# ---------------------------------------------------------------------------
VTK_ORDER = [0, 1, 3, 2, 4, 5, 7, 6]  # the structured mesh's element (mesh.json)


def homogeneous_uniaxial_error():
    """Every step end of the reference's run: the element evaluated from the
    previous step's committed history at the step's %.17g displacements (one
    return per step: each step is a monotonic branch along one deviatoric
    direction, on which the radial return from the step's start is exact),
    then the stress recovery; its max equivalent plastic strain and von
    Mises stress against the run's."""
    g = json.load(open(os.path.join(HERE, "golden", "homogeneous_uniaxial.json")))
    p = g["material"]["plasticity"]
    m = O.material(g["material"]["youngs_modulus"], g["material"]["poisson_ratio"], sy0=p["yield_stress"],
                   H=p["hardening_modulus"], Hkin=p["kinematic_hardening_modulus"], Q=p["saturation_stress"],
                   delta=p["saturation_rate"])
    assert g["kinematics"] == "finite_logarithmic"
    states, alpha = O.virgin_states(8), np.zeros(9)
    worst = 0.0
    for step in g["steps"]:
        nodes = np.array(step["nodes"])[VTK_ORDER]
        x0 = nodes[:, 1:4]
        ue = nodes[:, 4:7].reshape(-1)
        ev = O.element(x0, ue, m, states, alpha, kinematics=2, want_tangent=False)
        states, alpha = ev["states"], ev["alpha"]
        st = O.element_stress_batch(x0[None], ue[None], m, states[None], alpha[None], kinematics=2)
        fields = st["stress"][0]
        _, eq_plastic, von_mises = step["elements"][0]
        worst = max(worst, rel(fields[25], eq_plastic), rel(fields[24], von_mises))
    return worst


def test_reference_homogeneous_uniaxial_run():
    assert homogeneous_uniaxial_error() <= 1e-10
    check_plant(homogeneous_uniaxial_error, O.PLANT_VOCE_YIELD, 1e-10)
