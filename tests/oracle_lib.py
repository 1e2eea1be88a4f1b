"""ctypes access to the oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the CPU restatement of the reference's element-and-assembly
path (oracle/*.hpp); this module only marshals arrays. State records are 39
doubles per integration point and material records 54 doubles, laid out as
oracle/oracle_capi.cpp documents.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

STATE = 39
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_element_flops.restype = ctypes.c_longlong
        _lib.oracle_pattern.restype = ctypes.c_longlong
        _lib.oracle_gather_force.restype = ctypes.c_double
    return _lib


# Planted mistakes (oracle/om.hpp, enum Plant): a known-answer test runs its
# check once more with one planted, to show that the check catches it.
PLANT_VOCE_ENERGY_SIGN = 1
PLANT_KINEMATIC_ENERGY = 2
PLANT_BACKWARD_EULER_CHI = 3
PLANT_SLIP_RADIUS = 4
PLANT_NORMAL_FORCE = 5
PLANT_VOCE_YIELD = 6


class planted:
    """with planted(PLANT_...): the oracle computes with that mistake."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().oracle_set_plant(int(self.which))
        return self

    def __exit__(self, *exc):
        lib().oracle_set_plant(0)
        return False


def hardening(mat, a):
    """(yield stress, its slope, isotropic hardening energy) at abar = a."""
    out = np.zeros(3)
    lib().oracle_hardening(dp(np.ascontiguousarray(mat, dtype=np.float64)), ctypes.c_double(a), dp(out))
    return out


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


class OracleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def _check(status, buf):
    if status != 0:
        raise OracleError(status, buf.value.decode(errors="replace"))


def hill_yield_matrix(r0=1.0, r45=1.0, r90=1.0, rd=(1.0, 0.0, 0.0), nd=(0.0, 0.0, 1.0),
                      calibration=0, s45=1.0, s90=1.0, sb=1.0, L=1.5, M=1.5,
                      F=0.5, G=0.5, H=0.5, N=1.5):
    vals = np.array([r0, r45, r90, s45, s90, sb, L, M, F, G, H, N], dtype=np.float64)
    p = np.zeros(36)
    coeffs = np.zeros(6)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_hill_yield_matrix(calibration, dp(vals), dp(np.array(rd, float)),
                                        dp(np.array(nd, float)), dp(p), dp(coeffs), buf, 512)
    _check(st, buf)
    return p.reshape(6, 6), coeffs


def von_mises_yield_matrix():
    p = np.zeros(36)
    lib().oracle_von_mises_yield_matrix(dp(p))
    return p.reshape(6, 6)


def material(E, nu, sy0=0.0, H=0.0, Hkin=0.0, Q=0.0, delta=0.0, hill=None,
             backstresses=(), integration=0):
    """The 54-double material record. hill: dict of hill_yield_matrix args."""
    m = np.zeros(54)
    m[0], m[1], m[2], m[3], m[4], m[5], m[6] = E, nu, sy0, H, Hkin, Q, delta
    if hill is not None:
        m[7] = 1
        m[18:] = hill_yield_matrix(**hill)[0].ravel()
    else:
        m[18:] = von_mises_yield_matrix().ravel()
    m[8] = len(backstresses)
    for i, (c, g) in enumerate(backstresses):
        m[9 + i] = c
        m[13 + i] = g
    m[17] = integration
    return m


def rule_points(in_plane=2, thickness=0, axis=2):
    return lib().oracle_rule_points(in_plane, thickness, axis)


def rule(in_plane=2, thickness=0, axis=2):
    n = rule_points(in_plane, thickness, axis)
    xi, eta, zeta, w = (np.zeros(n) for _ in range(4))
    lib().oracle_rule(in_plane, thickness, axis, dp(xi), dp(eta), dp(zeta), dp(w))
    return xi, eta, zeta, w


def virgin_states(points):
    return np.zeros((points, STATE))


def plastic_return(mat, strain, state=None, want_tangent=True):
    strain = np.ascontiguousarray(strain, dtype=np.float64)
    state = np.zeros(STATE) if state is None else np.ascontiguousarray(state, dtype=np.float64)
    stress, tangent, out, sc = np.zeros(6), np.zeros(36), np.zeros(STATE), np.zeros(3)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_plastic_return(dp(mat), dp(strain), dp(state), int(want_tangent), dp(stress),
                                     dp(tangent), dp(out), dp(sc), buf, 512)
    _check(st, buf)
    return dict(stress=stress, tangent=tangent.reshape(6, 6), state=out, energy=sc[0],
                yielding=bool(sc[1]), symmetric=bool(sc[2]))


def log_strain(green, t=None):
    green = np.ascontiguousarray(green, dtype=np.float64)
    axes, mu, strain, proj, first = np.zeros(9), np.zeros(3), np.zeros(6), np.zeros(36), np.zeros(9)
    s, curv = np.zeros(6), np.zeros(36)
    tt = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_log_strain(dp(green), dp(tt), dp(axes), dp(mu), dp(strain), dp(proj), dp(first),
                                 dp(s), dp(curv), buf, 512)
    _check(st, buf)
    return dict(axes=axes.reshape(3, 3), mu=mu, strain=strain, projection=proj.reshape(6, 6),
                first=first.reshape(3, 3), s=s, curvature=curv.reshape(6, 6))


def element(x0, ue, mat, states=None, alpha=None, rule=(2, 0, 2), kinematics=2, want_tangent=True):
    """One incompatible-mode Hex8 evaluation. x0: (8, 3); ue: (24,)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(24)
    ue = np.ascontiguousarray(ue, dtype=np.float64).reshape(24)
    npts = rule_points(*rule)
    states = virgin_states(npts) if states is None else np.ascontiguousarray(states, dtype=np.float64)
    assert states.shape == (npts, STATE)
    a_in = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    f, k, a_out = np.zeros(24), np.zeros(576), np.zeros(9)
    s_out, sc = np.zeros((npts, STATE)), np.zeros(6)
    buf = ctypes.create_string_buffer(1024)
    st = lib().oracle_element(dp(x0), dp(ue), dp(states), dp(a_in), dp(mat), rule[0], rule[1], rule[2],
                              kinematics, int(want_tangent), dp(f), dp(k), dp(a_out), dp(s_out), dp(sc),
                              buf, 1024)
    _check(st, buf)
    return dict(force=f, tangent=k.reshape(24, 24), alpha=a_out, states=s_out, energy=sc[0],
                yielding=int(sc[1]), iterations=int(sc[2]), symmetric=bool(sc[3]), passes=int(sc[4]),
                path=int(sc[5]))


def element_flops(x0, ue, mat, states=None, alpha=None, rule=(2, 0, 2), kinematics=2, want_tangent=True):
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(24)
    ue = np.ascontiguousarray(ue, dtype=np.float64).reshape(24)
    npts = rule_points(*rule)
    states = virgin_states(npts) if states is None else np.ascontiguousarray(states, dtype=np.float64)
    a_in = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    return lib().oracle_element_flops(dp(x0), dp(ue), dp(states), dp(a_in), dp(mat), rule[0], rule[1],
                                      rule[2], kinematics, int(want_tangent))


def element_flops_split(x0, ue, mat, states=None, alpha=None, rule=(2, 0, 2), kinematics=2, want_tangent=True):
    """(FLOPs of the local iteration, FLOPs of the coupling blocks and condensation)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(24)
    ue = np.ascontiguousarray(ue, dtype=np.float64).reshape(24)
    npts = rule_points(*rule)
    states = virgin_states(npts) if states is None else np.ascontiguousarray(states, dtype=np.float64)
    a_in = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    parts = np.zeros(2, dtype=np.int64)
    f = lib().oracle_element_flops_split
    f.restype = ctypes.c_longlong
    total = f(dp(x0), dp(ue), dp(states), dp(a_in), dp(mat), rule[0], rule[1], rule[2], kinematics,
              int(want_tangent), parts.ctypes.data_as(ctypes.c_void_p))
    return int(total), int(parts[0]), int(parts[1])


def element_batch(x0, ue, mat, states, alpha=None, rule=(2, 0, 2), kinematics=2, want_tangent=True,
                  threads=0):
    """x0: (E, 8, 3); ue: (E, 24); states: (E, P, 39); alpha: (E, 9) or None."""
    ne = x0.shape[0]
    npts = rule_points(*rule)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    ue = np.ascontiguousarray(ue, dtype=np.float64)
    states = np.ascontiguousarray(states, dtype=np.float64)
    a_in = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    f = np.zeros((ne, 24))
    k = np.zeros((ne, 24, 24)) if want_tangent else None
    a_out = np.zeros((ne, 9))
    s_out = np.zeros((ne, npts, STATE))
    sc = np.zeros((ne, 6))
    status = np.zeros(ne, dtype=np.int32)
    lib().oracle_element_batch(ctypes.c_longlong(ne), dp(x0), dp(ue), dp(states), dp(a_in), dp(mat),
                               rule[0], rule[1], rule[2], kinematics, int(want_tangent), threads, dp(f),
                               dp(k), dp(a_out), dp(s_out), dp(sc), ip(status))
    return dict(force=f, tangent=k, alpha=a_out, states=s_out, scalars=sc, status=status)


STRESS_FIELDS = 28  # cauchy 6, piola_kirchhoff 6, kirchhoff 6, logarithmic_strain 6, von_mises,
#                    max_equivalent_plastic_strain, max_strain, min_jacobian


def element_stress_batch(x0, ue, mat, states, alpha=None, rule=(2, 0, 2), kinematics=2, threads=0):
    """incompatible_elastoplastic_stress for a batch at committed states
    (E, P, 39) and modes (E, 9) or None: (E, 28) measures and statuses."""
    ne = x0.shape[0]
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    ue = np.ascontiguousarray(ue, dtype=np.float64)
    states = np.ascontiguousarray(states, dtype=np.float64)
    a_in = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    out = np.zeros((ne, STRESS_FIELDS))
    status = np.zeros(ne, dtype=np.int32)
    lib().oracle_element_stress_batch(ctypes.c_longlong(ne), dp(x0), dp(ue), dp(states), dp(a_in), dp(mat),
                                      rule[0], rule[1], rule[2], kinematics, threads, dp(out), ip(status))
    return dict(stress=out, status=status)


def standard_force(x0, ue, mat, states=None, rule=(2, 0, 2), kinematics=2):
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(24)
    ue = np.ascontiguousarray(ue, dtype=np.float64).reshape(24)
    npts = rule_points(*rule)
    states = virgin_states(npts) if states is None else np.ascontiguousarray(states, dtype=np.float64)
    f = np.zeros(24)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_standard_force(dp(x0), dp(ue), dp(states), dp(mat), rule[0], rule[1], rule[2],
                                     kinematics, dp(f), buf, 512)
    _check(st, buf)
    return f


def linear_stiffness(x0, E, nu, rule=(2, 0, 2)):
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(24)
    k = np.zeros(576)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_linear_stiffness(dp(x0), ctypes.c_double(E), ctypes.c_double(nu), rule[0], rule[1],
                                       rule[2], dp(k), buf, 512)
    _check(st, buf)
    return k.reshape(24, 24)


def pattern(nn, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    ne = conn.shape[0]
    nnz = lib().oracle_pattern(nn, ne, ip(conn), None, None, None)
    outer = np.zeros(3 * nn + 1, dtype=np.int32)
    inner = np.zeros(nnz, dtype=np.int32)
    bo = np.zeros(ne * 64, dtype=np.int32)
    lib().oracle_pattern(nn, ne, ip(conn), ip(outer), ip(inner), ip(bo))
    return outer, inner, bo


def scatter(nn, conn, ke):
    """Column-major values of K (the reference's valuePtr order)."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    ke = np.ascontiguousarray(ke, dtype=np.float64)
    outer, inner, _ = pattern(nn, conn)
    values = np.zeros(inner.shape[0])
    lib().oracle_scatter(nn, conn.shape[0], ip(conn), dp(ke), dp(values))
    return values


def gather_force(nn, conn, fe):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    fe = np.ascontiguousarray(fe, dtype=np.float64)
    fint = np.zeros(3 * nn)
    gross = lib().oracle_gather_force(nn, conn.shape[0], ip(conn), dp(fe), dp(fint))
    return fint, gross


def add_node_blocks(nn, conn, values, nodes, blocks):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64).copy()
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    blocks = np.ascontiguousarray(blocks, dtype=np.float64)
    lib().oracle_add_node_blocks(nn, conn.shape[0], ip(conn), dp(values), len(nodes), ip(nodes), dp(blocks))
    return values


def stiffness_gross(nn, conn, values, x, free):
    """stiffness_gross on the column-major values: (norm, product (3 nn))."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    free = np.ascontiguousarray(free, dtype=np.int32)
    product = np.zeros(3 * nn)
    L = lib()
    L.oracle_stiffness_gross.restype = ctypes.c_double
    norm = L.oracle_stiffness_gross(nn, conn.shape[0], ip(conn), dp(values), dp(x), ip(free), free.size, dp(product))
    return norm, product


def free_block(nn, conn, values, free):
    """PartitionedFactor::build + assemble: the free-free block of the
    column-major values for ascending free DOFs, as (outer, inner, values,
    position)."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    free = np.ascontiguousarray(free, dtype=np.int32)
    cap = values.size
    outer = np.zeros(free.size + 1, np.int32)
    inner = np.zeros(cap, np.int32)
    out = np.zeros(cap)
    position = np.zeros(cap, np.int64)
    L = lib()
    L.oracle_free_block.restype = ctypes.c_longlong
    nnz = L.oracle_free_block(nn, conn.shape[0], ip(conn), dp(values), ip(free), free.size, ctypes.c_longlong(cap),
                              ip(outer), ip(inner), dp(out), position.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    assert nnz >= 0
    return outer, inner[:nnz].copy(), out[:nnz].copy(), position


def surface_raycast(surface, grid):
    """The height map of a structured surface (ny_s, nx_s, 3) on grid (x0,
    y0, nx, ny, h), before the nearest fill: (z (ny, nx), -inf where nothing
    is hit; hit; the count of folded triangles)."""
    surface = np.ascontiguousarray(surface, dtype=np.float64)
    nys, nxs = surface.shape[:2]
    x0, y0, nx, ny, h = grid
    z = np.zeros((ny, nx))
    hit = np.zeros((ny, nx), np.uint8)
    L = lib()
    L.oracle_surface_raycast.restype = ctypes.c_longlong
    L.oracle_surface_raycast.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_longlong, ctypes.c_longlong,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_longlong, ctypes.c_longlong,
                                         ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_ubyte)]
    folded = L.oracle_surface_raycast(dp(surface), nxs, nys, x0, y0, nx, ny, h, dp(z),
                                      hit.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return z, hit.astype(bool), int(folded)


def tool_record(shape=1, centre=(0, 0, 0), centre0=None, radius=0.0, normal=(0, 0, 1), axis=(0, 0, 1),
                friction=0.0, tangential_ratio=1.0):
    centre0 = centre if centre0 is None else centre0
    return np.array([shape, *centre, *centre0, radius, *normal, *axis, friction, tangential_ratio],
                    dtype=np.float64)


def contact_candidates(tool, coords, u0, nodes, margin):
    """ToolContact::begin_increment for one tool: candidate slave indices."""
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    out = np.zeros(nodes.shape[0], np.int32)
    n = lib().oracle_contact_candidates(dp(tool), dp(coords), dp(u0), nodes.shape[0], ip(nodes),
                                        ctypes.c_double(margin), ip(out))
    return out[:n].copy()


def contact_anticipate(tool, coords, u, du, nodes, kappa_a):
    """ToolContact::anticipate for one tool over the listed slaves:
    (residual 3 nn, blocks (ns, 9), active (ns,), added)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    nn = coords.shape[0] // 3
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    du = np.ascontiguousarray(du, dtype=np.float64).reshape(-1)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    kappa_a = np.ascontiguousarray(kappa_a, dtype=np.float64)
    ns = nodes.shape[0]
    residual, blocks, active = np.zeros(3 * nn), np.zeros(9 * ns), np.zeros(ns, np.int32)
    added = lib().oracle_contact_anticipate(dp(tool), dp(coords), dp(u), dp(du), ns, ip(nodes), dp(kappa_a),
                                            dp(residual), dp(blocks), ip(active))
    return residual, blocks.reshape(ns, 9), active, added


def contact(tool, coords, u, start, nodes, kappa_a, history=None, want_tangent=True, friction_symmetric=False):
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    nn = coords.shape[0] // 3
    u = np.ascontiguousarray(u, dtype=np.float64)
    start = np.ascontiguousarray(start, dtype=np.float64)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    ns = nodes.shape[0]
    kappa_a = np.ascontiguousarray(kappa_a, dtype=np.float64)
    if history is None:
        hp, hf, hr = np.zeros(ns, np.int32), np.zeros(3 * ns), np.zeros(3 * ns)
    else:
        hp, hf, hr = (np.ascontiguousarray(history[0], dtype=np.int32),
                      np.ascontiguousarray(history[1], dtype=np.float64).reshape(-1),
                      np.ascontiguousarray(history[2], dtype=np.float64).reshape(-1))
    residual = np.zeros(3 * nn)
    blocks = np.zeros(9 * ns)
    active = np.zeros(ns, np.int32)
    tp, tf, tr = np.zeros(ns, np.int32), np.zeros(3 * ns), np.zeros(3 * ns)
    sc = np.zeros(10)
    buf = ctypes.create_string_buffer(512)
    st = lib().oracle_contact(dp(tool), int(friction_symmetric), nn, dp(coords), dp(u), dp(start), ns, ip(nodes),
                              dp(kappa_a), ip(hp), dp(hf), dp(hr), int(want_tangent), dp(residual), dp(blocks),
                              ip(active), ip(tp), dp(tf), dp(tr), dp(sc), buf, 512)
    _check(st, buf)
    return dict(residual=residual, blocks=blocks.reshape(ns, 3, 3), active=active, t_present=tp,
                t_force=tf.reshape(ns, 3), t_relative=tr.reshape(ns, 3), gross=sc[0], round_off=sc[1],
                force_norm=sc[2], resultant=sc[3:6], n_active=int(sc[6]), n_slipping=int(sc[7]),
                max_penetration=sc[8], symmetric=bool(sc[9]))
