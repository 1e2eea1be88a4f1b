"""The plain CPU-runnable forming case (the reference's C1 configuration)
through the device path inside a Newton loop, against the oracle through the
same loop.

Blank 20 x 20 x 1 mm, 20 x 20 x 2 incompatible-mode Hex8 (800 elements,
1 323 nodes), edges clamped, E = 70 GPa, nu = 0.33, J2 with Voce hardening
sigma_y = 95 + 150 (1 - exp(-10 eps_p)) MPa, logarithmic strain; a rigid
sphere R = 5 mm pushed 1 mm at the centre in 10 increments against every
top-surface node, Coulomb friction 0.1 with the exact tangent. Each slave
node's spring is the reference's kappa_j A_j = s E sqrt(A_j)
(src/fem/RigidTool.cpp:298-321, include/sparlab/fem/RigidTool.hpp:28-35;
A_j the node's share of its top faces' area, exactly a quarter of each flat
bilinear face's, which the reference's 3 x 3 face quadrature integrates
exactly), s = 1/7 so that an interior node's spring is k = 1e4 N/mm. Then the
deck's two release steps as the reference runs them (src/fem/Forming.cpp:
801-900: a step's start imbalance on its free DOFs ramped out over 10
increments): the unload (the tool removed, the clamp held), then the release
onto 3-2-1 supports (a bottom corner node fixed in x, y, z, a second in y, z,
a third in z; the clamp let go), at whose end the statically determinate
supports carry no more than round-off.

The Newton driver is test code (the reference's driver is src/fem/
Forming.cpp:891-1140 and stays out of the device path): residual = system
residual + contact residual, tangent = assembled K + the contact node blocks
(Forming.cpp:747, 773, 988), solved on the free DOFs with scipy, with a
backtracking step on the residual norm. The device
side is NonlinearSystem.evaluate (element kernels, the reference's
column-major CSC values) + contact + add_node_blocks; the oracle side is its
element loop, scatter, force gather, contact and add_node_block. Bars: the
residual norms of every iteration agree to 1e-6 relative while they are above
1e-7 of the increment's first residual (below it both sides are at round-off),
the iteration counts agree, and the top surface's heights after the release
agree within 1e-6 mm, and so do the released part's height maps on the
reference top-node grid (the device's ray-cast of the device side's surface
against the oracle's of the oracle side's, same hit nodes). (The reference
halves an increment whose evaluation fails; this driver's steps are small
enough that none does.)
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

R = 0.005
STEP = 1.0e-4          # 1 mm in 10 increments
PENALTY = 1.0 / 7.0    # s: kappa A = s E sqrt(A) = 1e7 N/m for a 1 mm^2 node (k = 1e4 N/mm)
FRICTION = 0.1


def c1_model():
    coords, conn = W.structured_hex_mesh(20, 20, 2, 0.02, 0.02, 0.001, -0.01, -0.01, -0.001)
    m = G.Material(70.0e9, 0.33, yield_stress=95.0e6, saturation_stress=150.0e6, saturation_rate=10.0)
    top = np.nonzero(np.isclose(coords[:, 2], 0.0))[0].astype(np.int32)
    edge = np.isclose(np.abs(coords[:, 0]), 0.01) | np.isclose(np.abs(coords[:, 1]), 0.01)
    free = np.ones(3 * coords.shape[0], dtype=bool)
    for k in range(3):
        free[3 * np.nonzero(edge)[0] + k] = False
    return coords, conn, m, top, free


def slave_springs(coords, conn, top, youngs):
    """kappa_j A_j = s E_j sqrt(A_j) per top node (RigidTool.cpp:298-321):
    A_j the sum over the top faces at the node of its shape function's
    integral, a quarter of the face's area for a flat bilinear face."""
    on_top = np.isclose(coords[:, 2], 0.0)
    area = np.zeros(coords.shape[0])
    for e in conn:
        face = [a for a in e if on_top[a]]
        if len(face) != 4:
            continue
        p = coords[face]
        # the face's corners in order round it: a rectangle in the plane z = 0
        w = np.ptp(p[:, 0]) * np.ptp(p[:, 1])
        area[face] += 0.25 * w
    a = area[top]
    assert (a > 0).all()
    return PENALTY * youngs * np.sqrt(a)


def supports_321(coords):
    """The release's free DOFs: all but those of 3-2-1 supports at bottom
    corners: (-x, -y) fixed in x, y, z; (+x, -y) in y, z; (-x, +y) in z."""
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    bottom = np.isclose(z, z.min())
    def node(sx, sy):
        return int(np.nonzero(bottom & np.isclose(x, sx * np.abs(x).max()) & np.isclose(y, sy * np.abs(y).max()))[0][0])
    free = np.ones(3 * coords.shape[0], dtype=bool)
    fixed = [3 * node(-1, -1) + k for k in range(3)] + [3 * node(1, -1) + 1, 3 * node(1, -1) + 2,
                                                         3 * node(-1, 1) + 2]
    free[fixed] = False
    return free, np.array(fixed)


class ElementFailure(RuntimeError):
    """An element evaluation failed (the reference's SolverError)."""


class DeviceSide:
    def __init__(self, coords, conn, m, top, kappa_a):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.coords, self.conn, self.top = coords, conn, top
        self.sys = G.NonlinearSystem(coords, conn, m)
        t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.dev)
        self.t = t
        self.d_coords = t(coords.ravel())
        self.d_nodes = t(top, torch.int32)
        self.d_kappa = t(kappa_a)
        self.hist = (t(np.zeros(top.size), torch.int32), t(np.zeros(3 * top.size)), t(np.zeros(3 * top.size)))
        row_ptr, _ = self.sys.pattern()
        self.first = t(self.sys.diagonal_blocks()[top], torch.int32)
        self.row_len = t((row_ptr[3 * top + 1] - row_ptr[3 * top]).astype(np.int32), torch.int32)
        self.trial_hist = None

    def evaluate(self, u, centre, centre0, start, tool_on=True):
        torch = self.torch
        try:
            ev = self.sys.evaluate(u.reshape(-1, 3), want_tangent=True, column_major=True)
        except G.SolverError as err:
            raise ElementFailure(err.element) from err
        r = ev.internal.copy()
        values = ev.tangent_values.copy()
        if tool_on:
            tool = G.make_tool("sphere", centre, centre0, radius=R, friction=FRICTION)
            out = G.contact_evaluate_device(tool, self.d_coords, self.t(u.ravel()), self.t(start.ravel()), self.d_nodes,
                                            self.d_kappa, history=self.hist)
            d_values = self.t(values)
            G.add_node_blocks_device(out["blocks"], out["active"], self.first, self.row_len, d_values, column_major=True)
            torch.cuda.synchronize()
            values = d_values.cpu().numpy()
            self.contact_residual = out["residual"].cpu().numpy()
            r = r + self.contact_residual
            self.trial_hist = (out["trial_present"].clone(), out["trial_force"].reshape(-1).clone(),
                               out["trial_relative"].reshape(-1).clone())
        return r, values

    def commit(self, tool_on=True):
        self.sys.commit()
        if tool_on and self.trial_hist is not None:
            self.hist = self.trial_hist

    def device_solver(self, free, tolerance, outer, inner, symmetric, precond="diagonal"):
        """K_ff du = -r_f on the device: the partition's free-free block of
        the tangent's values and the block solve (conjugate gradients for a
        symmetric tangent, checked; BiCGStab on the row-major values, K_ff's
        rows, otherwise)."""
        torch = self.torch
        free_idx = np.nonzero(free)[0].astype(np.int32)
        self.sys.set_block_preconditioner(precond)
        self.sys.partition_device(self.t(free_idx, torch.int32))
        n = outer.size - 1
        if not hasattr(self, "solves"):
            self.solves = []
        checked = []  # the first symmetric solve of this solver checks the block's symmetry

        def solve(values, r):
            if not symmetric:  # the same entries in row order (the pattern is symmetric)
                rows = sp.csc_matrix((values, inner, outer), shape=(n, n)).tocsr()
                rows.sort_indices()
                values = rows.data
            block = self.sys.free_block_device(self.t(values))
            if symmetric and not checked:
                checked.append(True)
                ptr, index = self.sys.partition_pattern_device(self.dev)
                k = sp.csr_matrix((block.cpu().numpy(), index.cpu().numpy(), ptr.cpu().numpy()),
                                  shape=(free_idx.size, free_idx.size))
                assert abs(k - k.T).max() <= 1e-12 * abs(k).max()
            x, its, rel = self.sys.solve_free_block_device(block, self.t(-r[free_idx]), tolerance=tolerance,
                                                           max_iterations=200000, symmetric=symmetric)
            self.solves.append((symmetric, its, rel))
            return x.cpu().numpy()
        return solve


class OracleSide:
    def __init__(self, coords, conn, m, top, kappa_a):
        self.coords, self.conn, self.top = coords, conn, top
        self.kappa_a = kappa_a
        self.om = oracle_material(m)
        self.nn, self.ne = coords.shape[0], conn.shape[0]
        self.states = np.zeros((self.ne, 8, 39))
        self.alpha = np.zeros((self.ne, 9))
        self.hist = (np.zeros(top.size, np.int32), np.zeros((top.size, 3)), np.zeros((top.size, 3)))
        self.tool = None
        self.trial = None

    def evaluate(self, u, centre, centre0, start, tool_on=True):
        x0 = self.coords[self.conn]
        ue = u.reshape(-1, 3)[self.conn].reshape(self.ne, 24)
        ref = O.element_batch(x0, ue, self.om, self.states, alpha=self.alpha, kinematics=2, threads=8)
        if (ref["status"] != 0).any():
            raise ElementFailure(int(np.nonzero(ref["status"])[0][0]))
        r, _ = O.gather_force(self.nn, self.conn, ref["force"])
        values = O.scatter(self.nn, self.conn, ref["tangent"].reshape(self.ne, 576))
        hist = self.hist
        if tool_on:
            tool = O.tool_record(1, centre, centre0, radius=R, friction=FRICTION)
            c = O.contact(tool, self.coords, u.ravel(), start.ravel(), self.top, self.kappa_a, history=hist)
            act = c["active"] == 1
            values = O.add_node_blocks(self.nn, self.conn, values, self.top[act], c["blocks"][act].reshape(-1, 9))
            self.contact_residual = c["residual"]
            r = r + c["residual"]
            hist = (c["t_present"], c["t_force"], c["t_relative"])
        self.trial = (ref["states"], ref["alpha"], hist)
        return r, values

    def commit(self, tool_on=True):
        self.states, self.alpha, hist = self.trial
        if tool_on:
            self.hist = hist


def newton(side, u0, free, outer, inner, centre, centre0, tool_on=True, max_it=25, load=None, solve=None):
    """One increment: Newton on the free DOFs until the residual falls below
    1e-9 of the increment's first residual (or to its round-off floor); a
    step whose evaluation fails or whose residual is not smaller is halved
    (up to 16 times). The correction is scipy's direct solve, or solve(values,
    r) when given. Returns the converged u, the residual norms and the step
    lengths taken."""
    n = u0.size
    u = u0.copy()
    def evaluate(x):
        r, values = side.evaluate(x, centre, centre0, u0, tool_on)
        return (r if load is None else r + load), values

    r, values = evaluate(u)
    norms = [float(np.linalg.norm(r[free]))]
    steps = []
    for it in range(max_it):
        rn = norms[-1]
        if it > 0 and (rn <= 1e-9 * norms[0] or (it > 2 and rn >= 0.5 * norms[-2] and rn <= 1e-6 * norms[0])):
            break
        if solve is None:
            K = sp.csc_matrix((values, inner, outer), shape=(n, n))
            du = spla.spsolve(K[free][:, free].tocsc(), -r[free])
        else:
            du = solve(values, r)
        step = 1.0
        for _ in range(16):
            trial = u.copy()
            trial[free] += step * du
            try:
                rt, vt = evaluate(trial)
            except ElementFailure:
                step *= 0.5
                continue
            if np.linalg.norm(rt[free]) < rn:
                u, r, values = trial, rt, vt
                break
            step *= 0.5
        else:
            raise ElementFailure(-1)
        steps.append(step)
        norms.append(float(np.linalg.norm(r[free])))
    side.commit(tool_on)
    return u, norms, steps


def height_maps(coords, top, u_d, u_o):
    """(z, hit) of the deformed top surface on the reference top-node grid,
    from the device (u_d) and from the oracle (u_o)."""
    import torch
    xs, ys = np.unique(coords[top, 0]), np.unique(coords[top, 1])
    order = np.lexsort((coords[top, 0], coords[top, 1]))  # y-major, then x
    grid = (xs[0], ys[0], xs.size, ys.size, xs[1] - xs[0])

    def surface(u):
        cur = coords[top] + u.reshape(-1, 3)[top]
        return np.ascontiguousarray(cur[order].reshape(ys.size, xs.size, 3))

    z_d, hit_d = G.surface_raycast_device(torch.as_tensor(surface(u_d), device="cuda"), grid)
    torch.cuda.synchronize()
    z_o, hit_o, folded = O.surface_raycast(surface(u_o), grid)
    assert folded == 0
    return z_d.cpu().numpy(), hit_d.cpu().numpy(), z_o, hit_o


def release_step(dev, ora, u_d, u_o, free, outer, inner, centre0, bar, solve=None):
    """A release step (Forming.cpp:801-900, 10 increments, no tool): the
    imbalance of the start state on the step's free DOFs, R_0f, ramped out:
    increment i solves R_f(u) = (1 - i / 10) R_0f. Both sides through the
    same loop (the device side by `solve` when given); residual histories
    compared while above `bar` of an increment's first residual. Returns the
    released u of both sides and the device side's start imbalance."""
    r0_d, _ = dev.evaluate(u_d, centre0, centre0, u_d, tool_on=False)
    r0_o, _ = ora.evaluate(u_o, centre0, centre0, u_o, tool_on=False)
    assert np.abs(r0_d[free] - r0_o[free]).max() <= 1e-9 * np.abs(r0_o[free]).max()
    r0_d[~free] = 0.0
    r0_o[~free] = 0.0
    for i in range(1, 11):
        lam = 1.0 - i / 10.0
        u_d, h_d, s_d = newton(dev, u_d, free, outer, inner, centre0, centre0, tool_on=False, load=-lam * r0_d,
                               solve=solve)
        u_o, h_o, s_o = newton(ora, u_o, free, outer, inner, centre0, centre0, tool_on=False, load=-lam * r0_o)
        assert len(h_d) == len(h_o) and s_d == s_o, ("release", i, h_d, h_o, s_d, s_o)
        for a, b in zip(h_d, h_o):
            if b > bar * h_o[0]:
                assert abs(a - b) <= 1e-6 * b, ("release", i, h_d, h_o)
    r_d, _ = dev.evaluate(u_d, centre0, centre0, u_d, tool_on=False)
    assert np.abs(r_d[free]).max() <= 1e-6 * np.abs(r0_d).max()
    return u_d, u_o, r0_d


def unload_and_release(dev, ora, u_d, u_o, coords, clamp_free, outer, inner, centre0, bar, unload_solve=None,
                       release_solve=None):
    """The deck's unload (the tool removed, the clamp held) and its release
    onto 3-2-1 supports; at the end the statically determinate supports
    carry round-off only, and the part has sprung back free of them."""
    u_d, u_o, _ = release_step(dev, ora, u_d, u_o, clamp_free, outer, inner, centre0, bar, unload_solve)
    free, fixed = supports_321(coords)
    u_d, u_o, r0_d = release_step(dev, ora, u_d, u_o, free, outer, inner, centre0, bar, release_solve)
    r_d, _ = dev.evaluate(u_d, centre0, centre0, u_d, tool_on=False)
    reactions = r_d[fixed]
    assert np.abs(reactions).max() <= 1e-6 * np.abs(r0_d).max(), reactions
    return u_d, u_o


def test_c1_newton_histories_and_release_heights():
    coords, conn, m, top, free = c1_model()
    nn = coords.shape[0]
    outer, inner, _ = O.pattern(nn, conn)
    kappa_a = slave_springs(coords, conn, top, 70.0e9)
    dev = DeviceSide(coords, conn, m, top, kappa_a)
    ora = OracleSide(coords, conn, m, top, kappa_a)
    row_ptr, col_idx = dev.sys.pattern()
    assert np.array_equal(row_ptr, outer) and np.array_equal(col_idx, inner)
    u_d = np.zeros(3 * nn)
    u_o = np.zeros(3 * nn)
    centre0 = (0.0, 0.0, R)
    for inc in range(1, 11):
        centre = (0.0, 0.0, R - STEP * inc)
        u_d, h_d, s_d = newton(dev, u_d, free, outer, inner, centre, centre0)
        u_o, h_o, s_o = newton(ora, u_o, free, outer, inner, centre, centre0)
        assert len(h_d) == len(h_o) and s_d == s_o, (inc, h_d, h_o, s_d, s_o)
        for a, b in zip(h_d, h_o):
            if b > 1e-7 * h_o[0]:
                assert abs(a - b) <= 1e-6 * b, (inc, h_d, h_o)
        centre0 = centre
    # yielded, and the tool pressed the blank
    assert ora.states[..., 36].max() > 1e-3  # equivalent plastic strain
    assert u_o[3 * top + 2].min() < -5e-4
    # release onto 3-2-1 supports; the formerly clamped rim springs back
    rim = np.isclose(np.abs(coords[top, 0]), 0.01) | np.isclose(np.abs(coords[top, 1]), 0.01)
    u_d, u_o = unload_and_release(dev, ora, u_d, u_o, coords, free, outer, inner, centre0, 1e-7)
    assert np.abs(u_o[3 * top[rim] + 2]).max() > 1e-6  # the rim moved
    heights_d = coords[top, 2] + u_d[3 * top + 2]
    heights_o = coords[top, 2] + u_o[3 * top + 2]
    assert np.abs(heights_d - heights_o).max() <= 1e-9  # 1e-6 mm
    # the released part's height maps (formed_surface before its nearest
    # fill) on the reference top-node grid: the device side's by the device
    # ray-cast, the oracle side's by the oracle's
    z_d, hit_d, z_o, hit_o = height_maps(coords, top, u_d, u_o)
    assert np.array_equal(hit_d, hit_o) and hit_o.mean() > 0.75
    assert np.abs(z_d[hit_d] - z_o[hit_o]).max() <= 1e-9  # 1e-6 mm


@pytest.mark.parametrize("precond", ["diagonal", "multigrid"])
def test_c1_with_the_device_block_solve(precond):
    """The C1 case again with every Newton correction of the device side
    solved on the device: the partition's free-free block of the tangent and
    the block solve to 1e-10, BiCGStab for the loading (frictional contact
    with its exact tangent makes the tangent non-symmetric; the reference then
    factorises with LU) and conjugate gradients for the release (the released
    blank's tangent is symmetric; the reference: LLT) of the unload (clamp held)
    and of the release on the 3-2-1 supports' partition. The oracle side keeps
    scipy's direct solve. Bars: the same iteration counts and step lengths,
    residual norms within 1e-6 relative while above 1e-6 of the step's first
    residual (each iterative correction leaves a residual of 1e-10 of its
    right-hand side; late in an increment the right-hand side is so small that
    the round-off of b - K x itself is about 1e-11 of it, so a tighter
    tolerance is not always attainable), and the released heights within
    1e-6 mm. Both preconditioners: the block's diagonal and the multigrid
    V-cycle (built on the block's symmetric part for BiCGStab). The release
    onto 3-2-1 supports solves the free-floating blank's block (positive
    definite only through the six supports)."""
    coords, conn, m, top, free = c1_model()
    nn = coords.shape[0]
    outer, inner, _ = O.pattern(nn, conn)
    kappa_a = slave_springs(coords, conn, top, 70.0e9)
    dev = DeviceSide(coords, conn, m, top, kappa_a)
    ora = OracleSide(coords, conn, m, top, kappa_a)
    u_d = np.zeros(3 * nn)
    u_o = np.zeros(3 * nn)
    centre0 = (0.0, 0.0, R)
    loading_solve = dev.device_solver(free, 1e-10, outer, inner, symmetric=False, precond=precond)
    for inc in range(1, 11):
        centre = (0.0, 0.0, R - STEP * inc)
        u_d, h_d, s_d = newton(dev, u_d, free, outer, inner, centre, centre0, solve=loading_solve)
        u_o, h_o, s_o = newton(ora, u_o, free, outer, inner, centre, centre0)
        assert len(h_d) == len(h_o) and s_d == s_o, (inc, h_d, h_o, s_d, s_o)
        for a, b in zip(h_d, h_o):
            if b > 1e-6 * h_o[0]:
                assert abs(a - b) <= 1e-6 * b, (inc, h_d, h_o)
        centre0 = centre
    unload_solve = dev.device_solver(free, 1e-10, outer, inner, symmetric=True, precond=precond)
    u_d, u_o, _ = release_step(dev, ora, u_d, u_o, free, outer, inner, centre0, 1e-6, unload_solve)
    release_free, fixed = supports_321(coords)
    release_solve = dev.device_solver(release_free, 1e-10, outer, inner, symmetric=True, precond=precond)
    u_d, u_o, r0_d = release_step(dev, ora, u_d, u_o, release_free, outer, inner, centre0, 1e-6, release_solve)
    r_d, _ = dev.evaluate(u_d, centre0, centre0, u_d, tool_on=False)
    assert np.abs(r_d[fixed]).max() <= 1e-6 * np.abs(r0_d).max()
    assert any(not sym for sym, _, _ in dev.solves) and any(sym for sym, _, _ in dev.solves)
    assert all(rel <= 1e-10 for _, _, rel in dev.solves)
    heights_d = coords[top, 2] + u_d[3 * top + 2]
    heights_o = coords[top, 2] + u_o[3 * top + 2]
    assert np.abs(heights_d - heights_o).max() <= 1e-9  # 1e-6 mm
    print(f"device block solves with the {precond} preconditioner (symmetric, iterations, relative residual):",
          dev.solves)
