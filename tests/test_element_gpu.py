"""Parity of the HIP element kernels with the oracle (the CPU restatement of
src/fem/IncompatibleModes.cpp and src/material/*), on the same seeded inputs.

Tolerance (the reference tests' own norm, max |delta| / max |ref| per element):
1e-12 for f* and K*, 1e-10 for the mode parameters and the trial states.
Elements whose local iteration takes a different number of iterations on the
two sides are counted separately and must be rare (a convergence test decided
by round-off).
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material, rel, run_gpu
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

TOL_F = 1.0e-12
TOL_K = 1.0e-12
TOL_STATE = 1.0e-10

DISTORTED = np.array([[0.05, 1.10, 1.20, -0.05, 0.10, 0.95, 1.05, 0.00],
                      [0.00, 0.10, 1.15, 0.90, 0.05, -0.05, 1.05, 1.10],
                      [0.00, -0.05, 0.10, 0.05, 0.90, 1.05, 1.20, 1.00]]).T.copy()


def deformation(x, s):
    """test_incompatible_modes.cpp:382-392: stretch, shear, bending, twist."""
    ue = np.zeros(24)
    for a in range(8):
        p = x[a]
        ue[3 * a:3 * a + 3] = s * np.array([0.8 * p[0] + 0.6 * p[0] * p[2] + 0.2 * p[1],
                                            -0.3 * p[1] + 0.5 * p[1] * p[2] - 0.4 * p[0] * p[1],
                                            -0.4 * p[2] - 0.5 * p[0] * p[0] + 0.3 * p[0] * p[1] * p[2]])
    return ue


def materials():
    return {
        "elastic": G.Material(200.0e9, 0.3),
        "J2": G.Material(200.0e9, 0.3, yield_stress=250.0e6, hardening_modulus=2.0e9,
                         kinematic_hardening_modulus=1.0e9),
        "Hill48": G.Material(200.0e9, 0.3, yield_stress=250.0e6, hardening_modulus=1.5e9,
                             hill=dict(r0=1.9, r45=1.5, r90=2.3, rolling_direction=(1.0, 0.5, 0.0))),
        "Chaboche": G.Material(200.0e9, 0.3, yield_stress=250.0e6, saturation_stress=100.0e6,
                               saturation_rate=20.0, backstresses=((40.0e9, 400.0), (5.0e9, 20.0))),
        # backward-Euler integration of the backstresses (the other value of the flag)
        "ChabocheBE": G.Material(200.0e9, 0.3, yield_stress=250.0e6, saturation_stress=100.0e6,
                                 saturation_rate=20.0, backstresses=((40.0e9, 400.0), (5.0e9, 20.0)),
                                 kinematic_integration="backward_euler"),
        # four kinematic terms (three backstresses and Prager's modulus) on Hill48
        "Chaboche4": G.Material(200.0e9, 0.3, yield_stress=250.0e6, saturation_stress=100.0e6,
                                saturation_rate=20.0, kinematic_hardening_modulus=1.0e9,
                                hill=dict(r0=1.9, r45=1.5, r90=2.3),
                                backstresses=((40.0e9, 400.0), (5.0e9, 20.0), (2.0e9, 5.0))),
    }


def compare(gpu, ref, label):
    ok = (gpu["status"] == 0) & (ref["status"] == 0)
    assert np.array_equal(gpu["status"] != 0, ref["status"] != 0), f"{label}: failure sets differ"
    same_path = ok & (gpu["iterations"] == ref["scalars"][:, 2].astype(int))
    diverged = int(ok.sum() - same_path.sum())
    assert diverged <= max(1, int(0.001 * ok.sum())), f"{label}: {diverged} elements took other local paths"
    idx = np.nonzero(same_path)[0]
    ef = rel(gpu["force"][idx], ref["force"][idx])
    assert ef.max() <= TOL_F, f"{label}: force error {ef.max():.3e} at element {idx[ef.argmax()]}"
    if ref["tangent"] is not None:
        ek = rel(gpu["tangent"][idx], ref["tangent"][idx])
        assert ek.max() <= TOL_K, f"{label}: tangent error {ek.max():.3e} at element {idx[ek.argmax()]}"
    ea = np.abs(gpu["alpha"][idx] - ref["alpha"][idx]).max(axis=1) / np.maximum(
        np.abs(ref["alpha"][idx]).max(axis=1), 1e-300)
    assert ea.max() <= TOL_STATE or np.abs(gpu["alpha"][idx] - ref["alpha"][idx]).max() < 1e-18, \
        f"{label}: alpha error {ea.max():.3e}"
    es = rel(gpu["states"][idx], ref["states"][idx])
    assert es.max() <= TOL_STATE, f"{label}: state error {es.max():.3e}"
    ee = np.abs(gpu["energy"][idx] - ref["scalars"][idx, 0]) / np.maximum(np.abs(ref["scalars"][idx, 0]), 1e-300)
    assert ee.max() <= 1e-12, f"{label}: energy error {ee.max():.3e}"
    assert np.array_equal(gpu["yielding"][idx], ref["scalars"][idx, 1].astype(int)), f"{label}: yielding differs"
    return dict(force=ef.max(), tangent=0 if ref["tangent"] is None else ek.max(), diverged=diverged)


def single_hex_batch(scales, extra_z=0.2):
    """One distorted hex per scale (its own 8 nodes): coords, conn, u."""
    ne = len(scales)
    coords = np.concatenate([DISTORTED] * ne)
    conn = np.arange(8 * ne, dtype=np.int32).reshape(ne, 8)
    u = np.zeros((ne, 24))
    for i, s in enumerate(scales):
        ue = deformation(DISTORTED, 1.3 * s)
        for a in range(8):
            ue[3 * a + 2] += extra_z * s * DISTORTED[a, 0] * DISTORTED[a, 1]
        u[i] = ue
    return coords, conn, u.reshape(8 * ne, 3)


@pytest.mark.parametrize("law", ["elastic", "J2", "Hill48", "Chaboche", "ChabocheBE", "Chaboche4"])
@pytest.mark.parametrize("kinematics", [0, 1, 2])
@pytest.mark.parametrize("box", [(2, 0, 2), (2, 5, 2)], ids=["2x2x2", "2x2x5"])
def test_single_hex_laws(law, kinematics, box):
    """The laws x kinematics of test_incompatible_modes.cpp:433-508, from a
    committed first step as there, at several load levels; the 2 x 2 x 2 rule
    and the 2 x 2 x 5 box rule of the reference's fine-mesh decks."""
    m = materials()[law]
    rule = G.Rule(*box)
    base = {0: 4.0e-3, 1: 0.02, 2: 0.1}[kinematics] if law != "elastic" else {0: 1e-3, 1: 0.05, 2: 0.05}[kinematics]
    scales = base * np.array([0.25, 0.5, 1.0, 1.5])
    om = oracle_material(m)
    ne = len(scales)
    # committed states: the first step at deformation(x, s), as the reference test
    states = np.zeros((ne, rule.points, 39))
    alpha = np.zeros((ne, 9))
    for i, s in enumerate(scales):
        first = O.element(DISTORTED, deformation(DISTORTED, s), om, rule=box, kinematics=kinematics)
        if m.c.yield_stress > 0:
            states[i] = first["states"]
        alpha[i] = first["alpha"]
    coords, conn, u = single_hex_batch(scales)
    x0 = coords[conn]
    ue = u[conn].reshape(ne, 24)
    ref = O.element_batch(x0, ue, om, states, alpha=alpha, rule=box, kinematics=kinematics, threads=4)
    gpu = run_gpu(m, rule, coords, conn, u, states, alpha=alpha, kinematics=kinematics)
    assert (ref["status"] == 0).all()
    compare(gpu, ref, f"{law}/{kinematics}/{box}")


@pytest.mark.parametrize("box", [(2, 0, 2), (2, 5, 2), (2, 7, 2)], ids=["2x2x2", "2x2x5", "2x2x7"])
def test_microbench_sample_parity(box):
    """A 4096-element sample of the element microbenchmark (Hill48 + Chaboche,
    logarithmic strain), tangent and residual-only; with the 2 x 2 x 2 rule
    and the 2 x 2 x 5 and 2 x 2 x 7 box rules (their 20 and 28 points, each
    with a committed state drawn as the microbenchmark draws them)."""
    m = W.microbench_material()
    om = oracle_material(m)
    rule = G.Rule(*box)
    w = W.microbench(4096, seed=7, points=rule.points)
    x0, ue = W.element_arrays(w)
    ref = O.element_batch(x0, ue, om, w["states"], rule=box, kinematics=2, threads=8)
    gpu = run_gpu(m, rule, w["coords"], w["conn"], w["u"], w["states"], kinematics=2)
    res = compare(gpu, ref, f"microbench {box}")
    print("microbench parity", box, res)
    gpu_r = run_gpu(m, rule, w["coords"], w["conn"], w["u"], w["states"], kinematics=2, want_tangent=False)
    assert np.array_equal(gpu_r["force"], gpu["force"]), "residual-only force differs from the tangent pass"


def test_cost_order_changes_no_result():
    """The local solve's schedule (elements by expected work, or in index
    order) changes no bit of any output, with a mix of yielding, elastic and
    refused elements."""
    m = W.microbench_material()
    w = W.microbench(3000, seed=11)
    coords, u = w["coords"].copy(), w["u"].copy()
    # element 5: a displacement that turns it inside out (refused in its first pass)
    u[w["conn"][5, 4:]] += np.array([0.0, 0.0, -3.0])
    by_cost = run_gpu(m, G.Rule(), coords, w["conn"], u, w["states"], kinematics=2)
    by_index = run_gpu(m, G.Rule(), coords, w["conn"], u, w["states"], kinematics=2, index_order=True)
    assert by_cost["status"][5] != 0 and (by_cost["yielding"] > 0).any() and (by_cost["yielding"] == 0).any()
    for k in by_cost:
        assert np.array_equal(by_cost[k], by_index[k]), k


def test_cost_order_of_a_large_launch_changes_no_result():
    """A launch of many wavefronts per SIMD takes the elements in one
    descending run of expected work (a small one deals them round-robin): the
    outputs again equal the index-order ones bit for bit (compared on the
    device)."""
    import torch
    m = W.microbench_material()
    ne = 262144
    w = W.microbench(ne, seed=13)
    st = G.states_from_reference(w["states"], m)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")
    out = []
    for stages in (G.LOCAL_SOLVE | G.CONDENSATION, G.LOCAL_SOLVE | G.CONDENSATION | G.INDEX_ORDER):
        b = G.ElementBatch(t(w["coords"]), t(w["conn"], torch.int32), t(w["u"]), t(st), None, 8, True)
        G.evaluate_elements(m, G.Rule(), b, G.FINITE_LOGARITHMIC, True, stages=stages)
        torch.cuda.synchronize()
        out.append(b)
    a, b = out
    assert (a.yielding > 0).any()
    for name in ("force", "tangent", "state_out", "alpha_out", "energy", "yielding", "iterations", "status"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name


def test_refused_elements_report_status():
    """An inverted element is refused at set-up (MeshError), a step that
    inverts a point fails its first pass (SolverError); the others evaluate."""
    m = materials()["J2"]
    coords, conn, u = single_hex_batch([0.01, 0.01, 0.01])
    coords = coords.copy()
    # element 1: swap top and bottom faces -> det J < 0 at the centre
    top = coords[8 + 4:8 + 8].copy()
    coords[8 + 4:8 + 8] = coords[8 + 0:8 + 4]
    coords[8 + 0:8 + 4] = top
    # element 2: a displacement that turns the element inside out
    u = u.copy()
    u[16:24, 2] = -3.0 * DISTORTED[:, 2]
    states = np.zeros((3, 8, 39))
    gpu = run_gpu(m, G.Rule(), coords, conn, u, states, kinematics=2)
    assert gpu["status"][0] == 0
    assert gpu["status"][1] in (1, 2)
    assert gpu["status"][2] == 3


@pytest.mark.parametrize("material", ["J2", "Chaboche"])
def test_refused_elements_output_the_committed_state(material):
    """A refused element (at set-up, or in a pass) outputs zero force,
    tangent, energy and yielding, its committed modes and its committed
    history, whatever its output buffers held before (filled with NaN here)."""
    import torch
    m = materials()[material]
    coords, conn, u = single_hex_batch([0.01, 0.01, 0.01])
    coords = coords.copy()
    top = coords[8 + 4:8 + 8].copy()
    coords[8 + 4:8 + 8] = coords[8 + 0:8 + 4]
    coords[8 + 0:8 + 4] = top
    u = u.copy()
    u[16:24, 2] = -3.0 * DISTORTED[:, 2]
    rng = np.random.default_rng(5)
    st = np.zeros((m.state_fields, 3 * 8))
    st[:, 8:] = rng.uniform(0.0, 1e-3, (m.state_fields, 16))
    alpha = np.zeros((3, 9))
    alpha[1:] = rng.uniform(-1e-4, 1e-4, (2, 9))
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda:0")
    b = G.ElementBatch(t(coords), t(conn, torch.int32), t(u), t(st), t(alpha), 8, True)
    for name in ("state_out", "alpha_out", "force", "tangent", "energy"):
        getattr(b, name).fill_(float("nan"))
    b.yielding.fill_(-1)
    G.evaluate_elements(m, G.Rule(), b, G.FINITE_LOGARITHMIC, True)
    torch.cuda.synchronize()
    status = b.status.cpu().numpy()
    assert status[0] == 0 and status[1] in (1, 2) and status[2] == 3
    state_out = b.state_out.cpu().numpy()
    for e in (1, 2):
        assert (b.force[e] == 0).all() and (b.tangent[e] == 0).all()
        assert b.energy[e].item() == 0.0 and b.yielding[e].item() == 0
        assert np.array_equal(b.alpha_out[e].cpu().numpy(), alpha[e])
        assert np.array_equal(state_out[:, 8 * e:8 * e + 8], st[:, 8 * e:8 * e + 8])
    assert np.isfinite(b.force[0].cpu().numpy()).all() and np.isfinite(state_out[:, :8]).all()
