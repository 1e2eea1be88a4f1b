"""Parity of the device stress recovery (sparlab_system_stress_device) with
the oracle's incompatible_elastoplastic_stress
(src/fem/IncompatibleModes.cpp:602-664) on identical inputs: the same mesh,
displacement, committed history and modes (the oracle's own converged ones,
written into the system with set_history).

Bars: per element, the stresses (Cauchy, second Piola-Kirchhoff, Kirchhoff,
von Mises) within 1e-12 of the element's largest stress, the strains (log
strain, largest strain) within 1e-12 of their largest, det F within 1e-14,
the equivalent plastic strain within 1e-12 relative; a refused element gets
the oracle's outcome (MeshError) and zeros.
"""
import numpy as np
import pytest

import oracle_lib as O
import sparlab_gpu as G
from parity import oracle_material
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

LAWS = {
    "J2_Voce": lambda: G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0),
    "Hill48_Chaboche": W.microbench_material,
}


def jittered_mesh(seed):
    coords, conn = W.structured_hex_mesh(6, 5, 2, 0.006, 0.005, 0.002)
    rng = np.random.default_rng(seed)
    coords = coords + rng.uniform(-0.15, 0.15, coords.shape) * 0.001
    return coords, conn


@pytest.mark.parametrize("law", sorted(LAWS))
@pytest.mark.parametrize("kinematics", [0, 1, 2])
@pytest.mark.parametrize("box", [(2, 0, 2), (2, 7, 2)], ids=["2x2x2", "2x2x7"])
def test_stress_matches_oracle(device, law, kinematics, box):
    import torch
    m = LAWS[law]()
    npts = G.Rule(*box).points
    om = oracle_material(m)
    coords, conn = jittered_mesh(4)
    nn, ne = coords.shape[0], conn.shape[0]
    rng = np.random.default_rng(9)
    u = rng.normal(0, 3e-6, (nn, 3)) + 2e-2 * coords[:, [2, 0, 1]]
    x0 = coords[conn]
    ue = u[conn].reshape(ne, 24)
    ev = O.element_batch(x0, ue, om, np.zeros((ne, npts, O.STATE)), rule=box, kinematics=kinematics,
                         want_tangent=False, threads=8)
    assert (ev["status"] == 0).all() and (ev["scalars"][:, 1] > 0).any()
    states, alpha = ev["states"], ev["alpha"]
    ref = O.element_stress_batch(x0, ue, om, states, alpha, rule=box, kinematics=kinematics, threads=8)
    assert (ref["status"] == 0).all()
    sysm = G.NonlinearSystem(coords, conn, m, G.Rule(*box), kinematics)
    sysm.set_history(states, alpha)
    stress, status = sysm.stress_device(torch.as_tensor(u.reshape(-1), device=device))
    torch.cuda.synchronize()
    got, want = stress.cpu().numpy(), ref["stress"]
    assert (status.cpu().numpy() == 0).all()
    stress_cols = list(range(0, 18)) + [24]
    s_scale = np.abs(want[:, stress_cols]).max(axis=1)
    assert (np.abs(got[:, stress_cols] - want[:, stress_cols]).max(axis=1) <= 1e-12 * s_scale).all()
    strain_cols = list(range(18, 24)) + [26]
    e_scale = np.maximum(np.abs(want[:, strain_cols]).max(axis=1), 1e-300)
    assert (np.abs(got[:, strain_cols] - want[:, strain_cols]).max(axis=1) <= 1e-12 * e_scale).all()
    assert np.abs(got[:, 27] - want[:, 27]).max() <= 1e-14
    assert np.abs(got[:, 25] - want[:, 25]).max() <= 1e-12 * np.abs(want[:, 25]).max()
    assert (want[:, 25] > 0).any()
    # the host form returns the same, by name
    named = sysm.stress(u)
    assert np.array_equal(named["cauchy"], got[:, 0:6]) and np.array_equal(named["min_jacobian"], got[:, 27])


def test_stress_of_a_refused_element(device):
    import torch
    m = LAWS["J2_Voce"]()
    om = oracle_material(m)
    coords, conn = jittered_mesh(5)
    conn = conn.copy()
    conn[3] = conn[3][[4, 5, 6, 7, 0, 1, 2, 3]]  # top and bottom faces swapped: det J < 0
    nn, ne = coords.shape[0], conn.shape[0]
    u = 1e-3 * coords[:, [2, 0, 1]]
    ref = O.element_stress_batch(coords[conn], u[conn].reshape(ne, 24), om, np.zeros((ne, 8, O.STATE)),
                                 kinematics=2)
    assert ref["status"][3] == 2 and (np.delete(ref["status"], 3) == 0).all()  # MeshError
    sysm = G.NonlinearSystem(coords, conn, m)
    stress, status = sysm.stress_device(torch.as_tensor(u.reshape(-1), device=device))
    torch.cuda.synchronize()
    status = status.cpu().numpy()
    assert status[3] in (1, 2) and (np.delete(status, 3) == 0).all()
    assert (stress[3] == 0).all()
    with pytest.raises(G.SolverError):
        sysm.stress(u)
