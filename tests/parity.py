"""Shared helpers of the GPU-vs-oracle parity tests."""
import numpy as np

import oracle_lib as O
import sparlab_gpu as G


def oracle_material(m):
    """The oracle's 54-double record of a sparlab_gpu.Material."""
    c = m.c
    rec = np.zeros(54)
    rec[0:7] = [c.youngs_modulus, c.poisson_ratio, c.yield_stress, c.hardening_modulus,
                c.kinematic_hardening_modulus, c.saturation_stress, c.saturation_rate]
    rec[7] = c.criterion
    rec[8] = c.num_backstresses
    rec[9:13] = c.backstress_modulus[:]
    rec[13:17] = c.backstress_recovery[:]
    rec[17] = c.kinematic_integration
    rec[18:] = c.yield_matrix[:]
    return rec


def run_gpu(material, rule, coords, conn, u, states_ref, alpha=None, kinematics=2, want_tangent=True,
            device="cuda:0", index_order=False):
    """Evaluates a batch on the GPU; returns numpy outputs (states as 39-records).
    index_order: the local solve takes the elements in index order instead of
    by expected work."""
    import torch
    ne = conn.shape[0]
    npts = rule.points
    st = G.states_from_reference(states_ref, material)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    b = G.ElementBatch(t(coords), t(conn, torch.int32), t(u), t(st),
                       None if alpha is None else t(alpha), npts, want_tangent)
    stages = G.LOCAL_SOLVE | G.CONDENSATION | (G.INDEX_ORDER if index_order else 0)
    G.evaluate_elements(material, rule, b, kinematics, want_tangent, stages=stages)
    torch.cuda.synchronize()
    return dict(force=b.force.cpu().numpy(), tangent=None if b.tangent is None else b.tangent.cpu().numpy(),
                alpha=b.alpha_out.cpu().numpy(),
                states=G.states_to_reference(b.state_out.cpu().numpy(), material, ne, npts),
                energy=b.energy.cpu().numpy(), yielding=b.yielding.cpu().numpy(),
                iterations=b.iterations.cpu().numpy(), status=b.status.cpu().numpy(),
                local_path=b.local_path.cpu().numpy())


def rel(a, b):
    """max |a - b| / max |b| per leading index (the reference tests' norm)."""
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    den = np.abs(b).max(axis=1)
    den = np.where(den == 0.0, 1.0, den)
    return np.abs(a - b).max(axis=1) / den
