"""The C++ adapter of include/sparlab_gpu_system.hpp driven through the
reference forming driver's calls (tests/cpp/adapter_sequence.cpp: restart,
plastic/averaged/elastoplastic, tangent and residual-only evaluations,
symmetric, commit, max_plastic_strain, committed/internal per element and
all, the stress recovery, a restart of a second system), against the same
increments run through the Python host side of the same library.

Bars: every value the C++ calls returned equals the Python path's bit for
bit (f_int, energy, gross, yielding points, the committed PlasticState
records and modes, the stress measures) except the sum of |K| values
(summed in another order: 1e-12); max_plastic_strain is the committed
states' largest; the symmetry flag follows the material (a recovering
backstress makes it false); the driver's own checks pass (exit status 0).
Two laws (J2 + Voce; Hill48 + two Chaboche backstresses) and two rules
(2 x 2 x 2; the 2 x 2 x 5 box rule).
"""
import os
import subprocess

import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "adapter_sequence")
STEPS = 3


def law(which):
    if which == 0:
        return G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0)
    return G.Material(70e9, 0.33, yield_stress=95e6, saturation_stress=150e6, saturation_rate=10.0,
                      hill=dict(r0=0.7, r45=0.6, r90=0.8), backstresses=((6000e6, 60.0), (1500e6, 10.0)))


@pytest.mark.parametrize("thickness", [0, 5], ids=["2x2x2", "2x2x5"])
@pytest.mark.parametrize("which", [0, 1], ids=["J2", "Hill48_Chaboche"])
def test_adapter_sequence_matches_the_python_path(device, tmp_path, which, thickness):
    assert os.path.exists(DRIVER), "build() compiles tests/cpp/adapter_sequence.cpp"
    coords, conn = W.structured_hex_mesh(5, 4, 2, 0.01, 0.008, 0.002)
    rng = np.random.default_rng(3)
    coords = coords + rng.uniform(-0.15, 0.15, coords.shape) * 0.001
    nn, ne = coords.shape[0], conn.shape[0]
    shape = rng.normal(0, 2e-5, (nn, 3)) + 1e-4 * coords[:, [2, 0, 1]] / 0.002
    us = [(k + 1) * shape.reshape(-1) for k in range(STEPS)]
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([nn, ne, STEPS, which, thickness], dtype=np.int64).tofile(f)
        coords.astype(np.float64).tofile(f)
        conn.astype(np.int32).tofile(f)
        for u in us:
            u.astype(np.float64).tofile(f)
    print(f"adapter sequence: law {which}, thickness points {thickness}", flush=True)
    run = subprocess.run([DRIVER, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr, flush=True)
    assert run.returncode == 0, run.stderr
    got = np.fromfile(outp, dtype=np.float64)

    m = law(which)
    rule = G.Rule(2, thickness, 2)
    points = rule.points
    s = G.NonlinearSystem(coords, conn, m, rule)
    pos = 0

    def take(n):
        nonlocal pos
        out = got[pos:pos + n]
        pos += n
        return out

    plastic, symmetric, averaged, elastoplastic = take(4)
    assert plastic == 1.0 and averaged == 0.0 and elastoplastic == 1.0
    assert symmetric == (1.0 if which == 0 else 0.0)
    for k, u in enumerate(us):
        ev = s.evaluate(u, want_tangent=True, column_major=True)
        s.commit()
        energy, gross, yielding, top, sym_factor, abs_sum = take(6)
        internal = take(3 * nn)
        st, al = s.history()
        assert energy == ev.energy and gross == ev.gross and yielding == ev.yielding_points, k
        assert top == st[:, :, 36].max(), k
        assert sym_factor == symmetric
        assert abs(abs_sum - np.abs(ev.tangent_values).sum()) <= 1e-12 * abs_sum
        assert np.array_equal(internal, ev.internal), k
    assert yielding > 0
    records = take(ne * points * (39 + 0) + 9 * ne)
    per = points * 39 + 9
    records = records.reshape(ne, per)
    st, al = s.history()
    assert np.array_equal(records[:, :points * 39].reshape(ne, points, 39), st)
    assert np.array_equal(records[:, points * 39:], al)
    stress = take(ne * G.STRESS_FIELDS).reshape(ne, G.STRESS_FIELDS)
    import torch
    ref, status = s.stress_device(torch.as_tensor(us[-1], device=device))
    torch.cuda.synchronize()
    assert (status.cpu().numpy() == 0).all()
    assert np.array_equal(stress, ref.cpu().numpy())
    assert pos == got.size
