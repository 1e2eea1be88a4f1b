"""The deck reader of the precomp side (sparlab_gpu.precomp) on the
reference's own decks (tests/golden/decks, copied from its benchmarks by
tests/golden/copy_decks.py): the model the device path evaluates, as
sparlab_form's Config.cpp reads it, and the refusals of what it does not take.
No device calls (materials and rules are host-side C-ABI work).
"""
import copy
import json
import os

import numpy as np
import pytest

import sparlab_gpu as G
from sparlab_gpu import precomp as P

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def deck(name):
    with open(os.path.join(DECKS, name)) as f:
        return json.load(f)


def test_fine_mesh_deck_reads_as_the_reference_runs_it():
    """benchmarks/physics_audit/challenge/cases/kin_h0833: 48 x 48 x 1 IM Hex8
    (2 304 elements, as its summary.json counts), the 2 x 2 x 5 rule through z,
    logarithmic strain, Hill48 (r 0.75 / 0.7 / 0.8) + Voce + two backstresses."""
    m = P.deck_model(os.path.join(DECKS, "kin_h0833.json"))
    assert m.conn.shape == (2304, 8) and m.coords.shape == (49 * 49 * 2, 3)
    assert np.allclose(m.coords.min(axis=0), [-0.02, -0.02, -0.001]) and np.allclose(m.coords.max(axis=0),
                                                                                       [0.02, 0.02, 0.0])
    assert m.rule.points == 20 and (m.rule.c.stiffness_points, m.rule.c.thickness_points, m.rule.c.thickness_axis) == (2, 5, 2)
    assert m.kinematics == G.FINITE_LOGARITHMIC
    pl = deck("kin_h0833.json")["material"]["plasticity"]
    ref = G.Material(70e9, 0.33, yield_stress=pl["yield_stress"], hardening_modulus=pl["hardening_modulus"],
                     saturation_stress=pl["saturation_stress"], saturation_rate=pl["saturation_rate"],
                     hill=dict(r0=0.75, r45=0.7, r90=0.8),
                     backstresses=[(b["modulus"], b["recovery"]) for b in pl["backstresses"]])
    assert np.array_equal(m.material.yield_matrix, ref.yield_matrix)
    c, r = m.material.c, ref.c
    assert (c.num_backstresses, c.kinematic_integration, c.criterion) == (2, 0, 1)
    assert list(c.backstress_modulus) == list(r.backstress_modulus)
    assert list(c.backstress_recovery) == list(r.backstress_recovery)
    assert not m.material.symmetric_tangent


def test_standard_element_deck_is_refused_and_read_once_it_asks_for_the_modes():
    d = deck("homogeneous_uniaxial.json")
    with pytest.raises(P.DeckError, match="incompatible_modes"):
        P.deck_model(d)
    d["model"]["element_formulation"] = "incompatible_modes"
    m = P.deck_model(d)
    assert m.conn.shape == (1, 8) and m.rule.points == 8 and m.kinematics == G.FINITE_LOGARITHMIC
    assert m.material.c.criterion == 0 and m.material.c.yield_stress == 100e6
    assert m.material.symmetric_tangent


@pytest.mark.parametrize("edit, match", [
    (lambda d: d["model"].update(stress_state="plane_stress"), "three-dimensional"),
    (lambda d: d["mesh"].update(type="gmsh"), "structured_hex"),
    (lambda d: d["material"]["plasticity"].pop("anisotropy"), "anisotropy"),
    (lambda d: d["material"]["plasticity"]["anisotropy"].update(stress_ratios={"sigma_45": 1.0, "sigma_90": 1.0,
                                                                               "sigma_biaxial": 1.0}),
     "exactly one calibration"),
    (lambda d: d["material"]["plasticity"].update(yield_criterion="tresca"), "von_mises"),
    (lambda d: d["material"]["plasticity"].update(kinematic_integration="midpoint"), "backward_euler"),
    (lambda d: d["forming"].update(kinematics="corotational"), "forming.kinematics"),
    (lambda d: d["model"]["integration"].update(thickness_direction="w"), "thickness_direction"),
])
def test_decks_the_path_does_not_take_are_refused(edit, match):
    d = copy.deepcopy(deck("kin_h0833.json"))
    edit(d)
    with pytest.raises(P.DeckError, match=match):
        P.deck_model(d)


def test_hill_calibrations_and_frames():
    """The three Hill48 calibrations and the rolling angle of Config.cpp:191-256."""
    base = deck("kin_h0833.json")["material"]
    for aniso, hill in (({"r0": 1.9, "r45": 1.5, "r90": 2.3, "rolling_angle": 30.0},
                         dict(r0=1.9, r45=1.5, r90=2.3, rolling_direction=(np.cos(np.pi / 6), np.sin(np.pi / 6), 0.0))),
                        ({"stress_ratios": {"sigma_45": 1.05, "sigma_90": 0.98, "sigma_biaxial": 1.02}},
                         dict(calibration="stress_ratios", sigma45=1.05, sigma90=0.98, sigma_biaxial=1.02)),
                        ({"coefficients": {"F": 0.3, "G": 0.5, "H": 0.5, "L": 1.5, "M": 1.5, "N": 1.4}},
                         dict(calibration="coefficients", F=0.3, G=0.5, H=0.5, L=1.5, M=1.5, N=1.4))):
        mat = copy.deepcopy(base)
        mat["plasticity"]["anisotropy"] = aniso
        got = P.deck_material(mat)
        ref = G.Material(70e9, 0.33, yield_stress=mat["plasticity"]["yield_stress"],
                         hardening_modulus=mat["plasticity"]["hardening_modulus"],
                         saturation_stress=mat["plasticity"]["saturation_stress"],
                         saturation_rate=mat["plasticity"]["saturation_rate"], hill=hill,
                         backstresses=[(b["modulus"], b["recovery"]) for b in mat["plasticity"]["backstresses"]])
        assert np.array_equal(got.yield_matrix, ref.yield_matrix), aniso
