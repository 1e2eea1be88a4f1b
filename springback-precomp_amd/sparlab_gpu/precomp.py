"""The precomp side of the device path: forming decks into device systems,
many decks in one batched process, and the adjoint as a compensation hook.

The reference's precomp package writes one deck per forming run
(python/precomp/fea/deck.py:206-231, `deck_document`) and runs one
sparlab_form subprocess per deck (python/precomp/fea/runner.py:106-335,
`run_deck`, `simulate_many`); its data generator hands a list of jobs to a
`Simulator` (python/precomp/ml/generate.py:100-175). Here the decks' models
go to the GPU instead:

  * `deck_model(deck)` reads a deck document as sparlab_form's Config.cpp
    reads it (the mesh, the material and its plasticity, the model's
    integration rule and formulation, the forming kinematics) into a
    `DeckModel`: coordinates, connectivity, `Material`, `Rule`, kinematics.
    What this path does not evaluate is refused with DeckError (the
    reference's ConfigError): another element formulation than the
    incompatible-mode Hex8, a plane stress state, another mesh type.
  * `DeckBatch(decks)` puts many decks' models in one process: decks that
    share material, rule and kinematics become one `BatchedSystem` (the
    block-diagonal union evaluated by one captured launch sequence per
    Newton iteration), so a generator submits a whole design of
    experiments to one GPU process instead of one solver process per deck.
    A batched Newton driver (the reference's FormingAnalysis::run,
    src/fem/Forming.cpp, stays with the caller) calls `evaluate` once per
    iteration with every deck's displacement and `commit` per increment.
  * `shape_sensitivity(...)`: the derivative of a function of an
    equilibrium state with respect to the blank's nodal reference
    coordinates, by the adjoint of the equilibrium (a solve with the
    transposed free-free tangent on the device and the coordinate
    vector-Jacobian product of the internal force), the gradient a
    gradient-based shape compensation needs where the reference's
    displacement adjustment (python/precomp/compensation.py:273-410)
    iterates on predictions.
"""
import json
import os

import numpy as np

from . import BatchedSystem, Material, NonlinearSystem, Rule, FINITE, FINITE_LOGARITHMIC, SMALL_STRAIN, lib
from . import workloads as W
from .results import ResultDirectory, element_values


class DeckError(ValueError):
    """A deck this path does not take, or an inconsistent one (the reference's
    ConfigError, src/io/Config.cpp)."""


KINEMATICS = {"small_strain": SMALL_STRAIN, "finite": FINITE, "finite_logarithmic": FINITE_LOGARITHMIC}
AXES = {"x": 0, "y": 1, "z": 2}


def read_deck(deck):
    """A deck document: the dict itself, a deck.json path or its directory."""
    if isinstance(deck, dict):
        return deck
    path = os.fspath(deck)
    if os.path.isdir(path):
        path = os.path.join(path, "deck.json")
    with open(path, encoding="utf-8") as f:
        return json.load(f)


def _vector3(v, where):
    if not (isinstance(v, (list, tuple)) and len(v) == 3):
        raise DeckError(f"'{where}' must be a direction of three numbers [x, y, z]")
    return tuple(float(x) for x in v)


def deck_hill(node, where):
    """The Hill48 block of a plasticity section (Config.cpp:191-256,
    parse_anisotropy): exactly one calibration, the out-of-plane shear and the
    material frame (rolling_angle, or rolling_direction and sheet_normal)."""
    r_values = any(k in node for k in ("r0", "r45", "r90"))
    forms = int(r_values) + int("stress_ratios" in node) + int("coefficients" in node)
    if forms != 1:
        raise DeckError(f"'{where}' needs exactly one calibration of the Hill48 criterion: the r-values "
                        f"\"r0\", \"r45\", \"r90\", or \"stress_ratios\", or \"coefficients\" (got {forms})")
    hill = {}
    if r_values:
        hill.update(calibration="r_values", r0=float(node["r0"]), r45=float(node["r45"]), r90=float(node["r90"]))
    elif "stress_ratios" in node:
        sr = node["stress_ratios"]
        hill.update(calibration="stress_ratios", sigma45=float(sr["sigma_45"]), sigma90=float(sr["sigma_90"]),
                    sigma_biaxial=float(sr["sigma_biaxial"]))
    else:
        c = node["coefficients"]
        hill.update(calibration="coefficients", **{k: float(c[k]) for k in "FGHLMN"})
    if "out_of_plane_shear" in node:
        if "coefficients" in node:
            raise DeckError(f"'{where}.out_of_plane_shear' is given by L and M of the \"coefficients\"")
        lm = node["out_of_plane_shear"]
        if len(lm) != 2:
            raise DeckError(f"'{where}.out_of_plane_shear' must be the two numbers [L, M]")
        hill.update(L=float(lm[0]), M=float(lm[1]))
    if "rolling_angle" in node:
        if "rolling_direction" in node or "sheet_normal" in node:
            raise DeckError(f"'{where}.rolling_angle' excludes \"rolling_direction\" and \"sheet_normal\"")
        a = float(node["rolling_angle"]) * np.pi / 180.0
        hill.update(rolling_direction=(np.cos(a), np.sin(a), 0.0), sheet_normal=(0.0, 0.0, 1.0))
    else:
        if "rolling_direction" in node:
            hill["rolling_direction"] = _vector3(node["rolling_direction"], f"{where}.rolling_direction")
        if "sheet_normal" in node:
            hill["sheet_normal"] = _vector3(node["sheet_normal"], f"{where}.sheet_normal")
    return hill


def deck_material(mat):
    """A material section (Config.cpp:258-320, parse_material) as a Material."""
    if "youngs_modulus" not in mat or "poisson_ratio" not in mat:
        raise DeckError("a material needs \"youngs_modulus\" and \"poisson_ratio\"")
    kw = {}
    plastic = mat.get("plasticity")
    if plastic is not None:
        if not float(plastic.get("yield_stress", 0.0)) > 0.0:
            raise DeckError("'material.plasticity.yield_stress' must be a positive number")
        kw.update(yield_stress=float(plastic["yield_stress"]),
                  hardening_modulus=float(plastic.get("hardening_modulus", 0.0)),
                  kinematic_hardening_modulus=float(plastic.get("kinematic_hardening_modulus", 0.0)),
                  saturation_stress=float(plastic.get("saturation_stress", 0.0)),
                  saturation_rate=float(plastic.get("saturation_rate", 0.0)))
        criterion = plastic.get("yield_criterion", "von_mises")
        if criterion == "hill48":
            if "anisotropy" not in plastic:
                raise DeckError("'material.plasticity' has \"yield_criterion\": \"hill48\" but no \"anisotropy\" block")
            kw["hill"] = deck_hill(plastic["anisotropy"], "material.plasticity.anisotropy")
        elif criterion == "von_mises":
            if "anisotropy" in plastic:
                raise DeckError("'material.plasticity.anisotropy' needs \"yield_criterion\": \"hill48\"")
        else:
            raise DeckError(f"'material.plasticity.yield_criterion' is \"{criterion}\"; expected \"von_mises\" "
                            "or \"hill48\"")
        back = plastic.get("backstresses", [])
        if len(back) > 4:
            raise DeckError(f"'material.plasticity.backstresses' has {len(back)} entries; at most 4 are supported")
        kw["backstresses"] = [(float(b["modulus"]), float(b.get("recovery", 0.0))) for b in back]
        integration = plastic.get("kinematic_integration", "exponential")
        if integration not in ("exponential", "backward_euler"):
            raise DeckError(f"'material.plasticity.kinematic_integration' is \"{integration}\"; expected "
                            "\"exponential\" or \"backward_euler\"")
        kw["kinematic_integration"] = integration
    try:
        return Material(float(mat["youngs_modulus"]), float(mat["poisson_ratio"]), **kw)
    except RuntimeError as exc:
        raise DeckError(f"'material': {exc}") from exc


def deck_rule(model):
    """The model's Hex8 stiffness rule (Config.cpp:531-550): stiffness_points
    (2), thickness_points (0: the in-plane count) along thickness_direction (z)."""
    integ = model.get("integration", {})
    n = int(integ.get("stiffness_points", 2))
    m = int(integ.get("thickness_points", 0))
    axis = integ.get("thickness_direction", "z")
    if axis not in AXES:
        raise DeckError(f"'model.integration.thickness_direction' is \"{axis}\"; expected \"x\", \"y\" or \"z\"")
    if m < 0 or m > 7:
        raise DeckError(f"'model.integration.thickness_points' is {m}; it must be 1 to 7")
    return Rule(n, m, AXES[axis])


class DeckModel:
    """What the device path evaluates of one deck: the structured Hex8 mesh
    (coords (n, 3), conn (e, 8), src/mesh/StructuredMesh.cpp:62-113), the
    material, the integration rule and the forming kinematics."""

    def __init__(self, name, coords, conn, material, rule, kinematics, material_record, mesh_source=""):
        self.name = name
        self.coords, self.conn = coords, conn
        self.material, self.rule, self.kinematics = material, rule, kinematics
        self.material_record = material_record
        self.mesh_source = mesh_source  # Configuration::describe_mesh (src/io/Config.cpp:346-358)

    def system(self):
        """The model's device NonlinearSystem."""
        return NonlinearSystem(self.coords, self.conn, self.material, self.rule, self.kinematics)

    @property
    def group(self):
        """Decks of one group share one batched system: material, rule, kinematics."""
        r = self.rule.c
        return (json.dumps(self.material_record, sort_keys=True), r.stiffness_points, r.thickness_points,
                r.thickness_axis, self.kinematics)


def deck_model(deck):
    """A deck document (or deck.json path, or deck directory) as a DeckModel."""
    d = read_deck(deck)
    mesh = d.get("mesh", {})
    if mesh.get("type") != "structured_hex":
        raise DeckError(f"'mesh.type' is {mesh.get('type')!r}; the device path takes \"structured_hex\" meshes")
    for k in ("nx", "ny", "nz", "lx", "ly", "lz"):
        if k not in mesh:
            raise DeckError(f"'mesh.{k}' is missing")
    model = d.get("model", {})
    if model.get("stress_state", "three_dimensional") != "three_dimensional":
        raise DeckError("the device path evaluates three-dimensional Hex8 models (\"stress_state\": "
                        "\"three_dimensional\")")
    formulation = model.get("element_formulation", "standard")
    if formulation != "incompatible_modes":
        raise DeckError(f"'model.element_formulation' is \"{formulation}\"; the device path is the "
                        "incompatible-mode Hex8 (\"incompatible_modes\")")
    forming = d.get("forming", {})
    kin = forming.get("kinematics", "finite")
    if kin not in KINEMATICS:
        raise DeckError(f"'forming.kinematics' is \"{kin}\"; expected one of {sorted(KINEMATICS)}")
    coords, conn = W.structured_hex_mesh(int(mesh["nx"]), int(mesh["ny"]), int(mesh["nz"]), float(mesh["lx"]),
                                         float(mesh["ly"]), float(mesh["lz"]), float(mesh.get("x0", 0.0)),
                                         float(mesh.get("y0", 0.0)), float(mesh.get("z0", 0.0)))
    if "material" not in d:
        raise DeckError("the deck has no \"material\"")
    material = deck_material(d["material"])
    return DeckModel(d.get("name", "deck"), coords, conn, material, deck_rule(model), KINEMATICS[kin],
                     d["material"], f"structured_hex {int(mesh['nx'])} x {int(mesh['ny'])} x {int(mesh['nz'])}")


class DeckBatch:
    """Many decks in one GPU process: every group of decks sharing material,
    rule and kinematics is one BatchedSystem; `models[i]` is deck i's model.

    evaluate(us): one Newton iteration of every deck (us[i]: deck i's
    displacement, 3 per node, host or device); returns per deck (f_int,
    tangent values in the reference's column-major order or None) as device
    tensors, and the per-deck scalars (energy, gross, yielding, failures).
    commit(): every deck's trial history becomes its committed one.
    """

    def __init__(self, decks, graphs=True):
        self.models = [deck_model(d) for d in decks]
        groups = {}
        for i, m in enumerate(self.models):
            groups.setdefault(m.group, []).append(i)
        self.groups = []
        for members in groups.values():
            first = self.models[members[0]]
            b = BatchedSystem([(self.models[i].coords, self.models[i].conn) for i in members], first.material,
                              first.rule, first.kinematics)
            b.set_graphs(graphs)
            self.groups.append((members, b))

    def __len__(self):
        return len(self.models)

    def evaluate(self, us, want_tangent=True, stream=None):
        import torch
        if len(us) != len(self.models):
            raise DeckError(f"{len(us)} displacements for {len(self.models)} decks")
        out = [None] * len(self.models)
        scalars = [None] * len(self.models)
        for members, b in self.groups:
            dev = torch.device("cuda")
            u = torch.cat([torch.as_tensor(np.asarray(us[i]).reshape(-1) if not isinstance(us[i], torch.Tensor)
                                           else us[i].reshape(-1), dtype=torch.float64, device=dev)
                           for i in members])
            fint = torch.empty_like(u)
            values = torch.empty(b.system.nnz, dtype=torch.float64, device=dev) if want_tangent else None
            b.evaluate_device(u, fint, values, want_tangent=want_tangent, column_major=True, stream=stream)
            sc = b.result(stream)
            for p, i in enumerate(members):
                d0, d1 = 3 * b.node_offset[p], 3 * b.node_offset[p + 1]
                v0, v1 = b.value_offset[p], b.value_offset[p + 1]
                out[i] = (fint[d0:d1], None if values is None else values[v0:v1])
                scalars[i] = {k: sc[k][p] for k in sc}
        return out, scalars

    def commit(self):
        for _, b in self.groups:
            b.commit()

    def write_results(self, directories, us, step="form", termination="", **fields):
        """Every deck's result directory in sparlab_form's output contract
        (sparlab_gpu.results; read by python/precomp/fea/results.py:320-365):
        one completed step named `step` with deck i's node table at us[i]
        and its element table from the end-of-step stress recovery at us[i]
        on the committed history (src/fem/Forming.cpp:1255-1281; each group's
        union system recovers the stresses of all its decks at once), then
        summary.json. `fields` go into each step's summary entry. Returns
        the summaries."""
        import torch
        if len(directories) != len(self.models) or len(us) != len(self.models):
            raise DeckError(f"{len(directories)} directories and {len(us)} displacements for "
                            f"{len(self.models)} decks")
        host = [(us[i].detach().cpu().numpy() if isinstance(us[i], torch.Tensor) else np.asarray(us[i],
                                                                                              dtype=np.float64))
                .reshape(-1) for i in range(len(self.models))]
        version = lib().sparlab_version().decode()
        summaries = [None] * len(self.models)
        for members, b in self.groups:
            stress = b.system.stress(np.concatenate([host[i] for i in members]))
            eps, vm = element_values(stress, self.models[members[0]].material.plastic)
            for p, i in enumerate(members):
                m = self.models[i]
                e0, e1 = b.element_offset[p], b.element_offset[p + 1]
                d = ResultDirectory(directories[i], m.coords, m.conn, case=m.name, version=version,
                                    rule_points=m.rule.points, mesh_source=m.mesh_source)
                d.add_step(step, host[i], plastic_strain=eps[e0:e1], von_mises=vm[e0:e1], **fields)
                summaries[i] = d.close(termination=termination)
        return summaries


def shape_sensitivity(system, u, values, free, dj_du, tolerance=1e-10, max_iterations=200000):
    """dJ/dX at an equilibrium state: J a function of the displacement,
    f_int(u, X) restricted to the free DOFs = 0 (the prescribed DOFs fixed).

    system: the NonlinearSystem whose last evaluation is at u (device, 3 per
    node) with the tangent's column-major values `values` (device); free:
    ascending free DOF indices (int32 device tensor); dj_du: dJ/du on the
    free DOFs (device). With K_ff^T lambda = dJ/du_f (the column-major values'
    free-free block read by rows is K_ff^T; conjugate gradients for a
    symmetric tangent, BiCGStab otherwise), dJ/dX = -lambda^T df_int/dX, the
    coordinate vector-Jacobian product of the internal force with lambda
    (zero on the prescribed DOFs) as cotangent. Returns (dJ/dX (num_nodes,
    3) device tensor, solver iterations)."""
    import torch
    system.partition_device(free)
    block = system.free_block_device(values)
    lam_f, its, _ = system.solve_free_block_device(block, dj_du, tolerance=tolerance, max_iterations=max_iterations,
                                                    symmetric=system.material.symmetric_tangent)
    lam = torch.zeros(3 * system.num_nodes, dtype=torch.float64, device=u.device)
    lam[free.long()] = lam_f
    g = system.vjp_device(u, lam, coords=True)["coords"]
    return -g.reshape(-1, 3), its
