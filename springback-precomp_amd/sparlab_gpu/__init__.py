"""Python host side of the MI355X element-and-assembly path.

Mirrors the reference's element interface -- `elastoplastic_element`
(include/sparlab/fem/Elastoplastic.hpp:196-200) for a batch of elements, the
`PlasticityParameters`/`Material` record (include/sparlab/material/Material.hpp:
148-257) and the `PlasticState` history (include/sparlab/material/Plasticity.hpp:
185-201) -- over the C ABI of include/sparlab_gpu.h (libsparlab_gpu.so, built
in-tree by springback-precomp_amd/build.py). PyTorch provides device memory and
streams only; every computation runs in the library's HIP kernels. There is
no CPU fallback: without the built library or a GPU the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libsparlab_gpu.so")

SMALL_STRAIN, FINITE, FINITE_LOGARITHMIC = 0, 1, 2

ELEMENT_STATUS = {
    0: "ok",
    1: "MeshError: det J <= 0 at the element centre",
    2: "MeshError: det J <= 0 at an integration point",
    3: "SolverError: an element is inverted (det F <= 0)",
    4: "SolverError: a point is degenerate (1 + 2 mu <= 0 or a non-finite strain)",
    5: "SolverError: the return did not converge",
    6: "SolverError: the stiffness of the incompatible modes is singular",
    7: "SolverError: the incompatible modes did not converge in 25 local iterations",
    8: "SolverError: no step of the incompatible modes lowers their residual",
    9: "SolverError: the closest-point return met a non-finite trial stress",
    10: "internal error",
}


class SolverError(RuntimeError):
    """A per-element failure (the reference's SolverError / MeshError)."""

    def __init__(self, message, element=None, status=None):
        super().__init__(message)
        self.element = element
        self.status = status


class CMaterial(ctypes.Structure):
    _fields_ = [
        ("youngs_modulus", ctypes.c_double),
        ("poisson_ratio", ctypes.c_double),
        ("yield_stress", ctypes.c_double),
        ("hardening_modulus", ctypes.c_double),
        ("kinematic_hardening_modulus", ctypes.c_double),
        ("saturation_stress", ctypes.c_double),
        ("saturation_rate", ctypes.c_double),
        ("criterion", ctypes.c_int32),
        ("num_backstresses", ctypes.c_int32),
        ("backstress_modulus", ctypes.c_double * 4),
        ("backstress_recovery", ctypes.c_double * 4),
        ("kinematic_integration", ctypes.c_int32),
        ("hill_calibration", ctypes.c_int32),
        ("hill_r", ctypes.c_double * 3),
        ("hill_stress_ratio", ctypes.c_double * 3),
        ("hill_shear_lm", ctypes.c_double * 2),
        ("hill_coefficients", ctypes.c_double * 6),
        ("rolling_direction", ctypes.c_double * 3),
        ("sheet_normal", ctypes.c_double * 3),
        ("yield_matrix", ctypes.c_double * 36),
    ]


class CRule(ctypes.Structure):
    _fields_ = [("stiffness_points", ctypes.c_int32), ("thickness_points", ctypes.c_int32),
                ("thickness_axis", ctypes.c_int32)]


class CElementBatch(ctypes.Structure):
    _fields_ = [
        ("coords", ctypes.c_void_p), ("conn", ctypes.c_void_p), ("u", ctypes.c_void_p),
        ("num_nodes", ctypes.c_int64), ("num_elements", ctypes.c_int64),
        ("state_in", ctypes.c_void_p), ("alpha_in", ctypes.c_void_p),
        ("state_out", ctypes.c_void_p), ("alpha_out", ctypes.c_void_p),
        ("force", ctypes.c_void_p), ("tangent", ctypes.c_void_p), ("energy", ctypes.c_void_p),
        ("yielding", ctypes.c_void_p), ("iterations", ctypes.c_void_p), ("status", ctypes.c_void_p),
        ("fail_value", ctypes.c_void_p), ("workspace", ctypes.c_void_p), ("local_path", ctypes.c_void_p),
    ]


_lib = None


def workspace_doubles(num_elements, points):
    """Doubles of device scratch an element batch of `points`-point elements
    needs (sparlab_element_workspace_doubles: 86 per point, 108 per element
    and the local solve's schedule)."""
    rule = {8: (2, 0, 2), 1: (1, 0, 2), 27: (3, 0, 2), 64: (4, 0, 2)}.get(points)
    if rule is None:
        if points % 4 != 0:
            raise ValueError(f"no rule of {points} points")
        rule = (2, points // 4, 2)
    n = lib().sparlab_element_workspace_doubles(int(num_elements), ctypes.byref(CRule(*rule)))
    if n < 0:
        raise RuntimeError(f"sparlab_gpu error {n}: {lib().sparlab_last_error().decode()}")
    return n


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch's libamdhip64 (SONAME
        # libamdhip64.so.7) is loaded first, so the library's dependency of the
        # same SONAME binds to it and torch's device pointers and streams are
        # the library's.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run springback-precomp_amd/build.py "
                               "(or __graft_entry__.build()) first")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.sparlab_last_error.restype = ctypes.c_char_p
        _lib.sparlab_version.restype = ctypes.c_char_p
        _lib.sparlab_material_finalize.argtypes = [ctypes.POINTER(CMaterial)]
        _lib.sparlab_rule_points.argtypes = [ctypes.POINTER(CRule)]
        _lib.sparlab_state_fields.argtypes = [ctypes.POINTER(CMaterial)]
        _lib.sparlab_element_workspace_doubles.restype = ctypes.c_int64
        _lib.sparlab_element_workspace_doubles.argtypes = [ctypes.c_int64, ctypes.POINTER(CRule)]
        _lib.sparlab_elements_evaluate_device.argtypes = [
            ctypes.POINTER(CMaterial), ctypes.POINTER(CRule), ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(CElementBatch), ctypes.c_void_p]
        _lib.sparlab_elements_evaluate_stage_device.argtypes = [
            ctypes.POINTER(CMaterial), ctypes.POINTER(CRule), ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(CElementBatch), ctypes.c_int, ctypes.c_void_p]
        _lib.sparlab_elements_evaluate_noalloc_device.argtypes = [
            ctypes.POINTER(CMaterial), ctypes.POINTER(CRule), ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(CElementBatch), ctypes.c_void_p]
    return _lib


def _check(code):
    if code < 0:
        raise RuntimeError(f"sparlab_gpu error {code}: {lib().sparlab_last_error().decode()}")
    return code


# ---------------------------------------------------------------------------
# Material, rule, state layout
# ---------------------------------------------------------------------------
class Material:
    """Material + PlasticityParameters (include/sparlab/material/Material.hpp).

    Units SI. hill: None (von Mises) or a dict with r0, r45, r90 (or
    calibration/stress ratios/coefficients), rolling_direction, sheet_normal.
    backstresses: sequence of (C_i [Pa], gamma_i).
    """

    def __init__(self, youngs_modulus, poisson_ratio, yield_stress=0.0, hardening_modulus=0.0,
                 kinematic_hardening_modulus=0.0, saturation_stress=0.0, saturation_rate=0.0,
                 hill=None, backstresses=(), kinematic_integration="exponential"):
        c = CMaterial()
        c.youngs_modulus = youngs_modulus
        c.poisson_ratio = poisson_ratio
        c.yield_stress = yield_stress
        c.hardening_modulus = hardening_modulus
        c.kinematic_hardening_modulus = kinematic_hardening_modulus
        c.saturation_stress = saturation_stress
        c.saturation_rate = saturation_rate
        c.num_backstresses = len(backstresses)
        for i, (cm, g) in enumerate(backstresses):
            c.backstress_modulus[i] = cm
            c.backstress_recovery[i] = g
        c.kinematic_integration = 0 if kinematic_integration == "exponential" else 1
        c.hill_shear_lm[0] = c.hill_shear_lm[1] = 1.5
        c.rolling_direction[0] = 1.0
        c.sheet_normal[2] = 1.0
        if hill is not None:
            c.criterion = 1
            cal = hill.get("calibration", "r_values")
            c.hill_calibration = {"r_values": 0, "stress_ratios": 1, "coefficients": 2}[cal]
            c.hill_r[0], c.hill_r[1], c.hill_r[2] = hill.get("r0", 1.0), hill.get("r45", 1.0), hill.get("r90", 1.0)
            for i, k in enumerate(("sigma45", "sigma90", "sigma_biaxial")):
                c.hill_stress_ratio[i] = hill.get(k, 1.0)
            c.hill_shear_lm[0], c.hill_shear_lm[1] = hill.get("L", 1.5), hill.get("M", 1.5)
            for i, k in enumerate("FGHLMN"):
                c.hill_coefficients[i] = hill.get(k, 0.5 if i < 3 else 1.5)
            for i, v in enumerate(hill.get("rolling_direction", (1.0, 0.0, 0.0))):
                c.rolling_direction[i] = v
            for i, v in enumerate(hill.get("sheet_normal", (0.0, 0.0, 1.0))):
                c.sheet_normal[i] = v
        _check(lib().sparlab_material_finalize(ctypes.byref(c)))
        self.c = c

    @property
    def plastic(self):
        """The material yields (a positive yield stress: the library's rule,
        Plasticity::enabled in the reference): its elements are elastoplastic."""
        return self.c.yield_stress > 0.0

    @property
    def yield_matrix(self):
        return np.array(self.c.yield_matrix[:]).reshape(6, 6)

    @property
    def general(self):
        return self.c.criterion == 1 or self.c.num_backstresses > 0

    @property
    def kinematic_terms(self):
        return self.c.num_backstresses + (1 if self.c.kinematic_hardening_modulus > 0.0 else 0)

    @property
    def kin_fields(self):
        return self.kinematic_terms if self.general else 1

    @property
    def state_fields(self):
        return _check(lib().sparlab_state_fields(ctypes.byref(self.c)))

    @property
    def symmetric_tangent(self):
        return all(self.c.backstress_recovery[i] <= 0.0 for i in range(self.c.num_backstresses))


class Rule:
    """The Hex8 stiffness rule (src/elements/Hex8Kernels.hpp:125-130)."""

    def __init__(self, stiffness_points=2, thickness_points=0, thickness_axis=2):
        self.c = CRule(stiffness_points, thickness_points, thickness_axis)

    @property
    def points(self):
        return _check(lib().sparlab_rule_points(ctypes.byref(self.c)))


# The reference's PlasticState as 39 doubles: plastic strain (6), back stress
# beta (6), backstresses alpha_1..alpha_4 (24), equivalent plastic strain,
# thickness strain, loading flag.
REFERENCE_STATE = 39


def states_from_reference(ref, material):
    """(E, P, 39) PlasticState records -> (fields, E*P) library layout."""
    ref = np.asarray(ref, dtype=np.float64)
    e, p, _ = ref.shape
    nf = material.state_fields
    kf = material.kin_fields
    out = np.zeros((nf, e * p))
    flat = ref.reshape(e * p, REFERENCE_STATE)
    out[0:6] = flat[:, 0:6].T
    if material.general:
        for t in range(kf):
            out[6 + 6 * t:12 + 6 * t] = flat[:, 12 + 6 * t:18 + 6 * t].T
    else:
        out[6:12] = flat[:, 6:12].T
    out[6 + 6 * kf] = flat[:, 36]
    out[7 + 6 * kf] = flat[:, 38]
    return out


def states_to_reference(lib_states, material, num_elements, points):
    """(fields, E*P) library layout -> (E, P, 39) PlasticState records."""
    s = np.asarray(lib_states, dtype=np.float64)
    kf = material.kin_fields
    ref = np.zeros((num_elements * points, REFERENCE_STATE))
    ref[:, 0:6] = s[0:6].T
    if material.general:
        for t in range(kf):
            ref[:, 12 + 6 * t:18 + 6 * t] = s[6 + 6 * t:12 + 6 * t].T
            ref[:, 6:12] += s[6 + 6 * t:12 + 6 * t].T
    else:
        ref[:, 6:12] = s[6:12].T
    ref[:, 36] = s[6 + 6 * kf]
    ref[:, 38] = s[7 + 6 * kf]
    return ref.reshape(num_elements, points, REFERENCE_STATE)


# ---------------------------------------------------------------------------
# Element batches on the device
# ---------------------------------------------------------------------------
def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _require(t, name, length, dtype="float64", device=None, at_least=False):
    """Refuses (ValueError) a tensor the C ABI would read or write out of
    bounds: the wrong dtype, not contiguous, on another device, or fewer
    (or, unless at_least, other than) `length` elements. The library's
    device entry points take no sizes for these buffers."""
    import torch
    want = getattr(torch, dtype)
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != want or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {dtype} tensor")
    if t.device.type != "cuda" or (device is not None and t.device != device):
        raise ValueError(f"{name} must be on the device {device if device is not None else 'cuda'}")
    n = t.numel()
    if (n < length) if at_least else (n != length):
        raise ValueError(f"{name} holds {n} entries; {'at least ' if at_least else ''}{length} are needed")


class ElementBatch:
    """Device buffers of one element batch (torch tensors on a HIP device).

    coords (N, 3) f64, conn (E, 8) i32, u (N, 3) f64, state_in (F, E*P) f64,
    alpha_in (E, 9) f64 or None. Outputs are allocated once and reused.
    """

    def __init__(self, coords, conn, u, state_in, alpha_in, points, want_tangent=True):
        import torch
        self.coords, self.conn, self.u = coords, conn, u
        self.state_in, self.alpha_in = state_in, alpha_in
        ne = conn.shape[0]
        dev = coords.device
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.num_elements = ne
        self.points = points
        self.state_out = torch.empty_like(state_in)
        self.alpha_out = torch.empty((ne, 9), **f64)
        self.force = torch.empty((ne, 24), **f64)
        self.tangent = torch.empty((ne, 24, 24), **f64) if want_tangent else None
        self.energy = torch.empty((ne,), **f64)
        self.yielding = torch.empty((ne,), **i32)
        self.iterations = torch.empty((ne,), **i32)
        self.status = torch.empty((ne,), **i32)
        self.fail_value = torch.empty((ne,), **f64)
        self.local_path = torch.empty((ne,), **i32)  # SPARLAB_LOCAL_* bits per element
        # device scratch of the element kernels (86 doubles per point, 108 per element)
        self.workspace = torch.empty((workspace_doubles(ne, points),), **f64)
        for name in ("coords", "u", "state_in"):
            t = getattr(self, name)
            if not (t.dtype == torch.float64 and t.is_contiguous() and t.device == dev):
                raise ValueError(f"{name} must be a contiguous float64 tensor on {dev}")
        if not (conn.dtype == torch.int32 and conn.is_contiguous()):
            raise ValueError("conn must be a contiguous int32 tensor")
        self.c = CElementBatch(
            _ptr(coords), _ptr(conn), _ptr(u), coords.shape[0], ne, _ptr(state_in), _ptr(alpha_in),
            _ptr(self.state_out), _ptr(self.alpha_out), _ptr(self.force), _ptr(self.tangent),
            _ptr(self.energy), _ptr(self.yielding), _ptr(self.iterations), _ptr(self.status),
            _ptr(self.fail_value), _ptr(self.workspace), _ptr(self.local_path))


LOCAL_SOLVE = 1
CONDENSATION = 2
INDEX_ORDER = 4  # the local solve takes the elements in index order (no cost ordering)
# the per-element stress measures of NonlinearSystem.stress (name, offset, width)
STRESS_FIELDS = 28
STRESS_LAYOUT = (("cauchy", 0, 6), ("piola_kirchhoff", 6, 6), ("kirchhoff", 12, 6), ("logarithmic_strain", 18, 6),
                 ("von_mises", 24, 1), ("max_equivalent_plastic_strain", 25, 1), ("max_strain", 26, 1),
                 ("min_jacobian", 27, 1))
# layout flags of the assembly of given element matrices
VALUES_COLUMN_MAJOR = 1
KE_COLUMN_MAJOR = 2


def evaluate_elements(material, rule, batch, kinematics=FINITE_LOGARITHMIC, want_tangent=True, stream=None,
                      stages=LOCAL_SOLVE | CONDENSATION):
    """Queues the evaluation of every element of `batch` (see ElementBatch);
    `stages` selects the local solve, the condensation or both."""
    import torch
    if want_tangent and batch.tangent is None:
        raise ValueError("the batch was allocated without a tangent buffer")
    s = stream if stream is not None else torch.cuda.current_stream(batch.coords.device)
    _check(lib().sparlab_elements_evaluate_stage_device(ctypes.byref(material.c), ctypes.byref(rule.c), kinematics,
                                                        int(want_tangent), ctypes.byref(batch.c), int(stages),
                                                        ctypes.c_void_p(s.cuda_stream)))
    return batch


NOALLOC_SCRATCH_BYTES = 128 << 20  # SPARLAB_NOALLOC_SCRATCH_BYTES


def evaluate_elements_noalloc(material, rule, batch, kinematics=FINITE_LOGARITHMIC, want_tangent=True, stream=None):
    """Queues the evaluation of every element of `batch` through
    sparlab_elements_evaluate_noalloc_device: no workspace (the batch's own is
    not used) and no device allocation; the library's per-device scratch array
    holds batches of up to NOALLOC_SCRATCH_BYTES of workspace. Calls that
    overlap on one GPU race."""
    import torch
    if want_tangent and batch.tangent is None:
        raise ValueError("the batch was allocated without a tangent buffer")
    s = stream if stream is not None else torch.cuda.current_stream(batch.coords.device)
    _check(lib().sparlab_elements_evaluate_noalloc_device(ctypes.byref(material.c), ctypes.byref(rule.c), kinematics,
                                                          int(want_tangent), ctypes.byref(batch.c),
                                                          ctypes.c_void_p(s.cuda_stream)))
    return batch


def raise_first_failure(batch):
    """Raises SolverError for the lowest-numbered failed element, if any."""
    st = batch.status.cpu().numpy()
    bad = np.nonzero(st)[0]
    if bad.size:
        e = int(bad[0])
        raise SolverError(f"Hex8 element {e}: {ELEMENT_STATUS.get(int(st[e]), 'unknown')}", e, int(st[e]))


# ---------------------------------------------------------------------------
# Sparsity pattern, nonlinear system, contact
# ---------------------------------------------------------------------------
class CEvaluation(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("gross", ctypes.c_double), ("yielding_points", ctypes.c_int64),
                ("failed_elements", ctypes.c_int64), ("first_failed_element", ctypes.c_int64),
                ("first_failure_status", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("first_failure_value", ctypes.c_double)]


class CTool(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("pad", ctypes.c_int32), ("centre", ctypes.c_double * 3),
                ("centre0", ctypes.c_double * 3), ("radius", ctypes.c_double), ("normal", ctypes.c_double * 3),
                ("axis", ctypes.c_double * 3), ("friction", ctypes.c_double),
                ("tangential_ratio", ctypes.c_double)]


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def pattern(num_nodes, conn):
    """The reference's sparsity pattern (host): row_ptr, col_idx, block_offset.
    Structurally symmetric: also the reference's column-major outer/inner."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    L = lib()
    L.sparlab_pattern_build.restype = ctypes.c_int64
    L.sparlab_pattern_build.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
    ne = conn.shape[0]
    nnz = _check(L.sparlab_pattern_build(num_nodes, ne, _np_ptr(conn), None, None, None))
    row_ptr = np.zeros(3 * num_nodes + 1, dtype=np.int32)
    col_idx = np.zeros(nnz, dtype=np.int32)
    bo = np.zeros(64 * ne, dtype=np.int32)
    _check(L.sparlab_pattern_build(num_nodes, ne, _np_ptr(conn), _np_ptr(row_ptr), _np_ptr(col_idx), _np_ptr(bo)))
    return row_ptr, col_idx, bo


class Evaluation:
    """NonlinearSystem::Evaluation (src/fem/NonlinearSystem.hpp:27-47) on the host."""

    def __init__(self, internal, values, scalars):
        self.internal = internal
        self.external = np.zeros_like(internal)
        self.residual = internal.copy()
        self.tangent_values = values
        self.energy = scalars.energy
        self.gross = scalars.gross
        self.yielding_points = scalars.yielding_points


class NonlinearSystem:
    """The device nonlinear system (sparlab_system_*): the element loop, the
    gather and the assembly of detail::NonlinearSystem with the Assembler's
    pattern, and the committed history."""

    def __init__(self, coords, conn, material, rule=None, kinematics=FINITE_LOGARITHMIC, _handle=None):
        import torch  # noqa: F401  (HIP runtime loaded through torch)
        self.owned = _handle is None
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 8)
        self.material = material
        self.rule = rule or Rule()
        self.kinematics = kinematics
        L = lib()
        L.sparlab_system_create.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.POINTER(CMaterial), ctypes.POINTER(CRule), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p)]
        L.sparlab_system_destroy.argtypes = [ctypes.c_void_p]
        L.sparlab_system_nnz.restype = ctypes.c_int64
        L.sparlab_system_nnz.argtypes = [ctypes.c_void_p]
        for name in ("sparlab_system_pattern",):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_system_diagonal_blocks.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_system_evaluate_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_system_result.argtypes = [ctypes.c_void_p, ctypes.POINTER(CEvaluation), ctypes.c_void_p]
        L.sparlab_system_evaluate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CEvaluation)]
        L.sparlab_system_commit.argtypes = [ctypes.c_void_p]
        L.sparlab_system_history.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_system_set_history.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_system_element_outputs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        if _handle is None:
            h = ctypes.c_void_p()
            _check(L.sparlab_system_create(_np_ptr(self.coords), self.coords.shape[0], _np_ptr(self.conn),
                                           self.conn.shape[0], ctypes.byref(material.c), ctypes.byref(self.rule.c),
                                           kinematics, ctypes.byref(h)))
        else:
            h = _handle
        self.h = h
        # kept for __del__, which may run after the module's globals are gone
        self._destroy = L.sparlab_system_destroy
        self.num_nodes = self.coords.shape[0]
        self.num_elements = self.conn.shape[0]
        self.points = self.rule.points
        self.nnz = L.sparlab_system_nnz(h)

    def __del__(self):
        h = getattr(self, "h", None)
        destroy = getattr(self, "_destroy", None)
        if h and getattr(self, "owned", True) and destroy is not None:
            destroy(h)
        self.h = None

    def pattern(self):
        row_ptr = np.zeros(3 * self.num_nodes + 1, dtype=np.int32)
        col_idx = np.zeros(self.nnz, dtype=np.int32)
        _check(lib().sparlab_system_pattern(self.h, _np_ptr(row_ptr), _np_ptr(col_idx)))
        return row_ptr, col_idx

    def diagonal_blocks(self):
        out = np.zeros(self.num_nodes, dtype=np.int32)
        _check(lib().sparlab_system_diagonal_blocks(self.h, _np_ptr(out)))
        return out

    def evaluate(self, u, lam=0.0, want_tangent=True, column_major=False):
        """NonlinearSystem::evaluate(u, lambda, want_tangent) from host arrays."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        fint = np.zeros(3 * self.num_nodes)
        values = np.zeros(self.nnz) if want_tangent else None
        ev = CEvaluation()
        rc = lib().sparlab_system_evaluate(self.h, _np_ptr(u), int(want_tangent), int(column_major), _np_ptr(fint),
                                           _np_ptr(values), ctypes.byref(ev))
        if rc == -4:
            raise SolverError(f"Hex8 element {ev.first_failed_element}: "
                              f"{ELEMENT_STATUS.get(ev.first_failure_status, 'unknown')}",
                              ev.first_failed_element, ev.first_failure_status)
        _check(rc)
        return Evaluation(fint, values, ev)

    def evaluate_device(self, u, fint, values=None, want_tangent=True, column_major=False, stream=None):
        """Queues an evaluation on device tensors (u, fint: 3 nn; values: nnz)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream(u.device)
        _check(lib().sparlab_system_evaluate_device(self.h, ctypes.c_void_p(u.data_ptr()), int(want_tangent),
                                                    int(column_major), ctypes.c_void_p(fint.data_ptr()),
                                                    None if values is None else ctypes.c_void_p(values.data_ptr()),
                                                    ctypes.c_void_p(s.cuda_stream)))

    def assemble_device(self, fe=None, ke=None, fint=None, values=None, gross=None, column_major=False,
                        stream=None, ke_column_major=False):
        """Gathers given element arrays (device tensors) into fint / values.
        column_major: the reference's column-major values; ke_column_major:
        every element matrix stored column-major (Eigen's storage)."""
        import torch
        L = lib()
        L.sparlab_system_assemble_device.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 4
        ref = next(t for t in (fe, ke) if t is not None)
        s = stream if stream is not None else torch.cuda.current_stream(ref.device)
        p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        layout = (VALUES_COLUMN_MAJOR if column_major else 0) | (KE_COLUMN_MAJOR if ke_column_major else 0)
        _check(L.sparlab_system_assemble_device(self.h, p(fe), p(ke), layout, p(fint), p(values), p(gross),
                                                ctypes.c_void_p(s.cuda_stream)))

    def result(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        ev = CEvaluation()
        rc = lib().sparlab_system_result(self.h, ctypes.byref(ev), ctypes.c_void_p(s.cuda_stream))
        if rc == -4:
            raise SolverError(f"Hex8 element {ev.first_failed_element}: "
                              f"{ELEMENT_STATUS.get(ev.first_failure_status, 'unknown')}",
                              ev.first_failed_element, ev.first_failure_status)
        _check(rc)
        return ev

    def commit(self):
        _check(lib().sparlab_system_commit(self.h))

    def history(self, trial=False):
        """(states (E, P, 39) reference records, alpha (E, 9))."""
        nf = self.material.state_fields
        st = np.zeros((nf, self.num_elements * self.points))
        al = np.zeros((self.num_elements, 9))
        _check(lib().sparlab_system_history(self.h, int(trial), _np_ptr(st), _np_ptr(al)))
        return states_to_reference(st, self.material, self.num_elements, self.points), al

    def set_history(self, states_ref=None, alpha=None):
        st = None if states_ref is None else np.ascontiguousarray(states_from_reference(states_ref, self.material))
        al = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        _check(lib().sparlab_system_set_history(self.h, _np_ptr(st), _np_ptr(al)))

    def set_coords(self, coords):
        """New nodal reference coordinates (host array, num_nodes x 3)."""
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        if c.shape[0] != self.num_nodes:
            raise ValueError("one coordinate triple per node")
        lib().sparlab_system_set_coords.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _check(lib().sparlab_system_set_coords(self.h, _np_ptr(c)))
        self.coords = c

    def set_material(self, material):
        lib().sparlab_system_set_material.argtypes = [ctypes.c_void_p, ctypes.POINTER(CMaterial)]
        _check(lib().sparlab_system_set_material(self.h, ctypes.byref(material.c)))
        self.material = material

    def vjp_device(self, u, v, coords=True, displacement=False, parameters=False, stream=None):
        """Vector-Jacobian products of the last evaluation's internal force at
        displacement u for the cotangent v (device tensors, 3 per node):
        dict with 'coords' (3 nn), 'displacement' (3 nn, = K^T v), 'parameters'
        (11: r0, r45, r90, C_1, gamma_1, ..., C_4, gamma_4) as requested."""
        import torch
        L = lib()
        L.sparlab_system_vjp_device.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 4
        s = stream if stream is not None else torch.cuda.current_stream(u.device)
        out = {}
        wrt = 0
        if coords:
            out["coords"] = torch.zeros(3 * self.num_nodes, dtype=torch.float64, device=u.device)
            wrt |= 1
        if displacement:
            out["displacement"] = torch.zeros(3 * self.num_nodes, dtype=torch.float64, device=u.device)
            wrt |= 2
        if parameters:
            out["parameters"] = torch.zeros(VJP_PARAMETERS, dtype=torch.float64, device=u.device)
            wrt |= 4
        p = lambda name: ctypes.c_void_p(out[name].data_ptr()) if name in out else None
        _check(L.sparlab_system_vjp_device(self.h, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(v.data_ptr()), wrt,
                                           p("coords"), p("displacement"), p("parameters"),
                                           ctypes.c_void_p(s.cuda_stream)))
        return out

    def stiffness_gross_device(self, values, x, free, column_major=True, product=None, stream=None):
        """stiffness_gross on device tensors: values (nnz), x (3 per node),
        free (int32 DOF indices). Returns (out (1,), product (3 per node) or
        None); out[0] = ||(|K| |x|)_free||."""
        import torch
        L = lib()
        L.sparlab_system_stiffness_gross_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _require(values, "values", self.nnz)
        _require(x, "x", 3 * self.num_nodes, device=values.device)
        _require(free, "free", free.numel() if isinstance(free, torch.Tensor) else 0, "int32", values.device)
        if product is not None:
            _require(product, "product", 3 * self.num_nodes, device=values.device)
        out = torch.empty(1, dtype=torch.float64, device=x.device)
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        _check(L.sparlab_system_stiffness_gross_device(self.h, p(values), int(column_major), p(x), p(free),
                                                       free.shape[0], p(out), p(product),
                                                       ctypes.c_void_p(s.cuda_stream)))
        return out, product

    def partition_device(self, free, stream=None):
        """PartitionedFactor::reset + build on the device for the free DOFs
        (int32 device tensor, strictly ascending): the free-free block's
        pattern, kept by the system. Returns its nonzero count; the library
        refuses (RuntimeError) a list that does not strictly ascend in range."""
        import torch
        L = lib()
        L.sparlab_system_partition_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]
        _require(free, "free", free.numel() if isinstance(free, torch.Tensor) else 0, "int32")
        nnz = ctypes.c_int64(0)
        s = stream if stream is not None else torch.cuda.current_stream(free.device)
        self._block = (-1, 0)  # a refused list leaves no partition
        _check(L.sparlab_system_partition_device(self.h, ctypes.c_void_p(free.data_ptr()), free.shape[0],
                                                 ctypes.byref(nnz), ctypes.c_void_p(s.cuda_stream)))
        self._block = (free.shape[0], nnz.value)
        return nnz.value

    def partition_pattern_device(self, device="cuda", stream=None):
        """The free-free block's pattern: (block_ptr (num_free + 1),
        block_index (nnz)) int32 device tensors, the reference's kff_ outer
        and inner index arrays."""
        import torch
        L = lib()
        L.sparlab_system_partition_pattern_device.argtypes = [ctypes.c_void_p] * 4
        nf, nnz = getattr(self, "_block", (-1, 0))
        ptr = torch.empty(max(nf + 1, 1), dtype=torch.int32, device=device)
        index = torch.empty(max(nnz, 1), dtype=torch.int32, device=device)
        s = stream if stream is not None else torch.cuda.current_stream(ptr.device)
        _check(L.sparlab_system_partition_pattern_device(self.h, ctypes.c_void_p(ptr.data_ptr()),
                                                         ctypes.c_void_p(index.data_ptr()),
                                                         ctypes.c_void_p(s.cuda_stream)))
        return ptr, index[:nnz]

    def free_block_device(self, values, out=None, stream=None):
        """PartitionedFactor::assemble: the free-free block's values (device
        tensor, nnz of the block) out of the assembled values, in their
        layout."""
        import torch
        L = lib()
        L.sparlab_system_free_block_device.argtypes = [ctypes.c_void_p] * 4
        nf, nnz = getattr(self, "_block", (-1, 0))  # without a partition the library refuses
        _require(values, "values", self.nnz)
        if out is None:
            out = torch.empty(max(nnz, 1), dtype=torch.float64, device=values.device)
        else:
            _require(out, "out", nnz, device=values.device, at_least=True)
        s = stream if stream is not None else torch.cuda.current_stream(values.device)
        _check(L.sparlab_system_free_block_device(self.h, ctypes.c_void_p(values.data_ptr()),
                                                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
        return out[:nnz]

    def solve_free_block_device(self, block_values, b, x=None, tolerance=1e-10, max_iterations=20000, symmetric=True,
                                stream=None):
        """K_ff x = b on the partition's block, preconditioned with its
        diagonal: conjugate gradients (symmetric) or BiCGStab (device tensors,
        num_free each; a non-symmetric block's values row-major, as K_ff's
        rows): (x, iterations, relative residual). Raises SolverError when
        the iteration breaks down or does not converge."""
        import torch
        L = lib()
        L.sparlab_system_solve_free_block_device.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
            ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
        nf, nnz = getattr(self, "_block", (-1, 0))  # without a partition the library refuses
        if nf >= 0:
            _require(block_values, "block_values", nnz, at_least=True)
            _require(b, "b", nf, device=block_values.device)
            if x is not None:
                _require(x, "x", nf, device=block_values.device)
        if x is None:
            x = torch.empty(max(b.shape[0], 1), dtype=torch.float64, device=b.device)[:b.shape[0]]
        its, rel = ctypes.c_int64(0), ctypes.c_double(0.0)
        s = stream if stream is not None else torch.cuda.current_stream(b.device)
        code = L.sparlab_system_solve_free_block_device(self.h, ctypes.c_void_p(block_values.data_ptr()),
                                                        int(bool(symmetric)), ctypes.c_void_p(b.data_ptr()),
                                                        ctypes.c_void_p(x.data_ptr()),
                                                        float(tolerance), int(max_iterations), ctypes.byref(its),
                                                        ctypes.byref(rel), ctypes.c_void_p(s.cuda_stream))
        if code == -4:  # SPARLAB_E_SOLVER
            raise SolverError(lib().sparlab_last_error().decode())
        _check(code)
        return x, its.value, rel.value

    def set_block_preconditioner(self, kind):
        """The preconditioner of solve_free_block_device: "diagonal" (the
        default) or "multigrid" (smoothed-aggregation algebraic multigrid,
        one V-cycle per application; its hierarchy is built at the first
        solve after a partition or new coordinates)."""
        code = {"diagonal": PRECONDITIONER_DIAGONAL, "multigrid": PRECONDITIONER_MULTIGRID}.get(kind, kind)
        lib().sparlab_system_set_block_preconditioner.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _check(lib().sparlab_system_set_block_preconditioner(self.h, int(code)))

    def block_preconditioner_levels(self):
        """The multigrid hierarchy of the partition: (rows, nonzeros) per
        level, empty before a multigrid solve."""
        L = lib()
        L.sparlab_system_block_preconditioner_levels.argtypes = [ctypes.c_void_p] * 4
        count = ctypes.c_int32(0)
        rows = np.zeros(MULTIGRID_MAX_LEVELS, dtype=np.int64)
        nnz = np.zeros(MULTIGRID_MAX_LEVELS, dtype=np.int64)
        _check(L.sparlab_system_block_preconditioner_levels(self.h, ctypes.byref(count), _np_ptr(rows), _np_ptr(nnz)))
        return rows[:count.value].tolist(), nnz[:count.value].tolist()

    def solve_free_block_segments_device(self, block_values, b, segment_rows, x=None, tolerance=1e-10,
                                         max_iterations=20000, symmetric=True, stream=None):
        """The batched form's block solve: the partition's block is
        block-diagonal over the segments (segment_rows: num_segments + 1
        ascending row offsets from 0 to num_free), each solved as
        solve_free_block_device solves its rows alone, all in one set of
        launches per iteration. Returns (x, iterations, relative residuals,
        status) with per-segment numpy arrays (status 0: converged; else the
        plain solve's SolverError code) and never raises for a segment's
        failure."""
        import torch
        L = lib()
        L.sparlab_system_solve_free_block_segments_device.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p]
        nf, nnz = getattr(self, "_block", (-1, 0))
        if nf >= 0:
            _require(block_values, "block_values", nnz, at_least=True)
            _require(b, "b", nf, device=block_values.device)
            if x is not None:
                _require(x, "x", nf, device=block_values.device)
        rows = np.ascontiguousarray(segment_rows, dtype=np.int64)
        nseg = rows.shape[0] - 1
        if nseg < 1:
            raise ValueError("at least one segment")
        if x is None:
            x = torch.empty(max(b.shape[0], 1), dtype=torch.float64, device=b.device)[:b.shape[0]]
        its = np.zeros(nseg, dtype=np.int64)
        rel = np.zeros(nseg, dtype=np.float64)
        status = np.zeros(nseg, dtype=np.int32)
        s = stream if stream is not None else torch.cuda.current_stream(b.device)
        code = L.sparlab_system_solve_free_block_segments_device(
            self.h, ctypes.c_void_p(block_values.data_ptr()), int(bool(symmetric)), nseg, _np_ptr(rows),
            ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()), float(tolerance), int(max_iterations),
            _np_ptr(its), _np_ptr(rel), _np_ptr(status), ctypes.c_void_p(s.cuda_stream))
        if code != -4:  # SPARLAB_E_SOLVER: a segment failed, reported in status
            _check(code)
        return x, its, rel, status

    def stress_device(self, u, stress=None, status=None, stream=None):
        """elastoplastic_stress of every element at displacement u (device
        tensor, 3 per node) from the committed history and modes: (stress
        (elements, STRESS_FIELDS), status (elements,) int32) device tensors."""
        import torch
        L = lib()
        L.sparlab_system_stress_device.argtypes = [ctypes.c_void_p] * 5
        if stress is None:
            stress = torch.empty((self.num_elements, STRESS_FIELDS), dtype=torch.float64, device=u.device)
        if status is None:
            status = torch.empty((self.num_elements,), dtype=torch.int32, device=u.device)
        s = stream if stream is not None else torch.cuda.current_stream(u.device)
        _check(L.sparlab_system_stress_device(self.h, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(stress.data_ptr()),
                                              ctypes.c_void_p(status.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
        return stress, status

    def stress(self, u):
        """The same from a host array; raises SolverError for the lowest
        failing element, as the reference's call throws."""
        import torch
        d_u = torch.as_tensor(np.ascontiguousarray(u, dtype=np.float64).reshape(-1), device="cuda")
        stress, status = self.stress_device(d_u)
        torch.cuda.synchronize()
        status = status.cpu().numpy()
        bad = np.nonzero(status)[0]
        if bad.size:
            e = int(bad[0])
            raise SolverError(f"Hex8 element {e}: {ELEMENT_STATUS.get(int(status[e]), 'unknown')}", e, int(status[e]))
        out = stress.cpu().numpy()
        return {name: out[:, off:off + width] if width > 1 else out[:, off] for name, off, width in STRESS_LAYOUT}

    def element_outputs(self, first=None, count=None):
        """The last evaluation's element forces (E, 24) and tangents (E, 24,
        24), of every element or of elements first .. first + count - 1."""
        if first is None:
            fe = np.zeros((self.num_elements, 24))
            ke = np.zeros((self.num_elements, 24, 24))
            _check(lib().sparlab_system_element_outputs(self.h, _np_ptr(fe), _np_ptr(ke)))
            return fe, ke
        L = lib()
        L.sparlab_system_element_outputs_range.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                                           ctypes.c_void_p, ctypes.c_void_p]
        fe = np.zeros((count, 24))
        ke = np.zeros((count, 24, 24))
        _check(L.sparlab_system_element_outputs_range(self.h, int(first), int(count), _np_ptr(fe), _np_ptr(ke)))
        return fe, ke


class BatchedSystem:
    """The batched form (sparlab_batch_*): many independent problems, each a
    NonlinearSystem of its own, evaluated together as one block-diagonal
    union system whose launches are captured in a HIP graph.

    problems: list of (coords (n_p, 3), conn (e_p, 8) with problem-local node
    indices). `system` is the union's NonlinearSystem (pattern, history,
    commit); problem p's DOFs are [3 node_offset[p], 3 node_offset[p + 1]) and
    its CSR values [value_offset[p], value_offset[p + 1]).
    """

    def __init__(self, problems, material, rule=None, kinematics=FINITE_LOGARITHMIC):
        import torch  # noqa: F401
        L = lib()
        L.sparlab_batch_create.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.POINTER(CMaterial), ctypes.POINTER(CRule),
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.sparlab_batch_destroy.argtypes = [ctypes.c_void_p]
        L.sparlab_batch_system.restype = ctypes.c_void_p
        L.sparlab_batch_system.argtypes = [ctypes.c_void_p]
        L.sparlab_batch_value_offsets.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_batch_evaluate_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.sparlab_batch_set_graphs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.sparlab_batch_result.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 6
        L.sparlab_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        rule = rule or Rule()
        coords = [np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3) for c, _ in problems]
        conns = [np.ascontiguousarray(k, dtype=np.int32).reshape(-1, 8) for _, k in problems]
        self.num_problems = len(problems)
        self.node_offset = np.concatenate([[0], np.cumsum([c.shape[0] for c in coords])]).astype(np.int64)
        self.element_offset = np.concatenate([[0], np.cumsum([k.shape[0] for k in conns])]).astype(np.int64)
        ucoords = np.ascontiguousarray(np.concatenate(coords))
        lconn = np.ascontiguousarray(np.concatenate(conns))
        h = ctypes.c_void_p()
        _check(L.sparlab_batch_create(self.num_problems, _np_ptr(self.node_offset), _np_ptr(self.element_offset),
                                      _np_ptr(ucoords), _np_ptr(lconn), ctypes.byref(material.c),
                                      ctypes.byref(rule.c), kinematics, ctypes.byref(h)))
        self.h = h
        self._destroy = L.sparlab_batch_destroy
        uconn = lconn + np.repeat(self.node_offset[:-1], [k.shape[0] for k in conns])[:, None].astype(np.int32)
        self.system = NonlinearSystem(ucoords, uconn, material, rule, kinematics,
                                      _handle=ctypes.c_void_p(L.sparlab_batch_system(h)))
        self.value_offset = np.zeros(self.num_problems + 1, dtype=np.int64)
        _check(L.sparlab_batch_value_offsets(h, _np_ptr(self.value_offset)))

    def __del__(self):
        h = getattr(self, "h", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy is not None:
            sys_ = getattr(self, "system", None)
            if sys_ is not None:
                sys_.h = None
            destroy(h)
            self.h = None

    def set_graphs(self, enabled):
        _check(lib().sparlab_batch_set_graphs(self.h, int(enabled)))

    def autotune(self, u, stream=None):
        """Picks the local solve's elements per wavefront on this batch at u (device tensor)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream(u.device)
        chosen = ctypes.c_int(0)
        _check(lib().sparlab_batch_autotune(self.h, ctypes.c_void_p(u.data_ptr()), ctypes.byref(chosen),
                                            ctypes.c_void_p(s.cuda_stream)))
        return chosen.value

    def evaluate_device(self, u, fint, values=None, want_tangent=True, column_major=False, stream=None):
        """Every problem's evaluation (device tensors over the union; the same
        tensors every call, as the captured graph keeps their addresses)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream(u.device)
        _check(lib().sparlab_batch_evaluate_device(self.h, ctypes.c_void_p(u.data_ptr()), int(want_tangent),
                                                   int(column_major), ctypes.c_void_p(fint.data_ptr()),
                                                   None if values is None else ctypes.c_void_p(values.data_ptr()),
                                                   ctypes.c_void_p(s.cuda_stream)))

    def result(self, stream=None):
        """Per-problem scalars of the last evaluation: dict of arrays (energy,
        gross, yielding, failed, first_failed)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        n = self.num_problems
        out = dict(energy=np.zeros(n), gross=np.zeros(n), yielding=np.zeros(n, np.int64),
                   failed=np.zeros(n, np.int64), first_failed=np.zeros(n, np.int64))
        rc = lib().sparlab_batch_result(self.h, _np_ptr(out["energy"]), _np_ptr(out["gross"]),
                                        _np_ptr(out["yielding"]), _np_ptr(out["failed"]),
                                        _np_ptr(out["first_failed"]), ctypes.c_void_p(s.cuda_stream))
        if rc not in (0, -4):
            _check(rc)
        return out

    def commit(self):
        self.system.commit()

    def free_segments(self, free):
        """The block rows of each problem when the union's free DOFs are
        `free` (ascending union DOF numbers): num_problems + 1 offsets, for
        system.solve_free_block_segments_device."""
        free = np.asarray(free, dtype=np.int64)
        return np.searchsorted(free, 3 * self.node_offset).astype(np.int64)


PRECONDITIONER_DIAGONAL = 0
PRECONDITIONER_MULTIGRID = 1
MULTIGRID_MAX_LEVELS = 12

VJP_PARAMETERS = 11
PARAMETER_NAMES = ("r0", "r45", "r90", "C1", "gamma1", "C2", "gamma2", "C3", "gamma3", "C4", "gamma4")


def material_parameters(material):
    """The differentiable parameters of a material (PARAMETER_NAMES order)."""
    c = material.c
    out = np.zeros(VJP_PARAMETERS)
    out[0:3] = list(c.hill_r)
    for i in range(c.num_backstresses):
        out[3 + 2 * i] = c.backstress_modulus[i]
        out[4 + 2 * i] = c.backstress_recovery[i]
    return out


def with_parameters(material, params):
    """A copy of `material` with the differentiable parameters replaced."""
    m = Material.__new__(Material)
    m.c = CMaterial()
    ctypes.pointer(m.c)[0] = material.c
    for i in range(3):
        m.c.hill_r[i] = float(params[i])
    for i in range(m.c.num_backstresses):
        m.c.backstress_modulus[i] = float(params[3 + 2 * i])
        m.c.backstress_recovery[i] = float(params[4 + 2 * i])
    _check(lib().sparlab_material_finalize(ctypes.byref(m.c)))
    return m


class InternalForce:
    """f_int(X, theta) of a NonlinearSystem at a fixed displacement and
    committed history as a torch.autograd.Function: forward evaluates the
    assembled internal force for the nodal reference coordinates X
    (num_nodes x 3) and the material parameters theta (11, PARAMETER_NAMES);
    backward returns v^T df/dX and v^T df/dtheta from the adjoint kernel.

    Usage: f = InternalForce.apply(X, theta, system, u) with X, theta float64
    tensors (requires_grad as needed) and u the displacement (device tensor).
    """

    @staticmethod
    def _function():
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, X, theta, system, u):
                system.set_coords(X.detach().cpu().numpy())
                system.set_material(with_parameters(system.material, theta.detach().cpu().numpy()))
                fint = torch.zeros(3 * system.num_nodes, dtype=torch.float64, device=u.device)
                system.evaluate_device(u, fint, None, want_tangent=False)
                system.result()
                ctx.system, ctx.u = system, u
                ctx.devices = (X.device, theta.device)
                return fint

            @staticmethod
            def backward(ctx, grad):
                g = ctx.system.vjp_device(ctx.u, grad.contiguous(), coords=True, parameters=True)
                gx = g["coords"].reshape(-1, 3).to(ctx.devices[0])
                return gx, g["parameters"].to(ctx.devices[1]), None, None

        return _F

    @classmethod
    def apply(cls, X, theta, system, u):
        return cls._function().apply(X, theta, system, u)


def make_tool(shape="sphere", centre=(0, 0, 0), centre0=None, radius=0.0, normal=(0, 0, 1), axis=(0, 0, 1),
              friction=0.0, tangential_ratio=1.0):
    t = CTool()
    t.shape = {"plane": 0, "sphere": 1, "cylinder": 2}[shape]
    centre0 = centre if centre0 is None else centre0
    for i in range(3):
        t.centre[i], t.centre0[i], t.normal[i], t.axis[i] = centre[i], centre0[i], normal[i], axis[i]
    t.radius, t.friction, t.tangential_ratio = radius, friction, tangential_ratio
    return t


def contact_buffers(num_slaves, device):
    """Device outputs and work space of one contact evaluation of
    `num_slaves` slave nodes (reusable across evaluations)."""
    import torch
    f64 = dict(dtype=torch.float64, device=device)
    i32 = dict(dtype=torch.int32, device=device)
    ns = num_slaves
    return dict(blocks=torch.zeros((ns, 3, 3), **f64), active=torch.zeros(ns, **i32),
                trial_present=torch.zeros(ns, **i32), trial_force=torch.zeros((ns, 3), **f64),
                trial_relative=torch.zeros((ns, 3), **f64), work=torch.zeros((max(ns, 1), 8), **f64),
                work_status=torch.zeros(max(ns, 1), **i32), scalars=torch.zeros(10, **f64))


def contact_evaluate_device(tool, coords, u, start, nodes, kappa_area, history=None, residual=None,
                            want_tangent=True, friction_symmetric=False, stream=None, buffers=None,
                            read_scalars=True):
    """ToolContact::evaluate for one tool on device tensors. Returns a dict of
    device tensors (residual, blocks, active, trial history) and, with
    `read_scalars`, the host scalars (which synchronises the stream).
    `buffers` (contact_buffers) reuses the outputs of an earlier call."""
    import torch
    dev = coords.device
    ns = nodes.shape[0]
    nn = coords.shape[0] if coords.dim() == 2 else coords.shape[0] // 3
    if residual is None:
        residual = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    bufs = buffers if buffers is not None else contact_buffers(ns, dev)
    out = dict(residual=residual, **{k: bufs[k] for k in ("blocks", "active", "trial_present", "trial_force",
                                                          "trial_relative")})
    work, work_status, scalars = bufs["work"], bufs["work_status"], bufs["scalars"]
    hp, hf, hr = (None, None, None) if history is None else history
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    L = lib()
    p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    L.sparlab_contact_evaluate_device.argtypes = [ctypes.POINTER(CTool), ctypes.c_int, ctypes.c_int] + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int64] + [ctypes.c_void_p] * 14 + [ctypes.c_void_p]
    _check(L.sparlab_contact_evaluate_device(ctypes.byref(tool), int(friction_symmetric), int(want_tangent),
                                             p(coords), p(u), p(start), ns, p(nodes), p(kappa_area), p(hp), p(hf),
                                             p(hr), p(residual), p(out["blocks"]), p(out["active"]),
                                             p(out["trial_present"]), p(out["trial_force"]),
                                             p(out["trial_relative"]), p(work), p(work_status), p(scalars),
                                             ctypes.c_void_p(s.cuda_stream)))
    if not read_scalars:
        return out
    sc = scalars.cpu().numpy()
    out.update(gross=sc[0], round_off=sc[1], resultant=sc[2:5], max_penetration=sc[5], n_slipping=int(sc[6]),
               symmetric=sc[7] == 0, n_active=int(sc[8]), through=int(sc[9]))
    if out["through"] > 0:
        raise SolverError("a node lies less than half the radius from the tool's centre or axis: "
                          "the tool is through the surface")
    return out


def contact_candidates_device(tool, coords, u0, nodes, margin, stream=None, buffers=None):
    """ToolContact::begin_increment for one tool (centre0 = c(t0), centre =
    c(t1)) on device tensors: (candidates int32 (num_slaves capacity), count
    int32[1]); the first count entries are the candidate slave indices in
    order. `buffers` (candidates, count, work) reuses an earlier call's."""
    import torch
    dev = coords.device
    ns = nodes.shape[0]
    if buffers is None:
        buffers = (torch.zeros(ns, dtype=torch.int32, device=dev), torch.zeros(1, dtype=torch.int32, device=dev),
                   torch.zeros(ns + ns // 256 + 1, dtype=torch.int32, device=dev))
    candidates, count, work = buffers
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    L = lib()
    L.sparlab_contact_candidates_device.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p,
                                                                            ctypes.c_double] + [ctypes.c_void_p] * 4
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _check(L.sparlab_contact_candidates_device(ctypes.byref(tool), p(coords), p(u0), ns, p(nodes), float(margin),
                                               p(candidates), p(count), p(work), ctypes.c_void_p(s.cuda_stream)))
    return candidates, count


def contact_anticipate_device(tool, coords, u, du, nodes, kappa_area, residual, stream=None, buffers=None):
    """ToolContact::anticipate for one tool over the listed slaves on device
    tensors: adds into `residual` (3 per node) and returns (blocks (n, 3, 3),
    active int32, added int32[1]); `buffers` reuses an earlier call's."""
    import torch
    dev = coords.device
    ns = nodes.shape[0]
    if buffers is None:
        buffers = (torch.zeros((ns, 3, 3), dtype=torch.float64, device=dev), torch.zeros(ns, dtype=torch.int32, device=dev),
                   torch.zeros(1, dtype=torch.int32, device=dev))
    blocks, active, added = buffers
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    L = lib()
    L.sparlab_contact_anticipate_device.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 7
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _check(L.sparlab_contact_anticipate_device(ctypes.byref(tool), p(coords), p(u), p(du), ns, p(nodes), p(kappa_area),
                                               p(residual), p(blocks), p(active), p(added),
                                               ctypes.c_void_p(s.cuda_stream)))
    return blocks, active, added


def add_node_blocks_device(blocks, active, first_value, row_len, values, column_major=False, stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream(values.device)
    L = lib()
    L.sparlab_add_node_blocks_device.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_int,
                                                                         ctypes.c_void_p, ctypes.c_void_p]
    _check(L.sparlab_add_node_blocks_device(ctypes.c_void_p(blocks.data_ptr()), ctypes.c_void_p(active.data_ptr()),
                                            ctypes.c_void_p(first_value.data_ptr()),
                                            ctypes.c_void_p(row_len.data_ptr()), active.shape[0], int(column_major),
                                            ctypes.c_void_p(values.data_ptr()), ctypes.c_void_p(s.cuda_stream)))


def surface_raycast_device(surface, grid, stream=None):
    """The released part's height map before the nearest fill (the
    reference's formed_surface without _fill_nearest): surface, a float64
    device tensor (ny_s, nx_s, 3) of the deformed tool-side surface's
    structured nodes; grid = (x0, y0, nx, ny, h). Returns (z, hit) device
    tensors (ny, nx): the highest hit per node (-inf where none) and the hit
    flags. Raises RuntimeError for a surface that folds over in plan view."""
    import torch
    L = lib()
    L.sparlab_surface_raycast_device.argtypes = [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
        ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
        ctypes.c_void_p]
    if surface.dtype != torch.float64 or surface.dim() != 3 or surface.shape[2] != 3 or not surface.is_contiguous():
        raise ValueError("surface must be a contiguous float64 (ny_s, nx_s, 3) tensor")
    x0, y0, nx, ny, h = grid
    z = torch.empty((int(ny), int(nx)), dtype=torch.float64, device=surface.device)
    hit = torch.empty((int(ny), int(nx)), dtype=torch.uint8, device=surface.device)
    folded = ctypes.c_int64(0)
    s = stream if stream is not None else torch.cuda.current_stream(surface.device)
    _check(L.sparlab_surface_raycast_device(ctypes.c_void_p(surface.data_ptr()), surface.shape[1], surface.shape[0],
                                            float(x0), float(y0), int(nx), int(ny), float(h),
                                            ctypes.c_void_p(z.data_ptr()), ctypes.c_void_p(hit.data_ptr()),
                                            ctypes.byref(folded), ctypes.c_void_p(s.cuda_stream)))
    return z, hit.bool()
