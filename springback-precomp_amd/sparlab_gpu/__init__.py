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
        ("fail_value", ctypes.c_void_p),
    ]


_lib = None


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run springback-precomp_amd/build.py "
                               "(or __graft_entry__.build()) first")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.sparlab_last_error.restype = ctypes.c_char_p
        _lib.sparlab_version.restype = ctypes.c_char_p
        _lib.sparlab_material_finalize.argtypes = [ctypes.POINTER(CMaterial)]
        _lib.sparlab_rule_points.argtypes = [ctypes.POINTER(CRule)]
        _lib.sparlab_state_fields.argtypes = [ctypes.POINTER(CMaterial)]
        _lib.sparlab_elements_evaluate_device.argtypes = [
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
            _ptr(self.fail_value))


def evaluate_elements(material, rule, batch, kinematics=FINITE_LOGARITHMIC, want_tangent=True, stream=None):
    """Queues the evaluation of every element of `batch` (see ElementBatch)."""
    import torch
    if want_tangent and batch.tangent is None:
        raise ValueError("the batch was allocated without a tangent buffer")
    s = stream if stream is not None else torch.cuda.current_stream(batch.coords.device)
    _check(lib().sparlab_elements_evaluate_device(ctypes.byref(material.c), ctypes.byref(rule.c), kinematics,
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
