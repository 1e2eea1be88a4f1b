"""Synthetic inputs of the element-kernel workloads (numpy, seeded).

The element microbenchmark: independent incompatible-mode Hex8 elements, each
the unit cube [0, 1]^3 (VTK order) with every node jittered by U(-0.1, 0.1)
per coordinate; displacements N(0, 1e-3) per component; AA5754-O-like
Hill48 (r0 0.7, r45 0.6, r90 0.8; RD = x, ND = z) with two Armstrong-Frederick
backstresses (C1 = 6000 MPa, gamma1 = 60; C2 = 1500 MPa, gamma2 = 10,
exponential integration) and Voce isotropic hardening
sigma_y = 95 + 150 (1 - exp(-10 abar)) MPa; E = 70 GPa, nu = 0.33; SI units.
The committed history of every point lies inside its yield surface: a random
accumulated plastic strain abar, random deviatoric backstresses of up to half
their saturation (2/3) C_i / gamma_i, and a plastic strain that puts the
relative stress of the undeformed point at a fraction rho ~ U(0, rho_max) of
the yield stress along a random deviatoric direction. The committed mode
parameters are zero.
"""
import numpy as np

from . import Material

E_MODULUS = 70.0e9
POISSON = 0.33
SIGMA_Y0 = 95.0e6
VOCE_Q = 150.0e6
VOCE_RATE = 10.0
HILL = dict(r0=0.7, r45=0.6, r90=0.8, rolling_direction=(1.0, 0.0, 0.0), sheet_normal=(0.0, 0.0, 1.0))
BACKSTRESSES = ((6000.0e6, 60.0), (1500.0e6, 10.0))

# rho_max and the abar range put about 40 % of the elements past yield at
# one point at least (measured and reported by the benchmark).
RHO_MAX = 0.55
ABAR_MAX = 1.0

UNIT_CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)


def microbench_material():
    return Material(E_MODULUS, POISSON, yield_stress=SIGMA_Y0, saturation_stress=VOCE_Q,
                    saturation_rate=VOCE_RATE, hill=HILL, backstresses=BACKSTRESSES)


def _random_deviators(rng, n):
    """n random unit deviatoric tensors (tensorial Voigt, |t| = 1)."""
    v = rng.normal(size=(n, 6))
    tr = v[:, 0:3].sum(axis=1, keepdims=True) / 3.0
    v[:, 0:3] -= tr
    nrm = np.sqrt((v[:, 0:3] ** 2).sum(axis=1) + 2.0 * (v[:, 3:6] ** 2).sum(axis=1))
    return v / nrm[:, None]


def microbench(ne, seed=20240607, points=8, rho_max=RHO_MAX, abar_max=ABAR_MAX):
    """Inputs of `ne` independent elements.

    Returns dict with coords (8 ne, 3), conn (ne, 8) int32, u (8 ne, 3) and the
    committed history `states` as (ne, points, 39) PlasticState records
    (plastic strain, beta, alpha_1..alpha_4, abar, thickness strain, loading).
    """
    rng = np.random.default_rng(seed)
    coords = (UNIT_CUBE[None, :, :] + rng.uniform(-0.1, 0.1, size=(ne, 8, 3))).reshape(8 * ne, 3)
    conn = np.arange(8 * ne, dtype=np.int32).reshape(ne, 8)
    u = rng.normal(0.0, 1.0e-3, size=(8 * ne, 3))
    n = ne * points
    g = E_MODULUS / (2.0 * (1.0 + POISSON))
    p = microbench_material().yield_matrix
    abar = rng.uniform(0.0, abar_max, size=n)
    sy = SIGMA_Y0 + VOCE_Q * -np.expm1(-VOCE_RATE * abar)
    alphas = []
    for c, gam in BACKSTRESSES:
        r = rng.uniform(0.0, 0.5, size=n)
        alphas.append(_random_deviators(rng, n) * (r * (2.0 / 3.0) * c / gam)[:, None])
    direction = _random_deviators(rng, n)
    eq = np.sqrt(np.einsum("ni,ij,nj->n", direction, p, direction))
    rho = rng.uniform(0.0, rho_max, size=n)
    xi_pre = direction * (rho * sy / eq)[:, None]
    beta = alphas[0] + alphas[1]
    eps_t = -(xi_pre + beta) / (2.0 * g)          # tensorial, deviatoric
    eps = eps_t.copy()
    eps[:, 3:6] *= 2.0                               # engineering shears
    states = np.zeros((n, 39))
    states[:, 0:6] = eps
    states[:, 6:12] = beta
    states[:, 12:18] = alphas[0]
    states[:, 18:24] = alphas[1]
    states[:, 36] = abar
    states[:, 38] = (rho > 0.9 * rho_max).astype(np.float64)
    return dict(coords=coords, conn=conn, u=u, states=states.reshape(ne, points, 39))


def element_arrays(w):
    """Per-element views for the oracle: x0 (ne, 8, 3) and ue (ne, 24)."""
    conn = w["conn"]
    return w["coords"][conn], w["u"][conn].reshape(conn.shape[0], 24)


def structured_hex_mesh(nx, ny, nz, lx, ly, lz, x0=0.0, y0=0.0, z0=0.0):
    """The reference's make_structured_hex_mesh (src/mesh/StructuredMesh.cpp:62-113):
    node n = k nnx nny + j nnx + i, elements x fastest, VTK corner order."""
    nnx, nny, nnz = nx + 1, ny + 1, nz + 1
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    k, j, i = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    coords = np.stack([x0 + i.ravel() * dx, y0 + j.ravel() * dy, z0 + k.ravel() * dz], axis=1).astype(np.float64)
    layer = nnx * nny
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (ek * layer + ej * nnx + ei).ravel()
    conn = np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx,
                     n0 + layer, n0 + layer + 1, n0 + layer + nnx + 1, n0 + layer + nnx], axis=1).astype(np.int32)
    return coords, conn


def assembly_workload(nx=100, ny=100, nz=4, seed=99, contact_nodes=10000):
    """The coloured-assembly + contact workload: a structured nx x ny x nz Hex8
    plate with random FP64 element forces and non-symmetric tangents (seeded),
    and penalty contact of top-surface nodes with a sphere R = 5 mm at random
    seeded offsets (SI units: a 0.1 x 0.1 x 0.004 m plate)."""
    rng = np.random.default_rng(seed)
    coords, conn = structured_hex_mesh(nx, ny, nz, 0.1, 0.1, 0.004, -0.05, -0.05, -0.004)
    ne = conn.shape[0]
    fe = rng.normal(size=(ne, 24))
    ke = rng.normal(size=(ne, 24, 24))
    nnx, nny = nx + 1, ny + 1
    top = np.arange(nz * nnx * nny, (nz + 1) * nnx * nny, dtype=np.int32)
    nodes = np.sort(rng.choice(top, size=min(contact_nodes, top.size), replace=False)).astype(np.int32)
    return dict(coords=coords, conn=conn, fe=fe, ke=ke, contact_nodes=nodes)


FAMILIES = ("dome", "truncated_cone", "elliptic_cone", "pyramid")


def _family_depth(family, x, y, half, depth, shape):
    """Downward deflection profile of a part family over the footprint
    (centred at 0, half-width `half`), scaled to `depth` at the centre."""
    if family == "dome":
        r = np.hypot(x, y) / (0.9 * half)
        return -depth * np.clip(1.0 - r * r, 0.0, None)
    if family == "truncated_cone":
        r = np.hypot(x, y)
        top = shape * half
        return -depth * np.clip((0.9 * half - r) / (0.9 * half - top), 0.0, 1.0)
    if family == "elliptic_cone":
        r = np.hypot(x / (0.9 * half), y / (0.9 * half * (0.55 + 0.4 * shape)))
        return -depth * np.clip(1.0 - r, 0.0, 1.0)
    r = np.maximum(np.abs(x), np.abs(y)) / (0.9 * half)
    return -depth * np.clip(1.0 - r, 0.0, 1.0)


def batched_problems(count=64, seed=20241203, min_elements=2000, max_elements=40000, layers=2, step=0.03):
    """The batched form's workload: `count` sheet-forming problems from a
    scrambled Sobol draw (scipy.stats.qmc, seeded) over footprint
    15-20 mm, depth 2-4 mm, a family shape parameter and the mesh size, a
    quarter of them in each part family (dome, truncated cone, elliptic cone,
    pyramid); 1 mm sheets meshed `layers` Hex8 through the thickness with
    element counts spread over [min_elements, max_elements]. Each problem's
    displacement is one forming increment, `step` times its family's
    deflection profile (every node of a column moves alike), from a virgin
    history. Returns a list of dicts: coords (n, 3), conn (e, 8) local, u (n, 3),
    family, footprint, depth."""
    from scipy.stats import qmc
    pts = qmc.Sobol(d=4, scramble=True, seed=seed).random(count)
    out = []
    for i in range(count):
        family = FAMILIES[i % len(FAMILIES)]
        footprint = 0.015 + 0.005 * pts[i, 0]
        depth = 0.002 + 0.002 * pts[i, 1]
        shape = pts[i, 2]
        target = min_elements + (max_elements - min_elements) * pts[i, 3]
        n = max(2, int(round(np.sqrt(target / layers))))
        coords, conn = structured_hex_mesh(n, n, layers, footprint, footprint, 0.001,
                                           -footprint / 2, -footprint / 2, -0.001)
        w = _family_depth(family, coords[:, 0], coords[:, 1], footprint / 2, depth, shape)
        u = np.zeros_like(coords)
        u[:, 2] = step * w
        out.append(dict(coords=coords, conn=conn, u=u, family=family, footprint=footprint, depth=depth))
    return out


def differentiation_workload(nx=200, ny=250, nz=4, lx=0.020, ly=0.025, thickness=0.001, depth=0.003,
                             turns=3.0, progress=0.5, step=0.004):
    """The differentiation workload: an nx x ny x nz Hex8 sheet (200 x 250 x 4
    = 200 000 elements over 20 x 25 x 1 mm) mid-way through a single-point
    incremental forming path, a spiral of `turns` turns from the rim towards
    the centre. The committed displacement u1 is the shape at `progress` of
    the path (a cone-like pocket of `depth`, deepening along the spiral, plus
    the tool's local indentation); u2 is the next increment (`step` further).
    Returns coords, conn, u1, u2 (node-major, SI units)."""
    coords, conn = structured_hex_mesh(nx, ny, nz, lx, ly, thickness, -lx / 2, -ly / 2, -thickness)
    x, y = coords[:, 0], coords[:, 1]
    radius = 0.45 * min(lx, ly)

    def shape(t):
        # tool at radius r(t) (rim -> centre), angle 2 pi turns t; pocket depth grows with t
        rt = radius * (1.0 - 0.8 * t)
        ang = 2.0 * np.pi * turns * t
        tx, ty = rt * np.cos(ang), rt * np.sin(ang)
        r = np.hypot(x, y)
        pocket = -depth * t * np.clip(1.0 - r / radius, 0.0, 1.0)
        dent = -0.05 * depth * np.exp(-((x - tx) ** 2 + (y - ty) ** 2) / (0.1 * radius) ** 2)
        u = np.zeros_like(coords)
        u[:, 2] = pocket + dent
        return u

    u1 = shape(progress)
    u2 = shape(progress + step)
    return coords, conn, u1.reshape(-1), u2.reshape(-1)
