"""Benchmark: FP64 incompatible-mode Hex8 f_int + K_t element evaluations/s.

Workload (one GPU): the element microbenchmark -- 1,000,000 independent
incompatible-mode Hex8 elements (jittered unit cubes, displacements
N(0, 1e-3)), logarithmic strain, Hill48 (r 0.7/0.6/0.8) with two Chaboche
backstresses and Voce hardening, committed history inside the yield surface
(sparlab_gpu/workloads.py). One step evaluates every element: the local Newton
iteration of the modes, the condensed force f* (24) and tangent K* (24 x 24),
the points' trial history. Inputs are resident in HBM before timing.

Roofline: the element evaluation is FP64-compute-bound. Its algorithmic
FLOPs per element are counted by the oracle's op-counting instantiation
(add/sub/mul/div/sqrt and each libm call = 1) on the first 1000 elements of the
same inputs, split at the end of the local iteration. The dominant kernel is
the local solve: achieved = its FLOPs per element x elements / its HIP-event
time over the timed steps (events on the launch stream between the two
kernels); the condensation kernel and the whole evaluation are reported beside.

CPU baseline: the oracle's element loop (OpenMP schedule(static), one element
per call, the reference's NonlinearSystem.hpp:208-237 structure) on the same
workload (all of it unless --cpu-sample bounds it), on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "springback-precomp_amd"))

METRIC = "Hex8 Fint+Kt element evals/s (FP64, Hill48/Chaboche) batched vs plain; % roofline"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_1M_microbench_and_assembly.md")


def measured_traffic(kernel, elements):
    """HBM bytes per launch of `kernel` on the `elements`-element workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by
    the copy calibration recorded there), or None when it holds no figure for
    this workload."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "TRAFFIC" and parts[1] == kernel and \
                        parts[2] == f"elements={elements}" and parts[3].startswith("bytes="):
                    return float(parts[3][len("bytes="):])
    except OSError:
        return None
    return None


def cpu_baseline_leg(workload, sample, threads):
    """The oracle on the host: evals/s on `threads` cores over the first
    `sample` elements, and the algorithmic FLOPs per element (op-counting
    oracle, split at the end of the local iteration) on its first 1000."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from parity import oracle_material
    from sparlab_gpu import workloads as W

    m = W.microbench_material()
    om = oracle_material(m)
    x0 = workload["coords"][workload["conn"][:sample]]
    ue = workload["u"][workload["conn"][:sample]].reshape(sample, 24)
    st = workload["states"][:sample]
    t0 = time.perf_counter()
    r = O.element_batch(x0, ue, om, st, kinematics=2, want_tangent=True, threads=threads)
    dt = time.perf_counter() - t0
    nflop = min(sample, 1000)
    parts = np.array([O.element_flops_split(x0[i], ue[i], om, states=st[i], kinematics=2) for i in range(nflop)],
                     dtype=np.float64)
    return dict(evals_per_s=sample / dt, seconds=dt, failures=int((r["status"] != 0).sum()),
                flops_per_elem=float(parts[:, 0].mean()), flops_local=float(parts[:, 1].mean()),
                flops_condense=float(parts[:, 2].mean()), flops_sample=nflop,
                plastic_fraction=float((r["scalars"][:, 1] > 0).mean()))


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_single_core_leg(workload, sample):
    """The oracle's element loop on one pinned host core over the first
    `sample` elements: microseconds per evaluation, the figure comparable
    with a single-threaded element loop."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from parity import oracle_material
    from sparlab_gpu import workloads as W

    om = oracle_material(W.microbench_material())
    x0 = workload["coords"][workload["conn"][:sample]]
    ue = workload["u"][workload["conn"][:sample]].reshape(sample, 24)
    st = workload["states"][:sample]
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[0]})
    try:
        t0 = time.perf_counter()
        O.element_batch(x0, ue, om, st, kinematics=2, want_tangent=True, threads=1)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, set(cpus))
    return dict(evals_per_s=sample / dt, us_per_eval=dt / sample * 1e6, seconds=dt, core=cpus[0])


def box_rule_leg(G, W, torch, dev, steps, ne=262144):
    """The element microbenchmark with the 2 x 2 x 5 box rule of the
    reference's fine-mesh decks (20 points per element, 32-lane teams):
    `ne` elements, device evals/s and the algorithmic FLOP rate (op-counting
    oracle on 200 elements), beside the oracle on one host core (the figure
    comparable with the reference's 265 us per 20-point log-strain evaluation,
    docs/forming.md:470)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from parity import oracle_material
    box = (2, 5, 2)
    rule = G.Rule(*box)
    m = W.microbench_material()
    w = W.microbench(ne, seed=20241204, points=rule.points)
    st = G.states_from_reference(w["states"], m)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    b = G.ElementBatch(t(w["coords"]), t(w["conn"], torch.int32), t(w["u"]), t(st), None, rule.points, True)
    del st
    stream = torch.cuda.current_stream(dev)
    G.evaluate_elements(m, rule, b, G.FINITE_LOGARITHMIC, True, stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for k in range(steps):
        ev[k][0].record(stream)
        G.evaluate_elements(m, rule, b, G.FINITE_LOGARITHMIC, True, stream, stages=G.LOCAL_SOLVE)
        ev[k][1].record(stream)
        G.evaluate_elements(m, rule, b, G.FINITE_LOGARITHMIC, True, stream, stages=G.CONDENSATION)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    ms_local = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    ms_cond = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    ms = ms_local + ms_cond
    failures = int((b.status != 0).sum().item())
    plastic = float((b.yielding > 0).double().mean().item())
    om = oracle_material(m)
    x0 = w["coords"][w["conn"][:2000]]
    ue = w["u"][w["conn"][:2000]].reshape(2000, 24)
    sts = w["states"][:2000]
    parts = np.array([O.element_flops_split(x0[i], ue[i], om, states=sts[i], rule=box, kinematics=2)
                      for i in range(200)], dtype=np.float64)
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[0]})
    try:
        t0 = time.perf_counter()
        O.element_batch(x0, ue, om, sts, rule=box, kinematics=2, want_tangent=True, threads=1)
        one = (time.perf_counter() - t0) / 2000 * 1e6
    finally:
        os.sched_setaffinity(0, set(cpus))
    flops = float(parts[:, 0].mean())
    return {"workload": f"element microbench with the 2 x 2 x 5 box rule: {ne} elements, 20 points each "
                        "(32-lane teams), log strain, Hill48 + 2 Chaboche + Voce",
            "elements": ne, "integration_points": rule.points, "ms": ms, "local_solve_ms": ms_local,
            "condensation_ms": ms_cond, "evals_per_s": ne / (ms * 1e-3), "failures": failures,
            "plastic_fraction": plastic, "flops_per_elem": flops, "flops_sample": 200,
            "achieved_TFLOPs": flops * ne / (ms * 1e-3) / 1e12,
            "frac": flops * ne / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "cpu_single_core_us_per_eval": one}


def batched_leg(G, W, torch, dev, steps):
    """The batched form beside the plain drop-in on the 64-problem forming
    workload (workloads.batched_problems: 4 part families, 2 000-40 000 Hex8
    each, one forming increment from a virgin history): the plain drop-in
    evaluates each problem's own system in turn, the batched form evaluates
    the union in one captured launch sequence. Each step is one Newton
    iteration's evaluation of all problems (element kernels, force gather,
    CSR assembly, scalars)."""
    probs = W.batched_problems()
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1), device=dev)
    ne = int(sum(p["conn"].shape[0] for p in probs))
    stream = torch.cuda.current_stream(dev)
    # plain: one system per problem, evaluated one after another
    plain = []
    for p in probs:
        s = G.NonlinearSystem(p["coords"], p["conn"], m, rule)
        u = t(p["u"])
        plain.append((s, u, torch.zeros_like(u), torch.zeros(s.nnz, dtype=torch.float64, device=dev)))
    for s, u, f, v in plain:
        s.evaluate_device(u, f, v, stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        for s, u, f, v in plain:
            s.evaluate_device(u, f, v, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_plain = e0.elapsed_time(e1) / steps
    plain_ok = True
    for s, _, _, _ in plain:
        try:
            s.result(stream)
        except G.SolverError:
            plain_ok = False
    del plain
    # batched: the union, autotuned, graph-captured
    b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
    u = t(np.concatenate([p["u"] for p in probs]))
    fint = torch.zeros_like(u)
    values = torch.zeros(b.system.nnz, dtype=torch.float64, device=dev)
    chunk = b.autotune(u, stream)
    b.evaluate_device(u, fint, values, stream=stream)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        b.evaluate_device(u, fint, values, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_batched = e0.elapsed_time(e1) / steps
    sc = b.result(stream)
    return {"workload": "64 forming problems (dome, truncated cone, elliptic cone, pyramid; Sobol over footprint "
                        "15-20 mm, depth 2-4 mm), 1 mm sheet, 2 Hex8 layers, 2 000-40 000 elements each, one "
                        "forming increment from a virgin history, Hill48 + 2 Chaboche + Voce, log strain",
            "problems": len(probs), "elements": ne, "steps": steps,
            "plain_ms": ms_plain, "batched_ms": ms_batched,
            "plain_evals_per_s": ne / (ms_plain * 1e-3), "batched_evals_per_s": ne / (ms_batched * 1e-3),
            "speedup": ms_plain / ms_batched, "elements_per_wavefront": chunk,
            "failures": int(sc["failed"].sum()) + (0 if plain_ok else 1),
            "yielding_points": int(sc["yielding"].sum())}


def rim_free_dofs(coords):
    """Every DOF but those of the rim's nodes (a sheet clamped at its edge)."""
    x, y = coords[:, 0], coords[:, 1]
    rim = np.isclose(x, x.min()) | np.isclose(x, x.max()) | np.isclose(y, y.min()) | np.isclose(y, y.max())
    return np.nonzero(np.repeat(~rim, 3))[0].astype(np.int32)


def newton_leg(G, W, torch, dev, tolerance=1e-8, max_iterations=20000, precond="both", form="both"):
    """One Newton iteration of all 64 forming problems (the batched form's
    workload, rims clamped): evaluation (f_int + row-major K_t), the free-free
    block, r = f_int on the free DOFs, the block solve K_ff du = -r (BiCGStab:
    the Hill48 + Chaboche tangent is not symmetric) to `tolerance`, u += du.
    Plain: each problem's own system, one after another. Batched: the union
    evaluated in its captured graph and every problem's solve in one
    segmented solve. Each problem's du must equal its plain solve's bit for
    bit (x, iterations and residual: the same sums in the same order). Run
    with the diagonal preconditioner and with the multigrid V-cycle (each
    problem's own hierarchy; the batched form sets up and applies all of
    them in shared launches); the first iteration of each form, which builds
    the hierarchies on the host once per partition, is timed apart. precond
    and form restrict the leg to one preconditioner or one form (for a
    profile of it alone)."""
    probs = W.batched_problems()
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    stream = torch.cuda.current_stream(dev)
    frees = [rim_free_dofs(p["coords"]) for p in probs]

    def run(precond):
        out = {}
        plain_x = plain_its = None
        if form in ("both", "plain"):
            plain_x, plain_its = run_plain(precond, out)
        if form in ("both", "batched"):
            nfree = run_batched(precond, out, plain_x, plain_its)
        else:
            nfree = int(sum(f.size for f in frees))
        if "plain_ms" in out and "batched_ms" in out:
            out["speedup"] = out["plain_ms"] / out["batched_ms"]
        return out, nfree

    def run_plain(precond, out):
        plain = []
        for p, fr in zip(probs, frees):
            s = G.NonlinearSystem(p["coords"], p["conn"], m, rule)
            s.set_block_preconditioner(precond)
            u0 = torch.as_tensor(p["u"].reshape(-1), device=dev)
            d_free = torch.as_tensor(fr, device=dev)
            nnz = s.partition_device(d_free)
            plain.append(dict(s=s, u0=u0, u=u0.clone(), f=torch.zeros_like(u0),
                              v=torch.zeros(s.nnz, dtype=torch.float64, device=dev), free=d_free.long(),
                              blk=torch.empty(max(nnz, 1), dtype=torch.float64, device=dev),
                              x=torch.zeros(fr.size, dtype=torch.float64, device=dev)))

        def plain_iteration():
            its = []
            for q in plain:
                q["u"].copy_(q["u0"])
                q["s"].evaluate_device(q["u"], q["f"], q["v"], column_major=False, stream=stream)
                q["s"].result(stream)
                q["s"].free_block_device(q["v"], out=q["blk"], stream=stream)
                rhs = -q["f"][q["free"]]
                try:
                    _, it, _ = q["s"].solve_free_block_device(q["blk"], rhs, x=q["x"], tolerance=tolerance,
                                                              max_iterations=max_iterations, symmetric=False,
                                                              stream=stream)
                except G.SolverError:
                    it = -1
                q["u"][q["free"]] += q["x"]
                its.append(it)
            return its

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_iteration()  # warm-up (with the multigrid hierarchies' build)
        torch.cuda.synchronize()
        out.setdefault("first_iteration_ms", {})["plain"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        plain_its = plain_iteration()
        torch.cuda.synchronize()
        out["plain_ms"] = (time.perf_counter() - t0) * 1e3
        plain_x = [q["x"].cpu().numpy() for q in plain]
        del plain
        torch.cuda.empty_cache()
        return plain_x, plain_its

    def run_batched(precond, out, plain_x, plain_its):
        b = G.BatchedSystem([(p["coords"], p["conn"]) for p in probs], m, rule)
        b.system.set_block_preconditioner(precond)
        u0 = torch.as_tensor(np.concatenate([p["u"] for p in probs]).reshape(-1), device=dev)
        u = u0.clone()
        fint = torch.zeros_like(u)
        values = torch.zeros(b.system.nnz, dtype=torch.float64, device=dev)
        free = np.concatenate([f + 3 * b.node_offset[i] for i, f in enumerate(frees)]).astype(np.int32)
        d_free = torch.as_tensor(free, device=dev)
        nnz = b.system.partition_device(d_free)
        blk = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
        x = torch.zeros(free.size, dtype=torch.float64, device=dev)
        rows = b.free_segments(free)
        lfree = d_free.long()
        b.autotune(u, stream)

        def batched_iteration():
            u.copy_(u0)
            b.evaluate_device(u, fint, values, column_major=False, stream=stream)
            sc = b.result(stream)
            b.system.free_block_device(values, out=blk, stream=stream)
            rhs = -fint[lfree]
            _, its, _, status = b.system.solve_free_block_segments_device(blk, rhs, rows, x=x, tolerance=tolerance,
                                                                          max_iterations=max_iterations,
                                                                          symmetric=False, stream=stream)
            u[lfree] += x
            return its, status, sc

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batched_iteration()
        torch.cuda.synchronize()
        out.setdefault("first_iteration_ms", {})["batched"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        its, status, sc = batched_iteration()
        torch.cuda.synchronize()
        out["batched_ms"] = (time.perf_counter() - t0) * 1e3
        out.update({"iterations_min": int(np.min(its)), "iterations_median": float(np.median(its)),
                    "iterations_max": int(np.max(its)), "failed_solves": int((status != 0).sum()),
                    "failed_elements": int(sc["failed"].sum())})
        if plain_x is not None:
            xb = x.cpu().numpy()
            out["identical_to_plain"] = bool(
                all(np.array_equal(xb[rows[p]:rows[p + 1]], plain_x[p]) for p in range(len(probs))) and
                [int(i) if st == 0 else -1 for i, st in zip(its, status)] == [int(i) for i in plain_its])
        del b, values, blk
        torch.cuda.empty_cache()
        return int(free.size)

    out = {"workload": "one Newton iteration of the 64 forming problems (rims clamped): evaluation, free-free "
                       f"block, BiCGStab to {tolerance:g}, update",
           "problems": len(probs), "tolerance": tolerance}
    if precond in ("both", "diagonal"):
        diag, out["free_dofs"] = run("diagonal")
        out.update(diag)
    if precond in ("both", "multigrid"):
        mg, out["free_dofs"] = run("multigrid")
        out["multigrid"] = dict(mg, note="the same iteration with each problem's block solve preconditioned by its "
                                     "own multigrid V-cycle; first_iteration_ms includes the hierarchies' host "
                                     "build (once per partition)")
    return out


def differentiation_leg(G, W, torch, dev):
    """The differentiation form on the 200 x 250 x 4 sheet mid-way through a
    spiral incremental-forming path: the forward evaluation and the adjoint
    (v^T df_int/dX and the Hill48/Chaboche parameter products) for a seeded
    N(0, 1) cotangent, with a central-difference check of the coordinate
    product along a random direction (h = 1e-6 of the element size)."""
    coords, conn, u1, u2 = W.differentiation_workload()
    m = W.microbench_material()
    s = G.NonlinearSystem(coords, conn, m, G.Rule(2, 0, 2))
    ua = torch.as_tensor(u1, device=dev)
    u = torch.as_tensor(u2, device=dev)
    f = torch.zeros_like(u)
    s.evaluate_device(ua, f, None, want_tangent=False)
    s.result()
    s.commit()
    rng = np.random.default_rng(2024)
    v = torch.as_tensor(rng.normal(size=u.numel()), device=dev)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    s.evaluate_device(u, f, None, want_tangent=False)
    ev = s.result()
    e0.record()
    s.evaluate_device(u, f, None, want_tangent=False)
    e1.record()
    g = s.vjp_device(u, v, coords=True, parameters=True)
    e2.record()
    torch.cuda.synchronize()
    ms_fwd, ms_vjp = e0.elapsed_time(e1), e1.elapsed_time(e2)
    gx = g["coords"].cpu().numpy()
    d = rng.normal(size=coords.shape) * (0.020 / 200)
    h = 1e-6
    vn = v.cpu().numpy()

    def f_at(c):
        s.set_coords(c)
        s.evaluate_device(u, f, None, want_tangent=False)
        s.result()
        return f.cpu().numpy()

    fd = vn @ (f_at(coords + h * d) - f_at(coords - h * d)) / (2.0 * h)
    an = gx @ d.reshape(-1)
    s.set_coords(coords)
    # end-of-step stress recovery of every element, after the step's commit
    s.evaluate_device(u, f, None, want_tangent=False)
    s.result()
    s.commit()
    stress, status = s.stress_device(u)
    e0.record()
    for _ in range(5):
        s.stress_device(u, stress, status)
    e1.record()
    torch.cuda.synchronize()
    ms_stress = e0.elapsed_time(e1) / 5
    stress_failures = int((status != 0).sum().item())
    return {"workload": "200 x 250 x 4 Hex8 sheet (20 x 25 x 1 mm), mid-way through a 3-turn spiral incremental "
                        "forming path (committed increment, then the next), Hill48 + 2 Chaboche + Voce, log strain",
            "elements": int(conn.shape[0]), "yielding_points": int(ev.yielding_points),
            "forward_ms": ms_fwd, "vjp_ms": ms_vjp,
            "vjp": "coordinates + 11 material parameters", "fd_direction_rel_error": abs(fd - an) / abs(an),
            "stress_recovery_ms": ms_stress, "stress_recovery_failures": stress_failures}


def assembly_leg(G, W, torch, dev, reps):
    """The assembly + contact workload: a 100 x 100 x 4 Hex8 plate (51 005
    nodes, 153 015 DOFs, nnz 10 600 317) with random FP64 element forces and
    non-symmetric 24 x 24 tangents (seeded), assembled into the reference's
    column-major values; 10 000 top-surface slave nodes against a sphere
    R = 5 mm at a seeded offset pressed 0.2 mm into the plate, Coulomb
    friction 0.1 with the exact tangent, kappa A = 1e7 N/m per node; the
    active nodes' 3 x 3 blocks added into the assembled values.

    The CSR assembly is HBM-bound. Its algorithmic bytes: every element
    matrix read once (4 608 B), one 4-byte contribution code per element
    block (64 per element, the reference's block_offset table), every value
    written once (8 B x nnz)."""
    w = W.assembly_workload()
    coords, conn, fe, ke = w["coords"], w["conn"], w["fe"], w["ke"]
    nn, ne = coords.shape[0], conn.shape[0]
    sysm = G.NonlinearSystem(coords, conn, G.Material(70e9, 0.33), G.Rule(), G.FINITE_LOGARITHMIC)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    d_fe, d_ke = t(fe), t(ke)
    del fe, ke
    values = torch.zeros(sysm.nnz, dtype=torch.float64, device=dev)
    fint = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    gross = torch.zeros(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(77)
    radius = 0.005
    cx, cy = rng.uniform(-0.03, 0.03, 2)
    tool = G.make_tool("sphere", (cx, cy, radius - 2e-4), (cx - 1e-4, cy, radius - 2e-4), radius=radius,
                       friction=0.1)
    nodes = w["contact_nodes"]
    d_nodes = t(nodes, torch.int32)
    d_coords, d_u, d_start = t(coords.ravel()), torch.zeros(3 * nn, dtype=torch.float64, device=dev), \
        torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    d_kappa = torch.full((nodes.size,), 1e7, dtype=torch.float64, device=dev)
    bufs = G.contact_buffers(nodes.size, dev)
    residual = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    first = t(sysm.diagonal_blocks()[nodes], torch.int32)
    row_ptr, _ = sysm.pattern()
    row_len = t((row_ptr[3 * nodes + 1] - row_ptr[3 * nodes]).astype(np.int32), torch.int32)

    def contact():
        return G.contact_evaluate_device(tool, d_coords, d_u, d_start, d_nodes, d_kappa, residual=residual,
                                         stream=stream, buffers=bufs, read_scalars=False)

    def blocks():
        G.add_node_blocks_device(bufs["blocks"], bufs["active"], first, row_len, values, column_major=True,
                                 stream=stream)

    # the increment set-up and the anticipation over the same slaves (a
    # correction that pushes every node up by 0.3 mm)
    du = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    du[2::3] = 3e-4

    cand_bufs = G.contact_candidates_device(tool, d_coords, d_u, d_nodes, 2e-4, stream=stream)
    cand_bufs = (*cand_bufs, torch.zeros(nodes.size + nodes.size // 256 + 1, dtype=torch.int32, device=dev))
    ant_bufs = G.contact_anticipate_device(tool, d_coords, d_u, du, d_nodes, d_kappa, residual, stream=stream)

    # stiffness_gross over every DOF, on the assembled values
    kx = torch.as_tensor(np.random.default_rng(5).normal(size=3 * nn), device=dev)
    kfree = torch.arange(3 * nn, dtype=torch.int32, device=dev)
    kprod = torch.empty(3 * nn, dtype=torch.float64, device=dev)

    def gross_cm():
        sysm.stiffness_gross_device(values, kx, kfree, column_major=True, product=kprod, stream=stream)

    def gross_rm():
        sysm.stiffness_gross_device(values, kx, kfree, column_major=False, product=kprod, stream=stream)

    # the free-free block of the Newton system: the plate clamped at both
    # ends in x (every DOF of those nodes fixed)
    ends = np.isclose(coords[:, 0], coords[:, 0].min()) | np.isclose(coords[:, 0], coords[:, 0].max())
    free_mask = np.repeat(~ends, 3)
    d_pfree = t(np.nonzero(free_mask)[0], torch.int32)
    block_nnz = sysm.partition_device(d_pfree, stream=stream)
    block_values = torch.empty(max(block_nnz, 1), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        sysm.partition_device(d_pfree, stream=stream)
    partition_ms = (time.perf_counter() - t0) * 1e3 / 5

    def free_block():
        sysm.free_block_device(values, out=block_values, stream=stream)

    # the block solve on the plate's own elastic tangent (symmetric positive
    # definite with the ends clamped), a seeded right-hand side
    k_values = torch.empty(sysm.nnz, dtype=torch.float64, device=dev)
    sysm.evaluate_device(torch.zeros(3 * nn, dtype=torch.float64, device=dev),
                         torch.empty(3 * nn, dtype=torch.float64, device=dev), k_values, column_major=True,
                         stream=stream)
    k_block = sysm.free_block_device(k_values, stream=stream)
    rhs = torch.as_tensor(np.random.default_rng(9).normal(size=d_pfree.shape[0]) * 1e3, device=dev)
    x_sol = torch.empty_like(rhs)
    sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, stream=stream)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, solve_its, solve_rel = sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, stream=stream)
    torch.cuda.synchronize()
    solve_ms = (time.perf_counter() - t0) * 1e3
    # the same system by BiCGStab (the method for a non-symmetric tangent)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, bicg_its, bicg_rel = sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, symmetric=False,
                                                         stream=stream)
    torch.cuda.synchronize()
    bicg_ms = (time.perf_counter() - t0) * 1e3
    # the same two solves preconditioned by the multigrid V-cycle: the first
    # solve after the partition builds the hierarchy (host, once), every
    # solve sets its values up from the block (device)
    sysm.set_block_preconditioner("multigrid")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, stream=stream)
    torch.cuda.synchronize()
    mg_first_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    _, mg_its, mg_rel = sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, stream=stream)
    torch.cuda.synchronize()
    mg_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    _, mgb_its, mgb_rel = sysm.solve_free_block_device(k_block, rhs, x=x_sol, tolerance=1e-10, symmetric=False,
                                                       stream=stream)
    torch.cuda.synchronize()
    mgb_ms = (time.perf_counter() - t0) * 1e3
    mg_rows, mg_nnz = sysm.block_preconditioner_levels()
    sysm.set_block_preconditioner("diagonal")
    del k_values, k_block

    def candidates():
        G.contact_candidates_device(tool, d_coords, d_u, d_nodes, 2e-4, stream=stream, buffers=cand_bufs)

    def anticipate():
        G.contact_anticipate_device(tool, d_coords, d_u, du, d_nodes, d_kappa, residual, stream=stream,
                                    buffers=ant_bufs)

    # the same element matrices stored column-major (Eigen's storage of the reference's Ke)
    d_ke_cm = d_ke.reshape(ne, 24, 24).transpose(1, 2).contiguous().reshape(-1)
    stages = (("values", lambda: sysm.assemble_device(ke=d_ke, values=values, column_major=True, stream=stream)),
              ("values_row", lambda: sysm.assemble_device(ke=d_ke, values=values, column_major=False, stream=stream)),
              ("values_ke_cm", lambda: sysm.assemble_device(ke=d_ke_cm, values=values, column_major=True,
                                                            ke_column_major=True, stream=stream)),
              ("force", lambda: sysm.assemble_device(fe=d_fe, fint=fint, gross=gross, stream=stream)),
              ("contact", contact), ("node_blocks", blocks), ("candidates", candidates), ("anticipate", anticipate),
              ("gross_cm", gross_cm), ("gross_rm", gross_rm), ("free_block", free_block))
    us = {}
    for name, fn in stages:
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        us[name] = e0.elapsed_time(e1) * 1e3 / reps
    n_active = int(bufs["active"].sum().item())
    alg = ne * (4608 + 64 * 4) + 8 * sysm.nnz
    gbs = alg / (us["values"] * 1e-6) / 1e9
    gbs_cm = alg / (us["values_ke_cm"] * 1e-6) / 1e9
    return {"workload": "100 x 100 x 4 Hex8 plate, random non-symmetric element matrices, column-major values; "
                        "10 000 slave nodes vs a sphere R = 5 mm, friction 0.1, exact tangent",
            "elements": int(ne), "nodes": int(nn), "nnz": int(sysm.nnz), "reps": reps,
            "values_us": us["values"], "values_bytes": alg, "values_GBps": gbs,
            "values_hbm_frac": gbs / HBM_PEAK_GBS, "values_row_major_us": us["values_row"],
            "values_ke_column_major": {"note": "element matrices stored column-major (Eigen's storage of the "
                                               "reference's Ke), column-major values",
                                       "us": us["values_ke_cm"], "GBps": gbs_cm, "hbm_frac": gbs_cm / HBM_PEAK_GBS},
            "force_us": us["force"], "contact_us": us["contact"],
            "node_blocks_us": us["node_blocks"], "contact_active_nodes": n_active,
            "contact_candidates_us": us["candidates"], "contact_anticipate_us": us["anticipate"],
            "stiffness_gross_us": us["gross_cm"], "stiffness_gross_row_major_us": us["gross_rm"],
            "stiffness_gross_GBps": (8 * sysm.nnz + 3 * 8 * 3 * nn) / (us["gross_cm"] * 1e-6) / 1e9,
            "free_block": {"note": "PartitionedFactor: the free-free block of the assembled values, the plate "
                                   "clamped at both x ends; partition = pattern set-up (host clock, one "
                                   "synchronization), values = the per-iteration copy (20 B per block entry: "
                                   "source index, value read, value written)",
                           "free_dofs": int(free_mask.sum()), "block_nnz": int(block_nnz),
                           "partition_ms": partition_ms, "values_us": us["free_block"],
                           "values_GBps": 20 * block_nnz / (us["free_block"] * 1e-6) / 1e9,
                           "values_hbm_frac": 20 * block_nnz / (us["free_block"] * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "block_solve": {"note": "K_ff x = b on the clamped plate's elastic tangent by Jacobi-preconditioned "
                                    "conjugate gradients to ||b - K x|| <= 1e-10 ||b|| (host clock, reads of the "
                                    "residual every 16 iterations included); per iteration 12 B per block entry "
                                    "(the product's values and indices) + 112 B per free DOF (p read twice, q "
                                    "written and read, x, r read and written, D^-1 read, z written and read, p "
                                    "written)",
                            "iterations": int(solve_its), "relative_residual": solve_rel, "ms": solve_ms,
                            "us_per_iteration": solve_ms * 1e3 / max(solve_its, 1),
                            "GBps": (12 * block_nnz + 112 * int(free_mask.sum())) * max(solve_its, 1)
                                    / (solve_ms * 1e-3) / 1e9,
                            "bicgstab": {"note": "the same system by BiCGStab (symmetric = 0, the method for a "
                                                 "non-symmetric tangent): two products per iteration",
                                         "iterations": int(bicg_its), "relative_residual": bicg_rel, "ms": bicg_ms,
                                         "us_per_iteration": bicg_ms * 1e3 / max(bicg_its, 1)},
                            "multigrid": {"note": "the same solves preconditioned by a smoothed-aggregation "
                                                  "multigrid V-cycle (rigid-body modes, Chebyshev degree 2, dense "
                                                  "coarsest level); ms includes the set-up from the block's values",
                                          "levels_rows": mg_rows, "levels_nonzeros": mg_nnz,
                                          "first_solve_ms": mg_first_ms, "iterations": int(mg_its),
                                          "relative_residual": mg_rel, "ms": mg_ms,
                                          "bicgstab_iterations": int(mgb_its), "bicgstab_relative_residual": mgb_rel,
                                          "bicgstab_ms": mgb_ms}}}


def copy_probe(torch, mib):
    """A plain device copy of `mib` MiB, five times: the known-size kernel
    whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the others."""
    dev = torch.device("cuda:0")
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"# copy probe: {mib} MiB read + {mib} MiB written per copy, 5 copies", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--elements", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=0, help="elements of the CPU leg (default: all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-assembly", action="store_true", help="skip the assembly + contact leg")
    ap.add_argument("--no-box", action="store_true", help="skip the 2 x 2 x 5 box-rule leg")
    ap.add_argument("--no-newton", action="store_true", help="skip the batched Newton-iteration leg")
    ap.add_argument("--newton-precond", choices=("both", "diagonal", "multigrid"), default="both",
                    help="the Newton-iteration leg's preconditioners (profiling one alone)")
    ap.add_argument("--newton-form", choices=("both", "plain", "batched"), default="both",
                    help="the Newton-iteration leg's forms (profiling one alone; the comparison needs both)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain MIB-MiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              "(element evaluations are independent and would shard by element ranges)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe > 0:
        copy_probe(torch, args.copy_probe)
        return
    import sparlab_gpu as G
    from sparlab_gpu import workloads as W

    dev = torch.device("cuda:0")
    ne = args.elements
    t_gen = time.perf_counter()
    w = W.microbench(ne)
    m = W.microbench_material()
    rule = G.Rule(2, 0, 2)
    npts = rule.points
    states = G.states_from_reference(w["states"], m)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    batch = G.ElementBatch(t(w["coords"]), t(w["conn"], torch.int32), t(w["u"]), t(states), None, npts, True)
    del states
    torch.cuda.synchronize()
    print(f"# inputs ready in {time.perf_counter() - t_gen:.1f} s ({ne} elements)", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        G.evaluate_elements(m, rule, batch, G.FINITE_LOGARITHMIC, True, stream)
    torch.cuda.synchronize()

    # Timed region: K steps, each the local solve then the condensation, with
    # HIP events on the launch stream between the two kernels.
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        G.evaluate_elements(m, rule, batch, G.FINITE_LOGARITHMIC, True, stream, stages=G.LOCAL_SOLVE)
        ev[k][1].record(stream)
        G.evaluate_elements(m, rule, batch, G.FINITE_LOGARITHMIC, True, stream, stages=G.CONDENSATION)
        ev[k][2].record(stream)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps
    ms_local = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps
    ms_cond = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps
    value = ne / (ms * 1e-3)
    # the last step's outputs (every step evaluates the same inputs)
    status = batch.status.cpu().numpy()
    failures = int((status != 0).sum())
    plastic = float((batch.yielding.cpu().numpy() > 0).mean())
    mean_iters = float(batch.iterations.cpu().numpy().mean())
    print(f"# {ms:.3f} ms/step (event), {wall / args.steps * 1e3:.3f} ms/step (wall); local solve {ms_local:.3f} ms, "
          f"condensation {ms_cond:.3f} ms; failures {failures}, plastic elements {plastic:.3f}, "
          f"mean local iterations {mean_iters:.2f}", file=sys.stderr, flush=True)

    cpu = None
    roof = None
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        sample = ne if args.cpu_sample <= 0 else min(args.cpu_sample, ne)
        cb = cpu_baseline_leg(w, sample, threads)
        cpu = {"value": cb["evals_per_s"], "unit": "element evals/s", "cores": threads, "kind": "port",
               "sample": f"the first {sample} elements of the same workload (all of it by default), the oracle's "
                         f"element loop with OpenMP schedule(static) on {threads} threads, {cb['seconds']:.1f} s wall",
               "host_cpu": host_cpu_model()}
        one = cpu_single_core_leg(w, min(ne, 20000))
        cpu["single_core"] = {"value": one["evals_per_s"], "unit": "element evals/s", "cores": 1,
                              "us_per_eval": one["us_per_eval"],
                              "sample": f"the first {min(ne, 20000)} elements, one thread pinned to core {one['core']}, "
                                        f"{one['seconds']:.1f} s wall"}
        fl, fc = cb["flops_local"], cb["flops_condense"]
        achieved = fl * ne / (ms_local * 1e-3) / 1e12
        roof = {"bound": "fp64", "pipe": "FP64 vector ALU (no MFMA instructions; MI355X FP64 vector peak = FP64 matrix peak)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic("local_solve_kernel", ne),
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                                "profiles/hbm_traffic_1M_microbench_and_assembly.md)",
                "kernel": "local_solve_kernel (the dominant kernel)",
                "flops_per_elem": fl, "flops_sample": cb["flops_sample"],
                "kernel_ms": ms_local,
                "condensation": {"kernel": "condense_kernel", "flops_per_elem": fc, "kernel_ms": ms_cond,
                                 "achieved": fc * ne / (ms_cond * 1e-3) / 1e12},
                "element_evaluation": {"flops_per_elem": cb["flops_per_elem"], "ms": ms,
                                       "achieved": cb["flops_per_elem"] * ne / (ms * 1e-3) / 1e12}}
        print(f"# cpu oracle: {cb['evals_per_s']:.0f} evals/s on {threads} threads, {one['us_per_eval']:.1f} us/eval on "
              f"one core ({cpu['host_cpu']}); {cb['flops_per_elem']:.0f} "
              f"FLOP/elem ({fl:.0f} local + {fc:.0f} condensation); oracle plastic fraction "
              f"{cb['plastic_fraction']:.3f}", file=sys.stderr)

    del batch
    torch.cuda.empty_cache()
    box = None
    if not args.no_box:
        box = box_rule_leg(G, W, torch, dev, max(1, min(args.steps, 5)))
        torch.cuda.empty_cache()
        print(f"# 2 x 2 x 5 rule: {box['elements']} elements, {box['ms']:.2f} ms ({box['evals_per_s'] / 1e6:.2f} M evals/s, "
              f"{box['achieved_TFLOPs']:.2f} TFLOP/s, {box['frac'] * 100:.1f} % of FP64 peak; local solve "
              f"{box['local_solve_ms']:.2f} ms, condensation {box['condensation_ms']:.2f} ms); oracle on one core "
              f"{box['cpu_single_core_us_per_eval']:.1f} us/eval", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batched:
        batched = batched_leg(G, W, torch, dev, max(1, min(args.steps, 5)))
        print(f"# batched form: {batched['elements']} elements in {batched['problems']} problems; plain "
              f"{batched['plain_ms']:.2f} ms, batched {batched['batched_ms']:.2f} ms per evaluation of all "
              f"({batched['speedup']:.2f}x, {batched['elements_per_wavefront']} elements per wavefront)",
              file=sys.stderr, flush=True)

    adjoint = None
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        adjoint = differentiation_leg(G, W, torch, dev)
        print(f"# differentiation: {adjoint['elements']} elements; forward {adjoint['forward_ms']:.2f} ms, "
              f"VJP {adjoint['vjp_ms']:.2f} ms; FD check {adjoint['fd_direction_rel_error']:.2e}; stress recovery "
              f"{adjoint['stress_recovery_ms']:.2f} ms", file=sys.stderr, flush=True)

    assembly = None
    if not args.no_assembly:
        torch.cuda.empty_cache()
        assembly = assembly_leg(G, W, torch, dev, 20)
        print(f"# assembly: {assembly['nnz']} values in {assembly['values_us']:.1f} us "
              f"({assembly['values_GBps']:.0f} GB/s algorithmic; row-major values {assembly['values_row_major_us']:.1f} us, "
              f"column-major K_e {assembly['values_ke_column_major']['us']:.1f} us = "
              f"{assembly['values_ke_column_major']['GBps']:.0f} GB/s), force gather {assembly['force_us']:.1f} us, "
              f"contact {assembly['contact_us']:.1f} us ({assembly['contact_active_nodes']} active), node blocks "
              f"{assembly['node_blocks_us']:.1f} us, candidates {assembly['contact_candidates_us']:.1f} us, "
              f"anticipation {assembly['contact_anticipate_us']:.1f} us, stiffness_gross {assembly['stiffness_gross_us']:.1f} "
              f"us ({assembly['stiffness_gross_GBps']:.0f} GB/s; row-major {assembly['stiffness_gross_row_major_us']:.1f} us), "
              f"free-free block: partition {assembly['free_block']['partition_ms']:.2f} ms, values "
              f"{assembly['free_block']['values_us']:.1f} us ({assembly['free_block']['values_GBps']:.0f} GB/s); block "
              f"solve {assembly['block_solve']['iterations']} CG iterations in {assembly['block_solve']['ms']:.1f} ms "
              f"({assembly['block_solve']['us_per_iteration']:.1f} us each, {assembly['block_solve']['GBps']:.0f} GB/s); "
              f"multigrid: CG {assembly['block_solve']['multigrid']['iterations']} iterations in "
              f"{assembly['block_solve']['multigrid']['ms']:.1f} ms, BiCGStab "
              f"{assembly['block_solve']['multigrid']['bicgstab_iterations']} in "
              f"{assembly['block_solve']['multigrid']['bicgstab_ms']:.1f} ms (first solve with the hierarchy's build "
              f"{assembly['block_solve']['multigrid']['first_solve_ms']:.0f} ms; levels "
              f"{assembly['block_solve']['multigrid']['levels_rows']})",
              file=sys.stderr, flush=True)

    newton = None
    if not args.no_newton:
        torch.cuda.empty_cache()
        newton = newton_leg(G, W, torch, dev, precond=args.newton_precond, form=args.newton_form)
        torch.cuda.empty_cache()
    if newton and "plain_ms" in newton and "batched_ms" in newton:
        print(f"# batched Newton iteration: {newton['problems']} problems, {newton['free_dofs']} free DOFs; plain "
              f"{newton['plain_ms']:.1f} ms, batched {newton['batched_ms']:.1f} ms ({newton['speedup']:.2f}x); "
              f"BiCGStab iterations {newton['iterations_min']}..{newton['iterations_max']} (median "
              f"{newton['iterations_median']:.0f}); identical to plain: {newton['identical_to_plain']}",
              file=sys.stderr, flush=True)
    if newton and "multigrid" in newton and "speedup" in newton["multigrid"]:
        mg = newton["multigrid"]
        print(f"# batched Newton iteration, multigrid: plain {mg['plain_ms']:.1f} ms, batched "
              f"{mg['batched_ms']:.1f} ms ({mg['speedup']:.2f}x); BiCGStab iterations {mg['iterations_min']}.."
              f"{mg['iterations_max']} (median {mg['iterations_median']:.0f}); first iteration (hierarchy build) "
              f"plain {mg['first_iteration_ms']['plain']:.0f} ms, batched {mg['first_iteration_ms']['batched']:.0f} "
              f"ms; identical to plain: {mg['identical_to_plain']}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "element evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"element microbench: {ne} incompatible-mode Hex8 (jittered unit cubes), "
                               "log strain, Hill48 + 2 Chaboche + Voce, f_int + K_t + trial history",
                   "elements": ne, "integration_points": npts, "plastic_fraction": plastic,
                   "mean_local_iterations": mean_iters, "failures": failures},
        "roofline": roof, "cpu_baseline": cpu, "box_rule": box, "batched": batched, "newton": newton, "differentiation": adjoint,
        "assembly": assembly,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
