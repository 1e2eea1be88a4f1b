"""Benchmark: FP64 incompatible-mode Hex8 f_int + K_t element evaluations/s.

Workload (one GPU): the element microbenchmark -- 1,000,000 independent
incompatible-mode Hex8 elements (jittered unit cubes, displacements
N(0, 1e-3)), logarithmic strain, Hill48 (r 0.7/0.6/0.8) with two Chaboche
backstresses and Voce hardening, committed history inside the yield surface
(sparlab_gpu/workloads.py). One step evaluates every element: the local Newton
iteration of the modes, the condensed force f* (24) and tangent K* (24 x 24),
the points' trial history. Inputs are resident in HBM before timing.

Roofline: the element evaluation is FP64-compute-bound; its algorithmic FLOPs
per element are counted by the oracle's op-counting instantiation
(add/sub/mul/div/sqrt and each libm call = 1) on a sample of the same inputs,
and divided by the HIP-event time of the element kernels.

CPU baseline: the oracle's element loop (OpenMP schedule(static), one element
per call, the reference's NonlinearSystem.hpp:208-237 structure) on a bounded
sample of the same workload, on this host's cores.
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


def cpu_baseline_leg(workload, sample, threads):
    """The oracle on the host: evals/s on `threads` cores, and the algorithmic
    FLOPs per element (op-counting oracle) on the same sample."""
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
    flops = [O.element_flops(x0[i], ue[i], om, states=st[i], kinematics=2) for i in range(nflop)]
    return dict(evals_per_s=sample / dt, seconds=dt, failures=int((r["status"] != 0).sum()),
                flops_per_elem=float(np.mean(flops)), flops_sample=nflop,
                plastic_fraction=float((r["scalars"][:, 1] > 0).mean()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--elements", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              "(element evaluations are independent; see DESIGN.md for how they would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
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
    status = batch.status.cpu().numpy()
    failures = int((status != 0).sum())
    plastic = float((batch.yielding.cpu().numpy() > 0).mean())
    mean_iters = float(batch.iterations.cpu().numpy().mean())

    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        G.evaluate_elements(m, rule, batch, G.FINITE_LOGARITHMIC, True, stream)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps
    value = ne / (ms * 1e-3)
    print(f"# {ms:.3f} ms/step (event), {wall / args.steps * 1e3:.3f} ms/step (wall); failures {failures}, "
          f"plastic elements {plastic:.3f}, mean local iterations {mean_iters:.2f}", file=sys.stderr, flush=True)

    cpu = None
    roof = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, 16)
        cb = cpu_baseline_leg(w, min(args.cpu_sample, ne), threads)
        cpu = {"value": cb["evals_per_s"], "unit": "element evals/s", "cores": threads, "kind": "port",
               "sample": f"first {min(args.cpu_sample, ne)} elements of the same workload, OpenMP "
                         f"static over elements, {cb['seconds']:.1f} s wall"}
        flops = cb["flops_per_elem"]
        achieved = flops * ne / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "pipe": "fp64 (vector ALU; MI355X FP64 vector = matrix peak)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "flops_per_elem": flops, "flops_sample": cb["flops_sample"],
                "kernel": "local_solve_kernel + condense_kernel (one element evaluation)"}
        print(f"# cpu oracle: {cb['evals_per_s']:.0f} evals/s on {threads} threads; "
              f"{flops:.0f} FLOP/elem; oracle plastic fraction {cb['plastic_fraction']:.3f}", file=sys.stderr)

    line = {
        "metric": METRIC, "value": value, "unit": "element evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"element microbench: {ne} incompatible-mode Hex8 (jittered unit cubes), "
                               "log strain, Hill48 + 2 Chaboche + Voce, f_int + K_t + trial history",
                   "elements": ne, "integration_points": npts, "plastic_fraction": plastic,
                   "mean_local_iterations": mean_iters, "failures": failures},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
