#!/usr/bin/env python3
"""Throughput of the per-step PM gravity cycle on one MI355X.

Workload (default): N = 2^24 particles on a 256^3 head grid (lmin = 8) with one
refinement level (lmax = 9, a cell refines at >= 8 particles), half the
particles uniform and half in 64 Gaussian clumps (sigma uniform in [1, 4] head
cells), equal masses, zero velocities.  A step is one fixed-dt KDK step of the
reference loop (particle_updating_A, reset, tree rebuild, grid_density,
potential to 1e-10 relative residual, grid_acceleration, particle_acceleration,
particle_updating_B).  dt is the reference CFL rule (CFL 0.5) evaluated after
the initial force pass and then frozen.

Reported value: particle-steps/s of the mixed path (FP32 particles, grids and
V-cycles; FP64 outer residual/correction and FP64 branch CG).  The plain FP64
path is timed on the same workload for the speed-up.  The roofline object
covers the dominant kernel (algorithmic bytes from the library's per-launch
accounting over HIP-event time on the context stream); cpu_baseline times the
compiled reference (oracle/_ref) on a bounded sample of the same workload.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    # name: (N, lmin, lmax, crit, kind)
    "c3": (1 << 24, 8, 9, 8, "clumps"),
    "c3-small": (1 << 20, 7, 8, 8, "clumps"),
    "c1": (10000, 5, 5, 8, "plummer"),
}


def make_particles(kind, n, lmin, seed=20261203):
    from quantizer_amd.workloads import clumps, plummer
    if kind == "clumps":
        return clumps(n, lmin, n_clumps=64, seed=seed)
    return plummer(n, seed=seed)


def run_path(parts, lmin, lmax, crit, precision, steps, warmup, prof_steps, inner, solver=None):
    from quantizer_amd import core
    t0 = time.time()
    sim = core.Simulation(parts, lmin, lmax, precision=precision, ref_criterion_ptcl=crit, inner_cycles=inner,
                          branch_solver=solver)
    sim.force()
    dt = sim.timestep(cfl=0.5)
    log(f"  prec={precision} setup+force {time.time() - t0:.2f}s dt={dt:.3e} stats={sim.stats()}")
    for _ in range(warmup):
        sim.step(dt)
    sim.sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        sim.step(dt)
    sim.sync()
    elapsed = time.perf_counter() - t1
    st = sim.stats()
    report = {}
    if prof_steps > 0:
        sim.timing(True)
        tp = time.perf_counter()
        for _ in range(prof_steps):
            sim.step(dt)
        sim.sync()
        tprof = time.perf_counter() - tp
        report = sim.timing_report()
        sim.timing(False)
        report["_wall_ms_per_step"] = {"ms": 1e3 * tprof / prof_steps}
    sim.close()
    return elapsed, dt, st, report


def roofline_from(report, prof_steps):
    kernels = {k: v for k, v in report.items() if not k.startswith("_") and not k.startswith("host:")}
    if not kernels:
        return None, {}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    per_launch_ms = d["ms"] / max(1, d["launches"])
    per_launch_bytes = d["bytes"] / max(1, d["launches"])
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    total_ms = sum(v["ms"] for v in kernels.values())
    total_bytes = sum(v["bytes"] for v in kernels.values())
    table = {k: {"ms_per_step": round(v["ms"] / prof_steps, 4), "launches_per_step": v["launches"] / prof_steps,
                 "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 and v["bytes"] > 0 else None}
             for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms"])}
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": per_launch_bytes, "ms_per_launch": per_launch_ms,
            "step_kernel_ms": round(total_ms / prof_steps, 3),
            "step_algorithmic_GBps": round(total_bytes / (total_ms * 1e-3) / 1e9, 1) if total_ms > 0 else None}
    return roof, table


def cpu_baseline(parts, lmin, lmax, crit, dt, budget_s):
    """Time the compiled reference (oracle/_ref/ref_harness) on this workload: one
    force pass plus one step when it fits the budget, else on a particle
    subsample of the same grid."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(ref):
        return None
    from tests.oracle_lib import run_reference  # noqa: F401  (checker module; only the binary is timed)
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "p.bin")
        parts.write_bin(inp)
        cmd = [ref, inp, td, str(lmin), str(lmax), str(crit), "1", repr(float(dt)), "--time", "--noenergy",
               "--threads", "1"]
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=budget_s)
        except subprocess.TimeoutExpired:
            return {"value": None, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                    "sample": f"timed out after {budget_s}s"}
        line = [l for l in p.stderr.splitlines() if l.startswith("REF_TIMING")]
        if p.returncode != 0 or not line:
            return None
        kv = dict(t.split("=") for t in line[0].split()[1:])
        sps = float(kv["s_per_step"])
        return {"value": parts.n / sps, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                "sample": f"compiled reference, 1 thread, N={parts.n}, lmin={lmin}, lmax={lmax}: "
                          f"initial force pass {float(kv['first_force_s']):.2f}s, then 1 timed step "
                          f"({sps:.2f} s/step)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=150.0)
    ap.add_argument("--cpu-n", type=int, default=1 << 24, help="particles in the CPU sample")
    ap.add_argument("--inner", type=int, default=0)
    ap.add_argument("--branch-solver", choices=["cg", "mg"], default=None,
                    help="refinement-node solver (default: the library default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    from quantizer_amd import core
    N, lmin, lmax, crit, kind = WORKLOADS[args.workload]
    solver = {None: None, "cg": core.BRANCH_CG, "mg": core.BRANCH_MG}[args.branch_solver]
    log(f"workload {args.workload}: N={N} lmin={lmin} lmax={lmax} crit={crit} {kind}")
    t = time.time()
    parts = make_particles(kind, N, lmin)
    log(f"  generated in {time.time() - t:.1f}s")

    el_mx, dt, st_mx, rep_mx = run_path(parts, lmin, lmax, crit, core.PREC_MIXED, args.steps, args.warmup,
                                        args.prof_steps, args.inner, solver)
    value = N * args.steps / el_mx
    ms_step = 1e3 * el_mx / args.steps
    log(f"  mixed: {ms_step:.3f} ms/step, {value:.4e} particle-steps/s")
    fp64 = None
    if not args.skip_fp64:
        el64, _, st64, rep64 = run_path(parts, lmin, lmax, crit, core.PREC_FP64, args.steps, args.warmup, 0,
                                        args.inner, solver)
        fp64 = {"value": N * args.steps / el64, "ms_per_step": 1e3 * el64 / args.steps,
                "head_vcycles": st64["head_vcycles"], "branch_iterations": st64["branch_iterations"]}
        log(f"  fp64: {fp64['ms_per_step']:.3f} ms/step")
    roof, table = roofline_from(rep_mx, max(1, args.prof_steps))
    cpu = None
    if not args.skip_cpu:
        # bounded sample: one step of the same workload (about 20 s on one core)
        ns = min(N, args.cpu_n)
        sample = make_particles(kind, ns, lmin, seed=7) if ns < N else parts
        cpu = cpu_baseline(sample, lmin, lmax, crit, dt, args.cpu_budget)
        if cpu and cpu.get("value"):
            log(f"  cpu reference: {cpu['value']:.4e} particle-steps/s ({cpu['sample']})")
    out = {
        "metric": "particle-steps/s at N=2^24, 256^3 mixed vs FP64 plain; adjoint/fwd time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {kind}, N={N}, head {1 << lmin}^3 (lmin={lmin}), lmax={lmax}, "
                               f"refine at >= {crit} particles/cell, tol 1e-10, fixed dt {dt:.4e} (CFL 0.5 at t=0)",
                   "n_particles": N, "lmin": lmin, "lmax": lmax, "parallelism": "single GPU"},
        "fp64_plain": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "adjoint_fwd_ratio": None,
        "solver": {"head_outer": st_mx["head_outer"], "head_vcycles": st_mx["head_vcycles"],
                   "head_residual": st_mx["head_residual"], "n_nodes": st_mx["n_nodes"],
                   "branch_points": st_mx["branch_points"], "branch_iterations": st_mx["branch_iterations"],
                   "branch_cg_levels": st_mx["branch_cg_levels"]},
        "roofline": roof,
        "kernels": table,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
