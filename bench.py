#!/usr/bin/env python3
"""Throughput of the per-step PM gravity cycle on one MI355X.

Workload (default, "c3"): N = 2^24 particles on a 256^3 head grid (lmin = 8)
with one refinement level (lmax = 9, a cell refines at >= 8 particles), half the
particles uniform and half in 64 Gaussian clumps (sigma uniform in [1, 4] head
cells), equal masses, zero velocities.  A step is one iteration of the
reference main loop (main.c:369-478): timestep (the CFL 0.5 rule, capped at the
t = 0 value as _MAX_dt_ caps it), then the KDK step (particle_updating_A, reset,
tree rebuild, grid_density, potential to 1e-10 relative residual,
grid_acceleration, particle_acceleration, particle_updating_B).  Inputs are
resident in HBM before the timed region starts.

mixed_fixed_dt: the mixed path with dt held at its t = 0 value (the earlier
definition of this workload, kept for comparison; the reference's own main
loop, run that way, files particles of the collapsing clumps outside their
node's box by the fourth step and then reads past its cell lists).

value: particle-steps/s of the mixed path (FP32 particle velocities and
accelerations, FP32 grids and V-cycles; FP64 positions, potentials, outer
residual/correction and refinement-node PCG).  The plain FP64 path is timed on
the same workload for the speed-up.

adjoint_fwd_ratio: Plummer N = 2^18, 64^3 head grid, 50 steps, checkpoints
every 10: (checkpointed forward + reverse sweep) time over plain forward time,
in steady state (a first adjoint run of the same length in the same context
allocates its checkpoints and step records); reverse_to_forward is the reverse
sweep alone over the plain forward.  Like the forward steps, whose state stays
in HBM, the reverse sweep is timed with its cotangents (seeds and gradients)
in device buffers; host_array_copies_ms is the PCIe time of the four (3, N)
host arrays a caller with host memory adds on top.
adjoint_amr: the same ratio with refinement levels: two Plummer spheres of
2^23 particles each, 256^3 head grid, 3 refinement levels, the adjoint of the
last 20 of 40 steps with checkpoints every 5, loss = final kinetic energy of
the lighter sphere.

roofline: the dominant kernel (the head's red-black half sweep, here its red
colour, over-relaxed; the finest-level launches carry most of its time) replayed back to back on the
headline grid between two HIP events on the library's stream; algorithmic
bytes per launch over that time.  traffic: HBM bytes per launch of the same
kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/, scaled
by the counters of a plain copy of known size (null when the summary is absent).

precision_vs_fp64: mixed against FP64 on the same workload: head potential,
grid and particle accelerations after the first force pass, and positions and
velocities after 20 steps.

dropin_resident: the reference's own host program (its main loop over the
stages, the same per-step time step) linked against the resident drop-in
library, on the same workload: the host-inclusive rate of the drop-in boundary.

cpu_baseline: the compiled reference (oracle/_ref) on the same workload, at 1
thread and at the host's thread count (1 warm-up and 3 timed steps each, the
same per-step time step), with the CPU model.

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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_rbgs_fine_sweep_headline.json")


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


def run_path(parts, lmin, lmax, crit, precision, steps, warmup, prof_steps, inner, solver=None, sweep_reps=0):
    from quantizer_amd import core
    t0 = time.time()
    sim = core.Simulation(parts, lmin, lmax, precision=precision, ref_criterion_ptcl=crit, inner_cycles=inner,
                          branch_solver=solver)
    sim.force()
    dt0 = sim.timestep(cfl=0.5)
    log(f"  prec={precision} setup+force {time.time() - t0:.2f}s dt0={dt0:.3e} stats={sim.stats()}")

    def step():
        # main.c:369-478: the CFL time step from the current state, then the
        # step; dt0 (the t = 0 step) is the cap, as _MAX_dt_ is in the reference
        dt = sim.timestep(0.5, dt0)
        sim.step(dt)
        return dt

    for _ in range(warmup):
        step()
    sim.sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        dt = step()
    sim.sync()
    elapsed = time.perf_counter() - t1
    st = sim.stats()
    report = {}
    if prof_steps > 0:
        sim.timing(True)
        for _ in range(prof_steps):
            step()
        sim.sync()
        report = sim.timing_report()
        sim.timing(False)
    sweep = sim.bench_sweep(sweep_reps) if sweep_reps > 0 else None
    sim.close()
    return elapsed, (dt0, dt), st, report, sweep


def fixed_dt_leg(parts, lmin, lmax, crit, steps, warmup):
    """The mixed path with dt held at its t = 0 value for every step: the
    workload's earlier definition, kept so that figures stay comparable.  Beyond
    a few steps that dt moves particles of the collapsing clumps across more
    than a refinement node's box margin per step (the reference's own main loop
    shrinks dt instead)."""
    from quantizer_amd import core
    sim = core.Simulation(parts, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit)
    sim.force()
    dt0 = sim.timestep(cfl=0.5)
    for _ in range(warmup):
        sim.step(dt0)
    sim.sync()
    t = time.perf_counter()
    for _ in range(steps):
        sim.step(dt0)
    sim.sync()
    el = time.perf_counter() - t
    sim.close()
    return {"ms_per_step": 1e3 * el / steps, "value": parts.n * steps / el, "dt": dt0}


def kernel_table(report, prof_steps):
    """Per-stage table of the eager profiling steps (HIP events around each
    launch, so small launches include their dispatch gap)."""
    kernels = {k: v for k, v in report.items() if not k.startswith("_") and not k.startswith("host:")}
    return {k: {"ms_per_step": round(v["ms"] / prof_steps, 4), "launches_per_step": v["launches"] / prof_steps,
                "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 and v["bytes"] > 0 else None}
            for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms"])}


def step_roofline(report, prof_steps, ms_step):
    """The whole step against HBM: algorithmic bytes of every kernel launched in
    a step (each input read once, each output written once; from the eager
    profiling steps that follow the timed ones) over the timed step's wall time.
    kernel_only_GBps divides the same bytes by the kernels' own time instead, so
    the gap between the two is launch gaps and host work."""
    kernels = {k: v for k, v in report.items() if not k.startswith("_") and not k.startswith("host:")}
    if not kernels:
        return None
    by = sum(v["bytes"] for v in kernels.values()) / prof_steps
    kms = sum(v["ms"] for v in kernels.values()) / prof_steps
    unacc = sum(v["ms"] for v in kernels.values() if v["bytes"] <= 0) / prof_steps
    achieved = by / (ms_step * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_step": by, "ms_per_step": ms_step,
            "kernel_ms_per_step_eager": round(kms, 3), "kernel_only_GBps": round(by / (kms * 1e-3) / 1e9, 1),
            "kernel_ms_without_bytes": round(unacc, 4)}


def traffic_from_profile():
    """HBM bytes per launch of the fine sweep from the committed counter summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    t = json.load(open(TRAFFIC_FILE))
    fs = t["copy_bytes_read"] / (1024.0 * t["copy_fetch_kib"])
    ws = t["copy_bytes_written"] / (1024.0 * t["copy_write_kib"])
    return 1024.0 * (t["sweep_fetch_kib"] * fs + t["sweep_write_kib"] * ws), {"FETCH_SIZE": fs, "WRITE_SIZE": ws}


def roofline(sweep):
    ms, by = sweep
    achieved = by / (ms * 1e-3) / 1e9
    traffic, scale = traffic_from_profile()
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_rbgs (head red-black half sweep, red colour over-relaxed, finest level: 257^3 grid)",
            "bytes_per_launch": by, "ms_per_launch": ms, "counter_scale": scale}


def timed_reverse(s, xb, vb):
    """Time adjoint_backward with the cotangents resident in HBM (uploaded to
    device buffers before the timed region, the gradients left there), the way
    the forward steps keep their state; then time the four host <-> device
    copies of the (3, N) arrays that a caller with host arrays adds."""
    import ctypes
    import numpy as np
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    h2d, d2h = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    vb = np.ascontiguousarray(vb, dtype=np.float64)
    nbytes = xb.nbytes
    bufs = []
    try:
        for _ in range(4):
            b = ctypes.c_void_p()
            if hip.hipMalloc(ctypes.byref(b), nbytes) != 0:
                raise RuntimeError("hipMalloc of the adjoint cotangents failed")
            bufs.append(b)
        gx, gv = np.empty_like(xb), np.empty_like(vb)

        def copy(dst, src, kind):
            if hip.hipMemcpy(dst, src, nbytes, kind) != 0:
                raise RuntimeError("hipMemcpy of the adjoint cotangents failed")

        copy(bufs[0], xb.ctypes.data, h2d)
        copy(bufs[1], vb.ctypes.data, h2d)
        s.sync()
        t = time.perf_counter()
        s.adjoint_backward_ptr(bufs[0].value, bufs[1].value, bufs[2].value, bufs[3].value)
        s.sync()
        t_ab = time.perf_counter() - t
        t = time.perf_counter()
        copy(bufs[0], xb.ctypes.data, h2d)
        copy(bufs[1], vb.ctypes.data, h2d)
        copy(gx.ctypes.data, bufs[2], d2h)
        copy(gv.ctypes.data, bufs[3], d2h)
        t_x = time.perf_counter() - t
    finally:
        for b in bufs:
            hip.hipFree(b)
    return t_ab, t_x, gx, gv


def adjoint_ratio(steps=50, every=10):
    """Head-only Plummer N = 2^18 on 64^3, FP32 mixed path."""
    import numpy as np
    from quantizer_amd import core
    from quantizer_amd.workloads import plummer, plummer_dt
    p = plummer(1 << 18, seed=20261203)
    dt = plummer_dt(0.01)
    with core.Simulation(p, 6, 6, precision=core.PREC_MIXED) as s:
        s.force()
        s.step(dt)
        s.sync()
        t = time.perf_counter()
        for _ in range(steps):
            s.step(dt)
        s.sync()
        t_fwd = time.perf_counter() - t
    with core.Simulation(p, 6, 6, precision=core.PREC_MIXED) as s:
        # steady state: a first run of the same length allocates the context's
        # checkpoints and step records, then the state is set back
        s.force()
        s.adjoint_forward(steps, dt, every)
        s.adjoint_backward(np.zeros((3, p.n)), np.zeros((3, p.n)))
        s.set_particles(p)
        s.force()
        s.sync()
        t = time.perf_counter()
        s.adjoint_forward(steps, dt, every)
        s.sync()
        t_af = time.perf_counter() - t
        xb = np.ones((3, p.n)) / p.n
        t_ab, t_x, _, _ = timed_reverse(s, xb, np.zeros_like(xb))
    return {"value": (t_af + t_ab) / t_fwd, "reverse_to_forward": t_ab / t_fwd,
            "forward_ms_per_step": 1e3 * t_fwd / steps,
            "checkpointed_forward_ms_per_step": 1e3 * t_af / steps, "reverse_ms_per_step": 1e3 * t_ab / steps,
            "host_array_copies_ms": 1e3 * t_x,
            "config": f"Plummer N=2^18, head 64^3 (lmin=6), {steps} steps, checkpoint every {every}, mixed"}


def adjoint_full_combination(n_each=1 << 23, lmin=8, lmax=11, crit=8, warm=40, steps=20, every=5):
    """Two Plummer spheres (2 x 2^23, radii 1 and 0.5, masses 2:1, 20a apart,
    closing at 0.5) in an 80a box, 256^3 head grid, 3 refinement levels, mixed
    path.  After `warm` forward steps: `steps` plain steps timed, then the same
    `steps` as a checkpointed forward (every `every`) plus the reverse sweep of
    the lighter sphere's final kinetic energy."""
    import numpy as np
    from quantizer_amd import core
    from quantizer_amd.workloads import two_plummer
    p = two_plummer(n_each, seed=20261203)
    w = np.where(np.arange(p.n) >= n_each, p.m, 0.0)
    with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit) as s:
        s.force()
        dt = s.timestep(0.5)
        for k in range(warm):
            s.step(dt)
            if k % 10 == 9:
                s.sync()
                log(f"    adjoint workload: {k + 1} of {warm} steps")
        s.sync()
        state = s.particles()
        st = s.stats()
    from quantizer_amd.workloads import Particles
    mid = Particles(*[np.ascontiguousarray(a) for a in state[:6]], p.m)
    with core.Simulation(mid, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit) as s:
        s.force()
        s.step(dt)
        s.sync()
        t = time.perf_counter()
        for _ in range(steps):
            s.step(dt)
        s.sync()
        t_fwd = time.perf_counter() - t
    log(f"    adjoint workload: forward {1e3 * t_fwd / steps:.2f} ms/step")
    with core.Simulation(mid, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit) as s:
        # steady state: a first run of the same length allocates the context's
        # checkpoints and step records, then the state is set back
        s.force()
        s.adjoint_forward(steps, dt, every)
        s.adjoint_backward(np.zeros((3, p.n)), np.zeros((3, p.n)))
        s.set_particles(mid)
        s.force()
        s.sync()
        t = time.perf_counter()
        s.adjoint_forward(steps, dt, every)
        s.sync()
        t_af = time.perf_counter() - t
        log(f"    adjoint workload: checkpointed forward {1e3 * t_af / steps:.2f} ms/step")
        vN = s.particles()[3:6]
        t_ab, t_x, gx, gv = timed_reverse(s, np.zeros_like(vN), w * vN)
        ok = bool(np.all(np.isfinite(gx)) and np.all(np.isfinite(gv)))
    return {"value": (t_af + t_ab) / t_fwd, "reverse_to_forward": t_ab / t_fwd,
            "forward_ms_per_step": 1e3 * t_fwd / steps,
            "checkpointed_forward_ms_per_step": 1e3 * t_af / steps, "reverse_ms_per_step": 1e3 * t_ab / steps,
            "host_array_copies_ms": 1e3 * t_x,
            "n_nodes": st["n_nodes"], "gradients_finite": ok, "dt": dt,
            "config": f"two Plummer spheres 2x2^{n_each.bit_length() - 1}, 80a box, head {1 << lmin}^3, "
                      f"lmax={lmax}, refine at >= {crit}; {warm} steps, then the adjoint of the last {steps} "
                      f"(checkpoint every {every}), loss = final kinetic energy of the lighter sphere, mixed"}


def precision_study(parts, lmin, lmax, crit, dt, steps=20, head_bc=0):
    """Mixed against the all-FP64 GPU path on the same workload: head potential
    and grid acceleration, particle accelerations after the first force pass,
    then the divergence of the trajectories over `steps` steps (dt None: the
    CFL 0.5 time step at every step, from the FP64 path and capped at its t = 0
    value, the same dt for both paths), and each path's relative drift of the
    total energy (grid-interpolated potential energy) over those steps."""
    import numpy as np
    from quantizer_amd import core

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))

    out = {}
    sims = [core.Simulation(parts, lmin, lmax, precision=p, ref_criterion_ptcl=crit, head_bc=head_bc)
            for p in (core.PREC_MIXED, core.PREC_FP64)]
    try:
        for s in sims:
            s.force()
        hm, hd = sims[0].head_fields(), sims[1].head_fields()
        pm, pd = sims[0].particles(), sims[1].particles()
        out["head_potential_rel_err"] = rel(hm[1], hd[1])
        out["head_acceleration_rel_err"] = max(rel(hm[k], hd[k]) for k in (2, 3, 4))
        out["particle_acceleration_rel_err"] = max(rel(pm[k], pd[k]) for k in (6, 7, 8))
        del hm, hd
        e0 = [s.observables(exact_pe=False)[2] for s in sims]
        dt0 = sims[1].timestep(cfl=0.5)
        for _ in range(steps):
            h = dt if dt is not None else sims[1].timestep(0.5, dt0)
            for s in sims:
                s.step(h)
        e1 = [s.observables(exact_pe=False)[2] for s in sims]
        pm, pd = sims[0].particles(), sims[1].particles()
        dx = np.sqrt(np.sum((pm[:3] - pd[:3]) ** 2, axis=0))
        cell = 1.0 / (1 << lmin)
        out[f"positions_after_{steps}_steps_max_head_cells"] = float(dx.max() / cell)
        out[f"positions_after_{steps}_steps_rms_head_cells"] = float(np.sqrt(np.mean(dx ** 2)) / cell)
        out[f"velocities_after_{steps}_steps_rel_err"] = max(rel(pm[k], pd[k]) for k in (3, 4, 5))
        out[f"energy_drift_{steps}_steps_mixed"] = float(abs(e1[0] - e0[0]) / abs(e0[0]))
        out[f"energy_drift_{steps}_steps_fp64"] = float(abs(e1[1] - e0[1]) / abs(e0[1]))
    finally:
        for s in sims:
            s.close()
    return out


def precision_configs():
    """The precision study on each configuration of the brief (mixed against the
    FP64 GPU path; the FP64 path itself is checked against the FP64 oracle by the
    GPU tests)."""
    from quantizer_amd import core
    from quantizer_amd.workloads import clumps, plummer, plummer_dt, two_plummer
    cases = [
        ("plain: Plummer 1e4, 32^3, zero faces, dt 0.01, 100 steps",
         lambda: plummer(10000), 5, 5, 8, core.BC_ZERO, plummer_dt(0.01), 100),
        ("AMR: Plummer 2^20, 128^3, lmax 9, dt 0.005, 200 steps",
         lambda: plummer(1 << 20), 7, 9, 8, core.BC_MONOPOLE, plummer_dt(0.005), 200),
        ("precision: 2^24 (half uniform, half 64 clumps), 256^3, lmax 9, CFL dt every step, 20 steps",
         lambda: clumps(1 << 24, 8, n_clumps=64), 8, 9, 8, core.BC_MONOPOLE, None, 20),
        ("differentiation: Plummer 2^18, 64^3, dt 0.01, 50 steps",
         lambda: plummer(1 << 18), 6, 6, 8, core.BC_MONOPOLE, plummer_dt(0.01), 50),
        ("full: two Plummer spheres 2 x 2^23, 80a box, 256^3, lmax 11, CFL dt every step, 40 steps",
         lambda: two_plummer(1 << 23), 8, 11, 8, core.BC_MONOPOLE, None, 40),
    ]
    res = {}
    for name, make, lmin, lmax, crit, bc, dt, steps in cases:
        t = time.time()
        res[name] = precision_study(make(), lmin, lmax, crit, dt, steps, head_bc=bc)
        log(f"  {name}: {res[name]} ({time.time() - t:.1f} s)")
    return res


def cpu_model():
    """The host CPU's model name, family and model numbers (lscpu, else
    /proc/cpuinfo)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        if "Model name" in kv:
            extra = ", ".join(f"{k.lower()} {kv[k]}" for k in ("CPU family", "Model", "Stepping") if k in kv)
            return kv["Model name"] + (f" ({extra})" if extra else "")
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Threads of this process's CPU share: the affinity mask, capped by
    OMP_NUM_THREADS when the environment sets it."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def time_harness(binary, parts, lmin, lmax, crit, dt, threads, steps, budget_s):
    """Run an oracle/_ref harness build with --time --cfl-steps: the initial
    force pass, then `steps` steps (timestep() with dt as its cap, then the
    step) of which the first is a warm-up; returns its REF_TIMING fields, or a
    dict with "error"."""
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "p.bin")
        parts.write_bin(inp)
        cmd = [binary, inp, td, str(lmin), str(lmax), str(crit), str(steps), repr(float(dt)), "--time",
               "--cfl-steps", "--noenergy", "--threads", str(threads)]
        t = time.time()
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=budget_s, cwd=td)
        except subprocess.TimeoutExpired:
            return {"error": f"timed out after {budget_s}s"}
        line = [l for l in p.stderr.splitlines() if l.startswith("REF_TIMING")]
        if p.returncode != 0 or not line:
            return {"error": f"exit {p.returncode}: {p.stderr[-300:]}"}
        kv = dict(t.split("=") for t in line[0].split()[1:])
        kv["wall_s"] = time.time() - t
        return kv


def cpu_baseline(parts, lmin, lmax, crit, dt, budget_s, steps=4):
    """The compiled reference (oracle/_ref/ref_harness, the reference's own
    sources) on this workload at 1 thread (the reference default,
    global_variables.c:250) and at the host's thread count: the initial force
    pass, one warm-up step, then steps - 1 timed steps."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(ref):
        return None
    res = {}
    for threads in sorted({1, host_threads()}):
        kv = time_harness(ref, parts, lmin, lmax, crit, dt, threads, steps, budget_s)
        if "error" in kv:
            res[threads] = {"error": kv["error"]}
            continue
        sps = float(kv["s_per_step"])
        res[threads] = {"value": parts.n / sps, "s_per_step": sps, "timed_steps": int(kv["timed"]),
                        "first_force_s": float(kv["first_force_s"]), "last_dt": float(kv["last_dt"])}
        log(f"  cpu reference, {threads} thread(s): {sps:.2f} s/step")
    best = max((t for t in res if "value" in res[t]), default=None)
    if best is None:
        return {"value": None, "unit": "particle-steps/s", "cores": None, "kind": "reference",
                "sample": f"failed: {res}"}
    return {"value": res[best]["value"], "unit": "particle-steps/s", "cores": best, "kind": "reference",
            "cpu_model": cpu_model(), "by_threads": {str(t): v for t, v in res.items()},
            "sample": f"compiled reference on the full workload (N={parts.n}, lmin={lmin}, lmax={lmax}, CFL 0.5 "
                      f"time step at every step): initial force pass, 1 warm-up step, {steps - 1} timed steps, "
                      f"at 1 and {best} threads"}


def resident_dropin(parts, lmin, lmax, crit, dt, budget_s, steps=4):
    """The reference's host program linked against libqzdropin_resident.so
    (oracle/_ref/ref_harness_resident): its step loop with every stage on this
    GPU, host calls included; 1 warm-up step, steps - 1 timed steps.  The
    host-side set-up (reading the input, the reference's tree construction) is
    outside the timed steps."""
    binary = os.path.join(ROOT, "oracle", "_ref", "ref_harness_resident")
    if not os.path.exists(binary):
        return None
    kv = time_harness(binary, parts, lmin, lmax, crit, dt, 1, steps, budget_s)
    if "error" in kv:
        return {"value": None, "error": kv["error"]}
    sps = float(kv["s_per_step"])
    return {"value": parts.n / sps, "unit": "particle-steps/s", "ms_per_step": 1e3 * sps,
            "timed_steps": int(kv["timed"]), "precision": "fp64", "last_dt": float(kv["last_dt"]),
            "sample": f"reference host program + resident drop-in, N={parts.n}, lmin={lmin}, lmax={lmax}, "
                      f"FP64 (the drop-in default), 1 warm-up and {steps - 1} timed steps"}


def pmc_probe(args):
    """Short run for hardware counters: one force pass and step of the headline
    workload, a few fine sweeps, and one plain copy of 512 MiB."""
    from quantizer_amd import core
    N, lmin, lmax, crit, kind = WORKLOADS[args.workload]
    parts = make_particles(kind, N, lmin)
    with core.Simulation(parts, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit) as s:
        s.force()
        s.step(s.timestep(cfl=0.5))
        ms, by = s.bench_sweep(3)
        cms = s.copy_probe(512 << 20)
    print(json.dumps({"sweep_ms": ms, "sweep_bytes": by, "copy_bytes": 2 * (512 << 20), "copy_ms": cms}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--sweep-reps", type=int, default=50)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-precision", action="store_true")
    ap.add_argument("--skip-adjoint-amr", action="store_true")
    ap.add_argument("--skip-resident", action="store_true")
    ap.add_argument("--skip-fixed-dt", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=420.0)
    ap.add_argument("--cpu-n", type=int, default=1 << 24, help="particles in the CPU sample")
    ap.add_argument("--inner", type=int, default=0)
    ap.add_argument("--branch-solver", choices=["cg", "mg"], default=None,
                    help="refinement-node solver (default: the library default)")
    ap.add_argument("--pmc-probe", action="store_true", help="short run for rocprofv3 --pmc passes")
    ap.add_argument("--precision-configs", action="store_true",
                    help="mixed against FP64 on every configuration; prints one JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.precision_configs:
        print(json.dumps(precision_configs()), flush=True)
        return
    from quantizer_amd import core
    N, lmin, lmax, crit, kind = WORKLOADS[args.workload]
    solver = {None: None, "cg": core.BRANCH_CG, "mg": core.BRANCH_MG}[args.branch_solver]
    log(f"workload {args.workload}: N={N} lmin={lmin} lmax={lmax} crit={crit} {kind}")
    t = time.time()
    parts = make_particles(kind, N, lmin)
    log(f"  generated in {time.time() - t:.1f}s")

    el_mx, (dt, dt_last), st_mx, rep_mx, sweep = run_path(parts, lmin, lmax, crit, core.PREC_MIXED, args.steps, args.warmup,
                                               args.prof_steps, args.inner, solver, args.sweep_reps)
    value = N * args.steps / el_mx
    ms_step = 1e3 * el_mx / args.steps
    log(f"  mixed: {ms_step:.3f} ms/step, {value:.4e} particle-steps/s, last dt {dt_last:.4e}")
    fp64 = None
    if not args.skip_fp64:
        el64, _, st64, _, _ = run_path(parts, lmin, lmax, crit, core.PREC_FP64, args.steps, args.warmup, 0,
                                       args.inner, solver)
        fp64 = {"value": N * args.steps / el64, "ms_per_step": 1e3 * el64 / args.steps,
                "head_vcycles": st64["head_vcycles"], "branch_iterations": st64["branch_iterations"]}
        log(f"  fp64: {fp64['ms_per_step']:.3f} ms/step")
    fixed = None
    if not args.skip_fixed_dt:
        fixed = fixed_dt_leg(parts, lmin, lmax, crit, args.steps, args.warmup)
        log(f"  mixed, dt fixed at its t=0 value: {fixed['ms_per_step']:.3f} ms/step")
    roof = roofline(sweep) if sweep else None
    if roof:
        log(f"  roofline: {roof['achieved']} GB/s ({roof['frac']:.3f} of peak), traffic {roof['traffic']}")
    prec = None
    if not args.skip_precision and not args.skip_fp64:
        prec = precision_study(parts, lmin, lmax, crit, None)
        log(f"  precision study (mixed vs FP64): {prec}")
    adj = None
    if not args.skip_adjoint:
        adj = adjoint_ratio()
        log(f"  adjoint/forward: {adj['value']:.3f} ({adj})")
    adj_amr = None
    if not args.skip_adjoint_amr:
        adj_amr = adjoint_full_combination()
        log(f"  adjoint/forward with refinement: {adj_amr['value']:.3f} ({adj_amr})")
    resident = None
    if not args.skip_resident:
        resident = resident_dropin(parts, lmin, lmax, crit, dt, args.cpu_budget)
        if resident and resident.get("value"):
            log(f"  resident drop-in (host program + GPU): {resident['ms_per_step']:.2f} ms/step")
    cpu = None
    if not args.skip_cpu:
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
                               f"refine at >= {crit} particles/cell, tol 1e-10, CFL 0.5 time step at every step "
                               f"(cap {dt:.4e}, the t=0 value)",
                   "last_timed_dt": dt_last,
                   "n_particles": N, "lmin": lmin, "lmax": lmax, "parallelism": "single GPU"},
        "fp64_plain": fp64,
        "mixed_fixed_dt": fixed,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "precision_vs_fp64": prec,
        "adjoint_fwd_ratio": adj["value"] if adj else None,
        "adjoint": adj,
        "adjoint_amr": adj_amr,
        "solver": {k: st_mx[k] for k in ("head_outer", "head_vcycles", "head_residual", "n_nodes", "branch_points",
                                         "branch_intr", "branch_iterations", "branch_cg_levels")},
        "roofline": roof,
        "roofline_step": step_roofline(rep_mx, max(1, args.prof_steps), ms_step) if rep_mx else None,
        "kernels": kernel_table(rep_mx, max(1, args.prof_steps)) if rep_mx else None,
        "dropin_resident": resident,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
