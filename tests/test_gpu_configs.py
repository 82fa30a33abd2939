"""The five workloads of the step at their own sizes.

* AMR (Plummer 2^20, 128^3 head, lmax 9, >= 8 per cell, dt 0.005): one force
  pass and 3 steps against the FP64 oracle: node geometry bit-exact; rho, phi
  and a at every listed node point; head fields; particle accelerations;
  positions and velocities after the steps.  This size reaches the fused node
  multigrid levels (k_bmg_down / k_bmg_up) and the node set-up run ahead on
  its host thread, so those paths meet the oracle here.
* Precision study (2^24: half uniform, half in 64 clumps; 256^3 head, lmax 9):
  one force pass against the oracle, both precisions.
* Differentiation (Plummer 2^18, 64^3, dt 0.01): the checkpointed gradient
  against the oracle's FP64 reverse mode over 20 steps (checkpoints every 10),
  and central differences at eps = 1e-3 a over the full 50 steps.
* Full combination (two Plummer spheres 2 x 2^23, 80a box, 256^3, lmax 11):
  property checks against the GPU's own forward after 40 steps (mass and
  momentum; directional central differences of the lighter sphere's final
  kinetic energy over the 20-step adjoint window, checkpoints every 5).  The
  oracle takes more than a minute for one force pass at this size.

Each test prints the errors it measured ("errors:" lines).
"""
import time

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import Particles, clumps, plummer, plummer_dt, two_plummer

pytestmark = pytest.mark.gpu

SEED = 20261203
GEOM = ["lv", "lvrel", "idx", "parent", "ts_x", "ts_y", "ts_z", "rd_x", "rd_y", "rd_z", "min_x", "min_y", "min_z",
        "max_x", "max_y", "max_z", "ncell", "nintr", "nbdry", "nsim", "nptcl_full", "nptcl_outs", "nchild", "gcap"]
PREC_NAME = {core.PREC_FP64: "fp64", core.PREC_MIXED: "mixed"}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def listed_index(node, kind):
    xyz = node[kind + "_xyz"]
    X = node["rd_x"] + 1
    XY = X * (node["rd_y"] + 1)
    return (xyz[:, 0] - node["ts_x"]) + (xyz[:, 1] - node["ts_y"]) * X + (xyz[:, 2] - node["ts_z"]) * XY


def node_errors(gn, on):
    """Geometry must match bit for bit; returns the max relative errors of rho,
    phi and a over every listed point of every node."""
    assert len(gn) == len(on) and len(on) > 0
    err = {"rho": 0.0, "pot": 0.0, "a": 0.0}
    for a, b in zip(gn, on):
        for k in GEOM:
            assert a[k] == b[k], (k, a[k], b[k])
        for kind in ("intr", "bdry", "sim"):
            assert np.array_equal(a[kind + "_xyz"], b[kind + "_xyz"]), kind
        idx = np.concatenate([listed_index(b, k) for k in ("intr", "bdry", "sim")])
        err["rho"] = max(err["rho"], rel(a["rho"][idx], b["rho"][idx]))
        err["pot"] = max(err["pot"], rel(a["pot"][idx], b["pot"][idx]))
        for f in ("ax", "ay", "az"):
            err["a"] = max(err["a"], rel(a[f][idx], b[f][idx]))
    return err


# ------------------------------------------------------------------ AMR, 2^20
C2 = dict(n=1 << 20, lmin=7, lmax=9, crit=8, dt=plummer_dt(0.005), steps=3)


@pytest.fixture(scope="module")
def c2_oracle(oracle):
    t = time.time()
    p = plummer(C2["n"], seed=SEED)
    o = oracle.OracleSim(p, C2["lmin"], C2["lmax"], C2["crit"])
    o.force()
    ref = {"nodes": o.nodes(), "head": o.head(), "force": o.particles()}
    for _ in range(C2["steps"]):
        o.step(C2["dt"])
    ref["steps"] = o.particles()
    o.close()
    print(f"  oracle AMR 2^20: {time.time() - t:.1f} s", flush=True)
    return p, ref


@pytest.mark.parametrize("prec,bars", [
    (core.PREC_FP64, dict(rho=1e-12, pot=1e-10, a=1e-9, head_pot=1e-10, head_a=1e-9, part_a=1e-9, x=1e-10, v=1e-8)),
    (core.PREC_MIXED, dict(rho=1e-6, pot=1e-6, a=1e-6, head_pot=1e-6, head_a=1e-6, part_a=1e-6, x=1e-6, v=1e-4)),
])
def test_amr_2e20_force_and_steps_vs_oracle(c2_oracle, prec, bars):
    print(f"case: AMR Plummer 2^20 lmin 7 lmax 9 prec={PREC_NAME[prec]}", flush=True)
    p, ref = c2_oracle
    with core.Simulation(p, C2["lmin"], C2["lmax"], precision=prec, ref_criterion_ptcl=C2["crit"]) as s:
        s.force()
        gn, hg, pf, st = s.nodes(), s.head_fields(), s.particles(), s.stats()
        for _ in range(C2["steps"]):
            s.step(C2["dt"])
        ps = s.particles()
    assert st["branch_cg_levels"] == 0, st
    err = node_errors(gn, ref["nodes"])
    ho = ref["head"]
    err["head_rho"] = rel(hg[0], ho[0])
    err["head_pot"] = rel(hg[1], ho[1])
    err["head_a"] = max(rel(hg[k], ho[k]) for k in (2, 3, 4))
    err["part_a"] = max(rel(pf[k], ref["force"][k]) for k in (6, 7, 8))
    err["x"] = float(np.max(np.abs(ps[:3] - ref["steps"][:3])))
    err["v"] = max(rel(ps[k], ref["steps"][k]) for k in (3, 4, 5))
    print(f"errors: AMR 2^20 {PREC_NAME[prec]}: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["rho"] < bars["rho"] and err["head_rho"] < bars["rho"]
    for k in ("pot", "a", "head_pot", "head_a", "part_a", "x", "v"):
        assert err[k] < bars[k], (k, err[k])


def test_amr_2e20_reaches_fused_node_levels():
    """The configuration the oracle checks above runs the fused node levels, and
    the default path (graphs, set-up run ahead) gives what eager launches with
    the set-up in place give, bit for bit."""
    print("case: AMR Plummer 2^20 fused node levels reached", flush=True)
    p = plummer(C2["n"], seed=SEED)
    out = []
    for eager in (False, True):
        with core.Simulation(p, C2["lmin"], C2["lmax"], precision=core.PREC_FP64,
                             ref_criterion_ptcl=C2["crit"]) as s:
            s.timing(eager)
            s.force()
            out.append(s.particles())
            if eager:
                rep = s.timing_report()
            s.timing(False)
    assert rep.get("bmg_down", {}).get("launches", 0) > 0, sorted(rep)
    assert rep.get("bmg_up", {}).get("launches", 0) > 0
    np.testing.assert_array_equal(out[0], out[1])


# ------------------------------------------------------- precision study, 2^24
C3 = dict(n=1 << 24, lmin=8, lmax=9, crit=8)


@pytest.fixture(scope="module")
def c3_oracle(oracle):
    t = time.time()
    p = clumps(C3["n"], C3["lmin"], n_clumps=64, seed=SEED)
    o = oracle.OracleSim(p, C3["lmin"], C3["lmax"], C3["crit"])
    o.force()
    ref = {"nodes": o.nodes(), "head": o.head(), "force": o.particles()[6:9].copy()}
    o.close()
    print(f"  oracle 2^24 clumps force pass: {time.time() - t:.1f} s", flush=True)
    return p, ref


@pytest.mark.parametrize("prec,bars", [
    (core.PREC_FP64, dict(rho=1e-12, pot=1e-10, a=1e-9, head_pot=1e-10, head_a=1e-9, part_a=1e-9)),
    (core.PREC_MIXED, dict(rho=1e-6, pot=1e-6, a=1e-6, head_pot=1e-6, head_a=1e-6, part_a=1e-6)),
])
def test_precision_2e24_force_vs_oracle(c3_oracle, prec, bars):
    print(f"case: 2^24 clumps lmin 8 lmax 9 force pass prec={PREC_NAME[prec]}", flush=True)
    p, ref = c3_oracle
    with core.Simulation(p, C3["lmin"], C3["lmax"], precision=prec, ref_criterion_ptcl=C3["crit"]) as s:
        s.force()
        gn, hg, pf, st = s.nodes(), s.head_fields(), s.particles(), s.stats()
    assert st["branch_cg_levels"] == 0, st
    err = node_errors(gn, ref["nodes"])
    ho = ref["head"]
    err["head_rho"] = rel(hg[0], ho[0])
    err["head_pot"] = rel(hg[1], ho[1])
    err["head_a"] = max(rel(hg[k], ho[k]) for k in (2, 3, 4))
    err["part_a"] = max(rel(pf[6 + k], ref["force"][k]) for k in range(3))
    print(f"errors: 2^24 clumps {PREC_NAME[prec]}: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["rho"] < bars["rho"] and err["head_rho"] < bars["rho"]
    for k in ("pot", "a", "head_pot", "head_a", "part_a"):
        assert err[k] < bars[k], (k, err[k])


# ----------------------------------------------------- differentiation, 2^18
C4 = dict(n=1 << 18, lmin=6, dt=plummer_dt(0.01), every=10)


def c4_target(p, nsteps):
    """The same run from positions perturbed by N(0, 0.01 a) (a 40 a box)."""
    rng = np.random.default_rng(SEED + 1)
    t = p.copy()
    for a in (t.x, t.y, t.z):
        a += rng.normal(0.0, 0.01 / 40, p.n)
    with core.Simulation(t, C4["lmin"], C4["lmin"], precision=core.PREC_FP64) as s:
        s.force()
        for _ in range(nsteps):
            s.step(C4["dt"])
        return s.particles()[:3]


@pytest.fixture(scope="module")
def c4_setup():
    p = plummer(C4["n"], seed=SEED)
    return p, {20: c4_target(p, 20), 50: c4_target(p, 50)}


@pytest.fixture(scope="module")
def c4_oracle(oracle, c4_setup):
    """The oracle's FP64 reverse mode over 20 steps; the seed is the FP64 GPU
    run's final-position cotangent, so both precisions are checked on one seed."""
    p, targets = c4_setup
    with core.Simulation(p, C4["lmin"], C4["lmin"], precision=core.PREC_FP64) as s:
        s.force()
        for _ in range(20):
            s.step(C4["dt"])
        xN = s.particles()[:3]
    xbN = 2.0 * (xN - targets[20]) / p.n
    t = time.time()
    o = oracle.OracleSim(p, C4["lmin"], C4["lmin"])
    o.force()
    wx, wv = o.adjoint(20, C4["dt"], xbN, np.zeros_like(xbN))
    o.close()
    print(f"  oracle reverse mode, 2^18, 20 steps: {time.time() - t:.1f} s", flush=True)
    return xbN, wx, wv


@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-7), (core.PREC_MIXED, 5e-3)])
def test_differentiation_2e18_gradient_vs_oracle_reverse_mode(c4_setup, c4_oracle, prec, tol):
    print(f"case: 2^18 64^3 adjoint 20 steps ckpt 10 prec={PREC_NAME[prec]}", flush=True)
    p, _ = c4_setup
    xbN, wx, wv = c4_oracle
    with core.Simulation(p, C4["lmin"], C4["lmin"], precision=prec) as s:
        s.force()
        s.adjoint_forward(20, C4["dt"], C4["every"])
        gx, gv = s.adjoint_backward(xbN, np.zeros_like(xbN))
    ex, ev = rel(gx, wx), rel(gv, wv)
    print(f"errors: 2^18 gradient {PREC_NAME[prec]} vs oracle reverse mode: x {ex:.2e}, v {ev:.2e}", flush=True)
    assert ex < tol and ev < tol


def test_differentiation_2e18_central_differences_eps_1e3a(c4_setup):
    """Central differences of L = mean |x_50 - x_target|^2 at eps = 1e-3 a (and
    1e-3 Plummer velocity units), 64 seeded coordinates, FP64 forward, against
    the adjoint with checkpoints every 10 over the 50 steps.

    CIC weights have kinks at cell faces: a pair of perturbed runs whose
    particle sits in different head cells at some step differences across a
    kink.  Every component outside that band must agree to 1e-5; a component
    that misses it must be shown to be in the band (its pair of runs is
    replayed and the perturbed particle's cell compared step by step)."""
    print("case: 2^18 64^3 central differences eps 1e-3 a, 50 steps, fp64", flush=True)
    p, targets = c4_setup
    target, nsteps = targets[50], 50
    lmin, dt = C4["lmin"], C4["dt"]
    eps_x, eps_v = 1e-3 / 40.0, 1e-3 * np.sqrt(40.0)
    names = ["x", "y", "z", "vx", "vy", "vz"]
    with core.Simulation(p, lmin, lmin, precision=core.PREC_FP64) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, C4["every"])
        xN = s.particles()[:3]
        gx, gv = s.adjoint_backward(2.0 * (xN - target) / p.n, np.zeros((3, p.n)))
    sim = core.Simulation(p, lmin, lmin, precision=core.PREC_FP64)

    def run(parts, track=None):
        sim.set_particles(parts)
        sim.force()
        cells = []
        for _ in range(nsteps):
            sim.step(dt)
            if track is not None:
                cells.append(np.floor(sim.particles()[:3, track] * (1 << lmin)).astype(int))
        x = sim.particles()[:3]
        return float(np.mean(np.sum((x - target) ** 2, axis=0))), cells

    try:
        base, _ = run(p)
        assert base == float(np.mean(np.sum((xN - target) ** 2, axis=0)))  # a reused context replays exactly
        rng = np.random.default_rng(SEED + 2)
        worst_out, in_band = 0.0, []
        for _ in range(64):
            k, i = int(rng.integers(0, 6)), int(rng.integers(0, p.n))
            eps = eps_x if k < 3 else eps_v
            plus, minus = p.copy(), p.copy()
            getattr(plus, names[k])[i] += eps
            getattr(minus, names[k])[i] -= eps
            fd = (run(plus)[0] - run(minus)[0]) / (2 * eps)
            g = float((gx if k < 3 else gv)[k % 3, i])
            err = abs(fd - g) / max(abs(g), 1e-30)
            if err > 1e-5:
                cp, cm = run(plus, track=i)[1], run(minus, track=i)[1]
                crossed = any(np.any(a != b) for a, b in zip(cp, cm))
                assert crossed, (names[k], i, fd, g, err)  # a miss outside the band is an error
                in_band.append((names[k], i, err))
            else:
                worst_out = max(worst_out, err)
    finally:
        sim.close()
    print(f"errors: 2^18 central differences eps 1e-3 a: max rel {worst_out:.2e} over "
          f"{64 - len(in_band)} components; {len(in_band)} in the cell-crossing band {in_band}", flush=True)
    assert len(in_band) <= 8


# ---------------------------------------------------- full combination, 2^24
C5 = dict(n_each=1 << 23, lmin=8, lmax=11, crit=8, warm=40, window=20, every=5)


def test_full_combination_2e24_properties():
    """Two Plummer spheres, 3 refinement levels, mixed path.  After 40 steps:
    total mass and the deposited mass; total momentum against its scale; then
    the adjoint of the lighter sphere's final kinetic energy over 20 steps
    (checkpoints every 5) against central differences of the GPU's own
    forward along 8 directions: bulk x shifts of either sphere's positions
    and velocities (the approach axis; bulk y and z shifts have zero
    derivative by symmetry), and smooth random displacement and velocity
    fields over all particles and over the lighter sphere.

    One particle's coordinate moves this loss by about 1e-13 of itself, below
    what the FP32 forward resolves, so the check runs along directions that
    move many particles.  The forward's own noise on L is about 1e-8 of L
    (FP32 fields, and refinement flags that flip for cells at the threshold
    when every particle shifts), so a central difference that changes L by a
    fraction r carries a relative error up to 1e-8 / r: the bar is 1e-3 + 1e-8 / r."""
    print("case: two spheres 2 x 2^23 lmin 8 lmax 11, 40 steps + adjoint of 20, mixed", flush=True)
    n_each = C5["n_each"]
    p = two_plummer(n_each, seed=SEED)
    lmin, lmax, crit = C5["lmin"], C5["lmax"], C5["crit"]
    w = np.where(np.arange(p.n) >= n_each, p.m, 0.0)  # the lighter sphere
    H = 1.0 / (1 << lmin)
    kw = dict(precision=core.PREC_MIXED, ref_criterion_ptcl=crit)
    with core.Simulation(p, lmin, lmax, **kw) as s:
        s.force()
        dt = s.timestep(0.5)
        head = s.head_fields()
        mass_dep = float(np.sum(head[0], dtype=np.float64)) * H ** 3 / (4.0 * np.pi)
        for _ in range(C5["warm"]):
            s.step(dt)
        st = s.particles()
        assert s.stats()["n_levels"] == 4
    mid = Particles(*[np.ascontiguousarray(a) for a in st[:6]], p.m.copy())
    mom = np.array([np.sum(p.m * st[3 + k]) for k in range(3)])
    scale = float(np.sum(p.m * np.sqrt(np.sum(st[3:6] ** 2, axis=0))))
    err = {"deposited_mass": abs(mass_dep - 1.0), "momentum": float(np.max(np.abs(mom))) / scale}

    def loss(parts, sim):
        sim.set_particles(parts)
        sim.force()
        for _ in range(C5["window"]):
            sim.step(dt)
        v = sim.particles()[3:6]
        return float(0.5 * np.sum(w * np.sum(v * v, axis=0)))

    with core.Simulation(mid, lmin, lmax, **kw) as s:
        s.force()
        s.adjoint_forward(C5["window"], dt, C5["every"])
        vN = s.particles()[3:6]
        gx, gv = s.adjoint_backward(np.zeros_like(vN), w * vN)
        L0 = float(0.5 * np.sum(w * np.sum(vN * vN, axis=0)))
        assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gv))
        two = np.arange(p.n) >= n_each
        rng = np.random.default_rng(SEED + 3)
        pos = np.stack([mid.x, mid.y, mid.z])
        eps_x, eps_v = 1e-3 / 80.0, 1e-3 * np.sqrt(80.0)

        def smooth_field():
            k = rng.normal(0.0, 2 * np.pi * 4, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            d = np.zeros((3, p.n))
            d[int(rng.integers(0, 3))] = np.cos(k @ pos + ph)
            return d

        dirs = []
        for sel, kind in ((two, "x"), (two, "v"), (~two, "x"), (~two, "v")):
            d = np.zeros((3, p.n))
            d[0, sel] = 1.0
            dirs.append((f"bulk {'lighter' if sel is two else 'heavier'} {kind}x", kind, d))
        dirs.append(("smooth field x", "x", smooth_field()))
        dirs.append(("smooth field v", "v", smooth_field()))
        dirs.append(("smooth field x, lighter sphere", "x", smooth_field() * two))
        dirs.append(("smooth field v, lighter sphere", "v", smooth_field() * two))
        worst = 0.0
        for name, kind, d in dirs:
            eps = eps_x if kind == "x" else eps_v
            g = float(np.sum((gx if kind == "x" else gv) * d))
            vals = []
            for sgn in (1.0, -1.0):
                q = mid.copy()
                for ax, arr in enumerate((q.x, q.y, q.z) if kind == "x" else (q.vx, q.vy, q.vz)):
                    arr += sgn * eps * d[ax]
                vals.append(loss(q, s))
            fd = (vals[0] - vals[1]) / (2 * eps)
            e = abs(fd - g) / max(abs(g), 1e-30)
            r = abs(vals[0] - vals[1]) / L0
            print(f"  direction {name}: adjoint {g:.6e}, central difference {fd:.6e}, rel {e:.2e}, "
                  f"relative loss change {r:.2e}", flush=True)
            assert e < 1e-3 + 1e-8 / r, name
            worst = max(worst, e / (1e-3 + 1e-8 / r))
        err["directional_derivatives_over_bar"] = worst
    print("errors: two spheres mixed: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["deposited_mass"] < 1e-6
    assert err["momentum"] < 1e-5
