"""Refinement nodes on the GPU vs the oracle: bit-exact geometry, fields at the
listed grid points, particle accelerations and short trajectories."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt, clumps

pytestmark = pytest.mark.gpu

GEOM = ["lv", "lvrel", "idx", "parent", "ts_x", "ts_y", "ts_z", "rd_x", "rd_y", "rd_z", "min_x", "min_y", "min_z",
        "max_x", "max_y", "max_z", "ncell", "nintr", "nbdry", "nsim", "nptcl_full", "nptcl_outs", "nchild", "gcap"]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def listed_index(node, kind):
    xyz = node[kind + "_xyz"]
    gx = xyz[:, 0] - node["ts_x"]
    gy = xyz[:, 1] - node["ts_y"]
    gz = xyz[:, 2] - node["ts_z"]
    X = node["rd_x"] + 1
    XY = X * (node["rd_y"] + 1)
    return gx + gy * X + gz * XY


CASES = [
    ("plummer", 6000, 4, 6, 8),
    ("plummer", 20000, 5, 8, 4),
    ("clumps", 30000, 5, 7, 8),
]


def make(kind, n, lmin):
    if kind == "plummer":
        return plummer(n, seed=21)
    return clumps(n, lmin, n_clumps=6, seed=21)


SOLVERS = [core.BRANCH_CG, core.BRANCH_MG]


def check_solver(st, solver):
    """The multigrid path must finish every tree level without the CG hand-off."""
    if solver == core.BRANCH_MG:
        assert st["branch_cg_levels"] == 0, st


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("kind,n,lmin,lmax,crit", CASES)
def test_amr_geometry_and_fields_fp64(oracle, kind, n, lmin, lmax, crit, solver):
    print(f"case: amr fp64 {kind} N={n} lmin={lmin} lmax={lmax} crit={crit} solver={solver}", flush=True)
    p = make(kind, n, lmin)
    o = oracle.OracleSim(p, lmin, lmax, crit)
    o.force()
    on = o.nodes()
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, ref_criterion_ptcl=crit,
                         branch_solver=solver) as s:
        s.force()
        gn = s.nodes()
        hg = s.head_fields()
        pg = s.particles()
        st = s.stats()
    assert len(gn) == len(on) and len(on) > 0
    for a, b in zip(gn, on):
        for k in GEOM:
            assert a[k] == b[k], (k, a[k], b[k])
        for kind_ in ("intr", "bdry", "sim"):
            assert np.array_equal(a[kind_ + "_xyz"], b[kind_ + "_xyz"]), kind_
        idx = np.concatenate([listed_index(b, k) for k in ("intr", "bdry", "sim")])
        assert rel(a["rho"][idx], b["rho"][idx]) < 1e-12
        assert rel(a["pot"][idx], b["pot"][idx]) < 1e-10
        for f in ("ax", "ay", "az"):
            assert rel(a[f][idx], b[f][idx]) < 1e-9
    ho = o.head()
    assert rel(hg[0], ho[0]) < 1e-12
    assert rel(hg[1], ho[1]) < 1e-10
    po = o.particles()
    for k in (6, 7, 8):
        assert rel(pg[k], po[k]) < 1e-9
    assert st["n_nodes"] == len(on)
    check_solver(st, solver)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("kind,n,lmin,lmax,crit", CASES)
def test_amr_mixed_precision_fields(oracle, kind, n, lmin, lmax, crit, solver):
    print(f"case: amr mixed {kind} N={n} lmin={lmin} lmax={lmax} solver={solver}", flush=True)
    p = make(kind, n, lmin)
    o = oracle.OracleSim(p, lmin, lmax, crit)
    o.force()
    on = o.nodes()
    with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=crit,
                         branch_solver=solver) as s:
        s.force()
        gn = s.nodes()
        pg = s.particles()
        check_solver(s.stats(), solver)
    assert len(gn) == len(on)
    for a, b in zip(gn, on):
        idx = np.concatenate([listed_index(b, k) for k in ("intr", "bdry", "sim")])
        # FP32 fields after the FP64 refinement: 1e-5 is the bar; measured about 1e-7
        assert rel(a["rho"][idx], b["rho"][idx]) < 1e-6
        assert rel(a["pot"][idx], b["pot"][idx]) < 1e-6
        for f in ("ax", "ay", "az"):
            assert rel(a[f][idx], b[f][idx]) < 1e-6
    po = o.particles()
    for k in (6, 7, 8):
        assert rel(pg[k], po[k]) < 1e-6


@pytest.mark.parametrize("solver", SOLVERS)
def test_amr_steps_fp64(oracle, solver):
    print(f"case: amr fp64 3 steps plummer N=8000 lmin=4 lmax=6 solver={solver}", flush=True)
    p = plummer(8000, seed=9)
    dt = plummer_dt(0.005)
    o = oracle.OracleSim(p, 4, 6, 8)
    o.force()
    with core.Simulation(p, 4, 6, precision=core.PREC_FP64, ref_criterion_ptcl=8, branch_solver=solver) as s:
        s.force()
        for _ in range(3):
            o.step(dt)
            s.step(dt)
        pg = s.particles()
    po = o.particles()
    assert np.max(np.abs(pg[:3] - po[:3])) < 1e-10
    for k in range(3, 9):
        assert rel(pg[k], po[k]) < 1e-6


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("prec", [core.PREC_FP64, core.PREC_MIXED])
def test_graph_replays_match_eager_launches(prec, solver):
    """The head solve and the branch solvers replay captured HIP graphs; with
    timing on the same kernels are launched eagerly.  Several steps (many
    replays, with host work and syncs between them) must agree bit for bit."""
    print(f"case: graph vs eager prec={prec} solver={solver} clumps N=20000 lmin=5 lmax=7", flush=True)
    p = clumps(20000, 5, n_clumps=4, seed=3)
    out = []
    for eager in (False, True):
        with core.Simulation(p, 5, 7, precision=prec, ref_criterion_ptcl=8, branch_solver=solver) as s:
            s.timing(eager)
            s.force()
            dt = s.timestep()
            for _ in range(4):
                s.step(dt)
            out.append((s.particles(), s.head_fields()))
            s.timing(False)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
