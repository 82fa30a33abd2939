"""Fused node multigrid levels (k_bmg_down / k_bmg_up) against the per-level
node kernels they replace (QZ_NODE_MG_FUSED=0): same PCG iterations, node
fields and particle accelerations agree to rounding."""
import os

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run(p, lmin, lmax, prec, fused, steps):
    old = os.environ.get("QZ_NODE_MG_FUSED")
    os.environ["QZ_NODE_MG_FUSED"] = "1" if fused else "0"
    try:
        with core.Simulation(p, lmin, lmax, precision=prec, ref_criterion_ptcl=8,
                             branch_solver=core.BRANCH_MG) as s:
            s.timing(True)  # eager launches, per-kernel launch counts
            s.force()
            dt = plummer_dt(0.005)
            for _ in range(steps):
                s.step(dt)
            rep = s.timing_report()
            return s.nodes(), s.particles(), s.stats(), rep
    finally:
        if old is None:
            del os.environ["QZ_NODE_MG_FUSED"]
        else:
            os.environ["QZ_NODE_MG_FUSED"] = old


@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-12), (core.PREC_MIXED, 1e-6)])
def test_fused_node_levels_match_per_level_kernels(prec, tol):
    print(f"case: fused vs per-level node levels prec={prec} plummer N=2^20 lmin=7 lmax=9, 1 step", flush=True)
    p = plummer(1 << 20, seed=5)
    nf, pf, sf, rf = run(p, 7, 9, prec, True, 1)
    nu, pu, su, ru = run(p, 7, 9, prec, False, 1)
    # the case must reach the fused kernels (nodes with levels above their tail)
    assert rf.get("bmg_down", {}).get("launches", 0) > 0, sorted(rf)
    assert rf.get("bmg_up", {}).get("launches", 0) > 0
    assert "bmg_down" not in ru and "bmg_up" not in ru
    assert sf["branch_cg_levels"] == 0 and su["branch_cg_levels"] == 0
    assert sf["branch_iterations"] == su["branch_iterations"]
    assert len(nf) == len(nu) > 0
    worst = 0.0
    for a, b in zip(nf, nu):
        assert a["nintr"] == b["nintr"]
        n = b["gcap"]
        for f in ("pot", "ax", "ay", "az"):
            worst = max(worst, rel(a[f][:n], b[f][:n]))
    print(f"  node fields: max rel diff {worst:.3e}", flush=True)
    assert worst < tol
    for k in range(9):
        assert rel(pf[k], pu[k]) < tol
