"""The first tree level's node multigrid set-up runs on a host thread and a
second stream during the head density and solve (QZ_NODE_SETUP_AHEAD=0 builds
it in place, in the node solve).  With the set-up ahead, the head
accelerations also run on that stream, beside the node solves.  The same
tables reach the same kernels, so fields and trajectories agree bit for bit."""
import os

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import clumps, plummer, plummer_dt

pytestmark = pytest.mark.gpu


def run(p, lmin, lmax, ahead, steps):
    old = os.environ.get("QZ_NODE_SETUP_AHEAD")
    os.environ["QZ_NODE_SETUP_AHEAD"] = "1" if ahead else "0"
    try:
        with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=8,
                             branch_solver=core.BRANCH_MG) as s:
            s.force()
            dt = plummer_dt(0.005)
            for _ in range(steps):
                s.step(dt)
            return s.particles(), s.nodes(), s.stats()
    finally:
        if old is None:
            del os.environ["QZ_NODE_SETUP_AHEAD"]
        else:
            os.environ["QZ_NODE_SETUP_AHEAD"] = old


@pytest.mark.parametrize("kind,n,lmin,lmax", [("plummer", 200000, 6, 8), ("clumps", 300000, 6, 9)])
def test_setup_ahead_matches_setup_in_place(kind, n, lmin, lmax):
    print(f"case: node set-up ahead vs in place {kind} N={n} lmin={lmin} lmax={lmax}, 4 steps", flush=True)
    p = plummer(n, seed=17) if kind == "plummer" else clumps(n, lmin, n_clumps=8, seed=17)
    pa, na, sa = run(p, lmin, lmax, True, 4)
    pi, ni, si = run(p, lmin, lmax, False, 4)
    assert sa["n_nodes"] == si["n_nodes"] > 0
    assert sa["branch_iterations"] == si["branch_iterations"]
    np.testing.assert_array_equal(pa, pi)
    for a, b in zip(na, ni):
        for f in ("pot", "ax", "ay", "az"):
            np.testing.assert_array_equal(a[f], b[f])
