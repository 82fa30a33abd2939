"""Marks, zones and child boxes of the levels below the head are found on the
device (byte marks over each level's box pool, union-find over the marked
cells, a stable sort by zone); QZ_HOST_TREE=1 finds them on the host from the
flagged cells instead.  Both restate tree_construction.c:508-926 and
1092-1661, so they must build the same tree: every node header and grid-point
list equal, and then the same fields, accelerations and trajectories bit for
bit.  The oracle pins the device path's geometry on the AMR golden cases and
at the AMR configuration's own size; this test reaches the two-sphere
configuration (three refinement levels, 2 x 2^23 particles), where the oracle
does not finish in test time."""
import ctypes
import os

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import clumps, plummer, plummer_dt, two_plummer

pytestmark = pytest.mark.gpu


def headers_and_lists(s):
    lib = core.load_library()
    out = []
    for i in range(lib.qz_num_nodes(s._h)):
        hdr = np.zeros(24, dtype=np.int32)
        core._check(lib.qz_node_header(s._h, i, hdr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "qz_node_header")
        d = dict(zip(core.HDR_FIELDS, [int(v) for v in hdr]))
        lists = [np.zeros((max(d[k], 1), 3), dtype=np.int32) for k in ("nintr", "nbdry", "nsim")]
        core._check(lib.qz_node_grid_lists(s._h, i, *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) for a in lists]),
                    "qz_node_grid_lists")
        out.append((hdr, [a[: d[k]] for a, k in zip(lists, ("nintr", "nbdry", "nsim"))]))
    return out


def run(p, lmin, lmax, host, steps, dt, fields):
    old = os.environ.get("QZ_HOST_TREE")
    os.environ["QZ_HOST_TREE"] = "1" if host else "0"
    try:
        with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=8,
                             branch_solver=core.BRANCH_MG) as s:
            s.force()
            tree0 = headers_and_lists(s)
            for _ in range(steps):
                s.step(dt)
            return tree0, s.particles(), (s.nodes() if fields else headers_and_lists(s)), s.stats()
    finally:
        if old is None:
            del os.environ["QZ_HOST_TREE"]
        else:
            os.environ["QZ_HOST_TREE"] = old


def same_tree(ta, tb):
    assert len(ta) == len(tb) > 0
    for (ha, la), (hb, lb) in zip(ta, tb):
        np.testing.assert_array_equal(ha, hb)
        for a, b in zip(la, lb):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kind,n,lmin,lmax", [("plummer", 300000, 6, 9), ("clumps", 400000, 6, 9)])
def test_device_tree_matches_host_tree(kind, n, lmin, lmax):
    print(f"case: device vs host tree {kind} N={n} lmin={lmin} lmax={lmax}, force + 3 steps", flush=True)
    p = plummer(n, seed=23) if kind == "plummer" else clumps(n, lmin, n_clumps=12, seed=23)
    dt = plummer_dt(0.005)
    t0d, pd, nd, sd = run(p, lmin, lmax, False, 3, dt, True)
    t0h, ph, nh, sh = run(p, lmin, lmax, True, 3, dt, True)
    same_tree(t0d, t0h)
    assert sd["n_levels"] == sh["n_levels"] == lmax - lmin + 1
    assert sd["n_nodes"] == sh["n_nodes"]
    np.testing.assert_array_equal(pd, ph)
    assert len(nd) == len(nh)
    for a, b in zip(nd, nh):
        for f in ("intr_xyz", "bdry_xyz", "sim_xyz", "rho", "pot", "ax", "ay", "az"):
            np.testing.assert_array_equal(a[f], b[f])
    print(f"errors: none; {sd['n_nodes']} nodes on {sd['n_levels']} levels equal", flush=True)


def test_device_tree_matches_host_tree_two_spheres():
    print("case: device vs host tree, two Plummer spheres 2 x 2^23, 256^3, lmax 11, force + 1 step", flush=True)
    p = two_plummer(1 << 23, seed=20261203)
    with core.Simulation(p, 8, 11, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
        s.force()
        dt = s.timestep(0.5)
    t0d, pd, t1d, sd = run(p, 8, 11, False, 1, dt, False)
    t0h, ph, t1h, sh = run(p, 8, 11, True, 1, dt, False)
    same_tree(t0d, t0h)
    same_tree(t1d, t1h)
    assert sd["n_levels"] == sh["n_levels"] == 4
    np.testing.assert_array_equal(pd, ph)
    print(f"errors: none; {sd['n_nodes']} nodes on 4 levels equal", flush=True)
