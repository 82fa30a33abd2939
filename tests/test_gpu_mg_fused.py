"""Fused tile kernels of the head levels below the top (k_mg_down / k_mg_up)
against the per-level kernel sequence they replace (QZ_HEAD_MG_FUSED=0)."""
import os

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt

pytestmark = pytest.mark.gpu


def run(p, lmin, prec, fused, steps):
    old = os.environ.get("QZ_HEAD_MG_FUSED")
    os.environ["QZ_HEAD_MG_FUSED"] = "1" if fused else "0"
    try:
        with core.Simulation(p, lmin, lmin, precision=prec, head_bc=core.BC_ZERO) as s:
            s.force()
            dt = plummer_dt(0.01)
            for _ in range(steps):
                s.step(dt)
            return s.head_fields(), s.particles(), s.stats()
    finally:
        if old is None:
            del os.environ["QZ_HEAD_MG_FUSED"]
        else:
            os.environ["QZ_HEAD_MG_FUSED"] = old


@pytest.mark.parametrize("lmin", [6, 7])
@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-13), (core.PREC_MIXED, 1e-6)])
def test_fused_levels_match_per_level_kernels(lmin, prec, tol):
    print(f"case: fused vs per-level head levels lmin={lmin} prec={prec} N=20000, 3 steps", flush=True)
    p = plummer(20000, seed=31)
    hf, pf, sf = run(p, lmin, prec, True, 3)
    hu, pu, su = run(p, lmin, prec, False, 3)
    # the same update formulas and summation orders: the solves take the same
    # V-cycles and the fields agree to rounding
    assert sf["head_vcycles"] == su["head_vcycles"]
    assert sf["head_residual"] < 1e-10
    for k in range(5):
        d = float(np.max(np.abs(hf[k] - hu[k])) / max(np.max(np.abs(hu[k])), 1e-300))
        print(f"  head field {k}: max rel diff {d:.3e}", flush=True)
        assert d < tol
    for k in range(9):
        d = float(np.max(np.abs(pf[k] - pu[k])) / max(np.max(np.abs(pu[k])), 1e-300))
        assert d < tol
