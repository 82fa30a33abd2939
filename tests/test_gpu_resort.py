"""Re-sort of near-sorted keys: after a drift from a sorted state only the
particles whose key changed are radix-sorted, then merged with the rest by
(key, slot).  That is the permutation a stable radix sort of all keys gives, so
every step equals the full-sort path (QZ_RESORT=0) bit for bit, with 32-bit
keys (lmax 8) and 64-bit keys (lmax 12).  The reflection tests run the re-sort
on particles that leave the box and reflect."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt

pytestmark = pytest.mark.gpu


def run(p, lmin, lmax, steps, dt, monkeypatch, mode):
    monkeypatch.setenv("QZ_RESORT", mode)
    with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
        s.force()
        s.timing(True)  # kernel timer: the sort path of each step is named in the report
        for _ in range(steps):
            s.step(dt)
        rep = s.timing_report()
        s.timing(False)
        return s.particles(), s.stats(), rep


@pytest.mark.parametrize("lmax", [8, 12])
def test_resort_matches_full_sort(lmax, monkeypatch):
    print(f"case: re-sort vs full radix sort, Plummer N=60000 lmin=5 lmax={lmax}, 6 steps", flush=True)
    p = plummer(60000, seed=31)
    # a large time step, so that many particles change cell each step
    dt = plummer_dt(0.05)
    a, sa, ra = run(p, 5, lmax, 6, dt, monkeypatch, "0")
    b, sb, rb = run(p, 5, lmax, 6, dt, monkeypatch, "1")
    assert ra["sort"]["launches"] == 6 and "resort" not in ra
    assert rb["resort"]["launches"] == 6 and "sort" not in rb
    np.testing.assert_array_equal(a, b)
    assert sa["n_nodes"] == sb["n_nodes"] and sa["branch_iterations"] == sb["branch_iterations"]
