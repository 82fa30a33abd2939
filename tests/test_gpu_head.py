"""Head-only PM step on the GPU vs the oracle (configs shaped like the 32^3 / 64^3 head-only runs)."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec,tol_rho,tol_pot,tol_acc", [
    (core.PREC_FP64, 1e-12, 1e-10, 1e-9),
    (core.PREC_MIXED, 1e-6, 1e-6, 1e-6),
])
def test_head_force_pass_matches_oracle(oracle, prec, tol_rho, tol_pot, tol_acc):
    print(f"case: head force pass prec={prec} N=4000 lmin=5", flush=True)
    p = plummer(4000, seed=11)
    o = oracle.OracleSim(p, 5, 5)
    o.force()
    ho = o.head()
    with core.Simulation(p, 5, 5, precision=prec) as s:
        s.force()
        hg = s.head_fields()
        pg = s.particles()
        st = s.stats()
    po = o.particles()
    assert rel(hg[0], ho[0]) < tol_rho
    assert rel(hg[1], ho[1]) < tol_pot
    for k in (2, 3, 4):
        assert rel(hg[k], ho[k]) < tol_acc
    for k in (6, 7, 8):
        assert rel(pg[k], po[k]) < tol_acc
    assert st["head_residual"] < 1e-10


@pytest.mark.parametrize("prec,tol_x,tol_v", [(core.PREC_FP64, 1e-10, 1e-7), (core.PREC_MIXED, 1e-6, 1e-4)])
def test_head_steps_match_oracle(oracle, prec, tol_x, tol_v):
    print(f"case: head 5 steps prec={prec} N=3000 lmin=5", flush=True)
    p = plummer(3000, seed=5)
    dt = plummer_dt(0.01)
    o = oracle.OracleSim(p, 5, 5)
    o.force()
    with core.Simulation(p, 5, 5, precision=prec) as s:
        s.force()
        for _ in range(5):
            o.step(dt)
            s.step(dt)
        pg = s.particles()
    po = o.particles()
    for k in range(3):
        assert np.max(np.abs(pg[k] - po[k])) < tol_x
    for k in range(3, 6):
        assert rel(pg[k], po[k]) < tol_v


def test_clustered_particles_cell_ranges(oracle):
    """Particles packed into two far-apart corner regions of a 128^3 head grid
    (runs of millions of empty cells between occupied ones): density and
    accelerations against the oracle."""
    print("case: clustered cell ranges N=3000 lmin=7", flush=True)
    rng = np.random.default_rng(21)
    n = 3000
    x = np.concatenate([rng.uniform(0.01, 0.03, n // 2), rng.uniform(0.97, 0.99, n - n // 2)])
    y = np.concatenate([rng.uniform(0.01, 0.03, n // 2), rng.uniform(0.97, 0.99, n - n // 2)])
    z = np.concatenate([rng.uniform(0.01, 0.03, n // 2), rng.uniform(0.97, 0.99, n - n // 2)])
    zero = np.zeros(n)
    from quantizer_amd.workloads import Particles
    p = Particles(x, y, z, zero.copy(), zero.copy(), zero.copy(), np.full(n, 1.0 / n))
    o = oracle.OracleSim(p, 7, 7)
    o.force()
    ho = o.head()
    with core.Simulation(p, 7, 7, precision=core.PREC_FP64) as s:
        s.force()
        hg = s.head_fields()
        pg = s.particles()
    assert rel(hg[0], ho[0]) < 1e-12
    for k in (6, 7, 8):
        assert rel(pg[k], o.particles()[k]) < 1e-6


@pytest.mark.parametrize("prec,tol_x,tol_drift", [(core.PREC_FP64, 1e-8, 1e-6), (core.PREC_MIXED, 1e-4, 1e-4)])
def test_plummer_32_zero_boundary_energy_drift(oracle, prec, tol_x, tol_drift):
    """The small plain-method configuration: Plummer N = 10^4, 32^3 head grid,
    zero-potential faces, dt = 0.01 (Plummer units), 100 steps.  Positions
    against the oracle, and the relative energy drift E(t)/E(0) - 1 (exact
    pairwise potential energy) matching the oracle's."""
    print(f"case: Plummer 1e4 32^3 zero faces 100 steps prec={prec}", flush=True)
    p = plummer(10000)
    lmin, nsteps, dt = 5, 100, plummer_dt(0.01)
    o = oracle.OracleSim(p, lmin, lmin, head_bc=core.BC_ZERO)
    o.force()
    e0_ref = o.observables()[2]
    for _ in range(nsteps):
        o.step(dt)
    drift_ref = o.observables()[2] / e0_ref - 1.0
    x_ref = o.particles()[:3]
    with core.Simulation(p, lmin, lmin, precision=prec, head_bc=core.BC_ZERO) as s:
        s.force()
        e0 = s.observables()[2]
        for _ in range(nsteps):
            s.step(dt)
        drift = s.observables()[2] / e0 - 1.0
        x = s.particles()[:3]
    assert np.max(np.abs(x - x_ref)) < tol_x
    assert abs(drift - drift_ref) < tol_drift, (drift, drift_ref)
