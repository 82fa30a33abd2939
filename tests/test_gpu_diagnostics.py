"""Integer keys (ptcl_idx_to_box_idx), the CFL time step and the energy and
momentum diagnostics on the GPU against the oracle: keys bit-exact, including
points on cell faces, below 0 and at or above 1; time steps and grid energies to
the accuracy of the fields they use; FP64 sums to round-off."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, two_plummer

pytestmark = pytest.mark.gpu


def test_box_keys_match_oracle(oracle):
    print("case: box keys lv 7 with faces, negatives, >= 1", flush=True)
    rng = np.random.default_rng(3)
    lv, ts, rd = 7, np.array([-3, 5, 0], dtype=np.int32), np.array([140, 130, 135], dtype=np.int32)
    pts = list(rng.uniform(-0.02, 1.02, size=(4000, 3)))
    faces = np.arange(0, 129) / 128.0  # exactly on cell faces
    for f in faces[::7]:
        pts.append([f, f, f])
        pts.append([f, 0.5, -0.0])
    pts += [[-1e-300, 0.25, 0.75], [1.0, 1.0 - 2 ** -53, 0.0], [-0.5 / 128, 0.5, 0.5]]
    pts = np.array(pts)
    got = core.box_keys(pts[:, 0], pts[:, 1], pts[:, 2], lv, ts, rd)
    want = np.array([oracle.oracle_key(lv, ts, rd, *p) for p in pts], dtype=np.int32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [(5, 5), (4, 6)])
def test_timestep_matches_oracle(oracle, levels):
    lmin, lmax = levels
    print(f"case: CFL timestep levels={levels} two spheres N=3000", flush=True)
    p = two_plummer(1500, seed=34)
    o = oracle.OracleSim(p, lmin, lmax)
    o.set_tolerance(1e-12)
    o.force()
    want = o.timestep(0.5)
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, tol=1e-12) as s:
        s.force()
        got = s.timestep(0.5)
        capped = s.timestep(0.5, dt_max=0.1 * want)
    assert got == pytest.approx(want, rel=1e-8)
    assert capped == 0.1 * want


@pytest.mark.parametrize("levels", [(5, 5), (4, 6)])
@pytest.mark.parametrize("prec", [core.PREC_FP64, core.PREC_MIXED])
def test_timestep_folded_into_the_kick_is_exact(levels, prec):
    """After a step the kick has folded the next CFL time step into per-block
    minima (with each particle's finest level from the gather); a time step
    taken from them must equal the one computed afresh from the particles bit
    for bit: uncapped, capped, and after a change of CFL number."""
    lmin, lmax = levels
    print(f"case: folded CFL timestep levels={levels} prec={prec} two spheres N=3000", flush=True)
    p = two_plummer(1500, seed=35)
    with core.Simulation(p, lmin, lmax, precision=prec, tol=1e-12) as s:
        s.force()
        dt0 = s.timestep(0.5)
        for _ in range(3):
            s.step(s.timestep(0.5, dt0))
        folded = s.timestep(0.5)  # the kick's minima
        fresh = s.timestep(0.5)  # those are used once: computed from the particles
        assert folded == fresh
        s.step(fresh)
        cap = 0.5 * fresh
        c_folded, c_fresh = s.timestep(0.5, dt_max=cap), s.timestep(0.5, dt_max=cap)
        assert c_folded == c_fresh == min(cap, s.timestep(0.5))
        s.step(c_folded)
        o_first = s.timestep(0.3)  # the kick folded CFL 0.5: computed afresh
        o_fresh = s.timestep(0.3)
        assert o_first == o_fresh
        s.step(o_first)
        o_folded, o_fresh = s.timestep(0.3), s.timestep(0.3)  # now folded at 0.3
        assert o_folded == o_fresh


@pytest.mark.parametrize("levels", [(5, 5), (4, 6)])
def test_observables_match_oracle(oracle, levels):
    lmin, lmax = levels
    print(f"case: observables levels={levels} two spheres N=3000", flush=True)
    p = two_plummer(1500, seed=35)
    o = oracle.OracleSim(p, lmin, lmax)
    o.set_tolerance(1e-12)
    o.force()
    want = o.observables()
    want_grid = o.grid_pe()
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, tol=1e-12) as s:
        s.force()
        got = s.observables(exact_pe=True)
        got_grid = s.observables(exact_pe=False)
    np.testing.assert_allclose(got[:3], want[:3], rtol=1e-12)
    np.testing.assert_allclose(got[3:], want[3:], rtol=1e-12, atol=1e-14)
    assert got_grid[1] == pytest.approx(want_grid, rel=1e-8)
