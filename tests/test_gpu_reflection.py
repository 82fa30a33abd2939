"""The reflexive boundary (particle_updating.c:692-739, the reference default
bdry_cond_type = 1) on the GPU against the oracle: particles start just inside
every face of the box moving outward, so most of them cross a face within the
first steps and are reflected (x -> -x or 2 - x, v and a flipped).  Forward
trajectories and the adjoint of a reflecting trajectory, on the head grid and
with refinement levels (a clump in a corner refines nodes that touch the
simulation boundary)."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import Particles, plummer, plummer_dt

pytestmark = pytest.mark.gpu

NSTEPS = 5
DT = plummer_dt(0.01)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def face_particles(seed=41):
    """A Plummer core, 1200 particles within 3e-4 of the six faces moving
    outward at 2-10 code velocity units (0.8-4 e-4 box units per step), and a
    400-particle clump in the (0, 0, 0) corner moving outward."""
    rng = np.random.default_rng(seed)
    core_p = plummer(3000, seed=seed)
    nf = 1200
    pos = rng.uniform(0.05, 0.95, size=(nf, 3))
    vel = rng.normal(0.0, 1.0, size=(nf, 3))
    axis = rng.integers(0, 3, size=nf)
    side = rng.integers(0, 2, size=nf)
    depth = rng.uniform(0.0, 3e-4, size=nf)
    speed = rng.uniform(2.0, 10.0, size=nf)
    for i in range(nf):
        a = axis[i]
        pos[i, a] = depth[i] if side[i] == 0 else 1.0 - 1e-12 - depth[i]
        vel[i, a] = -speed[i] if side[i] == 0 else speed[i]
    nc = 400
    cpos = rng.uniform(1e-4, 6e-3, size=(nc, 3))
    cvel = -rng.uniform(1.0, 6.0, size=(nc, 3))
    x = np.concatenate([core_p.x, pos[:, 0], cpos[:, 0]])
    y = np.concatenate([core_p.y, pos[:, 1], cpos[:, 1]])
    z = np.concatenate([core_p.z, pos[:, 2], cpos[:, 2]])
    vx = np.concatenate([core_p.vx, vel[:, 0], cvel[:, 0]])
    vy = np.concatenate([core_p.vy, vel[:, 1], cvel[:, 1]])
    vz = np.concatenate([core_p.vz, vel[:, 2], cvel[:, 2]])
    n = x.shape[0]
    return Particles(x, y, z, vx, vy, vz, np.full(n, 1.0 / n))


def reflected_count(p0, final):
    """Particles with at least one velocity component whose sign flipped from
    an outward start near a face: the signature of a reflection (gravity over
    5 steps changes these velocities by far less than their size)."""
    v0 = np.stack([p0.vx, p0.vy, p0.vz])
    x0 = np.stack([p0.x, p0.y, p0.z])
    near = (x0 < 1e-3) | (x0 > 1.0 - 1e-3)
    flipped = near & (np.sign(final[3:6]) == -np.sign(v0)) & (np.abs(v0) > 1.0)
    return int(np.count_nonzero(np.any(flipped, axis=0)))


LEVELS = [(5, 5), (5, 7)]


@pytest.mark.parametrize("levels", LEVELS)
@pytest.mark.parametrize("prec,tol_x,tol_v", [(core.PREC_FP64, 1e-10, 1e-8), (core.PREC_MIXED, 1e-6, 1e-4)])
def test_reflexive_trajectories_match_oracle(oracle, levels, prec, tol_x, tol_v):
    lmin, lmax = levels
    print(f"case: reflexive faces N=4600 levels={levels} prec={prec} {NSTEPS} steps", flush=True)
    p = face_particles()
    o = oracle.OracleSim(p, lmin, lmax, 8)
    o.force()
    for _ in range(NSTEPS):
        o.step(DT)
    po = o.particles()
    with core.Simulation(p, lmin, lmax, precision=prec, ref_criterion_ptcl=8) as s:
        s.force()
        for _ in range(NSTEPS):
            s.step(DT)
        pg = s.particles()
        if lmax > lmin:
            assert s.stats()["n_nodes"] > 0
    n_refl = reflected_count(p, po)
    assert n_refl > 1000, n_refl  # the case exercises the reflection
    assert reflected_count(p, pg) == n_refl
    assert np.all((pg[:3] >= 0.0) & (pg[:3] <= 1.0))
    ex = float(np.max(np.abs(pg[:3] - po[:3])))
    ev = max(rel(pg[k], po[k]) for k in (3, 4, 5))
    ea = max(rel(pg[k], po[k]) for k in (6, 7, 8))
    print(f"errors: reflexive levels={levels} prec={prec}: {n_refl} reflected, x {ex:.2e}, v {ev:.2e}, a {ea:.2e}",
          flush=True)
    assert ex < tol_x and ev < tol_v and ea < tol_v


@pytest.mark.parametrize("levels", LEVELS)
@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-7), (core.PREC_MIXED, 5e-3)])
def test_reflecting_trajectory_adjoint_matches_oracle(oracle, levels, prec, tol):
    """Loss = sum of w . x_N (a fixed random weight per coordinate) over a
    5-step reflecting trajectory, checkpoints every 2: the gradient with
    respect to the initial positions and velocities against the oracle's FP64
    reverse mode, whose reflected axes flip the cotangents' signs."""
    lmin, lmax = levels
    print(f"case: reflecting adjoint N=4600 levels={levels} prec={prec}", flush=True)
    p = face_particles()
    rng = np.random.default_rng(43)
    wN = rng.standard_normal((3, p.n)) / p.n
    with core.Simulation(p, lmin, lmax, precision=prec, ref_criterion_ptcl=8, tol=1e-11) as s:
        s.force()
        s.adjoint_forward(NSTEPS, DT, 2)
        fin = s.particles()
        gx, gv = s.adjoint_backward(wN, np.zeros_like(wN))
    assert reflected_count(p, fin) > 1000
    o = oracle.OracleSim(p, lmin, lmax, 8)
    o.set_tolerance(1e-11)
    o.force()
    wx, wv = o.adjoint(NSTEPS, DT, wN, np.zeros_like(wN))
    ex, ev = rel(gx, wx), rel(gv, wv)
    # a reflected coordinate's gradient has the opposite sign of an unreflected
    # one: the sign pattern must agree everywhere the oracle's is clear
    big = np.abs(wx) > 1e-3 * np.max(np.abs(wx))
    assert np.array_equal(np.sign(gx[big]), np.sign(wx[big]))
    print(f"errors: reflecting adjoint levels={levels} prec={prec}: x {ex:.2e}, v {ev:.2e}", flush=True)
    assert ex < tol and ev < tol
