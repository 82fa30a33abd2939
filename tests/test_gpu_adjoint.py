"""Adjoint of the step on the GPU (head grid, and with refinement levels): force
VJP and checkpointed trajectory gradients against the FP64 CPU reverse mode
(oracle) and central differences."""
import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import clumps, plummer, plummer_dt, two_plummer

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-8), (core.PREC_MIXED, 2e-4)])
def test_force_vjp_matches_oracle(oracle, prec, tol):
    print(f"case: force vjp prec={prec} N=4000 lmin=5", flush=True)
    p = plummer(4000, seed=12)
    rng = np.random.default_rng(4)
    ab = rng.standard_normal((3, p.n))
    o = oracle.OracleSim(p, 5, 5)
    want = o.force_vjp(ab)
    with core.Simulation(p, 5, 5, precision=prec) as s:
        got = s.force_vjp(ab)
    assert rel(got, want) < tol


def _target(parts, lmin, nsteps, dt, seed, prec):
    rng = np.random.default_rng(seed)
    t = parts.copy()
    for a in (t.x, t.y, t.z):
        a += rng.normal(0.0, 0.01 / 40, parts.n)  # N(0, 0.01 a) in a 40 a box
    with core.Simulation(t, lmin, lmin, precision=prec) as s:
        s.force()
        for _ in range(nsteps):
            s.step(dt)
        return s.particles()[:3]


@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-7), (core.PREC_MIXED, 5e-3)])
def test_trajectory_adjoint_matches_oracle_reverse_mode(oracle, prec, tol):
    print(f"case: trajectory adjoint prec={prec} N=3000 lmin=5 10 steps ckpt 4", flush=True)
    p = plummer(3000, seed=13)
    lmin, nsteps, dt = 5, 10, plummer_dt(0.01)
    target = _target(p, lmin, nsteps, dt, 5, core.PREC_FP64)
    with core.Simulation(p, lmin, lmin, precision=prec) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, 4)
        xN = s.particles()[:3]
        xbN = 2.0 * (xN - target) / p.n
        gx, gv = s.adjoint_backward(xbN, np.zeros_like(xbN))
        x_after = s.particles()[:3]
    np.testing.assert_array_equal(x_after, xN)  # context left at the final state
    o = oracle.OracleSim(p, lmin, lmin)
    o.force()
    wx, wv = o.adjoint(nsteps, dt, xbN, np.zeros_like(xbN))
    assert rel(gx, wx) < tol
    assert rel(gv, wv) < tol


def test_trajectory_gradient_central_differences():
    """Central differences of the loss on 64 seeded coordinates (FP64 path)."""
    print("case: FD check N=2048 lmin=5 8 steps", flush=True)
    p = plummer(2048, seed=14)
    lmin, nsteps, dt = 5, 8, plummer_dt(0.01)
    target = _target(p, lmin, nsteps, dt, 6, core.PREC_FP64)

    def loss(parts):
        with core.Simulation(parts, lmin, lmin, precision=core.PREC_FP64) as s:
            s.force()
            for _ in range(nsteps):
                s.step(dt)
            x = s.particles()[:3]
        return float(np.mean(np.sum((x - target) ** 2, axis=0)))

    with core.Simulation(p, lmin, lmin, precision=core.PREC_FP64) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, 3)
        xN = s.particles()[:3]
        gx, gv = s.adjoint_backward(2.0 * (xN - target) / p.n, np.zeros((3, p.n)))
    rng = np.random.default_rng(7)
    eps = 1e-3 / 40 * 1e-3  # 1e-3 a scaled down so no particle changes cell during the pair
    bad = 0
    names = ["x", "y", "z", "vx", "vy", "vz"]
    for _ in range(64):
        k = int(rng.integers(0, 6))
        i = int(rng.integers(0, p.n))
        plus, minus = p.copy(), p.copy()
        getattr(plus, names[k])[i] += eps
        getattr(minus, names[k])[i] -= eps
        fd = (loss(plus) - loss(minus)) / (2 * eps)
        g = (gx if k < 3 else gv)[k % 3, i]
        if abs(fd - g) > 1e-4 * abs(g) + 1e-10:
            bad += 1
    # a perturbed particle can cross a cell face, where CIC weights have a kink
    assert bad <= 2


@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-8), (core.PREC_MIXED, 2e-4)])
def test_force_vjp_amr_matches_oracle(oracle, prec, tol):
    print(f"case: force vjp with levels prec={prec} clumps N=12000 lmin=4 lmax=6", flush=True)
    p = clumps(12000, 4, n_clumps=5, seed=103, sigma_cells=(0.4, 1.0))
    rng = np.random.default_rng(9)
    ab = rng.standard_normal((3, p.n))
    o = oracle.OracleSim(p, 4, 6)
    o.set_tolerance(1e-12)
    want = o.force_vjp(ab)
    with core.Simulation(p, 4, 6, precision=prec, tol=1e-11) as s:
        got = s.force_vjp(ab)
        assert s.stats()["n_nodes"] == o.stats()["nodes"]
    assert rel(got, want) < tol


def _ke_weight(p, n_each):
    return np.where(np.arange(p.n) >= n_each, p.m, 0.0)  # the lighter sphere


@pytest.mark.parametrize("levels", [(5, 5), (4, 6)])
@pytest.mark.parametrize("prec,tol", [(core.PREC_FP64, 1e-6), (core.PREC_MIXED, 5e-3)])
def test_trajectory_adjoint_two_spheres_matches_oracle(oracle, levels, prec, tol):
    """Final kinetic energy of the lighter of two Plummer spheres (particle
    masses 2:1), 6 steps with checkpoints every 4, against the oracle's reverse
    mode (head grid and two refinement levels)."""
    lmin, lmax = levels
    print(f"case: two-sphere trajectory adjoint prec={prec} levels={levels} N=3000 6 steps ckpt 4", flush=True)
    p = two_plummer(1500, seed=31)
    nsteps, dt = 6, plummer_dt(0.01, 80.0)
    w = _ke_weight(p, 1500)
    with core.Simulation(p, lmin, lmax, precision=prec, tol=1e-11) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, 4)
        fin = s.particles()
        vbN = w * fin[3:6]
        gx, gv = s.adjoint_backward(np.zeros_like(vbN), vbN)
        after = s.particles()
        if lmax > lmin:
            assert s.stats()["n_nodes"] > 0
    np.testing.assert_array_equal(after, fin)  # context left at the final state
    o = oracle.OracleSim(p, lmin, lmax)
    o.set_tolerance(1e-12)
    o.force()
    wx, wv = o.adjoint(nsteps, dt, np.zeros_like(vbN), vbN)
    assert rel(gx, wx) < tol
    assert rel(gv, wv) < tol


def test_trajectory_gradient_amr_central_differences():
    """Central differences of the final kinetic energy of the lighter sphere on
    seeded coordinates, two refinement levels (FP64 path)."""
    print("case: FD check with levels two spheres N=3000 lmin=4 lmax=6 4 steps", flush=True)
    p = two_plummer(1500, seed=32)
    lmin, lmax, nsteps, dt = 4, 6, 4, plummer_dt(0.01, 80.0)
    w = _ke_weight(p, 1500)

    def loss(parts):
        with core.Simulation(parts, lmin, lmax, precision=core.PREC_FP64, tol=1e-12) as s:
            s.force()
            for _ in range(nsteps):
                s.step(dt)
            v = s.particles()[3:6]
        return 0.5 * float(np.sum(w * np.sum(v * v, axis=0))), v

    L0, vN = loss(p)
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, tol=1e-12) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, 3)
        gx, gv = s.adjoint_backward(np.zeros_like(vN), w * vN)
    rng = np.random.default_rng(11)
    eps = 1e-5  # the FP64 sum of the loss limits central differences to ~1e-5 relative at 1e-6
    floor = 1e-13 * abs(L0) / eps  # resolution of a difference of two FP64 losses
    names = ["x", "y", "z", "vx", "vy", "vz"]
    bad = 0
    for _ in range(24):
        k = int(rng.integers(0, 6))
        i = int(rng.integers(0, p.n))
        plus, minus = p.copy(), p.copy()
        getattr(plus, names[k])[i] += eps
        getattr(minus, names[k])[i] -= eps
        fd = (loss(plus)[0] - loss(minus)[0]) / (2 * eps)
        g = (gx if k < 3 else gv)[k % 3, i]
        if abs(fd - g) > 2e-4 * abs(g) + floor:
            bad += 1
    # a perturbed particle can cross a cell face, where CIC weights have a kink
    assert bad <= 1


@pytest.mark.parametrize("levels", [(5, 5), (4, 6)])
def test_trajectory_adjoint_single_segment(oracle, levels):
    """All steps in one checkpoint segment (recorded by the forward itself)."""
    lmin, lmax = levels
    print(f"case: single-segment trajectory adjoint levels={levels} N=3000 3 steps ckpt 5", flush=True)
    p = two_plummer(1500, seed=33)
    nsteps, dt = 3, plummer_dt(0.01, 80.0)
    w = _ke_weight(p, 1500)
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, tol=1e-11) as s:
        s.force()
        s.adjoint_forward(nsteps, dt, 5)
        fin = s.particles()
        vbN = w * fin[3:6]
        gx, gv = s.adjoint_backward(np.zeros_like(vbN), vbN)
        np.testing.assert_array_equal(s.particles(), fin)
    o = oracle.OracleSim(p, lmin, lmax)
    o.set_tolerance(1e-12)
    o.force()
    wx, wv = o.adjoint(nsteps, dt, np.zeros_like(vbN), vbN)
    assert rel(gx, wx) < 1e-6
    assert rel(gv, wv) < 1e-6


def test_adjoint_backward_on_device_buffers_matches_host_arrays():
    """The cotangent arrays may live in device memory: the same reverse sweep on
    hipMalloc'd buffers gives the host-array result bit for bit (two contexts
    with the same history)."""
    import ctypes
    print("case: adjoint backward host arrays vs device buffers, two spheres N=2x1500 lmin=4 lmax=6 4 steps",
          flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    p = two_plummer(1500, seed=4)
    dt = plummer_dt(0.002, box=80.0)
    rng = np.random.default_rng(8)
    xbN, vbN = rng.standard_normal((3, p.n)), rng.standard_normal((3, p.n))
    with core.Simulation(p, 4, 6, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
        s.force()
        s.adjoint_forward(4, dt, 2)
        gx, gv = s.adjoint_backward(xbN, vbN)
    # a second context with the same history, the cotangents in device memory
    with core.Simulation(p, 4, 6, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
        s.force()
        s.adjoint_forward(4, dt, 2)
        nb = 8 * 3 * p.n
        bufs = []
        try:
            for _ in range(4):
                b = ctypes.c_void_p()
                assert hip.hipMalloc(ctypes.byref(b), nb) == 0
                bufs.append(b)
            assert hip.hipMemcpy(bufs[0], np.ascontiguousarray(xbN).ctypes.data, nb, 1) == 0
            assert hip.hipMemcpy(bufs[1], np.ascontiguousarray(vbN).ctypes.data, nb, 1) == 0
            s.adjoint_backward_ptr(bufs[0].value, bufs[1].value, bufs[2].value, bufs[3].value)
            dx, dv = np.empty((3, p.n)), np.empty((3, p.n))
            assert hip.hipMemcpy(dx.ctypes.data, bufs[2], nb, 2) == 0
            assert hip.hipMemcpy(dv.ctypes.data, bufs[3], nb, 2) == 0
        finally:
            for b in bufs:
                hip.hipFree(b)
    np.testing.assert_array_equal(dx, gx)
    np.testing.assert_array_equal(dv, gv)


@pytest.mark.parametrize("levels", [0, 2])
def test_overlapped_recomputation_matches_in_order(levels, monkeypatch):
    """The segment before the one being reversed is recomputed by a second engine
    on its own stream and host thread; it repeats the forward exactly, so the
    gradients equal those of the in-order sweep (QZ_ADJ_OVERLAP=0) bit for bit.
    Three segments, so two recomputations run beside a reverse sweep; a second
    backward run in the same context reuses the helper and both record sets."""
    lmin, lmax = 4, 4 + levels
    print(f"case: overlapped vs in-order recomputation, two spheres N=2x1500 lmin={lmin} lmax={lmax} 6 steps",
          flush=True)
    p = two_plummer(1500, seed=5)
    dt = plummer_dt(0.002, box=80.0)
    rng = np.random.default_rng(9)
    xbN, vbN = rng.standard_normal((3, p.n)), rng.standard_normal((3, p.n))
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("QZ_ADJ_OVERLAP", mode)
        with core.Simulation(p, lmin, lmax, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
            s.force()
            s.adjoint_forward(6, dt, 2)
            first = s.adjoint_backward(xbN, vbN)
            s.set_particles(p)
            s.force()
            s.adjoint_forward(6, dt, 2)
            out[mode] = (first, s.adjoint_backward(xbN, vbN))
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
