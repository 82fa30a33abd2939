"""The reference's own host program (initialisation, tree construction and
adaptation, main-loop order) linked against the drop-in libraries, which export
grid_density, potential, grid_acceleration, particle_acceleration, reset and
ptcl_idx_to_box_idx (and, resident, particle_updating_A/B) under the reference's
names and signatures.  Its particle output must match the same program linked
against the reference's own objects."""
import glob
import os
import subprocess

import numpy as np
import pytest

from quantizer_amd.workloads import plummer, plummer_dt, clumps
from tests import oracle_lib as OL

pytestmark = pytest.mark.gpu

BINS = {"dropin": os.path.join(OL.ORACLE_DIR, "_ref", "ref_harness_dropin"),
        "resident": os.path.join(OL.ORACLE_DIR, "_ref", "ref_harness_resident")}
CASES = [("plummer", 2000, 4, 4, 8, 3), ("plummer", 6000, 4, 6, 8, 3), ("clumps", 12000, 4, 6, 8, 2)]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("build", sorted(BINS))
@pytest.mark.parametrize("kind,n,lmin,lmax,crit,steps", CASES)
def test_reference_program_with_dropin(tmp_path, build, kind, n, lmin, lmax, crit, steps):
    for b in (OL.REF_BIN, BINS[build]):
        if not os.path.exists(b):
            pytest.skip(f"{os.path.basename(b)} not built (needs the reference sources at build time)")
    print(f"case: drop-in {build} {kind} N={n} lmin={lmin} lmax={lmax} steps={steps}", flush=True)
    p = plummer(n, seed=31) if kind == "plummer" else clumps(n, lmin, n_clumps=4, seed=31, sigma_cells=(0.4, 1.0))
    dt = plummer_dt(0.005)
    ref_dir, drop_dir = str(tmp_path / "ref"), str(tmp_path / build)
    OL.run_reference(p, ref_dir, lmin, lmax, crit, steps, dt, extra=("--noenergy",))
    OL.run_reference(p, drop_dir, lmin, lmax, crit, steps, dt, extra=("--noenergy",), binary=BINS[build])
    a_ref = OL.read_particles(os.path.join(ref_dir, "s0_acc.bin"), n, False)
    a_drp = OL.read_particles(os.path.join(drop_dir, "s0_acc.bin"), n, False)
    assert rel(a_drp, a_ref) < 1e-7
    k_ref = OL.read_particles(os.path.join(ref_dir, "sK_ptcl.bin"), n, True)
    k_drp = OL.read_particles(os.path.join(drop_dir, "sK_ptcl.bin"), n, True)
    assert np.max(np.abs(k_drp[:3] - k_ref[:3])) < 1e-10
    for c in range(3, 9):
        assert rel(k_drp[c], k_ref[c]) < 1e-6


ADAPTIVE_CASES = [("plummer", 2000, 4, 4, 8), ("plummer", 6000, 4, 6, 8), ("clumps", 12000, 4, 6, 8)]


@pytest.mark.parametrize("build", sorted(BINS))
@pytest.mark.parametrize("kind,n,lmin,lmax,crit", ADAPTIVE_CASES)
def test_reference_main_loop_with_dropin(tmp_path, build, kind, n, lmin, lmax, crit):
    """main.c's own loop order (main.c:341-573): the time step from timestep()
    at every step, observables() and output_snapshots() before the loop, every
    3 steps and after the last, garbage_collector(), output_main_parameters().
    The resident build answers timestep, observables and output_snapshots on
    the device (snapshots written asynchronously) with no per-step copies of
    the particles; the force-cycle build leaves them to the reference's host
    code.  Time steps, final particles, every snapshot file and
    Parameters.dat must match the program linked with the reference's own
    objects."""
    for b in (OL.REF_BIN, BINS[build]):
        if not os.path.exists(b):
            pytest.skip(f"{os.path.basename(b)} not built (needs the reference sources at build time)")
    print(f"case: main loop {build} {kind} N={n} lmin={lmin} lmax={lmax} 10 steps", flush=True)
    p = plummer(n, seed=31) if kind == "plummer" else clumps(n, lmin, n_clumps=4, seed=31, sigma_cells=(0.4, 1.0))
    runs = {}
    for name, binary in (("ref", OL.REF_BIN), (build, BINS[build])):
        top = tmp_path / name
        (top / "run").mkdir(parents=True)
        (top / "dumps").mkdir()
        inp = str(top / "p.bin")
        p.write_bin(inp)
        r = subprocess.run([binary, inp, str(top / "dumps"), str(lmin), str(lmax), str(crit), "10", "1.0",
                            "--adaptive", "3"], cwd=str(top / "run"), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        sims = glob.glob(str(top / "output" / "*"))
        assert len(sims) == 1, sims
        runs[name] = (top, sims[0])
    (rt, rs), (dtop, ds) = runs["ref"], runs[build]
    assert os.path.basename(rs) == os.path.basename(ds)  # the reference's folder naming
    dts_r = np.loadtxt(str(rt / "dumps" / "dts.txt"))
    dts_d = np.loadtxt(str(dtop / "dumps" / "dts.txt"))
    assert dts_r.shape == dts_d.shape == (10,)
    np.testing.assert_allclose(dts_d, dts_r, rtol=1e-9, atol=0)
    k_ref = OL.read_particles(str(rt / "dumps" / "sK_ptcl.bin"), n, True)
    k_drp = OL.read_particles(str(dtop / "dumps" / "sK_ptcl.bin"), n, True)
    assert np.max(np.abs(k_drp[:3] - k_ref[:3])) < 1e-10
    for c in range(3, 9):
        assert rel(k_drp[c], k_ref[c]) < 1e-6
    snaps = sorted(os.listdir(os.path.join(rs, "Data")))
    assert snaps == sorted(os.listdir(os.path.join(ds, "Data"))) == [f"Data_{k}.bin" for k in range(5)]
    for f in snaps:
        a = np.fromfile(os.path.join(ds, "Data", f))
        b = np.fromfile(os.path.join(rs, "Data", f))
        assert a.size == b.size == 7 * n + 7
        np.testing.assert_array_equal(a[:n], b[:n])  # masses
        assert np.max(np.abs(a[n:4 * n] - b[n:4 * n])) < 1e-10  # positions
        assert rel(a[4 * n:7 * n], b[4 * n:7 * n]) < 1e-6  # velocities
        np.testing.assert_allclose(a[7 * n:7 * n + 3], b[7 * n:7 * n + 3], rtol=1e-6)  # energies
        assert np.max(np.abs(a[7 * n + 3:7 * n + 6] - b[7 * n + 3:7 * n + 6])) < 1e-9  # momentum
        np.testing.assert_allclose(a[-1], b[-1], rtol=1e-9)  # time
    pr = open(os.path.join(rs, "Parameters.dat")).read().splitlines()[:38]
    pd = open(os.path.join(ds, "Parameters.dat")).read().splitlines()[:38]
    assert pd == pr
