"""Snapshot files: the library's writer against the reference's own
output_snapshots() on the same run (layout, unit conversion, observables)."""
import glob
import os
import subprocess

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer, plummer_dt
from tests import oracle_lib as OL

pytestmark = pytest.mark.gpu

UNITS = dict(dist=2.5, velocity=0.5, time_unit=3.0, cm=(0.1, 0.2, 0.3), cm_v=(0.01, 0.02, 0.03))


@pytest.mark.parametrize("lmin,lmax", [(4, 4), (4, 6)])
def test_snapshot_matches_reference_writer(tmp_path, lmin, lmax):
    if not os.path.exists(OL.REF_BIN):
        pytest.skip("compiled reference not built")
    print(f"case: snapshot plummer N=1500 lmin={lmin} lmax={lmax} 3 steps", flush=True)
    n, steps, dt = 1500, 3, plummer_dt(0.005)
    p = plummer(n, seed=41)
    run = tmp_path / "run"
    out = tmp_path / "dumps"
    run.mkdir()
    out.mkdir()
    inp = str(tmp_path / "p.bin")
    p.write_bin(inp)
    u = UNITS
    args = [str(v) for v in (u["dist"], u["velocity"], u["time_unit"], *u["cm"], *u["cm_v"])]
    r = subprocess.run([OL.REF_BIN, inp, str(out), str(lmin), str(lmax), "8", str(steps), repr(dt), "--snapshot", *args],
                       cwd=str(run), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref_file = glob.glob(str(tmp_path / "output" / "*" / "Data" / "Data_0.bin"))
    assert len(ref_file) == 1
    ref = np.fromfile(ref_file[0])
    with core.Simulation(p, lmin, lmax, precision=core.PREC_FP64, ref_criterion_ptcl=8) as s:
        s.force()
        for _ in range(steps):
            s.step(dt)
        mine_path = str(tmp_path / "mine.bin")
        s.write_snapshot(mine_path, steps * dt, **u)
    mine = np.fromfile(mine_path)
    assert mine.size == ref.size == 7 * n + 7
    np.testing.assert_array_equal(mine[:n], ref[:n])              # masses, unchanged
    np.testing.assert_allclose(mine[n:4 * n], ref[n:4 * n], rtol=0, atol=1e-9)        # positions
    v_m, v_r = mine[4 * n:7 * n], ref[4 * n:7 * n]                 # velocities: trajectory tolerance
    assert np.max(np.abs(v_m - v_r)) <= 1e-6 * np.max(np.abs(v_r))
    np.testing.assert_allclose(mine[7 * n:7 * n + 3], ref[7 * n:7 * n + 3], rtol=1e-6)  # energies
    np.testing.assert_allclose(mine[7 * n + 3:7 * n + 6], ref[7 * n + 3:7 * n + 6], rtol=0, atol=1e-9)  # momentum
    assert mine[-1] == ref[-1]                                    # time


def test_async_snapshot_is_the_state_at_the_call(tmp_path):
    """qz_write_snapshot_async returns before the file exists; steps run while
    it is written, and the file holds the state at the call, byte for byte
    what qz_write_snapshot writes (given the same observables block)."""
    print("case: async snapshot plummer N=3000 lmin=4 lmax=6", flush=True)
    n, dt = 3000, plummer_dt(0.005)
    p = plummer(n, seed=43)
    with core.Simulation(p, 4, 6, precision=core.PREC_MIXED, ref_criterion_ptcl=8) as s:
        s.force()
        for _ in range(2):
            s.step(dt)
        sync_path = str(tmp_path / "sync.bin")
        s.write_snapshot(sync_path, 2 * dt, **UNITS)
        sync = np.fromfile(sync_path)
        obs = sync[7 * n:7 * n + 6]
        apath, npath = str(tmp_path / "async.bin"), str(tmp_path / "async_noobs.bin")
        s.write_snapshot_async(apath, 2 * dt, obs6=obs, **UNITS)
        for _ in range(3):
            s.step(dt)  # runs while the first snapshot is written
        s.write_snapshot_async(npath, 5 * dt, obs6=None, **UNITS)  # waits for the first one
        s.snapshot_wait()
        later = str(tmp_path / "later.bin")
        s.write_snapshot(later, 5 * dt, observables=False, **UNITS)
    a = np.fromfile(apath)
    np.testing.assert_array_equal(a, sync)
    b = np.fromfile(npath)  # state after the 3 further steps, no observables block
    c = np.fromfile(later)
    assert b.size == 7 * n + 1
    np.testing.assert_array_equal(b[:7 * n], c[:7 * n])
    assert b[-1] == c[-1] == 5 * dt / UNITS["time_unit"]


REF_PARAMETER_NAMES = [
    "lmin", "lmax", "_MAX_NUMBER_OF_ITERATIONS_IN_POISSON_EQUATION_", "_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_",
    "_ERROR_THRESHOLD_IN_THE_POISSON_EQUATION_2", "check_poisson_error_method", "multigrid_cycle", "solverPreS",
    "solverfinddphic", "solverPostS", "_NiterPreS_", "_NiterPostS_", "_Niterfinddphic_", "_Iter_branches_solver_",
    "_w_SOR_", "_w_SOR_HEAD_", "head_pot_method", "_CFL_", "_MAX_dt_", "ref_criterion_ptcl", "n_exp",
    "force_stencil", "potential_energy_type", "GC_iter", "Ngrid", "Nparticles_initial", "Nparticles_final",
    "Number_timesteps", "Number_outputs", "fr_output", "MaxIterations", "BoxSize", "vtype", "Maxdt", "G",
    "boundary_type", "GL_total_mass_initial", "GL_total_mass_final"]


def test_parameters_file_in_the_reference_layout(tmp_path):
    """Parameters.dat (output_main_parameters.c:167-230): the 38 lines that
    example_plot.py parses (name, then a number it reads as int or float),
    in the reference's order, then the timing table."""
    print("case: Parameters.dat", flush=True)
    p = plummer(1000, seed=44)
    path = str(tmp_path / "Parameters.dat")
    times = np.linspace(0.0, 1.9, 20)
    with core.Simulation(p, 5, 7, precision=core.PREC_FP64, ref_criterion_ptcl=8) as s:
        s.write_parameters(path, time_unit=2.0, times=times, cfl=0.5, max_dt=1e-3, box_size=40.0, final_time=0.3,
                           total_mass_initial=1.0, total_mass_final=1.0, potential_energy_type=0, gc_iter=50,
                           n_particles_initial=1000, number_timesteps=12, number_outputs=4, fr_output=3,
                           max_iterations=100)
    lines = open(path).read().splitlines()
    names, values = [], []
    for line in lines[:38]:  # the parse example_plot.py applies
        w = line.split()
        names.append(w[0])
        v = float(w[1])
        values.append(int(v) if v == int(v) else v)
    assert names == REF_PARAMETER_NAMES
    got = dict(zip(names, values))
    assert got["lmin"] == 5 and got["lmax"] == 7 and got["Ngrid"] == 33
    assert got["Nparticles_final"] == 1000 and got["Number_outputs"] == 4 and got["vtype"] == 2
    assert got["Number_timesteps"] == 12 and got["fr_output"] == 3 and got["GC_iter"] == 50
    assert got["Maxdt"] == pytest.approx(0.15) and got["_CFL_"] == 0.5
    assert lines[40] == "TIMES in seconds" and lines[42].startswith("TOTAL ")
    assert float(lines[42].split()[1]) == pytest.approx(times.sum(), rel=1e-3)
    assert lines[43].startswith("Global variables = ") and lines[-1].startswith("Output Snapshots = ")
