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
