"""The reference's own host program (initialisation, tree construction and
adaptation, main-loop order) linked against the drop-in libraries, which export
grid_density, potential, grid_acceleration, particle_acceleration, reset and
ptcl_idx_to_box_idx (and, resident, particle_updating_A/B) under the reference's
names and signatures.  Its particle output must match the same program linked
against the reference's own objects."""
import os

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
