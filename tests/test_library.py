"""The C-ABI library loads and exports every function its headers declare (CPU only)."""
import ctypes
import glob
import os
import re

import pytest

from quantizer_amd import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# header -> the library that must export its functions
LIBS = {"qz.h": [core.LIB_PATH],
        "qz_dropin.h": [os.path.join(ROOT, "quantizer_amd", "libqzdropin_resident.so")]}
# functions of qz_dropin.h that only the resident build exports
RESIDENT_ONLY = {"particle_updating_A", "particle_updating_B", "tree_adaptation", "timestep", "observables",
                 "output_snapshots", "garbage_collector", "qz_dropin_sync_host"}


def declared_functions(header=None):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", header or "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("qz_create", "qz_step", "qz_grid_density", "qz_potential", "qz_timestep"):
        assert must in names


def test_every_header_has_a_library():
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        assert os.path.basename(h) in LIBS, h


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(core.LIB_PATH)
    missing = [n for n in declared_functions("qz.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_libraries_export_the_reference_entry_points():
    """Both drop-in builds load on their own (the reference globals are weak
    references) and export the reference's function names."""
    force = os.path.join(ROOT, "quantizer_amd", "libqzdropin.so")
    resident = LIBS["qz_dropin.h"][0]
    if not (os.path.exists(force) and os.path.exists(resident)):
        pytest.skip("drop-in libraries are built only against the reference headers")
    names = declared_functions("qz_dropin.h")
    assert {"grid_density", "potential", "ptcl_idx_to_box_idx", "particle_updating_A"} <= set(names)
    lf, lr = ctypes.CDLL(force), ctypes.CDLL(resident)
    assert not [n for n in names if not hasattr(lr, n)]
    assert not [n for n in names if n not in RESIDENT_ONLY and not hasattr(lf, n)]
    assert not [n for n in RESIDENT_ONLY if hasattr(lf, n)]


def test_product_fails_loudly_without_gpu():
    """No CPU fallback: creating a context without a device raises."""
    lib = core.load_library()
    n = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    from quantizer_amd.workloads import plummer
    with pytest.raises(core.QuantizerError, match="no HIP device"):
        core.Simulation(plummer(100, seed=1), 4, 4)
    assert lib is not None
