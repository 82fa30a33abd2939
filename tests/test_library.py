"""The C-ABI library loads and exports every function its headers declare (CPU only)."""
import ctypes
import glob
import os
import re

import pytest

from quantizer_amd import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
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


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(core.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


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
