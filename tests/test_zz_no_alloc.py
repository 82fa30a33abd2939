"""qz_observables_noalloc (partial sums in the library's fixed device array)
against qz_observables on the same particles, through the C ABI.

The inputs are FP64 device arrays the test allocates itself; the entry point
allocates nothing.  Two calls in a row, the second on fewer particles, so a
result left in the shared array by the first call cannot pass for the second.
The sums run in a different particle order than the context's (which keeps its
particles sorted by cell), so agreement is to FP64 rounding: 1e-12 relative.
"""
import ctypes

import numpy as np
import pytest

from quantizer_amd import core
from quantizer_amd.workloads import plummer

pytestmark = pytest.mark.gpu

RTOL = 1e-12


class DeviceArrays:
    """FP64 device copies of host arrays (hipMalloc through the HIP runtime)."""

    def __init__(self, arrays):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64)
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
            self.ptrs.append(p)
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def addresses(self):
        return [p.value for p in self.ptrs]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def particles(n, seed):
    p = plummer(n, seed=seed)
    rng = np.random.default_rng(seed)
    p.m[:] = p.m * (0.5 + rng.random(n))  # unequal masses, so m enters every sum
    return p


def with_context(p):
    with core.Simulation(p, 4, 4, precision=core.PREC_FP64) as s:
        return s.observables(exact_pe=True)


def check(out, ref, p):
    # momenta are near zero in the centre-of-mass frame: compare them on the
    # scale of sum m |v|
    pscale = float(np.sum(p.m * (np.abs(p.vx) + np.abs(p.vy) + np.abs(p.vz))))
    np.testing.assert_allclose(out[:3], ref[:3], rtol=RTOL, atol=0.0)
    np.testing.assert_allclose(out[3:], ref[3:], rtol=0.0, atol=RTOL * pscale)


def test_noalloc_matches_observables_on_two_calls():
    big, small = particles(20000, 11), particles(3000, 12)
    refs = [with_context(big), with_context(small)]
    for p, ref in zip((big, small), refs):
        d = DeviceArrays([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m])
        try:
            out = core.observables_noalloc(d.addresses(), p.n, 1.0, with_pe=True)
            check(out, ref, p)
            # without the pair sum: E_pot = 0 and the other values unchanged
            kin = core.observables_noalloc(d.addresses(), p.n, 1.0, with_pe=False)
            assert kin[1] == 0.0 and kin[2] == kin[0]
            np.testing.assert_array_equal(kin[[0, 3, 4, 5]], out[[0, 3, 4, 5]])
        finally:
            d.free()


def test_noalloc_writes_device_output_and_handles_empty_input():
    p = particles(5000, 13)
    ref = with_context(p)
    d = DeviceArrays([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m, np.zeros(6)])
    try:
        lib = core.load_library()
        addr = d.addresses()
        out_dev = ctypes.cast(ctypes.c_void_p(addr[7]), ctypes.POINTER(ctypes.c_double))
        assert lib.qz_observables_noalloc(*[ctypes.c_void_p(a) for a in addr[:7]], p.n, 1.0, 1, out_dev) == 0
        out = np.empty(6)
        assert d.hip.hipMemcpy(out.ctypes.data, addr[7], 48, 2) == 0  # device to host
        check(out, ref, p)
        np.testing.assert_array_equal(core.observables_noalloc(addr[:7], 0, 1.0, with_pe=True), np.zeros(6))
    finally:
        d.free()


def test_noalloc_refuses_more_particles_than_its_scratch_holds():
    p = particles(100, 14)
    d = DeviceArrays([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m])
    try:
        # refused before any launch: the arrays are never read
        with pytest.raises(core.QuantizerError, match="fixed scratch"):
            core.observables_noalloc(d.addresses(), (1 << 24) + 1, 1.0, with_pe=True)
        with pytest.raises(core.QuantizerError, match="fixed scratch"):
            core.observables_noalloc(d.addresses(), -1, 1.0, with_pe=True)
    finally:
        d.free()
