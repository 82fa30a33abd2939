"""ctypes binding of oracle/liboracle.so and readers for the reference harness dumps.

Test infrastructure: only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "ref_harness")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)


def _lib():
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    lib = ctypes.CDLL(ORACLE_SO)
    lib.oq_create.restype = ctypes.c_void_p
    lib.oq_create.argtypes = [ctypes.c_int] * 4 + [_dp] * 7 + [ctypes.c_int]
    for name in ("oq_destroy", "oq_reset", "oq_grid_density", "oq_grid_acceleration", "oq_particle_acceleration"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
        getattr(lib, name).restype = None
    for name in ("oq_build_tree", "oq_force", "oq_potential"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.oq_update_A.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.oq_update_A.restype = ctypes.c_int
    lib.oq_update_B.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.oq_update_B.restype = None
    lib.oq_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.oq_step.restype = ctypes.c_int
    lib.oq_get_particles.argtypes = [ctypes.c_void_p, _dp]
    lib.oq_set_particles.argtypes = [ctypes.c_void_p, _dp]
    lib.oq_get_head.argtypes = [ctypes.c_void_p, _dp]
    lib.oq_set_head_pot.argtypes = [ctypes.c_void_p, _dp]
    lib.oq_num_nodes.argtypes = [ctypes.c_void_p]
    lib.oq_num_nodes.restype = ctypes.c_int
    lib.oq_node_header.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
    lib.oq_node_ints.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip]
    lib.oq_node_fields.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.oq_stats.argtypes = [ctypes.c_void_p, _ip]
    lib.oq_observables.argtypes = [ctypes.c_void_p, _dp]
    lib.oq_key.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.oq_key.restype = ctypes.c_int
    lib.oq_head_solve.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.oq_head_solve.restype = ctypes.c_int
    lib.oq_set_tolerance.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.oq_cfl_dt.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double]
    lib.oq_cfl_dt.restype = ctypes.c_double
    lib.oq_timestep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    lib.oq_timestep.restype = ctypes.c_double
    lib.oq_grid_pe.argtypes = [ctypes.c_void_p]
    lib.oq_grid_pe.restype = ctypes.c_double
    lib.oq_force_vjp.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.oq_force_vjp.restype = ctypes.c_int
    lib.oq_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
    lib.oq_adjoint.restype = ctypes.c_int
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


HDR_FIELDS = ["lv", "lvrel", "idx", "parent", "ts_x", "ts_y", "ts_z", "rd_x", "rd_y", "rd_z",
              "min_x", "min_y", "min_z", "max_x", "max_y", "max_z", "ncell", "nintr", "nbdry", "nsim",
              "nptcl_full", "nptcl_outs", "nchild", "gcap"]


class OracleSim:
    """The C restatement as a Python object (FP64, sequential)."""

    def __init__(self, parts, lmin, lmax, ref_crit=8, head_bc=0):
        self.lmin, self.lmax = lmin, lmax
        self.n = parts.n
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in parts.arrays()]
        self._keep = arrs
        self.h = lib().oq_create(lmin, lmax, ref_crit, self.n, *[_d(a) for a in arrs], head_bc)

    def close(self):
        if self.h:
            lib().oq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tolerance(self, tol):
        lib().oq_set_tolerance(self.h, tol)

    def force(self):
        rc = lib().oq_force(self.h)
        if rc:
            raise RuntimeError("oracle force pass failed (Poisson did not converge)")

    def step(self, dt):
        rc = lib().oq_step(self.h, dt)
        if rc:
            raise RuntimeError("oracle step failed")

    def particles(self):
        out = np.empty(9 * self.n)
        lib().oq_get_particles(self.h, _d(out))
        return out.reshape(9, self.n)

    def set_particles(self, arr9):
        a = np.ascontiguousarray(arr9, dtype=np.float64).reshape(-1)
        lib().oq_set_particles(self.h, _d(a))

    def set_head_pot(self, pot):
        a = np.ascontiguousarray(pot, dtype=np.float64).reshape(-1)
        lib().oq_set_head_pot(self.h, _d(a))

    def head(self):
        n1 = (1 << self.lmin) + 1
        out = np.empty(5 * n1 ** 3)
        lib().oq_get_head(self.h, _d(out))
        return out.reshape(5, n1, n1, n1)  # [field, z, y, x]

    def stats(self):
        out = np.zeros(4, dtype=np.int32)
        lib().oq_stats(self.h, _i(out))
        return dict(vcycles=int(out[0]), cg_iters=int(out[1]), cg_solves=int(out[2]), nodes=int(out[3]))

    def observables(self):
        out = np.empty(6)
        lib().oq_observables(self.h, _d(out))
        return out

    def timestep(self, cfl=0.5, dt_max=1e300):
        return float(lib().oq_timestep(self.h, cfl, dt_max))

    def grid_pe(self):
        return float(lib().oq_grid_pe(self.h))

    def force_vjp(self, abar):
        a = np.ascontiguousarray(abar, dtype=np.float64).reshape(-1)
        out = np.zeros(3 * self.n)
        if lib().oq_force_vjp(self.h, _d(a), _d(out)):
            raise RuntimeError("oracle force_vjp failed")
        return out.reshape(3, self.n)

    def adjoint(self, nsteps, dt, xbar_n, vbar_n):
        xb = np.ascontiguousarray(xbar_n, dtype=np.float64).reshape(-1)
        vb = np.ascontiguousarray(vbar_n, dtype=np.float64).reshape(-1)
        ox = np.zeros(3 * self.n)
        ov = np.zeros(3 * self.n)
        if lib().oq_adjoint(self.h, nsteps, dt, _d(xb), _d(vb), _d(ox), _d(ov)):
            raise RuntimeError("oracle adjoint failed")
        return ox.reshape(3, self.n), ov.reshape(3, self.n)

    def nodes(self):
        res = []
        for i in range(lib().oq_num_nodes(self.h)):
            hdr = np.zeros(24, dtype=np.int32)
            lib().oq_node_header(self.h, i, _i(hdr))
            d = dict(zip(HDR_FIELDS, [int(v) for v in hdr]))
            sizes = {0: 3 * d["ncell"], 1: d["ncell"], 2: 3 * d["nintr"], 3: d["nintr"], 4: 3 * d["nbdry"],
                     5: d["nbdry"], 6: 3 * d["nsim"], 7: d["nsim"], 8: d["rd_x"] * d["rd_y"] * d["rd_z"]}
            names = {0: "cell_xyz", 1: "cell_box", 2: "intr_xyz", 3: "intr_idx", 4: "bdry_xyz", 5: "bdry_idx",
                     6: "sim_xyz", 7: "sim_idx", 8: "box"}
            for w, sz in sizes.items():
                a = np.zeros(max(sz, 1), dtype=np.int32)
                lib().oq_node_ints(self.h, i, w, _i(a))
                a = a[:sz]
                if w in (0, 2, 4, 6):
                    a = a.reshape(-1, 3)
                d[names[w]] = a
            f = np.empty(5 * d["gcap"])
            lib().oq_node_fields(self.h, i, _d(f))
            f = f.reshape(5, d["gcap"])
            for k, nm in enumerate(("rho", "pot", "ax", "ay", "az")):
                d[nm] = f[k]
            res.append(d)
        return res


def oracle_key(lv, ts, rd, x, y, z):
    ts = np.ascontiguousarray(ts, dtype=np.int32)
    rd = np.ascontiguousarray(rd, dtype=np.int32)
    return lib().oq_key(lv, _i(ts), _i(rd), x, y, z)


def oracle_head_solve(lmin, phi, rho, node_mass, tol=1e-10, max_iter=5000):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    it = lib().oq_head_solve(lmin, _d(phi), _d(rho), node_mass, tol, max_iter)
    return phi, it


# ---------------------------------------------------------------- reference dumps
def read_head(path, lmin):
    n1 = (1 << lmin) + 1
    a = np.fromfile(path, dtype=np.float64)
    return a.reshape(5, n1, n1, n1)


def read_particles(path, n, with_xv):
    a = np.fromfile(path, dtype=np.float64)
    return a.reshape(9 if with_xv else 3, n)


def read_nodes(path):
    raw = open(path, "rb").read()
    off = 0

    def ints(k):
        nonlocal off
        a = np.frombuffer(raw, dtype=np.int32, count=k, offset=off)
        off += 4 * k
        return a.copy()

    def dbls(k):
        nonlocal off
        a = np.frombuffer(raw, dtype=np.float64, count=k, offset=off)
        off += 8 * k
        return a.copy()

    total = int(ints(1)[0])
    res = []
    for _ in range(total):
        hdr = ints(24)
        d = dict(zip(HDR_FIELDS, [int(v) for v in hdr]))
        nc, ni, nb, ns = d["ncell"], d["nintr"], d["nbdry"], d["nsim"]
        cx, cy, cz = ints(nc), ints(nc), ints(nc)
        d["cell_xyz"] = np.stack([cx, cy, cz], axis=1)
        d["cell_box"] = ints(nc)
        for nm, k in (("intr", ni), ("bdry", nb), ("sim", ns)):
            x, y, z = ints(k), ints(k), ints(k)
            d[nm + "_xyz"] = np.stack([x, y, z], axis=1)
            d[nm + "_idx"] = ints(k)
        g = d["gcap"]
        for nm in ("rho", "pot", "ax", "ay", "az"):
            d[nm] = dbls(g)
        res.append(d)
    return res


def run_reference(parts, outdir, lmin, lmax, ref_crit, nsteps, dt, extra=(), binary=REF_BIN):
    """Run oracle/_ref/ref_harness (the compiled reference) on ``parts``; ``binary``
    selects one of the harness builds linked against the drop-in libraries."""
    os.makedirs(outdir, exist_ok=True)
    inp = os.path.join(outdir, "particles.bin")
    parts.write_bin(inp)
    cmd = [binary, inp, outdir, str(lmin), str(lmax), str(ref_crit), str(nsteps), repr(float(dt)), *extra]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"reference harness failed: {p.stderr}")
    return p.stderr


def cfl_dt(v3, a3, H, cfl=0.5, dt_max=1e300):
    """The oracle's CFL rule on explicit arrays (3 x n velocities and accelerations, n cell sizes)."""
    v = np.ascontiguousarray(v3, dtype=np.float64).reshape(-1)
    a = np.ascontiguousarray(a3, dtype=np.float64).reshape(-1)
    h = np.ascontiguousarray(H, dtype=np.float64)
    return float(lib().oq_cfl_dt(h.shape[0], _d(v), _d(a), _d(h), cfl, dt_max))
