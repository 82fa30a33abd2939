"""Python host mirror of the QUANTIZER per-step PM interface over libqzhip.so.

The method names and their order follow the reference step loop
(source/main.c:369-535): ``particle_updating_A(dt)``, ``reset()``,
``tree_adaptation()``, ``grid_density()``, ``potential()`` (=
``potential_head_node()`` + ``potential_branches()``), ``grid_acceleration()``,
``particle_acceleration()``, ``particle_updating_B(dt)``.  The reference
returns ``_SUCCESS_``/``_FAILURE_`` (0/1, common.h:81-82) and prints a reason;
here a failure raises :class:`QuantizerError` carrying ``qz_last_error()``,
and success returns ``SUCCESS`` (0).

All compute runs in the HIP library; this module only moves arguments.  There
is no CPU fallback: if the library is missing or no GPU is present, creating a
:class:`Simulation` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqzhip.so")

SUCCESS = 0
FAILURE = 1

PREC_FP64 = 0
PREC_MIXED = 1
BC_MONOPOLE = 0
BC_ZERO = 1
BRANCH_CG = 0
BRANCH_MG = 1


class QuantizerError(RuntimeError):
    pass


class QzConfig(ctypes.Structure):
    _fields_ = [
        ("lmin", ctypes.c_int),
        ("lmax", ctypes.c_int),
        ("n_particles", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("head_bc", ctypes.c_int),
        ("ref_criterion_ptcl", ctypes.c_int),
        ("n_exp", ctypes.c_int),
        ("G", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("tol_change", ctypes.c_double),
        ("max_iterations", ctypes.c_int),
        ("inner_cycles", ctypes.c_int),
        ("pre_sweeps", ctypes.c_int),
        ("post_sweeps", ctypes.c_int),
        ("branch_solver", ctypes.c_int),
    ]


class QzStats(ctypes.Structure):
    _fields_ = [
        ("head_outer", ctypes.c_int),
        ("head_vcycles", ctypes.c_int),
        ("head_residual", ctypes.c_double),
        ("head_change", ctypes.c_double),
        ("n_nodes", ctypes.c_int),
        ("n_levels", ctypes.c_int),
        ("branch_points", ctypes.c_longlong),
        ("branch_intr", ctypes.c_longlong),
        ("branch_iterations", ctypes.c_int),
        ("branch_iter_points", ctypes.c_longlong),
        ("branch_cg_levels", ctypes.c_int),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_LIB = None

EXPORTED = [
    "qz_default_config", "qz_create", "qz_destroy", "qz_last_error", "qz_set_particles", "qz_load_particles",
    "qz_get_particles",
    "qz_particle_updating_A", "qz_reset", "qz_tree_adaptation", "qz_grid_density", "qz_potential",
    "qz_potential_head_node", "qz_potential_branches", "qz_grid_acceleration", "qz_particle_acceleration",
    "qz_particle_updating_B", "qz_force", "qz_step", "qz_get_head_fields", "qz_set_head_potential",
    "qz_num_nodes", "qz_node_header", "qz_node_grid_lists", "qz_node_fields", "qz_get_stats", "qz_observables",
    "qz_timestep", "qz_force_vjp", "qz_adjoint_forward", "qz_adjoint_backward", "qz_keys", "qz_timing_enable", "qz_timing_report", "qz_sync",
    "qz_bench_sweep", "qz_copy_probe", "qz_write_snapshot",
]


class QzUnits(ctypes.Structure):
    _fields_ = [("dist", ctypes.c_double), ("velocity", ctypes.c_double), ("time", ctypes.c_double),
                ("cm", ctypes.c_double * 3), ("cm_v", ctypes.c_double * 3)]


class QzRunInfo(ctypes.Structure):
    _fields_ = [("cfl", ctypes.c_double), ("max_dt", ctypes.c_double), ("box_size", ctypes.c_double),
                ("final_time", ctypes.c_double), ("total_mass_initial", ctypes.c_double),
                ("total_mass_final", ctypes.c_double), ("potential_energy_type", ctypes.c_int),
                ("gc_iter", ctypes.c_int), ("n_particles_initial", ctypes.c_int),
                ("number_timesteps", ctypes.c_int), ("number_outputs", ctypes.c_int), ("fr_output", ctypes.c_int),
                ("max_iterations", ctypes.c_int), ("times", ctypes.POINTER(ctypes.c_double))]


def load_library(path: str = LIB_PATH):
    """Load libqzhip.so and declare the C ABI of include/qz.h."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise QuantizerError(f"HIP library not built: {path} (run quantizer_amd/build.py)")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.qz_default_config.argtypes = [ctypes.POINTER(QzConfig)]
    lib.qz_default_config.restype = None
    lib.qz_create.argtypes = [ctypes.POINTER(QzConfig)]
    lib.qz_create.restype = vp
    lib.qz_destroy.argtypes = [vp]
    lib.qz_destroy.restype = None
    lib.qz_last_error.restype = ctypes.c_char_p
    lib.qz_set_particles.argtypes = [vp] + [_dp] * 7
    lib.qz_load_particles.argtypes = [vp] + [_dp] * 7
    lib.qz_get_particles.argtypes = [vp] + [_dp] * 9
    for nm in ("qz_reset", "qz_tree_adaptation", "qz_grid_density", "qz_potential", "qz_potential_head_node",
               "qz_potential_branches", "qz_grid_acceleration", "qz_particle_acceleration", "qz_force", "qz_sync",
               "qz_num_nodes"):
        getattr(lib, nm).argtypes = [vp]
        getattr(lib, nm).restype = ctypes.c_int
    for nm in ("qz_particle_updating_A", "qz_particle_updating_B", "qz_step"):
        getattr(lib, nm).argtypes = [vp, ctypes.c_double]
        getattr(lib, nm).restype = ctypes.c_int
    lib.qz_get_head_fields.argtypes = [vp, _dp]
    lib.qz_set_head_potential.argtypes = [vp, _dp]
    lib.qz_node_header.argtypes = [vp, ctypes.c_int, _ip]
    lib.qz_node_grid_lists.argtypes = [vp, ctypes.c_int, _ip, _ip, _ip]
    lib.qz_node_fields.argtypes = [vp, ctypes.c_int, _dp]
    lib.qz_get_stats.argtypes = [vp, ctypes.POINTER(QzStats)]
    lib.qz_observables.argtypes = [vp, ctypes.c_int, _dp]
    lib.qz_timestep.argtypes = [vp, ctypes.c_double, ctypes.c_double, _dp]
    lib.qz_force_vjp.argtypes = [vp, _dp, _dp]
    lib.qz_adjoint_forward.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.qz_adjoint_backward.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.qz_keys.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip]
    lib.qz_timing_enable.argtypes = [vp, ctypes.c_int]
    lib.qz_timing_report.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int]
    lib.qz_timing_report.restype = ctypes.c_int
    lib.qz_observables_noalloc.argtypes = [vp] * 7 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    lib.qz_observables_noalloc.restype = ctypes.c_int
    _LIB = lib
    return lib


def last_error() -> str:
    return load_library().qz_last_error().decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise QuantizerError(f"{what}: {last_error()}")
    return SUCCESS


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


HDR_FIELDS = ["lv", "lvrel", "idx", "parent", "ts_x", "ts_y", "ts_z", "rd_x", "rd_y", "rd_z",
              "min_x", "min_y", "min_z", "max_x", "max_y", "max_z", "ncell", "nintr", "nbdry", "nsim",
              "nptcl_full", "nptcl_outs", "nchild", "gcap"]


class Simulation:
    """One device-resident PM simulation (head grid + refinement nodes + particles)."""

    def __init__(self, particles, lmin: int, lmax: int | None = None, precision: int = PREC_FP64,
                 head_bc: int = BC_MONOPOLE, ref_criterion_ptcl: int = 8, tol: float = 1e-10,
                 tol_change: float | None = None, inner_cycles: int = 0, max_iterations: int = 5000,
                 pre_sweeps: int = 2, post_sweeps: int = 2, G: float = 1.0,
                 branch_solver: int | None = None):
        lib = load_library()
        cfg = QzConfig()
        lib.qz_default_config(ctypes.byref(cfg))
        cfg.lmin = lmin
        cfg.lmax = lmin if lmax is None else lmax
        cfg.n_particles = particles.n
        cfg.precision = precision
        cfg.head_bc = head_bc
        cfg.ref_criterion_ptcl = ref_criterion_ptcl
        cfg.tol = tol
        cfg.tol_change = tol if tol_change is None else tol_change
        cfg.inner_cycles = inner_cycles
        cfg.max_iterations = max_iterations
        cfg.pre_sweeps = pre_sweeps
        cfg.post_sweeps = post_sweeps
        if branch_solver is not None:
            cfg.branch_solver = branch_solver
        cfg.G = G
        self.cfg = cfg
        self.lmin, self.lmax, self.n = cfg.lmin, cfg.lmax, particles.n
        self._h = lib.qz_create(ctypes.byref(cfg))
        if not self._h:
            raise QuantizerError(f"qz_create: {last_error()}")
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in particles.arrays()]
        _check(lib.qz_set_particles(self._h, *[_d(a) for a in arrs]), "qz_set_particles")

    def set_particles(self, particles):
        """Replace the particle state (qz_set_particles: the head potential history restarts)."""
        if particles.n != self.n:
            raise QuantizerError(f"set_particles: {particles.n} particles, the context holds {self.n}")
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in particles.arrays()]
        _check(load_library().qz_set_particles(self._h, *[_d(a) for a in arrs]), "qz_set_particles")

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None):
            load_library().qz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ step stages (main.c:391-478)
    def particle_updating_A(self, dt: float):
        return _check(load_library().qz_particle_updating_A(self._h, dt), "particle_updating_A")

    def reset(self):
        return _check(load_library().qz_reset(self._h), "reset")

    def tree_adaptation(self):
        return _check(load_library().qz_tree_adaptation(self._h), "tree_adaptation")

    def grid_density(self):
        return _check(load_library().qz_grid_density(self._h), "grid_density")

    def potential(self):
        return _check(load_library().qz_potential(self._h), "potential")

    def potential_head_node(self):
        return _check(load_library().qz_potential_head_node(self._h), "potential_head_node")

    def potential_branches(self):
        return _check(load_library().qz_potential_branches(self._h), "potential_branches")

    def grid_acceleration(self):
        return _check(load_library().qz_grid_acceleration(self._h), "grid_acceleration")

    def particle_acceleration(self):
        return _check(load_library().qz_particle_acceleration(self._h), "particle_acceleration")

    def particle_updating_B(self, dt: float):
        return _check(load_library().qz_particle_updating_B(self._h, dt), "particle_updating_B")

    def force(self):
        """Initial force pass (main.c:280-338): reset, tree, density, potential, accelerations."""
        return _check(load_library().qz_force(self._h), "force")

    def step(self, dt: float):
        """One fixed-dt KDK step (main.c:391-478)."""
        return _check(load_library().qz_step(self._h, dt), "step")

    def sync(self):
        return _check(load_library().qz_sync(self._h), "sync")

    def write_snapshot(self, path: str, actualtime: float, dist: float = 1.0, velocity: float = 1.0,
                       time_unit: float = 1.0, cm=(0.0, 0.0, 0.0), cm_v=(0.0, 0.0, 0.0),
                       observables: bool = True):
        """The state as a snapshot file in the reference's layout (output_snapshots.c)."""
        lib = load_library()
        u = QzUnits(dist, velocity, time_unit, (ctypes.c_double * 3)(*cm), (ctypes.c_double * 3)(*cm_v))
        lib.qz_write_snapshot.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(QzUnits), ctypes.c_double,
                                          ctypes.c_int]
        _check(lib.qz_write_snapshot(self._h, os.fsencode(path), ctypes.byref(u), actualtime, int(observables)),
               "qz_write_snapshot")

    def write_snapshot_async(self, path: str, actualtime: float, dist: float = 1.0, velocity: float = 1.0,
                             time_unit: float = 1.0, cm=(0.0, 0.0, 0.0), cm_v=(0.0, 0.0, 0.0), obs6=None):
        """qz_write_snapshot_async: returns once the state is copied on the device;
        the file is written by a host thread (snapshot_wait joins it).  obs6:
        E_kin, E_pot, E_tot, px, py, pz to write after the arrays, or None."""
        lib = load_library()
        self._units = QzUnits(dist, velocity, time_unit, (ctypes.c_double * 3)(*cm), (ctypes.c_double * 3)(*cm_v))
        lib.qz_write_snapshot_async.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(QzUnits),
                                                ctypes.c_double, _dp]
        o = None if obs6 is None else np.ascontiguousarray(obs6, dtype=np.float64)
        _check(lib.qz_write_snapshot_async(self._h, os.fsencode(path), ctypes.byref(self._units), actualtime,
                                           None if o is None else _d(o)), "qz_write_snapshot_async")

    def snapshot_wait(self):
        lib = load_library()
        lib.qz_snapshot_wait.argtypes = [ctypes.c_void_p]
        return _check(lib.qz_snapshot_wait(self._h), "qz_snapshot_wait")

    def write_parameters(self, path: str, time_unit: float = 1.0, times=None, **info):
        """Parameters.dat in the reference layout (output_main_parameters.c); info
        holds the QzRunInfo fields other than times."""
        lib = load_library()
        u = QzUnits(1.0, 1.0, time_unit, (ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(0, 0, 0))
        t = None if times is None else np.ascontiguousarray(times, dtype=np.float64)
        if t is not None and t.shape[0] != 20:
            raise QuantizerError("write_parameters: times holds 20 values")
        r = QzRunInfo(**info)
        r.times = _d(t) if t is not None else None
        lib.qz_write_parameters.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(QzUnits),
                                            ctypes.POINTER(QzRunInfo)]
        _check(lib.qz_write_parameters(self._h, os.fsencode(path), ctypes.byref(u), ctypes.byref(r)),
               "qz_write_parameters")

    # ------------------------------------------------------------ measurement
    def bench_sweep(self, reps: int = 50):
        """(ms, algorithmic bytes) per launch of the head's finest red half sweep."""
        lib = load_library()
        ms, by = ctypes.c_double(0), ctypes.c_double(0)
        lib.qz_bench_sweep.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        _check(lib.qz_bench_sweep(self._h, reps, ctypes.byref(ms), ctypes.byref(by)), "qz_bench_sweep")
        return ms.value, by.value

    def copy_probe(self, nbytes: int) -> float:
        """ms of one plain device copy of nbytes (kernel k_copy_probe)."""
        lib = load_library()
        ms = ctypes.c_double(0)
        lib.qz_copy_probe.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _dp]
        _check(lib.qz_copy_probe(self._h, nbytes, ctypes.byref(ms)), "qz_copy_probe")
        return ms.value

    # ------------------------------------------------------------ data access
    def particles(self) -> np.ndarray:
        """(9, N) array x,y,z,vx,vy,vz,ax,ay,az in the original particle order."""
        out = np.empty((9, self.n))
        _check(load_library().qz_get_particles(self._h, *[_d(out[q]) for q in range(9)]), "qz_get_particles")
        return out

    def head_fields(self) -> np.ndarray:
        """(5, n1, n1, n1) rho, pot, ax, ay, az of the head solve domain, indexed [field, z, y, x]."""
        n1 = (1 << self.lmin) + 1
        out = np.empty(5 * n1 ** 3)
        _check(load_library().qz_get_head_fields(self._h, _d(out)), "qz_get_head_fields")
        return out.reshape(5, n1, n1, n1)

    def set_head_potential(self, pot: np.ndarray):
        a = np.ascontiguousarray(pot, dtype=np.float64).reshape(-1)
        return _check(load_library().qz_set_head_potential(self._h, _d(a)), "qz_set_head_potential")

    def stats(self) -> dict:
        st = QzStats()
        _check(load_library().qz_get_stats(self._h, ctypes.byref(st)), "qz_get_stats")
        return {f: getattr(st, f) for f, _ in QzStats._fields_}

    def observables(self, exact_pe: bool = True) -> np.ndarray:
        out = np.empty(6)
        _check(load_library().qz_observables(self._h, 1 if exact_pe else 0, _d(out)), "qz_observables")
        return out

    def nodes(self) -> list:
        lib = load_library()
        res = []
        for i in range(lib.qz_num_nodes(self._h)):
            hdr = np.zeros(24, dtype=np.int32)
            _check(lib.qz_node_header(self._h, i, _i(hdr)), "qz_node_header")
            d = dict(zip(HDR_FIELDS, [int(v) for v in hdr]))
            ix = np.zeros((max(d["nintr"], 1), 3), dtype=np.int32)
            bx = np.zeros((max(d["nbdry"], 1), 3), dtype=np.int32)
            sx = np.zeros((max(d["nsim"], 1), 3), dtype=np.int32)
            _check(lib.qz_node_grid_lists(self._h, i, _i(ix), _i(bx), _i(sx)), "qz_node_grid_lists")
            d["intr_xyz"], d["bdry_xyz"], d["sim_xyz"] = ix[: d["nintr"]], bx[: d["nbdry"]], sx[: d["nsim"]]
            f = np.empty(5 * d["gcap"])
            _check(lib.qz_node_fields(self._h, i, _d(f)), "qz_node_fields")
            f = f.reshape(5, d["gcap"])
            for k, nm in enumerate(("rho", "pot", "ax", "ay", "az")):
                d[nm] = f[k]
            res.append(d)
        return res

    def timestep(self, cfl: float = 0.5, dt_max: float = 1e300) -> float:
        """CFL time step from the current velocities and accelerations (timestep.c:589-699)."""
        out = np.zeros(1)
        _check(load_library().qz_timestep(self._h, cfl, dt_max, _d(out)), "qz_timestep")
        return float(out[0])

    # ------------------------------------------------------------ adjoint
    def force_vjp(self, abar: np.ndarray) -> np.ndarray:
        """J_A(x)^T abar for the force pass at the current positions; abar is (3, N)."""
        a = np.ascontiguousarray(abar, dtype=np.float64).reshape(-1)
        out = np.empty(3 * self.n)
        _check(load_library().qz_force_vjp(self._h, _d(a), _d(out)), "qz_force_vjp")
        return out.reshape(3, self.n)

    def adjoint_forward(self, nsteps: int, dt: float, checkpoint_every: int):
        return _check(load_library().qz_adjoint_forward(self._h, nsteps, dt, checkpoint_every), "qz_adjoint_forward")

    def adjoint_backward(self, xbar_n: np.ndarray, vbar_n: np.ndarray):
        """Cotangents of the initial (x, v) from those of the final state, each (3, N)."""
        xb = np.ascontiguousarray(xbar_n, dtype=np.float64).reshape(-1)
        vb = np.ascontiguousarray(vbar_n, dtype=np.float64).reshape(-1)
        ox = np.empty(3 * self.n)
        ov = np.empty(3 * self.n)
        _check(load_library().qz_adjoint_backward(self._h, _d(xb), _d(vb), _d(ox), _d(ov)), "qz_adjoint_backward")
        return ox.reshape(3, self.n), ov.reshape(3, self.n)

    def adjoint_backward_ptr(self, xbar_n: int, vbar_n: int, xbar_0: int, vbar_0: int):
        """adjoint_backward on caller-owned buffers of 3 N doubles given by address,
        in host or device memory (device-resident cotangents stay off PCIe)."""
        vp = ctypes.c_void_p
        _check(load_library().qz_adjoint_backward(self._h, ctypes.cast(vp(xbar_n), _dp), ctypes.cast(vp(vbar_n), _dp),
                                                  ctypes.cast(vp(xbar_0), _dp), ctypes.cast(vp(vbar_0), _dp)),
               "qz_adjoint_backward")

    # ------------------------------------------------------------ timing
    def timing(self, on: bool = True, host_only: bool = False):
        mode = (2 if host_only else 1) if on else 0
        return _check(load_library().qz_timing_enable(self._h, mode), "qz_timing_enable")

    def timing_report(self) -> dict:
        lib = load_library()
        mx, nl = 128, 48
        names = ctypes.create_string_buffer(mx * nl)
        ms = np.zeros(mx)
        cnt = np.zeros(mx, dtype=np.int32)
        byt = np.zeros(mx)
        k = lib.qz_timing_report(self._h, names, nl, _d(ms), _i(cnt), _d(byt), mx)
        raw = names.raw
        res = {}
        for q in range(k):
            nm = raw[q * nl:(q + 1) * nl].split(b"\0", 1)[0].decode()
            res[nm] = {"ms": float(ms[q]), "launches": int(cnt[q]), "bytes": float(byt[q])}
        return res


def box_keys(x, y, z, lv: int, ts, rd) -> np.ndarray:
    """ptcl_idx_to_box_idx (ptcl_idx_to_box_idx.c:30-139) for many points, on the device."""
    lib = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    out = np.empty(x.shape[0], dtype=np.int32)
    ts = np.ascontiguousarray(ts, dtype=np.int32)
    rd = np.ascontiguousarray(rd, dtype=np.int32)
    _check(lib.qz_keys(_d(x), _d(y), _d(z), x.shape[0], lv, _i(ts), _i(rd), _i(out)), "qz_keys")
    return out


def observables_noalloc(dev_ptrs, n: int, G: float = 1.0, with_pe: bool = True) -> np.ndarray:
    """qz_observables_noalloc on FP64 device arrays x, y, z, vx, vy, vz, m (device
    addresses as ints, n particles): E_kin, E_pot, E_tot, px, py, pz."""
    out = np.empty(6)
    _check(load_library().qz_observables_noalloc(*[ctypes.c_void_p(p) for p in dev_ptrs], n, G,
                                                 1 if with_pe else 0, _d(out)), "qz_observables_noalloc")
    return out
