"""Seeded synthetic particle sets for the PM gravity step, in code units.

Code units follow the reference's ``input_code_units`` (source/input.c:280-304):
the simulation box is mapped to [0, 1)^3 by ``x' = (x - x_cm) / L_box + 1/2`` and
G = 1.  Physical inputs are given in Plummer units (G = M = a = 1), so with a
box of side ``L`` (in units of a) lengths scale by 1/L, times by L^-3/2 and
velocities by L^1/2 (G and the total mass stay 1).

Samplers:
* ``plummer``: Aarseth-Henon-Wielen inverse CDF for radii, ``r = a / sqrt(X^-2/3 - 1)``
  truncated at ``r_max``; isotropic velocities with speed ``q * sqrt(2) * (1 + r^2)^-1/4``
  where q is drawn from ``g(q) = q^2 (1 - q^2)^{7/2}`` by rejection.
* ``clumps``: half uniform in the box, half in Gaussian clumps (sigma in cells).
* ``two_plummer``: two Plummer spheres on a head-on approach.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Particles:
    """Structure-of-arrays particle set in code units (float64)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def arrays(self):
        return [self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m]

    def copy(self) -> "Particles":
        return Particles(*[a.copy() for a in self.arrays()])

    def write_bin(self, path: str) -> None:
        """Write the layout the reference harness reads: int32 N, then 7 float64 arrays."""
        with open(path, "wb") as f:
            np.array([self.n], dtype=np.int32).tofile(f)
            for a in self.arrays():
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def _plummer_unit(n: int, rng: np.random.Generator, r_max: float = 10.0):
    """Plummer sphere in G = M = a = 1 units, centred on its centre of mass."""
    r = np.empty(n)
    filled = 0
    while filled < n:
        k = (n - filled) * 2 + 16
        X = rng.uniform(0.0, 1.0, size=k)
        X = X[X > 0]
        rr = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= r_max]
        take = min(n - filled, rr.shape[0])
        r[filled:filled + take] = rr[:take]
        filled += take
    ct = rng.uniform(-1.0, 1.0, size=n)
    ph = rng.uniform(0.0, 2.0 * np.pi, size=n)
    st = np.sqrt(1.0 - ct * ct)
    x = r * st * np.cos(ph)
    y = r * st * np.sin(ph)
    z = r * ct
    # speed from the distribution function by rejection on q^2 (1-q^2)^3.5 (max 0.1 bounds it)
    q = np.empty(n)
    filled = 0
    while filled < n:
        k = (n - filled) * 12 + 16
        qq = rng.uniform(0.0, 1.0, size=k)
        yy = rng.uniform(0.0, 0.1, size=k)
        ok = qq[yy < qq * qq * (1.0 - qq * qq) ** 3.5]
        take = min(n - filled, ok.shape[0])
        q[filled:filled + take] = ok[:take]
        filled += take
    v = q * np.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    ct2 = rng.uniform(-1.0, 1.0, size=n)
    ph2 = rng.uniform(0.0, 2.0 * np.pi, size=n)
    st2 = np.sqrt(1.0 - ct2 * ct2)
    vx = v * st2 * np.cos(ph2)
    vy = v * st2 * np.sin(ph2)
    vz = v * ct2
    m = np.full(n, 1.0 / n)
    # centre of mass and of momentum (input.c:252-277)
    for a in (x, y, z, vx, vy, vz):
        a -= np.sum(m * a) / np.sum(m)
    return x, y, z, vx, vy, vz, m


def plummer(n: int, seed: int = 20261203, box: float = 40.0, r_max: float = 10.0) -> Particles:
    """Plummer sphere (G = M = a = 1) mapped to code units in a box of side ``box`` a."""
    rng = np.random.default_rng(seed)
    x, y, z, vx, vy, vz, m = _plummer_unit(n, rng, r_max)
    vs = np.sqrt(box)
    return Particles(x / box + 0.5, y / box + 0.5, z / box + 0.5, vx * vs, vy * vs, vz * vs, m)


def plummer_dt(dt_plummer: float, box: float = 40.0) -> float:
    """Convert a time step in Plummer units to code units."""
    return dt_plummer / box ** 1.5


def clumps(n: int, lmin: int, n_clumps: int = 64, seed: int = 20261203,
           sigma_cells=(1.0, 4.0)) -> Particles:
    """Half the particles uniform in [0,1)^3, half in ``n_clumps`` Gaussian clumps.

    Clump widths are uniform in ``sigma_cells`` head cells (cell = 2^-lmin), centres
    uniform in [4 sigma, 1 - 4 sigma]^3; samples outside the box are redrawn.  Equal
    masses summing to 1, zero velocities.
    """
    rng = np.random.default_rng(seed)
    cell = 1.0 / (1 << lmin)
    n_uni = n // 2
    n_clu = n - n_uni
    pos = np.empty((n, 3))
    pos[:n_uni] = rng.uniform(0.0, 1.0, size=(n_uni, 3))
    per = np.full(n_clumps, n_clu // n_clumps)
    per[: n_clu - per.sum()] += 1
    off = n_uni
    for c in range(n_clumps):
        sig = rng.uniform(sigma_cells[0], sigma_cells[1]) * cell
        ctr = rng.uniform(4 * sig, 1.0 - 4 * sig, size=3)
        k = int(per[c])
        got = 0
        while got < k:
            s = ctr + sig * rng.standard_normal(size=(2 * (k - got) + 8, 3))
            ok = np.all((s > 0.0) & (s < 1.0), axis=1)
            s = s[ok][: k - got]
            pos[off + got: off + got + s.shape[0]] = s
            got += s.shape[0]
        off += k
    z0 = np.zeros(n)
    m = np.full(n, 1.0 / n)
    return Particles(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), z0.copy(), z0.copy(), z0.copy(), m)


def two_plummer(n_each: int, seed: int = 20261203, box: float = 80.0) -> Particles:
    """Two Plummer spheres (a = 1 and 0.5, masses 2:1, total 1) 20a apart, closing at 0.5.

    Positions and velocities are in the centre-of-mass frame of the pair.
    """
    rng = np.random.default_rng(seed)
    a1, a2 = 1.0, 0.5
    m1, m2 = 2.0 / 3.0, 1.0 / 3.0
    p1 = _plummer_unit(n_each, rng)
    p2 = _plummer_unit(n_each, rng)

    def scaled(p, a, mt):
        x, y, z, vx, vy, vz, m = p
        # Plummer sphere of mass mt and radius a: lengths * a, velocities * sqrt(G mt / a)
        vsc = np.sqrt(mt / a)
        return [x * a, y * a, z * a, vx * vsc, vy * vsc, vz * vsc, m * mt]

    s1 = scaled(p1, a1, m1)
    s2 = scaled(p2, a2, m2)
    sep, vrel = 20.0, 0.5
    # COM frame: x1 = -sep * m2/M, x2 = +sep * m1/M; v1 = +vrel*m2, v2 = -vrel*m1 (closing)
    s1[0] = s1[0] - sep * m2
    s2[0] = s2[0] + sep * m1
    s1[3] = s1[3] + vrel * m2
    s2[3] = s2[3] - vrel * m1
    arrs = [np.concatenate([s1[i], s2[i]]) for i in range(7)]
    vs = np.sqrt(box)
    x, y, z, vx, vy, vz, m = arrs
    return Particles(x / box + 0.5, y / box + 0.5, z / box + 0.5, vx * vs, vy * vs, vz * vs, m)
