"""Seeded initial-condition generators and the 8-column text IC format.

The text format is the one NEXT's loader reads (load_particle.hpp:100-116):
one particle per line, whitespace separated ``x y z vx vy vz m type``.  The
reference generators (tools/icbuilder.py:4-7 writes the same columns) are
unseeded and have no distribution-function velocities, so these are written
from the published models instead:

* Plummer sphere: radii from the inverse cumulative mass profile
  r = a (u^{-2/3} - 1)^{-1/2} (truncated at ``rmax``), isotropic directions,
  speeds by von Neumann rejection on g(q) = q^2 (1 - q^2)^{7/2} with
  v = q v_esc(r), v_esc = sqrt(2GM/a) (1 + r^2/a^2)^{-1/4}  (Aarseth, Henon &
  Wielen 1974).  Centre of mass and its velocity are removed, and velocities
  are rescaled to virial ratio T/|W| = 1/2 with W from an FP64 direct sum.
* Uniform ball: r = R u^{1/3}, isotropic, at rest (cold collapse).
* Two-cluster merger: two Plummer spheres shifted and boosted.
* Ensemble sizes: N_i = round(exp(U[ln lo, ln hi])) from seed i.

All draws use numpy's PCG64 generator with the given seed, so a config
regenerates bit-for-bit on any machine.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np


@dataclass
class Particles:
    """Structure of arrays mirroring NEXT's ``Particle`` (particle.h:22-52)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray
    type: np.ndarray

    def __len__(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "Particles":
        return Particles(*(np.array(a, copy=True) for a in self.arrays()))

    def arrays(self):
        return (self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m, self.type)

    @staticmethod
    def from_columns(cols: np.ndarray, types: np.ndarray) -> "Particles":
        c = np.ascontiguousarray(cols, dtype=np.float64)
        return Particles(*(np.ascontiguousarray(c[:, k]) for k in range(7)),
                         np.ascontiguousarray(types, dtype=np.int32))

    @staticmethod
    def concat(parts) -> "Particles":
        return Particles(*(np.concatenate(col) for col in zip(*(p.arrays() for p in parts))))


def _isotropic(rng: np.random.Generator, n: int):
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    return sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t


def potential_energy(x, y, z, m, eps: float = 0.0, chunk: int = 2048) -> float:
    """FP64 direct-sum W = -sum_{i<j} m_i m_j / sqrt(r_ij^2 + eps^2)."""
    pos = np.stack([x, y, z], axis=1)
    n = pos.shape[0]
    w = 0.0
    for s in range(0, n, chunk):
        p = pos[s:s + chunk]
        d2 = ((p[:, None, :] - pos[None, :, :]) ** 2).sum(-1) + eps * eps
        idx = np.arange(s, min(s + chunk, n))
        d2[np.arange(len(idx)), idx] = np.inf
        w -= 0.5 * float((m[idx, None] * m[None, :] / np.sqrt(d2)).sum())
    return w


def plummer(n: int, seed: int, a: float = 1.0, mass: float = 1.0, rmax: float = 10.0,
            virial: float | None = 0.5, type_fraction_dm: float = 0.0) -> Particles:
    """Plummer sphere of n equal masses (G = 1), truncated at rmax * a."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = np.empty(n)
    filled = 0
    while filled < n:
        u = rng.uniform(0.0, 1.0, n - filled)
        u = u[u > 0.0]
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= rmax * a]
        r[filled:filled + rr.size] = rr
        filled += rr.size
    ux, uy, uz = _isotropic(rng, n)
    x, y, z = r * ux, r * uy, r * uz
    q = np.empty(n)
    filled = 0
    while filled < n:
        k = 2 * (n - filled) + 16
        qq = rng.uniform(0.0, 1.0, k)
        yy = rng.uniform(0.0, 0.1, k)
        ok = qq[yy < qq * qq * (1.0 - qq * qq) ** 3.5][: n - filled]
        q[filled:filled + ok.size] = ok
        filled += ok.size
    vesc = math.sqrt(2.0 * mass / a) * (1.0 + (r / a) ** 2) ** -0.25
    v = q * vesc
    wx, wy, wz = _isotropic(rng, n)
    vx, vy, vz = v * wx, v * wy, v * wz
    m = np.full(n, mass / n)
    for arr in (x, y, z, vx, vy, vz):
        arr -= arr.mean()
    if virial is not None and n > 1:
        t = 0.5 * float((m * (vx * vx + vy * vy + vz * vz)).sum())
        if n <= 65536:
            w = potential_energy(x, y, z, m)
        else:  # untruncated Plummer: W = -(3 pi / 32) G M^2 / a
            w = -(3.0 * math.pi / 32.0) * mass * mass / a
        if t > 0:
            f = math.sqrt(virial * abs(w) / t)
            vx, vy, vz = vx * f, vy * f, vz * f
    types = np.zeros(n, dtype=np.int32)
    if type_fraction_dm > 0:
        types[rng.uniform(0.0, 1.0, n) < type_fraction_dm] = 1
    return Particles(x, y, z, vx, vy, vz, m, types)


def uniform_sphere(n: int, seed: int, radius: float = 1.0, mass: float = 1.0) -> Particles:
    """Cold uniform ball: n equal masses, zero velocities."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = radius * rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0)
    ux, uy, uz = _isotropic(rng, n)
    zero = np.zeros(n)
    return Particles(r * ux, r * uy, r * uz, zero.copy(), zero.copy(), zero.copy(),
                     np.full(n, mass / n), np.zeros(n, dtype=np.int32))


def two_cluster(n_each: int, seed: int, offset: float = 5.0, vbulk=(0.3, 0.1, 0.0),
                mass_each: float = 0.5) -> Particles:
    """Two Plummer spheres at (+-offset, 0, 0); the +offset cluster moves with
    (-vx, +vy, vz) and the -offset one with (+vx, -vy, vz)."""
    a = plummer(n_each, seed, mass=mass_each, virial=None)
    b = plummer(n_each, seed + 1, mass=mass_each, virial=None)
    a.x += offset
    a.vx -= vbulk[0]; a.vy += vbulk[1]; a.vz += vbulk[2]
    b.x -= offset
    b.vx += vbulk[0]; b.vy -= vbulk[1]; b.vz += vbulk[2]
    return Particles.concat([a, b])


def ensemble_sizes(count: int, lo: int = 1024, hi: int = 65536) -> list[int]:
    """N_i = round(exp(U[ln lo, ln hi])) drawn from seed i."""
    out = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(i))
        out.append(int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))))
    return out


def write_ic(path: str | os.PathLike, p: Particles) -> None:
    """Write the 8-column text format at 17 significant digits (exact)."""
    cols = np.stack([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m], axis=1)
    with open(path, "w") as f:
        for row, t in zip(cols, p.type):
            f.write(" ".join(repr(float(v)) for v in row) + f" {int(t)}\n")


def read_ic(path: str | os.PathLike) -> Particles:
    """Read the 8-column text format with the loader's stop-at-first-failure
    rule: tokens are consumed eight at a time and reading ends at the first
    group that does not parse (7 reals then an integer)."""
    with open(path) as f:
        toks = f.read().split()
    cols, types = [], []
    for k in range(0, len(toks) - 7, 8):
        try:
            vals = [float(t) for t in toks[k:k + 7]]
            t = int(toks[k + 7])
        except ValueError:
            break
        cols.append(vals)
        types.append(t)
    if not cols:
        return Particles.from_columns(np.zeros((0, 7)), np.zeros(0, dtype=np.int32))
    return Particles.from_columns(np.array(cols), np.array(types, dtype=np.int32))
