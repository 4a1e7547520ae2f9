"""Seeded synthetic particle beds in the reference's particle-file format.

The reference reads its initial DEM state from a text file (DEM.cpp:74-127):
first line the element count, then per element 20 numbers
``index size radius x y z vx vy vz wx wy wz q0 q1 q2 q3 dq0 dq1 dq2 dq3``.
This module generates non-overlapping sphere packings by random sequential
addition (RSA) from a fixed seed and writes/reads that format.

PRNG: SplitMix64 seeds a xoshiro256** generator (Blackman & Vigna, 2018);
uniform doubles take the top 53 bits.  Everything is deterministic in the seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = (1 << 64) - 1


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256ss:
    """xoshiro256** seeded through SplitMix64."""

    def __init__(self, seed: int):
        s = seed & _M64
        st = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            st.append(z ^ (z >> 31))
        self.s = st

    def next_u64(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def normal(self) -> float:
        # Box-Muller on two uniforms
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass
class Bed:
    """A particle bed: centres (P,3) and radii (P,) in lattice units."""

    x: np.ndarray
    r: np.ndarray

    @property
    def count(self) -> int:
        return int(self.r.shape[0])

    def solid_volume(self) -> float:
        return float(np.sum(4.0 / 3.0 * math.pi * self.r ** 3))


def rsa_bed(seed: int, box_lo, box_hi, radius_fn, count: int, periodic=(False, False, False),
            period=(0.0, 0.0, 0.0), gap: float = 0.0, max_attempts: int = 2_000_000) -> Bed:
    """Random sequential addition of ``count`` spheres.

    ``radius_fn(rng)`` draws a radius.  Centres are uniform in
    ``[box_lo + r, box_hi - r]`` along non-periodic axes and in
    ``[box_lo, box_hi)`` along periodic ones, where overlap is tested against
    the nearest periodic image.  ``gap`` is an extra clearance between spheres.
    Stops early (returning fewer spheres) when ``max_attempts`` is exhausted.
    """
    rng = Xoshiro256ss(seed)
    lo = np.asarray(box_lo, dtype=np.float64)
    hi = np.asarray(box_hi, dtype=np.float64)
    per = np.asarray(period, dtype=np.float64)
    pbc = np.asarray(periodic, dtype=bool)
    radii = [radius_fn(rng) for _ in range(count)]
    rmax = max(radii) if radii else 1.0
    cell = 2.0 * rmax + gap
    ncell = np.maximum(1, np.floor((hi - lo) / cell).astype(int))
    cellw = (hi - lo) / ncell
    grid: dict = {}
    xs: list = []
    rs: list = []
    attempts = 0
    for r in radii:
        placed = False
        while attempts < max_attempts:
            attempts += 1
            c = np.empty(3)
            for k in range(3):
                if pbc[k]:
                    c[k] = rng.uniform(lo[k], hi[k])
                else:
                    c[k] = rng.uniform(lo[k] + r, hi[k] - r)
            ic = np.floor((c - lo) / cellw).astype(int)
            ic = np.clip(ic, 0, ncell - 1)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        key = [ic[0] + dx, ic[1] + dy, ic[2] + dz]
                        skip = False
                        for k in range(3):
                            if pbc[k]:
                                key[k] %= ncell[k]
                            elif key[k] < 0 or key[k] >= ncell[k]:
                                skip = True
                        if skip:
                            continue
                        for j in grid.get(tuple(key), ()):
                            d = c - xs[j]
                            for k in range(3):
                                if pbc[k]:
                                    d[k] -= per[k] * round(d[k] / per[k])
                            if float(d @ d) < (r + rs[j] + gap) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault(tuple(ic), []).append(len(xs))
                xs.append(c)
                rs.append(r)
                placed = True
                break
        if not placed:
            break
    return Bed(np.array(xs, dtype=np.float64).reshape(-1, 3), np.array(rs, dtype=np.float64))


def write_particle_file(path: str, bed: Bed, velocities: np.ndarray | None = None, sizes=None, omegas=None,
                        quats=None) -> None:
    """Write ``bed`` in the reference particle-file format (DEM.cpp:78-127).

    Optional per element: ``sizes`` (spheres per element, 1-4, laid out by the
    prototypes of DEM.cpp:404-433), angular velocities ``omegas`` and
    orientation quaternions ``quats`` (q0, q1, q2, q3); defaults 1, zero and
    the identity.
    """
    with open(path, "w") as fh:
        fh.write(f"{bed.count}\n")
        for i in range(bed.count):
            x, y, z = bed.x[i]
            vx = vy = vz = 0.0
            if velocities is not None:
                vx, vy, vz = velocities[i]
            size = 1 if sizes is None else int(sizes[i])
            w = (0.0, 0.0, 0.0) if omegas is None else tuple(float(c) for c in omegas[i])
            q = (1.0, 0.0, 0.0, 0.0) if quats is None else tuple(float(c) for c in quats[i])
            vals = [float(bed.r[i]), float(x), float(y), float(z), float(vx), float(vy), float(vz), *w, *q]
            fh.write(f"{i} {size} " + " ".join(repr(v) for v in vals) + " 0 0 0 0\n")


def read_particle_file(path: str):
    """Read the reference particle-file format; returns a dict of arrays.

    Fields follow DEM.cpp:85-120: index, size, radius, x(3), v(3), w(3), q(4), dq(4).
    """
    with open(path) as fh:
        toks = fh.read().split()
    n = int(toks[0])
    vals = np.array([float(t) for t in toks[1:1 + 20 * n]], dtype=np.float64).reshape(n, 20)
    return {
        "index": vals[:, 0].astype(np.int64),
        "size": vals[:, 1].astype(np.int64),
        "radius": vals[:, 2],
        "x": vals[:, 3:6],
        "v": vals[:, 6:9],
        "w": vals[:, 9:12],
        "q": vals[:, 12:16],
        "dq": vals[:, 16:20],
    }


def object_records(objects) -> np.ndarray:
    """Objects as (n, 9) records: index, size, r, x, y, z, vx, vy, vz (HYB_OBJECT_RECORD)."""
    obj = np.asarray(objects, dtype=np.float64).reshape(-1, 7)  # r, x, y, z, vx, vy, vz
    rec = np.zeros((obj.shape[0], 9))
    rec[:, 0] = np.arange(obj.shape[0])
    rec[:, 1] = 1.0
    rec[:, 2:] = obj
    return rec


def write_object_file(path: str, objects) -> None:
    """Object file (DEM.cpp:136-173): n, then per object index, size, r, x, y, z,
    vx, vy, vz and eleven rotation values the reference skips. ``objects`` rows
    are r, x, y, z, vx, vy, vz."""
    rec = object_records(objects)
    with open(path, "w") as fh:
        fh.write(f"{rec.shape[0]}\n")
        for r in rec:
            vals = " ".join(repr(float(v)) for v in r[2:])
            fh.write(f"{int(r[0])} 1 {vals} 0 0 0 1 0 0 0 0 0 0 0\n")


def write_empty_object_file(path: str) -> None:
    """Object file with zero objects (DEM.cpp:131-137 requires the file to exist)."""
    with open(path, "w") as fh:
        fh.write("0\n")
