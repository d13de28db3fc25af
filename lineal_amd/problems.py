"""Synthetic inputs of the benchmark and test workloads (seeded, reproducible).

* lognormal(dims, sigma, seed): per-cell coefficients c = exp(sigma * z) with
  z ~ N(0, 1) from numpy's PCG64 generator seeded with `seed`.
* diagonal_lines(dims): material ids (y + z) % 4, the reference test's pattern
  (test/aux/materials.hpp:217-233).
* spheres(dims, seed, ...): two-phase u8 material ids: random spheres of
  radius U[4, 12] (centres uniform) are painted solid (id 1) until the solid
  fraction reaches 35%; the rest is pore (id 0).
* uniform(n, seed): vector entries U[-1, 1].
"""

from __future__ import annotations

import numpy as np


def lognormal(dims, sigma=1.0, seed=1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    z = rng.standard_normal(n)
    z *= sigma
    np.exp(z, out=z)
    return z


def diagonal_lines(dims) -> np.ndarray:
    nx, ny, nz = dims
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[None, None, :]
    m = np.broadcast_to((y + z) % 4, (nx, ny, nz))
    return np.ascontiguousarray(m, dtype=np.uint8).reshape(-1)


def diagonal_lines_eff_diff(dims, mat_coeffs) -> float:
    """Mean coefficient over a cross-section (test/aux/materials.hpp:236-249)."""
    _, ny, nz = dims
    s = 0.0
    for y in range(ny):
        for z in range(nz):
            s += mat_coeffs[(y + z) % 4]
    return s / (ny * nz)


def spheres(dims, seed=3, solid_fraction=0.35, rmin=4.0, rmax=12.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(d) for d in dims)
    m = np.zeros((nx, ny, nz), np.uint8)
    target = solid_fraction * m.size
    solid = 0
    while solid < target:
        for _ in range(64):
            r = rng.uniform(rmin, rmax)
            c = rng.uniform(0, 1, 3) * np.array([nx, ny, nz])
            lo = np.maximum(np.floor(c - r).astype(int), 0)
            hi = np.minimum(np.ceil(c + r).astype(int) + 1, [nx, ny, nz])
            xs = np.arange(lo[0], hi[0])[:, None, None] + 0.5
            ys = np.arange(lo[1], hi[1])[None, :, None] + 0.5
            zs = np.arange(lo[2], hi[2])[None, None, :] + 0.5
            inside = (xs - c[0]) ** 2 + (ys - c[1]) ** 2 + (zs - c[2]) ** 2 <= r * r
            m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= inside.astype(np.uint8)
        solid = int(m.sum(dtype=np.int64))
    return m.reshape(-1)


def uniform(n, seed=2, dtype=np.float64) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, int(n)).astype(dtype)
