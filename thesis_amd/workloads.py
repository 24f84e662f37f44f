"""Synthetic workload definitions shared by the bench, the tests and the tuner.

All randomness comes from the counter-based generator the library uses for
boundary data, U(seed, idx) = (splitmix64(seed ^ idx*0x9E3779B97F4A7C15) >> 40) * 2^-24,
so every consumer derives the same workloads from a seed.
"""
from __future__ import annotations

import math

_M64 = (1 << 64) - 1


def splitmix64(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def uniform(seed: int, idx: int) -> float:
    """U(seed, idx) in [0, 1), 24 random bits."""
    return (splitmix64(seed ^ ((idx * 0x9E3779B97F4A7C15) & _M64)) >> 40) / 16777216.0


def batched_shapes(n: int = 512, seed: int = 0x5EED0004, lo: int = 16, hi: int = 128):
    """Uneven 3D problems: each edge round(lo * (hi/lo)^U) drawn independently
    (log-uniform on [lo, hi]); radius 1 or 2 with probability 1/2 each.
    Returns a list of (nx, ny, nz, radius)."""
    ratio = hi / lo
    out = []
    for i in range(n):
        e = [int(round(lo * ratio ** uniform(seed, 4 * i + a))) for a in range(3)]
        r = 1 if uniform(seed, 4 * i + 3) < 0.5 else 2
        out.append((e[0], e[1], e[2], r))
    return out


def batched_cells(shapes) -> int:
    return sum(nx * ny * nz for nx, ny, nz, _ in shapes)


def batched_interior_cells(shapes) -> int:
    return sum((nx - 2 * r) * (ny - 2 * r) * (nz - 2 * r) for nx, ny, nz, r in shapes)


def fourier_faces(n: int, seed: int, modes: int = 8, kmax: int = 4):
    """Dirichlet data for the 6 faces of an n^3 Laplace problem: each face is
    sum_{m<modes} a_m sin(pi kx s) sin(pi ky t) with integer kx, ky in
    [1, kmax] and a_m ~ N(0, 1) (Box-Muller on two U draws), s, t in [0, 1]
    along the face.  The sine basis makes neighbouring faces agree (at 0) on
    shared edges.  Returns a dict face -> list of (a, kx, ky)."""
    faces = {}
    idx = 0
    for f in ("x0", "x1", "y0", "y1", "z0", "z1"):
        terms = []
        for _ in range(modes):
            u1 = max(uniform(seed, idx), 1e-12)
            u2 = uniform(seed, idx + 1)
            a = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
            kx = 1 + int(uniform(seed, idx + 2) * kmax)
            ky = 1 + int(uniform(seed, idx + 3) * kmax)
            idx += 4
            terms.append((a, kx, ky))
        faces[f] = terms
    return faces
