"""Seeded synthetic periodic microstructures (phase ids, x-major [x][y][z]).

Spheres are placed sequentially from numpy's PCG64 stream (default_rng(seed)),
rejecting any candidate that overlaps an accepted one under the periodic
minimum-image distance; a voxel belongs to a sphere when its centre lies inside
it.  The achieved volume fraction is whatever the count and radius law give.
"""
from __future__ import annotations

import numpy as np


def random_spheres(n, count, rmin, rmax, seed=0, n_inclusion_phases=1, max_tries=200000, dtype=np.uint16):
    """Phase 0 matrix; inclusions get ids 1..n_inclusion_phases (cycled)."""
    n = np.array([n, n, n] if np.isscalar(n) else n, dtype=np.int64)
    rng = np.random.default_rng(seed)
    C = np.zeros((max(count, 1), 3))
    R = np.zeros(max(count, 1))
    m = 0
    tries = 0
    while m < count and tries < max_tries:
        tries += 1
        c = rng.uniform(0.0, 1.0, size=3) * n
        r = rng.uniform(rmin, rmax)
        if m:
            d = np.abs(C[:m] - c)
            d = np.minimum(d, n - d)
            if np.any(np.sqrt((d * d).sum(axis=1)) < r + R[:m]):
                continue
        C[m], R[m] = c, r
        m += 1
    centers, radii = list(C[:m]), list(R[:m])
    ms = np.zeros(tuple(n), dtype=dtype)
    for k, (c, r) in enumerate(zip(centers, radii)):
        phase = 1 + (k % n_inclusion_phases)
        lo = np.floor(c - r - 1).astype(np.int64)
        hi = np.ceil(c + r + 1).astype(np.int64)
        ax = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
        X, Y, Z = np.meshgrid(*ax, indexing="ij")
        inside = (X + 0.5 - c[0]) ** 2 + (Y + 0.5 - c[1]) ** 2 + (Z + 0.5 - c[2]) ** 2 <= r * r
        ms[X[inside] % n[0], Y[inside] % n[1], Z[inside] % n[2]] = phase
    return ms


def laminate(n, axis, widths, phases):
    """Layers normal to `axis`: widths[i] voxels of phase phases[i], repeated."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    ms = np.zeros(n, dtype=np.uint16)
    seq = np.concatenate([np.full(w, p, dtype=np.uint16) for w, p in zip(widths, phases)])
    reps = int(np.ceil(n[axis] / len(seq)))
    line = np.tile(seq, reps)[: n[axis]]
    shape = [1, 1, 1]
    shape[axis] = n[axis]
    ms[...] = line.reshape(shape)
    return ms


def volume_fractions(ms):
    vals, counts = np.unique(ms, return_counts=True)
    return {int(v): c / ms.size for v, c in zip(vals, counts)}
