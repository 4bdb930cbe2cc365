"""CPU oracle of the surfacer's field (test infrastructure only: tests import it as the checker).

A numpy restatement of the reference's density splat and normalisation (src/mpm_surface/field.py:35-63)
and of its trilinear velocity sample (field.py:66-85), in float32 as the reference's kernels compute.
Accumulation runs particle by particle in float64 and is rounded once, so it differs from any GPU atomic
order by float32 rounding only.

Pinned by tests/test_surface_oracle.py against what the mathematics fixes: the kernel's values, continuity
and integral (pi h^3 / 8 over R^3), a lone particle on a node, the total splatted weight of a particle far
from the grid's edge (a Riemann sum of that integral), the rest density of a lattice against the same
integral, velocity = the particles' common velocity, and exact trilinear sampling of a linear field.
"""

from __future__ import annotations

import numpy as np


def cubic(q: np.ndarray) -> np.ndarray:
    """field.py:24-32."""
    q = np.asarray(q, np.float32)
    t = np.float32(1) - q
    return np.where(q < 0.5, np.float32(1) - np.float32(6) * q * q + np.float32(6) * q * q * q,
                    np.where(q < 1, np.float32(2) * t * t * t, np.float32(0))).astype(np.float32)


def splat(x: np.ndarray, v: np.ndarray, origin, cell: float, h: float, dims) -> tuple[np.ndarray, np.ndarray]:
    """Density (D0,D1,D2) and momentum (D0,D1,D2,3) before normalisation (field.py:35-55)."""
    x = np.asarray(x, np.float32).reshape(-1, 3)
    v = np.asarray(v, np.float32).reshape(-1, 3)
    origin = np.asarray(origin, np.float32)
    inv_cell, cell, h = np.float32(1.0 / cell), np.float32(cell), np.float32(h)
    density = np.zeros(tuple(dims), np.float64)
    momentum = np.zeros(tuple(dims) + (3,), np.float64)
    r = h * inv_cell
    for p in range(len(x)):
        g = (x[p] - origin) * inv_cell
        lo = np.maximum(np.floor(g - r).astype(np.int64), 0)
        hi = np.minimum(np.ceil(g + r).astype(np.int64), np.asarray(dims) - 1)
        if np.any(hi < lo):
            continue
        ii, jj, kk = np.meshgrid(*[np.arange(lo[a], hi[a] + 1) for a in range(3)], indexing="ij")
        d = np.sqrt(((np.stack([ii, jj, kk], -1).astype(np.float32) - g) ** 2).sum(-1, dtype=np.float32)) * cell
        w = cubic(d / h)
        sel = w > 0
        np.add.at(density, (ii[sel], jj[sel], kk[sel]), w[sel])
        np.add.at(momentum, (ii[sel], jj[sel], kk[sel]), w[sel][:, None].astype(np.float64) * v[p])
    return density.astype(np.float32), momentum.astype(np.float32)


def normalize(density: np.ndarray, momentum: np.ndarray, inv_rest: float) -> tuple[np.ndarray, np.ndarray]:
    """Velocity = momentum / density where density > 0; density scaled by 1 / rest density
    (field.py:58-63)."""
    d = density.astype(np.float32)
    vel = np.where(d[..., None] > 0, momentum / np.where(d > 0, d, 1)[..., None], momentum).astype(np.float32)
    return (d * np.float32(inv_rest)).astype(np.float32), vel


def sample_velocity(g: np.ndarray, velocity: np.ndarray) -> np.ndarray:
    """Trilinear sample at points in grid units, the base cell clamped into the grid (field.py:66-85)."""
    g = np.asarray(g, np.float64).reshape(-1, 3)
    dims = np.asarray(velocity.shape[:3])
    base = np.clip(np.floor(g).astype(np.int64), 0, dims - 2)
    f = g - base
    out = np.zeros((len(g), 3))
    for a in range(2):
        for b in range(2):
            for c in range(2):
                w = (f[:, 0] if a else 1 - f[:, 0]) * (f[:, 1] if b else 1 - f[:, 1]) * (f[:, 2] if c else 1 - f[:, 2])
                out += w[:, None] * velocity[base[:, 0] + a, base[:, 1] + b, base[:, 2] + c]
    return out
