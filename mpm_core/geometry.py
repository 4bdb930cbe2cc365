"""Mesh helpers: closedness, transforms and their interpolation, volume sampling.

Behaviour follows the reference's ``mpm_core.geometry`` (src/mpm_core/geometry.py:1-155):

* ``decompose_transform`` splits a world matrix into translation, rotation quaternion (x, y, z, w)
  and symmetric stretch with M = [R S | t] (geometry.py:30-39); collider poses are built from it.
* ``interpolate_transform`` slerps the rotation and lerps translation and stretch (geometry.py:71-86).
* ``sample_mesh_volume`` puts particles on the cell-centred lattice ``(i + 0.5) * spacing`` inside a
  closed mesh, decided by z-ray crossing parity per lattice column, then jitters them by
  ``U(-0.25, 0.25) * spacing`` from ``numpy.random.default_rng(seed)`` (geometry.py:89-155). Emission
  counts and positions are part of a bake's result, so the lattice, the crossing test, the output order
  (column-major in x, then y, then z ascending) and the single jitter draw are kept as the reference
  defines them; the crossing search itself is vectorised over all triangles at once here.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

# Columns are shifted off the lattice by this fraction of the spacing so that rays avoid mesh edges
# and vertices (geometry.py:11); the same offsets keep emission counts identical to the reference's.
_RAY_NUDGE = np.array([1.234567e-4, 2.345678e-4])


def mesh_is_closed(tris: np.ndarray) -> bool:
    """True when every undirected edge belongs to exactly two triangles."""
    t = np.asarray(tris).reshape(-1, 3)
    if t.shape[0] == 0:
        return False
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1).astype(np.int64)
    key = e[:, 0] * (int(e.max()) + 1) + e[:, 1]
    _, counts = np.unique(key, return_counts=True)
    return bool(np.all(counts == 2))


def transform_points(matrix: np.ndarray, points: np.ndarray) -> np.ndarray:
    """Apply a 4x4 affine to (N,3) points in double precision, return float32."""
    m = np.asarray(matrix, np.float64)
    p = np.asarray(points, np.float64)
    return (p @ m[:3, :3].T + m[:3, 3]).astype(np.float32)


def _quat_from_matrix(R: np.ndarray) -> np.ndarray:
    """Rotation matrix -> unit quaternion (x, y, z, w).

    Each component's magnitude comes from the diagonal, its sign from the antisymmetric part
    (the branch-free form of the standard matrix-to-quaternion conversion).
    """
    d0, d1, d2 = R[0, 0], R[1, 1], R[2, 2]
    w = 0.5 * np.sqrt(max(0.0, 1.0 + d0 + d1 + d2))
    x = np.copysign(0.5 * np.sqrt(max(0.0, 1.0 + d0 - d1 - d2)), R[2, 1] - R[1, 2])
    y = np.copysign(0.5 * np.sqrt(max(0.0, 1.0 - d0 + d1 - d2)), R[0, 2] - R[2, 0])
    z = np.copysign(0.5 * np.sqrt(max(0.0, 1.0 - d0 - d1 + d2)), R[1, 0] - R[0, 1])
    q = np.array([x, y, z, w])
    return q / np.linalg.norm(q)


def _matrix_from_quat(q: np.ndarray) -> np.ndarray:
    x, y, z, w = q
    xx, yy, zz = x * x, y * y, z * z
    return np.array([[1 - 2 * (yy + zz), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (xx + zz), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (xx + yy)]])


def decompose_transform(m: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """World matrix -> (translation, rotation quaternion xyzw, symmetric stretch), M = [R S | t].

    A mirrored matrix keeps R a proper rotation; the reflection goes into S through the sign of the
    smallest singular value.
    """
    return decompose_transform_sv(m)[:3]


def decompose_transform_sv(m: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray, float]:
    """decompose_transform(m) and the smallest singular value of the linear part, from the same SVD."""
    A = np.asarray(m, np.float64)[:3, :3]
    U, sig, Vt = np.linalg.svd(A)
    smallest = float(sig[2])  # descending and non-negative before the reflection
    if np.linalg.det(U @ Vt) < 0:
        U[:, 2] *= -1.0
        sig[2] *= -1.0
    R = U @ Vt
    S = Vt.T @ (sig[:, None] * Vt)
    return np.asarray(m, np.float64)[:3, 3].copy(), _quat_from_matrix(R), S, smallest


def _slerp(q0: np.ndarray, q1: np.ndarray, s: float) -> np.ndarray:
    if np.dot(q0, q1) < 0:
        q1 = -q1  # the short way round
    angle = np.arccos(min(float(np.dot(q0, q1)), 1.0))
    if angle < 1e-6:
        q = q0 + s * (q1 - q0)
    else:
        q = (np.sin((1.0 - s) * angle) * q0 + np.sin(s * angle) * q1) / np.sin(angle)
    return q / np.linalg.norm(q)


def interpolate_transform(a: np.ndarray, b: np.ndarray, s: float) -> np.ndarray:
    """Pose a fraction ``s`` of the way from world matrix ``a`` to ``b`` (rotation slerped, rest lerped)."""
    if s == 0:
        return np.asarray(a, np.float32)
    if s == 1:
        return np.asarray(b, np.float32)
    t0, q0, S0 = decompose_transform(a)
    t1, q1, S1 = decompose_transform(b)
    out = np.eye(4)
    out[:3, :3] = _matrix_from_quat(_slerp(q0, q1, s)) @ (S0 + s * (S1 - S0))
    out[:3, 3] = t0 + s * (t1 - t0)
    return out.astype(np.float32)


def _column_crossings(tri: np.ndarray, xs: np.ndarray, ys: np.ndarray, first: np.ndarray, spacing: float,
                      nudge: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Every (column id, crossing z) of the nudged vertical rays through the triangles.

    Column id is ``i * len(ys) + j``. A ray crosses a triangle when all three signed edge functions
    (barycentric weights) of its xy point are >= 0; triangles with zero xy area are never crossed.
    """
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    lo = np.minimum(np.minimum(a, b), c)[:, :2]
    hi = np.maximum(np.maximum(a, b), c)[:, :2]
    area = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    i0 = np.maximum(np.ceil((lo[:, 0] - nudge[0]) / spacing - 0.5).astype(np.int64) - first[0], 0)
    i1 = np.minimum(np.floor((hi[:, 0] - nudge[0]) / spacing - 0.5).astype(np.int64) - first[0], len(xs) - 1)
    j0 = np.maximum(np.ceil((lo[:, 1] - nudge[1]) / spacing - 0.5).astype(np.int64) - first[1], 0)
    j1 = np.minimum(np.floor((hi[:, 1] - nudge[1]) / spacing - 0.5).astype(np.int64) - first[1], len(ys) - 1)
    ni = np.where((i1 >= i0) & (j1 >= j0) & (area != 0.0), i1 - i0 + 1, 0)
    nj = np.where(ni > 0, j1 - j0 + 1, 0)
    count = ni * nj
    total = int(count.sum())
    if total == 0:
        return np.zeros(0, np.int64), np.zeros(0)
    t = np.repeat(np.arange(len(tri)), count)  # triangle of each (triangle, column) pair
    local = np.arange(total) - np.repeat(np.cumsum(count) - count, count)
    ii = i0[t] + local // nj[t]
    jj = j0[t] + local % nj[t]
    px = xs[ii] + nudge[0]
    py = ys[jj] + nudge[1]
    A, B, Cc, ar = a[t], b[t], c[t], area[t]
    w0 = ((B[:, 0] - px) * (Cc[:, 1] - py) - (B[:, 1] - py) * (Cc[:, 0] - px)) / ar
    w1 = ((Cc[:, 0] - px) * (A[:, 1] - py) - (Cc[:, 1] - py) * (A[:, 0] - px)) / ar
    w2 = 1.0 - w0 - w1
    hit = (w0 >= 0) & (w1 >= 0) & (w2 >= 0)
    z = w0 * A[:, 2] + w1 * B[:, 2] + w2 * Cc[:, 2]
    return (ii * len(ys) + jj)[hit], z[hit]


def sample_mesh_volume(verts: np.ndarray, tris: np.ndarray, spacing: float,
                       seed: int | Sequence[int]) -> np.ndarray:
    """Jittered lattice points inside a closed mesh (see the module docstring)."""
    if not mesh_is_closed(tris):
        raise ValueError("emitter mesh is not closed (every edge must be shared by exactly two triangles)")
    v = np.asarray(verts, np.float64)
    tri = v[np.asarray(tris)]
    lo, hi = v.min(0), v.max(0)
    first = np.ceil(lo / spacing - 0.5).astype(np.int64)
    last = np.floor(hi / spacing - 0.5).astype(np.int64)
    if np.any(last < first):
        return np.zeros((0, 3), np.float32)
    xs, ys, zs = [(np.arange(first[a], last[a] + 1) + 0.5) * spacing for a in range(3)]
    cols, z = _column_crossings(tri, xs, ys, first, spacing, _RAY_NUDGE * spacing)
    if len(cols) == 0:
        return np.zeros((0, 3), np.float32)
    order = np.lexsort((z, cols))
    cols, z = cols[order], z[order]
    ucols, start = np.unique(cols, return_index=True)
    stop = np.r_[start[1:], len(cols)]
    chunks = []
    for col, s0, s1 in zip(ucols, start, stop):
        below = np.searchsorted(z[s0:s1], zs)  # crossings strictly below each lattice z
        zin = zs[(below & 1) == 1]
        if len(zin):
            chunks.append(np.column_stack([np.full(len(zin), xs[col // len(ys)]),
                                           np.full(len(zin), ys[col % len(ys)]), zin]))
    if not chunks:
        return np.zeros((0, 3), np.float32)
    pts = np.concatenate(chunks)
    pts += np.random.default_rng(seed).uniform(-0.25, 0.25, pts.shape) * spacing
    return pts.astype(np.float32)
