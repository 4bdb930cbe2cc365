"""The surfacer's marching cubes tables (mpm_surface_tables, built on the host, no device call).

The reference meshes with Warp's MarchingCubes (src/mpm_surface/field.py:130-138), which is absent here;
its tests ask for closed, outward-wound meshes of the right volume and component count
(tests/test_surface_field.py). The tables are pinned here by those properties, checked on meshes that a
plain numpy assembly of the tables makes from fields whose level sets are known: every one of the 256
corner configurations alone, random fields, and a sphere (Euler characteristic 2, volume, centre).
"""

import numpy as np
import pytest

from mpm_plugin_amd import _native as N

# edge e: lower corner and axis (the layout mpm_hip.h documents for mpm_surface_tables)
EDGE_LOWER = [0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3]
EDGE_AXIS = [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2]


@pytest.fixture(scope="module")
def tables():
    count = np.zeros(256, np.int32)
    edges = np.zeros(256 * 36, np.int8)
    N.check(N.lib().mpm_surface_tables(N.iptr(count), edges.ctypes.data_as(N.C.POINTER(N.C.c_int8))))
    return count, edges.reshape(256, 36).astype(np.int64)


def assemble(f, iso, tables):
    """Vertices (grid units) and triangles of the iso surface of f from the tables."""
    count, edges = tables
    inside = f > iso
    D = f.shape
    vid = -np.ones((3,) + D, np.int64)
    verts = []
    n = 0
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, -1), slice(1, None)
        cross = inside[tuple(lo)] != inside[tuple(hi)]
        idx = np.argwhere(cross)
        v0, v1 = f[tuple(lo)][cross], f[tuple(hi)][cross]
        pos = idx.astype(np.float64)
        pos[:, a] += (iso - v0) / (v1 - v0)
        view = vid[a][tuple(lo)]
        view[cross] = np.arange(n, n + len(idx))
        n += len(idx)
        verts.append(pos)
    verts = np.concatenate(verts) if verts else np.zeros((0, 3))
    cells = np.argwhere(np.ones([d - 1 for d in D], bool))
    cfg = np.zeros(len(cells), np.int64)
    for c in range(8):
        o = cells + [(c & 1), (c >> 1) & 1, (c >> 2) & 1]
        cfg |= inside[o[:, 0], o[:, 1], o[:, 2]].astype(np.int64) << c
    tris = []
    for m in range(12):
        sel = count[cfg] > m
        if not sel.any():
            break
        tri = np.zeros((int(sel.sum()), 3), np.int64)
        for s in range(3):
            e = edges[cfg[sel], 3 * m + s]
            lower = np.array(EDGE_LOWER)[e]
            o = cells[sel] + np.stack([lower & 1, (lower >> 1) & 1, (lower >> 2) & 1], 1)
            tri[:, s] = vid[np.array(EDGE_AXIS)[e], o[:, 0], o[:, 1], o[:, 2]]
        tris.append(tri)
    tris = np.concatenate(tris) if tris else np.zeros((0, 3), np.int64)
    assert (tris >= 0).all(), "a triangle uses an edge that does not cross iso"
    return verts, tris


def directed_edges_pair_up(tris):
    """Every directed edge appears once and its reverse once: closed and consistently wound."""
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    m = int(tris.max()) + 1
    fwd = e[:, 0] * m + e[:, 1]
    rev = e[:, 1] * m + e[:, 0]
    return len(np.unique(fwd)) == len(fwd) and np.array_equal(np.sort(fwd), np.sort(rev))


def signed_volume(verts, tris):
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6)


def test_counts(tables):
    count, _ = tables
    assert count[0] == 0 and count[255] == 0
    assert count.max() <= 12
    for c in range(8):  # one inside corner: one triangle; one outside corner: one triangle
        assert count[1 << c] == 1 and count[255 ^ (1 << c)] == 1


@pytest.mark.parametrize("cfg", range(1, 255))
def test_every_configuration_closes(tables, cfg):
    """The configuration's corners in an otherwise empty field: one closed, outward-wound surface."""
    f = np.zeros((4, 4, 4))
    for c in range(8):
        if (cfg >> c) & 1:
            f[1 + (c & 1), 1 + ((c >> 1) & 1), 1 + ((c >> 2) & 1)] = 1.0
    verts, tris = assemble(f, 0.5, tables)
    assert len(tris) > 0 and directed_edges_pair_up(tris)
    assert signed_volume(verts, tris) > 0


@pytest.mark.parametrize("seed", range(12))
def test_random_fields_close(tables, seed):
    rng = np.random.default_rng(seed)
    f = rng.random((10, 9, 11))
    f[[0, -1], :, :] = 0
    f[:, [0, -1], :] = 0
    f[:, :, [0, -1]] = 0
    verts, tris = assemble(f, 0.5, tables)
    assert directed_edges_pair_up(tris)
    # the enclosed volume is positive, and bounded by the inside nodes' dual cells grown by half a cell
    vol = signed_volume(verts, tris)
    assert 0 < vol < 8 * int((f > 0.5).sum())


def test_sphere(tables):
    n = 40
    g = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    c, r = np.array([19.3, 20.1, 18.7]), 12.5
    f = 1.0 - np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / r  # > 0 inside the sphere
    verts, tris = assemble(f, 0.0, tables)
    assert directed_edges_pair_up(tris)
    edges = np.unique(np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), 1), axis=0)
    assert len(verts) - len(edges) + len(tris) == 2  # a sphere's Euler characteristic
    assert signed_volume(verts, tris) == pytest.approx(4 / 3 * np.pi * r**3, rel=0.01)
    assert np.allclose(verts.mean(0), c, atol=0.05)
    # every vertex lies on the sphere to interpolation accuracy
    assert np.abs(np.linalg.norm(verts - c, axis=1) - r).max() < 0.05
