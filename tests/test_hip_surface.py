"""The HIP surfacer (mpm_surface_build) against the CPU oracle and the reference's surfacing tests.

Field: the splatted, normalised density and velocity equal the oracle's (oracle/surface.py, a restatement
of src/mpm_surface/field.py:35-63) to float32 atomic-order rounding. Mesh: Warp's marching cubes is not
available, so the mesh is checked by what the reference's tests ask of it (tests/test_surface_field.py:
closed, outward, volume within 10%, centred, components, floor contact, vertex velocity, detail, memory
budget) and by exact relations to the device's own field: one vertex per crossed grid edge, at the linear
interpolation of iso along it, with the trilinear velocity sample of field.py:66-85 (oracle).
"""

import numpy as np
import pytest
from helpers import box_particles, sphere_particles

from mpm_core.geometry import mesh_is_closed
from mpm_surface.field import field_grid, rest_density, surface_particles, surfacer
from mpm_surface.settings import SurfaceSettings
from oracle import surface as S

pytestmark = pytest.mark.gpu

SPACING = 0.005  # the helpers sample at dx / 2
DX = 0.01


def signed_volume(verts, tris):
    v = verts.astype(np.float64)
    a, b, c = v[tris[:, 0]], v[tris[:, 1]], v[tris[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6)


def components(tris):
    parent = list(range(int(tris.max()) + 1))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for a, b, c in tris:
        for u, v in ((a, b), (b, c)):
            parent[find(u)] = find(v)
    return len({find(i) for i in np.unique(tris)})


def consistently_wound(tris):
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]).astype(np.int64)
    m = int(tris.max()) + 1
    fwd, rev = e[:, 0] * m + e[:, 1], e[:, 1] * m + e[:, 0]
    return len(np.unique(fwd)) == len(fwd) and np.array_equal(np.sort(fwd), np.sort(rev))


def mesh(p, vel=None, settings=SurfaceSettings(), device="cuda:0"):
    return surface_particles(p, np.zeros_like(p) if vel is None else vel, SPACING, settings, device)


def device_field(p, v, settings):
    """Build through the surfacer directly: grid-unit vertices, triangles, velocities, and the field."""
    from mpm_plugin_amd import _native as N

    origin, dims, cell, h = field_grid(p, SPACING, settings)
    desc = N.SurfaceDesc()
    desc.dims[:] = dims
    desc.origin[:] = [float(o) for o in origin]
    desc.cell, desc.h, desc.inv_rest, desc.iso = cell, h, 1.0 / rest_density(SPACING, h), settings.iso
    s = surfacer("cuda:0")
    g, tris, vel = s.build(desc, p, v)
    density, velocity = s.field(dims)
    return dict(origin=origin, dims=dims, cell=cell, h=h, inv_rest=desc.inv_rest, g=g, tris=tris, vel=vel,
                density=density, velocity=velocity)


# a kernel reach of 3 cells (the gather's two waves per brick) and of 8 cells (one wave per brick)
BLOB_SETTINGS = {"reach3_two_waves": SurfaceSettings(smoothing=1.5, iso=0.4),
                 "reach8_one_wave": SurfaceSettings(detail=2.0, iso=0.4)}


@pytest.fixture(scope="module", params=sorted(BLOB_SETTINGS))
def blob(request):
    rng = np.random.default_rng(3)
    p = sphere_particles((0.31, 0.52, 0.47), 0.03, DX, seed=4)
    v = rng.normal(size=p.shape).astype(np.float32)
    settings = BLOB_SETTINGS[request.param]
    return p, v, device_field(p, v, settings) | {"iso": settings.iso}


def test_field_matches_oracle(blob):
    p, v, f = blob
    d, m = S.splat(p, v, np.asarray(f["origin"], np.float32), f["cell"], f["h"], f["dims"])
    dn, vel = S.normalize(d, m, f["inv_rest"])
    scale = float(dn.max())
    assert scale > 1.0
    np.testing.assert_allclose(f["density"], dn, rtol=0, atol=1e-5 * scale)
    dense = d > 1e-3 * float(d.max())
    np.testing.assert_allclose(f["velocity"][dense], vel[dense], rtol=1e-4, atol=1e-5 * float(np.abs(v).max()))
    assert (f["velocity"][f["density"] == 0] == 0).all()


def test_vertices_are_the_edge_crossings(blob):
    _, _, f = blob
    dens, iso, g = f["density"].astype(np.float64), f["iso"], f["g"].astype(np.float64)
    inside = f["density"] > np.float32(iso)
    crossed = sum(int((np.diff(inside.astype(np.int8), axis=a) != 0).sum()) for a in range(3))
    assert len(g) == crossed
    frac = g - np.floor(g)
    axis = np.argmax(frac, axis=1)
    assert ((frac > 0).sum(1) <= 1).all()  # on a grid edge: at most one fractional coordinate
    lo = np.floor(g).astype(np.int64)
    hi = lo.copy()
    hi[np.arange(len(g)), axis] += 1
    hi = np.minimum(hi, np.asarray(f["dims"]) - 1)
    v0, v1 = dens[lo[:, 0], lo[:, 1], lo[:, 2]], dens[hi[:, 0], hi[:, 1], hi[:, 2]]
    t = frac[np.arange(len(g)), axis]
    np.testing.assert_allclose(v0 + t * (v1 - v0), iso, atol=1e-5)
    np.testing.assert_allclose(f["vel"], S.sample_velocity(g, f["velocity"]), atol=1e-5)


def test_triangles_stay_in_their_cell(blob):
    _, _, f = blob
    g, tris = f["g"], f["tris"]
    assert tris.min() >= 0 and tris.max() < len(g)
    span = g[tris].max(1) - g[tris].min(1)
    assert (span <= 1.0 + 1e-6).all()
    assert consistently_wound(tris)
    edges = np.unique(np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), 1), axis=0)
    assert len(g) - len(edges) + len(tris) == 2  # one sphere-like surface


def test_sphere_closed_outward_right_size(device):
    m = mesh(sphere_particles((0.5, 0.5, 0.5), 0.1, DX), device=device)
    assert mesh_is_closed(m.tris) and consistently_wound(m.tris)
    vol = signed_volume(m.verts, m.tris)
    assert vol > 0 and vol == pytest.approx(4 / 3 * np.pi * 0.1**3, rel=0.10)
    assert np.allclose(m.verts.mean(0), 0.5, atol=0.005)


def test_two_blobs_two_components(device):
    p = np.concatenate([sphere_particles((0.3, 0.5, 0.5), 0.05, DX), sphere_particles((0.7, 0.5, 0.5), 0.05, DX)])
    assert components(mesh(p, device=device).tris) == 2


def test_resting_on_floor_closed(device):
    m = mesh(box_particles((0.4, 0.4, 0.0), (0.6, 0.6, 0.1), DX), device=device)
    assert mesh_is_closed(m.tris) and m.verts[:, 2].min() < 0.0  # closes just below z = 0, no hole


def test_empty(device):
    m = mesh(np.zeros((0, 3), np.float32), device=device)
    assert m.verts.shape == (0, 3) and m.tris.shape == (0, 3) and m.vel.shape == (0, 3)


def test_vertex_velocity(device):
    p = sphere_particles((0.5, 0.5, 0.5), 0.05, DX)
    m = mesh(p, vel=np.tile([1.0, 0, 0], (len(p), 1)).astype(np.float32), device=device)
    assert len(m.vel) == len(m.verts) and np.allclose(m.vel, [1.0, 0, 0], atol=1e-3)


def test_detail_increases_triangles(device):
    p = sphere_particles((0.5, 0.5, 0.5), 0.05, DX)
    assert len(mesh(p, settings=SurfaceSettings(detail=2), device=device).tris) > 2 * len(mesh(p, device=device).tris)


def test_too_many_cells_raises(device):
    p = np.array([[0, 0, 0], [4, 4, 0.1]], np.float32)
    with pytest.raises(ValueError, match="GPU memory"):
        surface_particles(p, np.zeros_like(p), 0.001, SurfaceSettings(detail=2), device)


def test_memory_budget_counts_bytes(device):
    p = sphere_particles((0.5, 0.5, 0.5), 0.05, DX)
    with pytest.raises(ValueError, match="GPU memory"):
        surface_particles(p, np.zeros_like(p), SPACING, SurfaceSettings(), device, max_bytes=1_000_000)


def test_large_field_chunked_scans(device):
    """Two blobs at opposite corners: over 16.7M grid edges, so the vertex-id scan runs in chunks."""
    p = np.concatenate([sphere_particles((0.05, 0.05, 0.05), 0.03, DX), sphere_particles((0.9, 0.92, 0.88), 0.03, DX)])
    origin, dims, cell, h = field_grid(p, SPACING, SurfaceSettings())
    assert 3 * dims[0] * dims[1] * dims[2] > 4096 * 4096
    m = mesh(p, device=device)
    assert mesh_is_closed(m.tris) and consistently_wound(m.tris) and components(m.tris) == 2
    far = m.verts[:, 0] > 0.5
    assert far.any() and (~far).any()
    vol = signed_volume(m.verts, m.tris)
    assert vol == pytest.approx(2 * 4 / 3 * np.pi * 0.03**3, rel=0.2)


def test_buffers_reused_across_sizes(device):
    """A large build, then a smaller one, then the large one again: same meshes as fresh builds."""
    big = sphere_particles((0.5, 0.5, 0.5), 0.08, DX)
    small = sphere_particles((0.2, 0.3, 0.4), 0.03, DX)
    a = mesh(big, device=device)
    b = mesh(small, device=device)
    c = mesh(big, device=device)
    assert mesh_is_closed(b.tris) and len(b.tris) < len(a.tris)
    assert mesh_is_closed(c.tris)
    # float atomics may order the splat differently: a node within rounding of iso may flip
    assert abs(len(a.tris) - len(c.tris)) <= 0.001 * len(a.tris)
    assert signed_volume(c.verts, c.tris) == pytest.approx(signed_volume(a.verts, a.tris), rel=1e-4)


def test_invalid_descriptions_rejected():
    from mpm_plugin_amd import _native as N

    s = surfacer("cuda:0")
    p = np.zeros((1, 3), np.float32)
    for field, value in (("cell", 0.0), ("h", float("nan")), ("inv_rest", -1.0)):
        desc = N.SurfaceDesc()
        desc.dims[:] = (4, 4, 4)
        desc.cell, desc.h, desc.inv_rest, desc.iso = 0.1, 0.2, 1.0, 0.5
        setattr(desc, field, value)
        with pytest.raises(ValueError):
            s.build(desc, p, p)
    desc = N.SurfaceDesc()
    desc.dims[:] = (1, 4, 4)
    desc.cell, desc.h, desc.inv_rest, desc.iso = 0.1, 0.2, 1.0, 0.5
    with pytest.raises(ValueError):
        s.build(desc, p, p)


def test_device_bounds_match_numpy():
    """mpm_surface_upload's bounds are numpy's column minima and maxima, so the field is placed exactly as
    the reference places it; a NaN coordinate gives NaN bounds, and surface_particles raises ValueError
    (as the reference's math.ceil of a NaN extent does)."""
    from mpm_plugin_amd import _native as N

    s = surfacer("cuda:0")
    rng = np.random.default_rng(7)
    for n in (1, 300, 1_000_003):
        x = (rng.normal(size=(n, 3)) * [1.0, 3.0, 0.1] + [5.0, -2.0, 0.3]).astype(np.float32)
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        N.check(s._lib.mpm_surface_upload(s._handle, n, N.fptr(x), N.fptr(x), N.fptr(lo), N.fptr(hi)))
        np.testing.assert_array_equal(lo, [x[:, a].min() for a in range(3)])
        np.testing.assert_array_equal(hi, [x[:, a].max() for a in range(3)])
    x[123, 1] = np.nan
    N.check(s._lib.mpm_surface_upload(s._handle, len(x), N.fptr(x), N.fptr(x), N.fptr(lo), N.fptr(hi)))
    assert np.isnan(lo).all() and np.isnan(hi).all()
    p = sphere_particles((0.5, 0.5, 0.5), 0.05, DX)
    p[7, 2] = np.nan
    with pytest.raises(ValueError):
        mesh(p)
