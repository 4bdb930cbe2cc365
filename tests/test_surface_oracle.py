"""Pins the surfacer's CPU oracle (oracle/surface.py) and the host half of mpm_surface.field against what
the mathematics fixes, independently of their formulas: the cubic B-spline's values, continuity and
integral, a lone particle on a node, a particle's total splatted weight, the rest density of a lattice,
the velocity of a uniformly moving particle set, and trilinear sampling of a linear field."""

import math

import numpy as np
import pytest
from scipy import integrate

from mpm_surface.field import field_grid, rest_density
from mpm_surface.settings import SurfaceSettings
from oracle import surface as S


def test_cubic_values_and_continuity():
    assert S.cubic(np.float32(0)) == 1.0
    assert S.cubic(np.float32(0.5)) == pytest.approx(0.25)
    assert S.cubic(np.float32(1.0)) == 0.0 and S.cubic(np.float32(1.5)) == 0.0
    eps = 1e-4
    for q in (0.5,):  # value and slope continuous at the knot: -1.5 from both sides
        left = (S.cubic(np.float32(q - eps)) - S.cubic(np.float32(q - 2 * eps))) / eps
        right = (S.cubic(np.float32(q + 2 * eps)) - S.cubic(np.float32(q + eps))) / eps
        assert left == pytest.approx(-1.5, abs=2e-2) and right == pytest.approx(-1.5, abs=2e-2)


def test_cubic_integral():
    """The 3D integral of w(|r| / h) is 4 pi h^3 int_0^1 w(q) q^2 dq = pi h^3 / 8."""
    val, _ = integrate.quad(lambda q: float(S.cubic(np.float32(q))) * q * q, 0, 1, points=[0.5])
    assert 4 * math.pi * val == pytest.approx(math.pi / 8, rel=1e-5)


def test_lone_particle_on_a_node():
    dims = (9, 9, 9)
    x = np.array([[4.0, 4.0, 4.0]], np.float32)  # exactly on node (4, 4, 4) with origin 0 and cell 1
    d, m = S.splat(x, np.array([[1.0, 2.0, 3.0]], np.float32), (0, 0, 0), 1.0, 2.0, dims)
    assert d[4, 4, 4] == 1.0
    assert d[5, 4, 4] == pytest.approx(0.25)  # q = 1/2
    assert d[5, 5, 4] == pytest.approx(2 * (1 - math.sqrt(2) / 2) ** 3, rel=1e-6)
    assert d[6, 4, 4] == 0.0  # q = 1: outside the support
    np.testing.assert_allclose(m[4, 4, 4], [1, 2, 3])


@pytest.mark.parametrize("h", [3.0, 4.5])
def test_total_weight_is_the_kernel_integral(h):
    """Far from the edges a particle's weights sum to the Riemann sum of pi h^3 / 8 (cell = 1)."""
    rng = np.random.default_rng(1)
    x = (12 + rng.random((5, 3))).astype(np.float32)
    d, _ = S.splat(x, np.zeros_like(x), (0, 0, 0), 1.0, h, (26, 26, 26))
    assert d.sum() / len(x) == pytest.approx(math.pi * h**3 / 8, rel=0.02)


def test_rest_density_matches_the_integral():
    for spacing, h in ((0.005, 0.02), (0.01, 0.06), (1.0, 3.0)):
        assert rest_density(spacing, h) == pytest.approx(math.pi * (h / spacing) ** 3 / 8, rel=0.02)


def test_lattice_normalizes_to_one():
    """A uniform lattice at the spacing reads 1 at interior nodes after normalisation."""
    spacing, h = 1.0, 4.0
    r = np.arange(0, 20, dtype=np.float32)
    x = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) + 0.25
    d, m = S.splat(x, np.tile([0.5, -1.0, 2.0], (len(x), 1)).astype(np.float32), (0, 0, 0), 1.0, h, (20, 20, 20))
    dn, vel = S.normalize(d, m, 1.0 / rest_density(spacing, h))
    assert dn[8:12, 8:12, 8:12] == pytest.approx(1.0, abs=0.02)
    occupied = d > 0
    np.testing.assert_allclose(vel[occupied], np.broadcast_to([0.5, -1.0, 2.0], vel[occupied].shape), rtol=1e-5)
    assert (vel[~occupied] == 0).all()


def test_trilinear_is_exact_on_linear_fields():
    dims = (5, 6, 7)
    g = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in dims], indexing="ij"), -1)
    a = np.array([[1.0, -2.0, 0.5], [0.3, 0.0, 1.0], [-1.0, 4.0, 2.0]])
    field = g @ a.T + [1.0, 2.0, 3.0]
    rng = np.random.default_rng(0)
    pts = rng.random((50, 3)) * (np.asarray(dims) - 1)
    pts[0] = [4.0, 5.0, 6.0]  # the far corner node: base cell clamped to dims - 2
    np.testing.assert_allclose(S.sample_velocity(pts, field), pts @ a.T + [1.0, 2.0, 3.0], atol=1e-12)


def test_field_grid_pads_beyond_the_kernel():
    p = np.array([[0.1, 0.2, 0.3], [0.4, 0.25, 0.35]], np.float32)
    s = SurfaceSettings(detail=2.0, smoothing=1.5)
    origin, dims, cell, h = field_grid(p, 0.01, s)
    assert cell == 0.005 and h == pytest.approx(0.03)
    top = origin + (np.asarray(dims) - 1) * cell
    assert (p.min(0) - origin >= h + 2 * cell - 1e-9).all() and (top - p.max(0) >= h + 2 * cell - 1e-6).all()
