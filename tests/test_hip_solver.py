"""The reference's solver behaviour tests (tests/test_warp_basic.py:12-102, test_warp_materials.py:9-58)
run against the HIP backend through the unchanged Solver interface."""

import numpy as np
import pytest
from helpers import ALL_WALLS, box_particles, make_solver, solver_with_box, solver_with_cube, solver_with_presets, \
    solver_with_sphere

from mpm_core.backends import create_solver
from mpm_core.materials import MaterialParams, list_presets, load_preset
from mpm_core.scene import Domain

pytestmark = pytest.mark.gpu
JELLY = load_preset("jelly").material


def test_free_fall(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = create_solver("hip", domain=d, materials=[JELLY], gravity=(0, 0, -9.81), device=device)
    p = box_particles((0.45, 0.45, 0.6), (0.55, 0.55, 0.7), d.dx)
    s.add_particles(p, np.zeros_like(p), np.zeros(len(p), np.int32), np.zeros(len(p), np.int32))
    z0 = p[:, 2].mean()
    for _ in range(6):
        s.advance(1 / 24)
    assert z0 - s.get_particles().positions[:, 2].mean() == pytest.approx(0.5 * 9.81 * 0.25 ** 2, rel=0.02)


def test_rests_on_floor(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = solver_with_cube(device, d, lo=(0.4, 0.4, 0.0), hi=(0.5, 0.5, 0.1))
    start = s.get_particles().positions.copy()
    for _ in range(24):
        s.advance(1 / 24)
    pos = s.get_particles().positions
    assert pos[:, 2].min() >= -0.25 * d.dx
    assert np.abs(pos - start).max() < 0.5 * d.dx


def test_non_cubic_domain(device):
    d = Domain((-0.5, -0.2, 0.0), (0.5, 0.05, 0.5), 64)
    assert d.grid_shape == (64, 16, 32)
    s = solver_with_cube(device, d, lo=(-0.05, -0.1, 0.3), hi=(0.05, 0.0, 0.4))
    for _ in range(24):
        s.advance(1 / 24)
    pos = s.get_particles().positions
    assert pos[:, 2].min() >= -0.25 * d.dx and np.isfinite(pos).all()


def test_open_side_removes(device):
    d = Domain((0, 0, 0), (1, 1, 1), 32, walls={**ALL_WALLS, "z-": False})
    s = solver_with_cube(device, d, lo=(0.4, 0.4, 0.1), hi=(0.6, 0.6, 0.2))
    removed = sum(s.advance(1 / 24).removed for _ in range(24))
    assert s.particle_count == 0 and removed > 0


def test_state_roundtrip(device):
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    a = solver_with_cube(device, d, lo=(0.4, 0.4, 0.3), hi=(0.6, 0.6, 0.5))
    a.advance(1 / 24)
    b = make_solver(device, d, [JELLY])
    b.load_state(a.save_state())
    a.advance(1 / 24)
    b.advance(1 / 24)
    assert np.allclose(a.get_particles().positions, b.get_particles().positions, atol=1e-5)


def test_is_occupied(device):
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    s = solver_with_cube(device, d, lo=(0.4, 0.4, 0.0), hi=(0.6, 0.6, 0.2))
    s.advance(1 / 24)
    assert s.is_occupied(np.array([[0.5, 0.5, 0.1], [0.9, 0.9, 0.9]], np.float32)).tolist() == [True, False]
    assert s.is_occupied(np.zeros((0, 3), np.float32)).shape == (0,)


def test_nan_is_visible_in_positions(device):
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    s = solver_with_cube(device, d, lo=(0.4, 0.4, 0.0), hi=(0.5, 0.5, 0.1))
    state = s.save_state()
    state["v"][0] = np.nan
    s.load_state(state)
    s.advance(1 / 24)
    assert not np.isfinite(s.get_particles().positions).all()


def test_material_id_out_of_range_raises(device):
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    s = make_solver(device, d, [JELLY])
    with pytest.raises(ValueError):
        s.add_particles(np.zeros((1, 3), np.float32) + 0.5, np.zeros((1, 3), np.float32), np.array([1], np.int32),
                        np.array([0], np.int32))


@pytest.mark.parametrize("zmin", [0.0, 0.0100, 0.1326, 1.0359])
def test_wall_friction_independent_of_domain_placement(device, zmin):
    a = np.radians(20)
    d = Domain((0, 0, zmin), (1, 1, zmin + 1), 64, wall_friction=0.5)
    s = make_solver(device, d, [load_preset("whipped_cream").material],
                    gravity=(9.81 * np.sin(a), 0, -9.81 * np.cos(a)))
    p = box_particles((0.45, 0.45, zmin), (0.55, 0.55, zmin + 0.1), d.dx)
    s.add_particles(p, np.zeros_like(p), np.zeros(len(p), np.int32), np.zeros(len(p), np.int32))
    for _ in range(24):
        s.advance(1 / 24)
    assert abs(s.get_particles().positions[:, 0].mean() - p[:, 0].mean()) < 0.005


@pytest.mark.parametrize("preset", list_presets())
def test_slab_at_rest(device, preset):
    d = Domain((0, 0, 0), (0.5, 0.5, 0.5), 64)
    s = solver_with_box(device, d, load_preset(preset).material, lo=(0, 0, 0), hi=(0.5, 0.5, 0.07))
    start = s.get_particles().positions.copy()
    for _ in range(24):
        s.advance(1 / 24)
    assert np.abs(s.get_particles().positions - start).max() < 0.5 * d.dx


def test_sand_column_collapses(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = solver_with_box(device, d, load_preset("sand_dry").material, lo=(0.44, 0.44, 0), hi=(0.56, 0.56, 0.33))
    for _ in range(48):
        s.advance(1 / 24)
    pos = s.get_particles().positions
    assert pos[:, 2].max() < 0.6 * 0.33
    assert 0.12 < pos[:, 0].max() - 0.5 < 0.45


def test_cream_holds_shape(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = solver_with_sphere(device, d, load_preset("whipped_cream").material, center=(0.5, 0.5, 0.11), radius=0.11)
    for _ in range(48):
        s.advance(1 / 24)
    z = s.get_particles().positions[:, 2]
    assert z.max() - z.min() > 0.5 * 0.22


def test_jelly_recovers(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = solver_with_box(device, d, load_preset("jelly").material, lo=(0.4, 0.4, 0.2), hi=(0.6, 0.6, 0.4))
    for _ in range(72):
        s.advance(1 / 24)
    z = s.get_particles().positions[:, 2]
    assert z.max() - z.min() == pytest.approx(0.2, rel=0.15)


def test_mixed_materials_one_domain(device):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    s = solver_with_presets(device, d, [("sand_dry", (0.3, 0.3, 0), (0.5, 0.5, 0.1)),
                                        ("snow_packed", (0.5, 0.5, 0), (0.7, 0.7, 0.1))])
    for _ in range(24):
        s.advance(1 / 24)
    assert np.isfinite(s.get_particles().positions).all()


@pytest.mark.filterwarnings("ignore")
def test_unknown_backend():
    with pytest.raises(KeyError, match="hip"):
        create_solver("nope", domain=None, materials=[], gravity=(0, 0, 0))
