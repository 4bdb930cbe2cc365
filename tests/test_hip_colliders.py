"""Mesh colliders on the GPU: parity with the CPU oracle (exact closest point, pseudonormal sign) and the
reference's collider behaviour tests (tests/test_warp_colliders.py:47-121) run against the HIP backend."""

import numpy as np
import pytest
from helpers import (add, box_particles, icosphere, make_solver, plane_mesh, rel_err, rot_x, rot_y, run, scale,
                     solver_with_box, solver_with_cube, translate, unit_cube_mesh)
from test_hip_parity import pair, seeded_state

from mpm_core.geometry import transform_points
from mpm_core.materials import load_preset
from mpm_core.scene import Domain, MeshData
from mpm_core.solver import ColliderGeometry

pytestmark = pytest.mark.gpu

DOMAIN = Domain((0, 0, 0), (1, 1, 1), 64)
DX = DOMAIN.dx

COLLIDER_CASES = {
    "sphere_moving": (ColliderGeometry(MeshData(*icosphere(1.0, 3)), closed=True, friction=0.5),
                      translate(0.5, 0.5, 0.30) @ rot_x(10) @ scale(0.12),
                      translate(0.52, 0.49, 0.26) @ rot_x(25) @ scale(0.12)),
    "box_scaled": (ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=0.3),
                   translate(0.5, 0.5, 0.2) @ rot_y(20) @ scale(0.3, 0.15, 0.1),
                   translate(0.5, 0.5, 0.22) @ rot_y(24) @ scale(0.3, 0.15, 0.1)),
    "open_plane": (ColliderGeometry(MeshData(*plane_mesh(0.8)), closed=False, friction=0.8),
                   translate(0.5, 0.5, 0.25) @ rot_x(8), translate(0.5, 0.5, 0.24) @ rot_x(8)),
}


def collider_pair(case, seed, mode=0):
    geom, a, b = COLLIDER_CASES[case]
    state = seeded_state(40000, 64, seed, strain=0.02)
    hip, ora = pair(state, mode=mode)
    for s in (hip, ora):
        s.set_colliders([geom])
        s.set_collider_transforms(a[None], b[None])
    return hip, ora


# Grid velocity after collider contact: 1e-4. The collider velocity is a central difference of the FP32
# pose with h = 1e-3 (colliders.py:68-74), which scales last-ulp differences of the interpolated pose by
# 1 / (2 h) = 500; P2G's own 1e-5 bound is checked by the parity tests without colliders.
@pytest.mark.parametrize("case", sorted(COLLIDER_CASES))
def test_grid_contact_matches_oracle(case):
    hip, ora = collider_pair(case, 50)
    dt, s, frame_dt = 1.0e-4, 0.5, 1.0 / 24
    hip.debug_p2g(dt)
    ora.frame_dt = frame_dt
    ora.p2g(dt)
    gv_h, _ = hip.debug_grid_update(dt, s, frame_dt)
    gv_o = ora.grid_update(dt, s)
    assert rel_err(gv_h, gv_o) <= 1e-4


@pytest.mark.parametrize("case", sorted(COLLIDER_CASES))
def test_particle_projection_matches_oracle(case):
    hip, ora = collider_pair(case, 51)
    dt, s, frame_dt = 1.0e-4, 0.75, 1.0 / 24
    ora.frame_dt = frame_dt
    ora.p2g(dt)
    gv = ora.grid_update(dt, s).astype(np.float32)
    hip.set_grid(gv, ora.gm.astype(np.float32))
    hip.debug_g2p(dt, s, frame_dt)
    ora.g2p(dt, s, grid_v=gv)
    a, b = hip.save_state(), ora.save_state()
    for k in ("x", "v", "F"):
        assert rel_err(a[k], b[k]) <= 1e-4, k


def below_blob(case, state):
    """The case's collider moved so that its top sits half a cell below the particle block."""
    geom, a, b = COLLIDER_CASES[case]
    top = max(transform_points(m, geom.mesh.verts)[:, 2].max() for m in (a, b))
    shift = translate(0.0, 0.0, float(state["x"][:, 2].min()) - 0.5 * DX - top)
    return geom, shift @ a, shift @ b


@pytest.mark.parametrize("case", sorted(COLLIDER_CASES))
def test_substeps_with_collider_match_oracle(case):
    """Four substeps of a block resting in the contact band of a moving collider.

    The block starts outside the collider: a particle that starts inside is projected onto the surface,
    and on the next substep the sign of its rounding-sized distance decides whether its inward velocity is
    removed, so such states differ at the percent level even between the FP32 and FP64 oracles.
    """
    state = seeded_state(40000, 64, 52, strain=0.02)
    state["v"][:, 2] -= 1.0
    geom, a, b = below_blob(case, state)
    hip, ora = pair(state)
    for s in (hip, ora):
        s.set_colliders([geom])
        s.set_collider_transforms(a[None], b[None])
    frame_dt = 1.0 / 24
    ora.frame_dt = frame_dt
    for k in range(4):
        s = (k + 1) / 4
        hip.substep(1.0e-4, s, frame_dt)
        ora.substep(1.0e-4, s)
    a_, b_ = hip.save_state(), ora.save_state()
    for k in ("x", "v", "C", "F"):
        assert rel_err(a_[k], b_[k]) <= 1e-4, k


def test_block_on_sphere_aggregates_match_oracle():
    """300 substeps of a jelly block dropping onto a sphere: chaotic in detail, so compare what the
    reference's multi-step criterion compares: particle count (exact), centre of mass, bounding box and
    kinetic energy within 1%."""
    from mpm_core.backends import create_solver
    from oracle import oracle as O

    d = Domain((0, 0, 0), (1, 1, 1), 32)
    mat = [load_preset("jelly").material]
    x = box_particles((0.35, 0.35, 0.45), (0.65, 0.65, 0.6), d.dx)
    geom = ColliderGeometry(MeshData(*icosphere(1.0, 3)), closed=True, friction=0.5)
    xf = translate(0.5, 0.5, 0.25) @ scale(0.15)
    hip = create_solver("hip", domain=d, materials=mat, gravity=(0, 0, -9.81), device="cuda:0")
    ora = O.OracleSolver(d, mat, (0, 0, -9.81))
    for s in (hip, ora):
        s.add_particles(x, np.zeros_like(x), np.zeros(len(x), np.int32), np.zeros(len(x), np.int32))
        s.set_colliders([geom])
        s.set_collider_transforms(xf[None], xf[None])
    for _ in range(12):
        hip.advance(1 / 120)
        ora.advance(1 / 120)
    a, b = hip.save_state(), ora.save_state()
    assert len(a["x"]) == len(b["x"])
    com_a, com_b = a["x"].mean(0), b["x"].mean(0)
    assert np.abs(com_a - com_b).max() <= 0.01 * np.ptp(b["x"], axis=0).max()
    for f in (np.min, np.max):
        assert np.abs(f(a["x"], 0) - f(b["x"], 0)).max() <= 0.01 * np.ptp(b["x"], axis=0).max()
    ke_a, ke_b = (a["v"].astype(np.float64) ** 2).sum(), (b["v"].astype(np.float64) ** 2).sum()
    assert abs(ke_a - ke_b) <= 0.01 * ke_b


def test_collider_count_mismatch_raises():
    s = make_solver("cuda:0", DOMAIN, [load_preset("jelly").material])
    s.set_colliders([ColliderGeometry(MeshData(*plane_mesh(0.8)), closed=False, friction=0.5)])
    with pytest.raises(ValueError):
        s.set_collider_transforms(np.zeros((2, 4, 4)), np.zeros((2, 4, 4)))


# ---- the reference's collider behaviour tests (tests/test_warp_colliders.py) --------------------------

def cube_on_plane_tilted(device, degrees, friction, preset):
    plane_xf = translate(0.5, 0.5, 0.3) @ rot_y(degrees)
    s = make_solver(device, DOMAIN, [load_preset(preset).material])
    local = box_particles((-0.05, -0.05, 0.25 * DX), (0.05, 0.05, 0.1), DX)
    add(s, transform_points(plane_xf, local))
    s.set_colliders([ColliderGeometry(MeshData(*plane_mesh(0.8)), closed=False, friction=friction)])
    s.set_collider_transforms(plane_xf[None], plane_xf[None])
    return s


def points_in_box(pos, box_xf):
    local = transform_points(np.linalg.inv(box_xf), pos)
    half = 0.5 - 0.25 * DX / np.linalg.norm(box_xf[:3, :3], axis=0)
    return (np.abs(local) < half).all(axis=1)


def test_rests_on_open_plane(device):
    s = solver_with_cube(device, DOMAIN, lo=(0.4, 0.4, 0.45), hi=(0.6, 0.6, 0.55))
    s.set_colliders([ColliderGeometry(MeshData(*plane_mesh(0.8)), closed=False, friction=0.5)])
    xf = translate(0.5, 0.5, 0.3)
    s.set_collider_transforms(xf[None], xf[None])
    run(s, 36)
    zmin = s.get_particles().positions[:, 2].min()
    assert 0.3 - 0.25 * DX <= zmin <= 0.3 + 1.0 * DX


def test_scaled_box_collider(device):
    s = solver_with_cube(device, DOMAIN, lo=(0.45, 0.45, 0.4), hi=(0.55, 0.55, 0.5))
    s.set_colliders([ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=0.5)])
    xf = translate(0.5, 0.5, 0.0) @ scale(0.6, 0.6, 0.4)  # top face at z = 0.2
    s.set_collider_transforms(xf[None], xf[None])
    run(s, 36)
    assert s.get_particles().positions[:, 2].min() == pytest.approx(0.2, abs=1.0 * DX)


@pytest.mark.parametrize("friction,slides", [(0.0, True), (1.0, False)])
def test_tilted_plane_friction(device, friction, slides):
    s = cube_on_plane_tilted(device, degrees=10, friction=friction, preset="whipped_cream")
    start = s.get_particles().positions.mean(0)
    run(s, 24)
    moved = np.linalg.norm(s.get_particles().positions.mean(0) - start)
    assert (moved > 0.05) if slides else (moved < 0.01)


def test_moving_box_clears_sand(device):
    s = solver_with_box(device, DOMAIN, load_preset("sand_dry").material, lo=(0.1, 0.3, 0), hi=(0.9, 0.7, 0.06))
    s.set_colliders([ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=0.5)])

    def box(x):
        return translate(x, 0.5, 0.05) @ scale(0.1, 0.5, 0.1)

    for f in range(24):
        s.set_collider_transforms(box(0.15 + 0.3 * f / 24)[None], box(0.15 + 0.3 * (f + 1) / 24)[None])
        s.advance(1 / 24)
    inside = points_in_box(s.get_particles().positions, box(0.45))
    assert inside.sum() <= 0.001 * s.particle_count


def test_spinning_sphere_does_not_disturb_sand(device):
    r = 0.12

    def displacement(deg_per_frame):
        def ball(deg):
            return translate(0.5, 0.5, 0.1) @ rot_x(deg) @ scale(r)

        s = solver_with_box(device, DOMAIN, load_preset("sand_dry").material, lo=(0.25, 0.25, 0), hi=(0.75, 0.75, 0.15))
        s.set_colliders([ColliderGeometry(MeshData(*icosphere(1.0, 5)), closed=True, friction=0.0)])
        s.set_collider_transforms(ball(0)[None], ball(0)[None])
        s.advance(1 / 24)
        start = s.get_particles().positions.copy()
        for f in range(6):
            s.set_collider_transforms(ball(deg_per_frame * f)[None], ball(deg_per_frame * (f + 1))[None])
            s.advance(1 / 24)
        return np.linalg.norm(s.get_particles().positions - start, axis=1).mean()

    assert displacement(90) < displacement(0) + 0.25 * DX
