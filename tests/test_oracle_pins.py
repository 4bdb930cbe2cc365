"""Closed-form pins of the oracle's transfer, update and contact rules (CPU).

Each test states a fact the mathematics of the reference's substep fixes, independent of how the oracle
computes it, and chooses inputs on which a plausible misreading fails:

* P2G's affine moment (kernels.py:126-140). For one particle the grid momenta p_i at nodes x_i satisfy
  sum_i p_i (x_i - x_p)^T = -dt V_p tau + m_p C dx^2 / 4, because the quadratic B-spline weights have
  sum w (x_i - x_p) = 0 and sum w (x_i - x_p)(x_i - x_p)^T = dx^2/4 I. tau comes from eigen-decompositions
  (tests/closed_forms.py), not from an SVD.
* G2P on an affine grid field u = A x + b (kernels.py:210-242): v = A x_p + b, C = A and F <- (I + dt A) F.
* The return maps inside G2P (kernels.py:243-263): the new F has the projected singular values of the
  updated F with the same singular vectors, and snow keeps Jp det F (plastic volume takes what the clamp
  removed from the elastic part).
* Grid contact with a plane (colliders.py:96-117): behind the surface an approaching node keeps no normal
  velocity relative to the collider; inside the band of one cell it keeps its approach; in both its tangential
  speed drops by friction x normal speed (Coulomb); beyond the band, or separating, nothing changes.
* Particle projection (colliders.py:137-147): a particle inside lands on the surface and loses only the inward
  part of its velocity relative to the collider.

tests/test_oracle_mutations.py plants misreadings of each of these in a copy of the oracle and checks that this
file fails on every one.
"""

from __future__ import annotations

import numpy as np
import pytest
from closed_forms import f32, material_tau, rotation
from helpers import icosphere, scale, translate, unit_cube_mesh

from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain, MeshData
from mpm_core.solver import ColliderGeometry
from oracle import oracle as O

# resolution 32: dx = 1/32 and the padded origin -3/32 are exact in float32, so the float32 constants the
# oracle receives are the exact ones
DOMAIN = Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 32)
DX = 1.0 / 32

MATS = [MaterialParams("jelly", "elastic", 1.0e4, 0.3, 1000.0),
        MaterialParams("cream", "von_mises", 5.0e3, 0.45, 1000.0, yield_stress=200.0),
        MaterialParams("sand", "drucker_prager", 3.5e4, 0.3, 1600.0, friction_angle_deg=30.0),
        MaterialParams("snow", "snow", 1.4e5, 0.2, 400.0, hardening=10.0, critical_compression=0.025,
                       critical_stretch=0.0075)]
MODELS = ["elastic", "von_mises", "drucker_prager", "snow"]


def one_particle(x, v, C, F, jp, mat):
    return {"x": np.asarray(x, np.float32).reshape(1, 3), "v": np.asarray(v, np.float32).reshape(1, 3),
            "C": np.asarray(C, np.float32).reshape(1, 3, 3), "F": np.asarray(F, np.float32).reshape(1, 3, 3),
            "Jp": np.array([jp], np.float32), "mat": np.array([mat], np.int32), "emitter": np.zeros(1, np.int32)}


def node_positions(o) -> np.ndarray:
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in o.gm.shape], indexing="ij"), -1)
    return np.asarray(o.d.origin, np.float64) + idx * np.float64(o.d.dx)


def particle_mass(m: MaterialParams) -> float:
    """float32(density) x float32(dx^3 / ppc) rounded to float32 (solver.py:105)."""
    return f32(np.float32(m.density) * np.float32(DX ** 3 / 8))


def expected_tau(mi: int, F64: np.ndarray, jp: float) -> np.ndarray:
    m = MATS[mi]
    mu, lam = m.lame()
    return material_tau(m.model, F64, f32(mu), f32(lam), f32(m.hardening), f32(m.max_hardening), f32(jp))


# ---- P2G: mass, momentum and the affine moment ---------------------------------------------------

@pytest.mark.parametrize("mi", range(4), ids=MODELS)
@pytest.mark.parametrize("case", ["stress_and_affine", "affine_only", "stress_only"])
def test_p2g_affine_moment(mi, case):
    rng = np.random.default_rng(100 + 7 * mi + len(case))
    o = O.OracleSolver(DOMAIN, MATS, (0.0, 0.0, 0.0), precision="f64")
    pos = node_positions(o)
    dt = 1.0e-3
    for _ in range(3):
        x = rng.uniform(0.3, 0.7, 3).astype(np.float32)
        v = rng.normal(size=3).astype(np.float32)
        C = rng.normal(size=(3, 3)).astype(np.float32) if case != "stress_only" else np.zeros((3, 3), np.float32)
        F = (rotation(rng) @ (np.eye(3) + 0.1 * rng.normal(size=(3, 3)))).astype(np.float32)
        if case == "affine_only":
            F = np.eye(3, dtype=np.float32)  # tau(I) = 0 in every model
        assert np.linalg.det(F) > 0
        jp = float(rng.uniform(0.9, 1.1))
        o.load_state(one_particle(x, v, C, F, jp, mi))
        gv, gm = o.p2g(dt)
        mp = particle_mass(MATS[mi])
        xp = x.astype(np.float64)
        assert gm.sum() == pytest.approx(mp, rel=1e-12)
        assert np.allclose(gv.reshape(-1, 3).sum(0), mp * v.astype(np.float64), rtol=1e-12, atol=1e-18)
        assert np.allclose((gm[..., None] * pos).reshape(-1, 3).sum(0), mp * xp, rtol=1e-12)
        moment = np.einsum("ijka,ijkb->ab", gv, pos - xp)
        tau = expected_tau(mi, F.astype(np.float64), jp)
        want = -dt * (DX ** 3 / 8) * tau + mp * C.astype(np.float64) * DX * DX / 4
        assert np.linalg.norm(moment - want) <= 1e-9 * np.linalg.norm(want), (moment, want)


def test_p2g_affine_moment_adds_over_particles():
    """The moment identity is linear in the particles: sum_i p_i x_i^T = sum_p (m v x^T + aff dx^2/4)."""
    rng = np.random.default_rng(11)
    o = O.OracleSolver(DOMAIN, MATS, (0.0, 0.0, 0.0), precision="f64")
    pos = node_positions(o)
    n, dt = 64, 1.0e-3
    F = np.stack([rotation(rng) @ (np.eye(3) + 0.05 * rng.normal(size=(3, 3))) for _ in range(n)]).astype(np.float32)
    st = {"x": rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32), "v": rng.normal(size=(n, 3)).astype(np.float32),
          "C": rng.normal(size=(n, 3, 3)).astype(np.float32), "F": F,
          "Jp": rng.uniform(0.9, 1.1, n).astype(np.float32), "mat": (np.arange(n) % 4).astype(np.int32),
          "emitter": np.zeros(n, np.int32)}
    o.load_state(st)
    gv, _ = o.p2g(dt)
    got = np.einsum("ijka,ijkb->ab", gv, pos)
    want = np.zeros((3, 3))
    for p in range(n):
        mi = int(st["mat"][p])
        mp = particle_mass(MATS[mi])
        tau = expected_tau(mi, F[p].astype(np.float64), float(st["Jp"][p]))
        want += mp * np.outer(st["v"][p].astype(np.float64), st["x"][p].astype(np.float64))
        want += -dt * (DX ** 3 / 8) * tau + mp * st["C"][p].astype(np.float64) * DX * DX / 4
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


# ---- G2P on an affine field: transfer, F update, return maps ---------------------------------------

def affine_g2p(mi: int, F0: np.ndarray, A: np.ndarray, b: np.ndarray, dt: float, jp: float, rng):
    o = O.OracleSolver(DOMAIN, MATS, (0.0, 0.0, 0.0), precision="f64")
    x0 = rng.uniform(0.35, 0.65, 3).astype(np.float32)
    o.load_state(one_particle(x0, np.zeros(3), np.zeros((3, 3)), F0, jp, mi))
    o.g2p(dt, grid_v=node_positions(o) @ A.T + b)
    return x0.astype(np.float64), o.save_state()


def test_g2p_affine_field_elastic_f_update():
    rng = np.random.default_rng(21)
    for _ in range(8):
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        F0 = (rotation(rng) @ (np.eye(3) + 0.2 * rng.normal(size=(3, 3)))).astype(np.float32)
        dt = 1.0e-2
        x0, s = affine_g2p(0, F0, A, b, dt, 1.0, rng)
        v = A @ x0 + b
        assert np.allclose(s["v"][0], v, rtol=1e-6, atol=1e-6)
        assert np.allclose(s["C"][0], A, rtol=1e-6, atol=1e-6)
        assert np.allclose(s["x"][0], x0 + dt * v, rtol=0, atol=1e-6)
        want = (np.eye(3) + dt * A) @ F0.astype(np.float64)
        assert np.linalg.norm(s["F"][0] - want) <= 1e-6 * np.linalg.norm(want)
        assert s["Jp"][0] == 1.0  # elastic never touches Jp


def svd_desc(M: np.ndarray):
    U, s, Vt = np.linalg.svd(M)
    return U, s, Vt.T


@pytest.mark.parametrize("mi", [1, 2, 3], ids=MODELS[1:])
def test_g2p_return_map_on_updated_f(mi):
    """F_new = U diag(P(sigma)) V^T of the updated F (kernels.py:242-263); P is pinned in test_oracle.py."""
    rng = np.random.default_rng(31 + mi)
    m = MATS[mi]
    mu, lam = m.lame()
    for trial in range(12):
        A = 0.5 * rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        sig0 = rng.uniform(0.85, 1.2, 3) if trial % 2 else np.array([1.2, 1.0, 0.9])
        F0 = (rotation(rng) @ np.diag(sig0) @ rotation(rng).T).astype(np.float32)
        jp = float(rng.uniform(0.7, 1.1))
        dt = 1.0e-2
        _, s = affine_g2p(mi, F0, A, b, dt, jp, rng)
        Fn = (np.eye(3) + dt * A) @ F0.astype(np.float64)
        U, sig, V = svd_desc(Fn)
        P = O.project(m.model, sig, mu=f32(mu), lam=f32(lam), yield_stress=f32(m.yield_stress),
                      alpha=f32(O.dp_alpha(m.friction_angle_deg)), crit_c=f32(m.critical_compression),
                      crit_s=f32(m.critical_stretch))
        want = U @ np.diag(P) @ V.T
        assert np.linalg.norm(s["F"][0] - want) <= 2e-6 * np.linalg.norm(want), (trial, s["F"][0], want)
        if m.model == "snow":
            # J = Jp det F is carried through the clamp: what leaves the elastic part enters Jp
            got_j = float(s["Jp"][0]) * np.linalg.det(s["F"][0].astype(np.float64))
            assert got_j == pytest.approx(jp * np.linalg.det(Fn), rel=2e-6)
            assert np.allclose(np.linalg.svd(s["F"][0].astype(np.float64), compute_uv=False),
                               np.clip(sig, 1 - m.critical_compression, 1 + m.critical_stretch), rtol=2e-6)
        else:
            assert s["Jp"][0] == np.float32(jp)


def test_snow_jp_on_pure_clamp():
    """Diagonal F outside the clamp range, zero velocity field: Jp_new = Jp sigma1 sigma2 sigma3 / clamped."""
    rng = np.random.default_rng(41)
    F0 = np.diag([1.2, 1.0, 0.5]).astype(np.float32)
    _, s = affine_g2p(3, F0, np.zeros((3, 3)), np.zeros(3), 1.0e-3, 0.9, rng)
    clamped = np.array([1.0075, 1.0, 0.975])
    assert np.allclose(np.diag(s["F"][0]), clamped, rtol=1e-6)
    assert float(s["Jp"][0]) == pytest.approx(0.9 * (1.2 * 1.0 * 0.5) / clamped.prod(), rel=1e-6)


# ---- grid contact with a plane (colliders.py:77-117) ------------------------------------------------

TOP = 16.25 * DX  # top face of the slab collider, between node planes k = 19 (z = 16 dx) and k = 20
NODE_I = (16, 16)


def slab(friction: float) -> ColliderGeometry:
    return ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=friction)


def slab_pose(top: float, dx_shift: float = 0.0) -> np.ndarray:
    """A 3 x 3 x 0.4 slab whose top face (outward normal +z) is at height ``top``."""
    return translate(0.5 + dx_shift, 0.5, top - 0.2) @ scale(3.0, 3.0, 0.4)


def contact(friction: float, vel: np.ndarray, k: int, start=None, end=None, s_frac: float = 0.5,
            frame_dt: float = 1.0):
    o = O.OracleSolver(DOMAIN, MATS[:1], (0.0, 0.0, 0.0))
    o.set_colliders([slab(friction)])
    pose = slab_pose(TOP)
    o.set_collider_transforms([pose if start is None else start], [pose if end is None else end])
    o.frame_dt = frame_dt
    o.gv[...] = 0
    o.gm[...] = 0
    node = (*NODE_I, k)
    o.gm[node] = 1.0
    o.gv[node] = vel
    return o.grid_update(1.0e-4, s_frac)[node].astype(np.float64)


def split(vel, vcol):
    rel = np.asarray(vel, np.float64) - vcol
    vn = rel[2]  # the slab's top normal is +z
    return vn, rel - vn * np.array([0.0, 0.0, 1.0])


@pytest.mark.parametrize("friction", [0.0, 0.3, 1.0e6])
def test_contact_behind_surface_removes_approach(friction):
    vel = np.array([1.0, 0.5, -2.0], np.float32)
    out = contact(friction, vel, 19)  # phi = -0.25 dx
    vn0, rt0 = split(vel, 0.0)
    vn, rt = split(out, 0.0)
    assert abs(vn) <= 1e-5
    speed = max(0.0, np.linalg.norm(rt0) - friction * abs(vn0))  # Coulomb: |rt'| = max(0, |rt| - mu |vn|)
    assert np.linalg.norm(rt) == pytest.approx(speed, abs=1e-5)  # float32 oracle: ulp-level normal
    if speed > 0:
        assert np.allclose(rt / np.linalg.norm(rt), rt0 / np.linalg.norm(rt0), atol=1e-5)


@pytest.mark.parametrize("friction", [0.0, 0.3, 1.0e6])
def test_contact_in_band_keeps_approach(friction):
    vel = np.array([1.0, 0.5, -2.0], np.float32)
    out = contact(friction, vel, 20)  # phi = +0.75 dx: in front of the surface, within the band of one cell
    vn0, rt0 = split(vel, 0.0)
    vn, rt = split(out, 0.0)
    assert vn == pytest.approx(vn0, abs=1e-5)
    assert np.linalg.norm(rt) == pytest.approx(max(0.0, np.linalg.norm(rt0) - friction * abs(vn0)), abs=1e-5)


@pytest.mark.parametrize("k", [21, 22])
def test_contact_beyond_band_untouched(k):
    vel = np.array([1.0, 0.5, -2.0], np.float32)
    assert np.array_equal(contact(0.3, vel, k), vel)  # phi = 1.75 dx and 2.75 dx


@pytest.mark.parametrize("k", [19, 20])
def test_contact_separating_untouched(k):
    vel = np.array([1.0, 0.5, 2.0], np.float32)
    assert np.array_equal(contact(0.3, vel, k), vel)


def test_contact_relative_to_moving_collider():
    """A slab moving at (0.2, 0, -0.1) per unit time: the rules hold for the velocity relative to it."""
    frame_dt = 0.5
    move = np.array([0.1, 0.0, -0.05])  # over one frame
    vcol = move / frame_dt
    start = slab_pose(TOP - 0.5 * move[2], -0.5 * move[0])
    end = slab_pose(TOP + 0.5 * move[2], 0.5 * move[0])  # at s = 0.5 the top face is at TOP
    vel = np.array([1.0, 0.5, -2.0], np.float32)
    vn0, rt0 = split(vel, vcol)
    for k, keep in ((19, False), (20, True)):
        out = contact(0.3, vel, k, start, end, 0.5, frame_dt)
        vn, rt = split(out, vcol)
        assert vn == pytest.approx(vn0 if keep else 0.0, abs=2e-3)
        assert np.linalg.norm(rt) == pytest.approx(np.linalg.norm(rt0) - 0.3 * abs(vn0), abs=2e-3)


# ---- particle projection (colliders.py:120-149) ----------------------------------------------------

def project(geom, pose, x, v, frame_dt=1.0, end=None):
    o = O.OracleSolver(DOMAIN, MATS[:1], (0.0, 0.0, 0.0))
    n = len(x)
    o.load_state({"x": np.asarray(x, np.float32), "v": np.asarray(v, np.float32),
                  "C": np.zeros((n, 3, 3), np.float32), "F": np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)),
                  "Jp": np.ones(n, np.float32), "mat": np.zeros(n, np.int32), "emitter": np.zeros(n, np.int32)})
    o.set_colliders([geom])
    o.set_collider_transforms([pose], [pose if end is None else end])
    o.frame_dt = frame_dt
    o.project_particles(0.5)
    s = o.save_state()
    return s["x"].astype(np.float64), s["v"].astype(np.float64)


def test_projection_onto_plane_face():
    x = np.array([[0.5, 0.5, TOP - 0.3 * DX], [0.4, 0.6, TOP - 0.9 * DX], [0.5, 0.5, TOP + 0.5 * DX]])
    v = np.array([[0.5, 0.2, -1.0], [0.5, 0.2, 1.0], [0.5, 0.2, -1.0]])
    xo, vo = project(slab(0.5), slab_pose(TOP), x, v)
    assert np.allclose(xo[:2, 2], TOP, atol=1e-6) and np.allclose(xo[:2, :2], x[:2, :2].astype(np.float32))
    assert np.allclose(vo[0], [0.5, 0.2, 0.0], atol=1e-7)  # inward part removed, no friction
    assert np.allclose(vo[1], v[1].astype(np.float32))  # already leaving: velocity kept
    assert np.array_equal(xo[2], x[2].astype(np.float32)) and np.array_equal(vo[2], v[2].astype(np.float32))


def test_projection_onto_sphere():
    rng = np.random.default_rng(51)
    geom = ColliderGeometry(MeshData(*icosphere(1.0, 5)), closed=True, friction=0.0)
    c, R = np.array([0.5, 0.5, 0.5]), 0.2
    d = rng.normal(size=(200, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.12, 0.195, 200)
    x = c + r[:, None] * d
    v = rng.normal(size=(200, 3))
    xo, vo = project(geom, translate(*c) @ scale(R), x, v)
    rad = np.linalg.norm(xo - c, axis=1)
    assert np.all(rad <= R + 1e-6) and np.all(rad >= R - 2.5e-4)  # on the facets of subdivision 5
    out = (xo - c) / rad[:, None]
    assert np.all(np.einsum("ij,ij->i", out, d) > 0.999)  # pushed out along the radius
    dv = vo - v.astype(np.float32)
    inward = np.einsum("ij,ij->i", v, d) < 0
    # approaching particles lose their radial approach and nothing else; leaving ones keep their velocity
    assert np.all(np.abs(np.einsum("ij,ij->i", vo[inward], d[inward])) < 2e-2 * np.linalg.norm(v[inward], axis=1))
    tang = dv[inward] - np.einsum("ij,ij->i", dv[inward], d[inward])[:, None] * d[inward]
    assert np.all(np.linalg.norm(tang, axis=1) < 2e-2 * np.linalg.norm(v[inward], axis=1))
    assert np.allclose(dv[~inward], 0.0)


def test_projection_relative_to_moving_collider():
    move = np.array([0.0, 0.0, 0.1])  # the slab rises by 0.1 over the frame (0.2 per unit time)
    start, end = slab_pose(TOP - 0.05), slab_pose(TOP + 0.05)
    x = np.array([[0.5, 0.5, TOP - 0.3 * DX]])
    v = np.array([[0.5, 0.2, 0.1]])  # rising slower than the slab: approaching it
    xo, vo = project(slab(0.0), start, x, v, frame_dt=0.5, end=end)
    assert xo[0, 2] == pytest.approx(TOP, abs=1e-6)
    assert np.allclose(vo[0], [0.5, 0.2, move[2] / 0.5], atol=1e-3)  # normal velocity = the slab's
