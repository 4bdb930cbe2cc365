"""Pins the CPU oracle (oracle/cpu_ref.cpp) with what the mathematics fixes, independent of its formulas:

* B-spline weights: partition of unity, zero first moment, constant second moment 1/4, closed values.
* 3x3 SVD: numpy.linalg.svd singular values, reconstruction, proper rotations, sign convention.
* Stresses: polar decomposition and Hencky strain from numpy.linalg.eigh of F^T F / F F^T.
* Return maps: the projected log strain lies on the yield surface (von Mises) or cone (Drucker-Prager)
  with its volume part kept; special branches by known answers.
* P2G: total mass and total momentum equal the particles' (sum w = 1, sum w dpos = 0).
* G2P: an affine grid velocity u = A x + b is reproduced exactly, v = A x_p + b and C = A.
* Grid update and walls: per-node closed forms.
* Colliders: analytic distances of spheres, boxes and planes; rigid-motion velocities.
* FP32 and FP64 oracles agree on a smooth state.
The transfer, F-update and contact rules that these invariants cannot see (P2G's affine moment, the F update
and snow's Jp, plane contact and particle projection) are pinned by closed forms in tests/test_oracle_pins.py;
tests/test_oracle_mutations.py plants misreadings in a copy of the oracle and checks the two files catch each.
"""

import math

import numpy as np
import pytest
from helpers import icosphere, plane_mesh, rot_x, scale, translate, unit_cube_mesh

from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain, MeshData
from mpm_core.solver import ColliderGeometry
from oracle import oracle as O

RNG = np.random.default_rng(1234)


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---- B-spline weights (kernels.py:16-22) --------------------------------------------------------

@pytest.mark.parametrize("fx", [0.5, 0.5 + 1e-7, 0.73, 1.0, 1.25, 1.4999999])
def test_bspline_moments(fx):
    w = np.array([O.bspline_w(fx, i) for i in range(3)])
    d = np.arange(3) - fx
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert (w * d).sum() == pytest.approx(0.0, abs=1e-15)
    assert (w * d * d).sum() == pytest.approx(0.25, abs=1e-15)
    assert (w >= 0).all()


def test_bspline_known_values():
    assert [O.bspline_w(1.0, i) for i in range(3)] == [0.125, 0.75, 0.125]
    assert [O.bspline_w(0.5, i) for i in range(3)] == [0.5, 0.5, 0.0]


# ---- SVD ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["random", "near_identity", "reflection", "degenerate", "rank_deficient"])
def test_svd_against_numpy(kind):
    rng = np.random.default_rng(hash(kind) % 2**32)
    for _ in range(200):
        if kind == "random":
            F = rng.normal(size=(3, 3))
        elif kind == "near_identity":
            F = rotation(rng) @ (np.eye(3) + 0.05 * rng.normal(size=(3, 3)))
        elif kind == "reflection":
            F = rotation(rng) @ np.diag([1.2, 0.9, -0.7]) @ rotation(rng).T
        elif kind == "degenerate":
            s = rng.uniform(0.5, 1.5)
            F = rotation(rng) @ np.diag([s, s * (1 + 1e-9), s * (1 - 1e-9)]) @ rotation(rng).T
        else:
            F = rotation(rng) @ np.diag([1.3, 0.4, 0.0]) @ rotation(rng).T
        U, s, V = O.svd3(F)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-12)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-12) and np.allclose(V @ V.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) == pytest.approx(1, abs=1e-12) and np.linalg.det(V) == pytest.approx(1, abs=1e-12)
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.abs(s), ref, atol=1e-12)
        assert s[0] >= 0 and s[1] >= 0 and abs(s[0]) >= abs(s[1]) - 1e-12 >= abs(s[2]) - 2e-12
        if abs(np.linalg.det(F)) > 1e-9:
            assert (s[2] < 0) == (np.linalg.det(F) < 0)


def test_svd_f32_accuracy():
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(500):
        F = (rotation(rng) @ (np.eye(3) + 0.2 * rng.normal(size=(3, 3)))).astype(np.float32)
        U, s, V = O.svd3(F, "f32")
        worst = max(worst, np.abs(U.astype(np.float64) @ np.diag(s) @ V.T - F).max())
    assert worst < 5e-6


# ---- stresses (kernels.py:30-60) ------------------------------------------------------------------

def polar(F):
    lam, Q = np.linalg.eigh(F.T @ F)
    S = Q @ np.diag(np.sqrt(lam)) @ Q.T
    return F @ np.linalg.inv(S)


@pytest.mark.parametrize("seed", range(5))
def test_fixed_corotated_from_polar(seed):
    rng = np.random.default_rng(seed)
    F = rotation(rng) @ (np.eye(3) + 0.2 * rng.normal(size=(3, 3)))
    assert np.linalg.det(F) > 0
    mu, lam = 3.0e4, 2.0e4
    R = polar(F)
    J = np.linalg.det(F)
    expect = 2 * mu * (F - R) @ F.T + lam * J * (J - 1) * np.eye(3)
    assert np.allclose(O.tau("elastic", F, mu, lam), expect, rtol=1e-9, atol=1e-6)


@pytest.mark.parametrize("seed", range(5))
def test_snow_hardening_scales_fixed_corotated(seed):
    rng = np.random.default_rng(seed)
    F = rotation(rng) @ (np.eye(3) + 0.1 * rng.normal(size=(3, 3)))
    mu, lam, xi, jp = 3.0e4, 2.0e4, 10.0, 0.95
    h = min(math.exp(xi * (1 - jp)), 5.0)
    expect = O.tau("elastic", F, mu * h, lam * h)
    assert np.allclose(O.tau("snow", F, mu, lam, hardening=xi, max_hardening=5.0, Jp=jp), expect, rtol=1e-12)
    capped = O.tau("snow", F, mu, lam, hardening=xi, max_hardening=1.2, Jp=0.5)
    assert np.allclose(capped, O.tau("elastic", F, mu * 1.2, lam * 1.2), rtol=1e-12)


@pytest.mark.parametrize("model", ["von_mises", "drucker_prager"])
def test_hencky_from_left_stretch(model):
    rng = np.random.default_rng(7)
    F = rotation(rng) @ (np.eye(3) + 0.2 * rng.normal(size=(3, 3)))
    mu, lam = 3.0e4, 2.0e4
    b_eig, Q = np.linalg.eigh(F @ F.T)
    eps = 0.5 * np.log(b_eig)
    expect = Q @ np.diag(2 * mu * eps + lam * eps.sum()) @ Q.T
    assert np.allclose(O.tau(model, F, mu, lam), expect, rtol=1e-9, atol=1e-6)


def test_hencky_diagonal_and_log_clamp():
    F = np.diag([2.0, 0.5, 1e-6])
    mu, lam = 1.0, 0.0
    eps = np.log([2.0, 0.5, 1e-4])  # singular values below 1e-4 are clamped before the log
    assert np.allclose(np.diag(O.tau("von_mises", F, mu, lam)), 2 * eps)


# ---- return maps (kernels.py:63-89, 258-262) ------------------------------------------------------

def test_von_mises_inside_yield_unchanged():
    sig = np.array([1.001, 0.999, 1.0])
    assert np.array_equal(O.project("von_mises", sig, mu=1e4, yield_stress=1e6), sig)


def test_von_mises_projects_onto_yield_surface():
    sig = np.array([1.3, 0.9, 0.8])
    mu, ys = 1.0e4, 500.0
    out = O.project("von_mises", sig, mu=mu, yield_stress=ys)
    e, eo = np.log(sig), np.log(out)
    dev = eo - eo.mean()
    assert np.linalg.norm(dev) == pytest.approx(ys / (2 * mu), rel=1e-12)
    assert eo.sum() == pytest.approx(e.sum(), abs=1e-12)  # volume kept
    assert np.allclose(dev / np.linalg.norm(dev), (e - e.mean()) / np.linalg.norm(e - e.mean()))


def test_drucker_prager_branches():
    mu, lam = 1.0e4, 1.5e4
    alpha = O.dp_alpha(30.0)
    # expansion: stress free
    assert np.array_equal(O.project("drucker_prager", np.array([1.1, 1.0, 1.0]), mu=mu, lam=lam, alpha=alpha),
                          [1.0, 1.0, 1.0])
    # pure compression, no deviatoric part: unchanged
    s = np.array([0.9, 0.9, 0.9])
    assert np.array_equal(O.project("drucker_prager", s, mu=mu, lam=lam, alpha=alpha), s)
    # inside the cone: unchanged
    s = np.array([0.90, 0.899, 0.901])
    assert np.array_equal(O.project("drucker_prager", s, mu=mu, lam=lam, alpha=alpha), s)
    # outside the cone: lands on it, volume kept
    s = np.array([1.2, 0.9, 0.7])
    out = O.project("drucker_prager", s, mu=mu, lam=lam, alpha=alpha)
    assert not np.allclose(out, s)
    e, eo = np.log(s), np.log(out)
    assert eo.sum() == pytest.approx(e.sum(), abs=1e-12)
    k = (3 * lam + 2 * mu) / (2 * mu) * alpha
    assert np.linalg.norm(eo - eo.mean()) == pytest.approx(-k * e.sum(), rel=1e-10)


def test_dp_alpha_closed_form():
    s = math.sin(math.radians(30.0))
    assert O.dp_alpha(30.0) == pytest.approx(math.sqrt(2 / 3) * 2 * s / (3 - s))
    assert O.dp_alpha(0.0) == 0.0


def test_snow_clamp():
    out = O.project("snow", np.array([1.2, 1.0, 0.5]), crit_c=0.025, crit_s=0.0075)
    assert np.allclose(out, [1.0075, 1.0, 0.975])


# ---- P2G / G2P / grid update invariants ------------------------------------------------------------

MATS = [MaterialParams("e", "elastic", 1.0e4, 0.3, 1000.0),
        MaterialParams("v", "von_mises", 5.0e3, 0.45, 1000.0, yield_stress=200.0),
        MaterialParams("d", "drucker_prager", 3.5e4, 0.3, 1600.0, friction_angle_deg=30.0),
        MaterialParams("s", "snow", 1.4e5, 0.2, 400.0, hardening=10.0, critical_compression=0.025,
                       critical_stretch=0.0075)]


def random_state(n, rng, lo=0.3, hi=0.7):
    F = np.stack([rotation(rng) @ (np.eye(3) + 0.03 * rng.normal(size=(3, 3))) for _ in range(n)])
    return {"x": rng.uniform(lo, hi, (n, 3)).astype(np.float32), "v": rng.normal(size=(n, 3)).astype(np.float32),
            "C": rng.normal(size=(n, 3, 3)).astype(np.float32), "F": F.astype(np.float32),
            "Jp": rng.uniform(0.9, 1.1, n).astype(np.float32), "mat": rng.integers(0, 4, n).astype(np.int32),
            "emitter": np.zeros(n, np.int32)}


def test_p2g_conserves_mass_and_momentum():
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    o = O.OracleSolver(d, MATS, (0, 0, 0), precision="f64")
    st = random_state(500, np.random.default_rng(0))
    o.load_state(st)
    gv, gm = o.p2g(1.0e-3)
    pm = np.array([o.m[i].p_mass for i in range(4)], np.float64)[st["mat"]]
    assert gm.sum() == pytest.approx(pm.sum(), rel=1e-12)
    assert np.allclose(gv.reshape(-1, 3).sum(0), (pm[:, None] * st["v"]).sum(0), rtol=1e-10, atol=1e-12)


def test_p2g_single_particle_at_node_centre():
    """A particle at rest with C = 0 and F = I at fx = (1,1,1): weights 1/8, 3/4, 1/8 per axis."""
    d = Domain((0, 0, 0), (1, 1, 1), 16)
    o = O.OracleSolver(d, MATS[:1], (0, 0, 0), precision="f64")
    dx = d.dx
    node = np.array([10, 7, 5])
    x = (np.asarray(o.d.origin, np.float64) + node * dx).astype(np.float32)[None]
    st = {"x": x, "v": np.array([[1.0, 2.0, 3.0]], np.float32), "C": np.zeros((1, 3, 3), np.float32),
          "F": np.eye(3, dtype=np.float32)[None], "Jp": np.ones(1, np.float32), "mat": np.zeros(1, np.int32),
          "emitter": np.zeros(1, np.int32)}
    o.load_state(st)
    gv, gm = o.p2g(1.0e-3)
    w1 = np.array([0.125, 0.75, 0.125])
    w = np.einsum("i,j,k->ijk", w1, w1, w1)
    pm = o.m[0].p_mass
    got = gm[node[0] - 1:node[0] + 2, node[1] - 1:node[1] + 2, node[2] - 1:node[2] + 2]
    assert np.allclose(got, w * pm, rtol=1e-6)
    assert np.allclose(gv[tuple(node)], 0.75 ** 3 * pm * np.array([1.0, 2.0, 3.0]), rtol=1e-6)


def test_g2p_reproduces_affine_velocity_field():
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    o = O.OracleSolver(d, MATS[:1], (0, 0, 0), precision="f64")
    rng = np.random.default_rng(3)
    st = random_state(300, rng)
    st["mat"][:] = 0
    o.load_state(st)
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in o.gv.shape[:3]], indexing="ij"), -1)
    pos = np.asarray(o.d.origin, np.float64) + idx * np.float64(o.d.dx)
    grid = pos @ A.T + b
    dt = 1.0e-7  # tiny advection so x stays put for the check
    o.g2p(dt, grid_v=grid)
    s = o.save_state()
    x0 = st["x"].astype(np.float64)
    # the oracle computes positions in float32 on the host side of the transfer; compare loosely there
    assert np.allclose(s["v"], x0 @ A.T + b, rtol=0, atol=2e-5)
    assert np.allclose(s["C"], np.broadcast_to(A, s["C"].shape), rtol=0, atol=2e-4)


def test_grid_update_closed_forms():
    d = Domain((0, 0, 0), (1, 1, 1), 16, wall_friction=0.5)
    o = O.OracleSolver(d, MATS[:1], (0.0, 0.0, -10.0), precision="f64")
    o.gv[...] = 0
    o.gm[...] = 0
    dt = 0.01
    free = (8, 8, 8)
    o.gm[free] = 2.0
    o.gv[free] = [2.0, 4.0, 6.0]
    lo_z = int(o.d.wall_lo_node[2])
    wall = (8, 8, lo_z)
    o.gm[wall] = 1.0
    o.gv[wall] = [3.0, 0.0, -4.0]  # into the floor after gravity: vz = -4.1
    gv = o.grid_update(dt)
    assert np.allclose(gv[free], [1.0, 2.0, 3.0 - 0.1])
    vn, vt = 4.1, 3.0
    assert np.allclose(gv[wall], [3.0 * max(0.0, 1 - 0.5 * vn / vt), 0.0, 0.0])
    away = (8, 8, lo_z + 1)
    o.gm[...] = 0
    o.gm[away] = 1.0
    o.gv[away] = [3.0, 0.0, -4.0]
    assert np.allclose(o.grid_update(dt)[away], [3.0, 0.0, -4.1])  # above the wall node: untouched


def test_f32_and_f64_oracles_agree():
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    st = random_state(2000, np.random.default_rng(9), 0.35, 0.65)
    st["C"] *= 0.1
    out = {}
    for prec in ("f32", "f64"):
        o = O.OracleSolver(d, MATS, (0, 0, -9.81), precision=prec)
        o.load_state(st)
        o.substep(1.0e-4)
        out[prec] = o.save_state()
    for k in ("x", "v", "C", "F", "Jp"):
        a, b = out["f32"][k].astype(np.float64), out["f64"][k].astype(np.float64)
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), k


# ---- colliders ------------------------------------------------------------------------------------

def test_sphere_distance_and_normal():
    geom = ColliderGeometry(MeshData(*icosphere(1.0, 5)), closed=True, friction=0.0)
    xf = translate(0.5, 0.5, 0.5) @ scale(0.2)
    rng = np.random.default_rng(4)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = rng.uniform(0.05, 0.23)
        p = 0.5 + r * d
        n, phi = O.collider_sdf(geom, xf, xf, 0.5, p, 2 / 64, 1 / 64)
        assert phi == pytest.approx(r - 0.2, abs=2.5e-4)  # faceted sphere: chord sag of subdivision 5
        if abs(r - 0.2) > 1e-3:
            assert np.dot(n, d) > 0.99


def test_box_distance_sign_and_far():
    geom = ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=0.0)
    xf = translate(0.5, 0.5, 0.2) @ scale(0.4, 0.3, 0.2)
    n, phi = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.32], 2 / 64, 1 / 64)
    assert phi == pytest.approx(0.02, abs=1e-6) and np.allclose(n, [0, 0, 1])
    n, phi = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.28], 2 / 64, 1 / 64)
    assert phi == pytest.approx(-0.02, abs=1e-6) and np.allclose(n, [0, 0, 1])
    # distance is measured in the local frame then mapped to the world (colliders.py:53-55)
    n, phi = O.collider_sdf(geom, xf, xf, 0.0, [0.69, 0.5, 0.2], 2 / 64, 1 / 64)
    assert phi == pytest.approx(-0.01, abs=1e-6) and np.allclose(n, [1, 0, 0])
    _, phi = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.9], 2 / 64, 1 / 64)
    assert phi == 1e30  # beyond the AABB margin


def test_open_plane_thickness():
    geom = ColliderGeometry(MeshData(*plane_mesh(0.8)), closed=False, friction=0.0)
    xf = translate(0.5, 0.5, 0.3)
    th = 2 / 64
    _, above = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.31], th, 1 / 64)
    _, behind = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.29], th, 1 / 64)
    _, deep = O.collider_sdf(geom, xf, xf, 0.0, [0.5, 0.5, 0.3 - 1.5 * th], th, 1 / 64)
    assert above == pytest.approx(0.01, abs=1e-7) and behind == pytest.approx(-0.01, abs=1e-7)
    assert deep == 1e30


def test_collider_velocity_rigid_motion():
    geom = ColliderGeometry(MeshData(*unit_cube_mesh(centered=True)), closed=True, friction=0.0)
    a = translate(0.2, 0.3, 0.4)
    b = translate(0.5, 0.3, 0.4)
    v = O.collider_velocity(geom, a, b, 0.5, [0.35, 0.3, 0.4], 24.0, 1 / 64)
    assert np.allclose(v, [0.3 * 24.0, 0.0, 0.0], rtol=1e-6)
    ra, rb = rot_x(0.0), rot_x(10.0)
    v = O.collider_velocity(geom, ra, rb, 0.5, [0.0, 0.0, 1.0], 1.0, 1 / 64)
    omega = math.radians(10.0)  # rad per frame about +x: the point at p moves with omega x p
    assert np.allclose(v, np.cross([omega, 0.0, 0.0], [0.0, 0.0, 1.0]), rtol=1e-5, atol=1e-8)


# ---- CFL, occupancy, removal ---------------------------------------------------------------------

def test_cfl_steps_known_answers():
    dt, steps = O.choose_dt(1 / 24, 0.0, 1 / 64, 8.0)
    assert steps == math.ceil((1 / 24) / (0.4 / 64 / 8.0) - 1e-9) and dt == pytest.approx((1 / 24) / steps)
    assert O.choose_dt(5e-5, 3.0, 1 / 128, 44.1) == (5e-5, 1)
    assert O.choose_dt(1 / 24, float("nan"), 1 / 64, 8.0) == O.choose_dt(1 / 24, 0.0, 1 / 64, 8.0)


def test_is_occupied_known_answer():
    d = Domain((0, 0, 0), (1, 1, 1), 32)
    o = O.OracleSolver(d, MATS[:1], (0, 0, 0))
    g = (np.stack(np.meshgrid(*[np.arange(0.4, 0.6, 1 / 64)] * 3, indexing="ij"), -1).reshape(-1, 3) + 1 / 128)
    o.add_particles(g.astype(np.float32), np.zeros_like(g, np.float32), np.zeros(len(g), np.int32),
                    np.zeros(len(g), np.int32))
    assert o.is_occupied(np.array([[0.5, 0.5, 0.5], [0.9, 0.9, 0.9], [-5, 0, 0]], np.float32)).tolist() == \
        [True, False, False]


def test_escaped_mask():
    d = Domain((0, 0, 0), (1, 1, 1), 10, walls={"z-": False})
    x = np.array([[0.5, 0.5, -0.05], [0.5, 0.5, -0.15], [0.5, 0.5, 1.5], [-0.5, 0.5, 0.5]], np.float32)
    assert O.escaped_mask(x, d).tolist() == [False, True, False, False]
