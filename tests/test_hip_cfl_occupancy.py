"""Integer results of the path against the oracle, bit for bit (GPU).

* The CFL step count and dt of Solver.advance (src/mpm_core/backends/warp/solver.py:176-185; max_speed
  kernels.py:266-269): from identical states, AdvanceStats.substeps equals the oracle's step count and dt
  equals frame_dt / steps, frame after frame, on a snow scene (its hardened wave speed forces hundreds of
  substeps per frame), a fast-moving scene and a scene whose speed grows under gravity.
* is_occupied (solver.py:244-255; kernels.py:272-298), which decides the inflow emission counts
  (bake.py:75-76): the mask over an emitter's whole lattice equals the oracle's. The volume sums are float
  atomics in any order on the GPU and sequential on the CPU, so a point may only differ where its node's
  volume lies within rounding of the half-cell threshold; the test computes that volume in FP64 and allows
  exactly those (none occur on these inputs in practice).
"""

from __future__ import annotations

import numpy as np
import pytest
from helpers import box_particles, icosphere, make_solver, scale, sphere_particles, translate

from mpm_core.geometry import sample_mesh_volume, transform_points
from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SNOW = MaterialParams("snow", "snow", 1.4e5, 0.2, 400.0, hardening=10.0, critical_compression=2.5e-2,
                      critical_stretch=7.5e-3)
JELLY = MaterialParams("jelly", "elastic", 5.0e4, 0.3, 1000.0)
SAND = MaterialParams("sand", "drucker_prager", 3.5e4, 0.3, 1600.0, friction_angle_deg=30.0)


def scene(kind: str):
    d = Domain((0, 0, 0), (1, 1, 1), 64)
    if kind == "snow":
        x = sphere_particles((0.5, 0.5, 0.4), 0.15, d.dx, seed=1)
        return d, [SNOW], x, np.tile(np.float32([-3.0, 0.0, 0.0]), (len(x), 1)), np.zeros(len(x), np.int32)
    if kind == "fast":
        x = box_particles((0.3, 0.3, 0.3), (0.5, 0.5, 0.5), d.dx, seed=2)
        rng = np.random.default_rng(2)
        v = (rng.normal(size=(len(x), 3)) * 4.0 + [9.0, 0.0, 0.0]).astype(np.float32)
        return d, [JELLY], x, v, np.zeros(len(x), np.int32)
    # falling mix: speed grows frame after frame, two materials present
    box = box_particles((0.2, 0.2, 0.5), (0.45, 0.45, 0.8), d.dx, seed=3)
    ball = sphere_particles((0.7, 0.6, 0.6), 0.12, d.dx, seed=4)
    x = np.concatenate([box, ball])
    mat = np.concatenate([np.zeros(len(box), np.int32), np.ones(len(ball), np.int32)])
    return d, [JELLY, SAND], x, np.zeros_like(x), mat


@pytest.mark.parametrize("kind,frames", [("snow", 2), ("fast", 3), ("falling", 6)])
def test_step_count_matches_oracle(kind, frames):
    d, mats, x, v, mat = scene(kind)
    hip = make_solver("cuda:0", d, mats)
    hip.add_particles(x, v, mat, mat)
    ora = O.OracleSolver(d, mats, (0.0, 0.0, -9.81))
    frame_dt = 1.0 / 24.0
    counts = []
    for _ in range(frames):
        ora.load_state(hip.save_state())  # identical states at the start of every frame
        want_dt, want_steps = ora.choose_dt(frame_dt)
        st = hip.advance(frame_dt)
        counts.append(st.substeps)
        assert st.substeps == want_steps
        assert st.dt == want_dt == frame_dt / want_steps
        assert st.removed == 0
    print(f"{kind}: substeps per frame {counts}", flush=True)
    assert max(counts) > 1
    if kind == "snow":
        assert min(counts) > 100


def volume_f64(d: Domain, x: np.ndarray) -> np.ndarray:
    """The is_occupied volume grid in FP64 (scatter of p_vol with quadratic B-spline weights)."""
    dims = tuple(s + 6 for s in d.grid_shape)
    origin = np.asarray(d.min, np.float64) - 3 * d.dx
    xp = (x.astype(np.float64) - origin) / d.dx
    base = np.floor(xp - 0.5).astype(np.int64)
    fx = xp - base
    w = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2]
    vol = np.zeros(dims)
    p_vol = d.dx ** 3 / 8
    for i in range(3):
        for j in range(3):
            for k in range(3):
                np.add.at(vol, (base[:, 0] + i, base[:, 1] + j, base[:, 2] + k),
                          w[i][:, 0] * w[j][:, 1] * w[k][:, 2] * p_vol)
    return vol


def borderline(d: Domain, x: np.ndarray, pts: np.ndarray) -> np.ndarray:
    vol = volume_f64(d, x)
    origin = np.asarray(d.min, np.float64) - 3 * d.dx
    g = np.rint((pts.astype(np.float64) - origin) / d.dx).astype(np.int64)
    g = np.clip(g, 0, np.array(vol.shape) - 1)
    thr = 0.5 * d.dx ** 3
    return np.abs(vol[g[:, 0], g[:, 1], g[:, 2]] - thr) <= 1e-5 * thr


@pytest.mark.parametrize("res", [32, 64])
def test_is_occupied_mask_matches_oracle(res):
    """An inflow emitter's lattice overlapping material already emitted: the occupancy mask that filters it."""
    d = Domain((0, 0, 0), (1, 1, 1), res)
    hip = make_solver("cuda:0", d, [JELLY])
    body = sphere_particles((0.5, 0.5, 0.45), 0.2, d.dx, seed=5)
    hip.add_particles(body, np.zeros_like(body), np.zeros(len(body), np.int32), np.zeros(len(body), np.int32))
    for _ in range(3):
        hip.advance(1 / 48)  # let the body deform so the volumes are not a clean lattice
    state = hip.save_state()
    ora = O.OracleSolver(d, [JELLY], (0.0, 0.0, -9.81))
    ora.load_state(state)
    v, t = icosphere(1.0, 3)
    xf = translate(0.55, 0.5, 0.5) @ scale(0.25)
    pts = sample_mesh_volume(transform_points(xf, v), t, d.dx / 2, [7, 1, 0, 1, 1])
    lattice = np.stack(np.meshgrid(*[np.arange(0.0, 1.0, d.dx / 3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    for q in (pts, lattice.astype(np.float32)):
        got, want = hip.is_occupied(q), ora.is_occupied(q)
        diff = got != want
        print(f"res {res}: {len(q)} points, {int(want.sum())} occupied, {int(diff.sum())} differ", flush=True)
        assert 0 < want.sum() < len(q)
        assert np.all(borderline(d, state["x"], q[diff]))
    # emission counts of the inflow filter (bake.py:75-76) are then equal
    assert int((~hip.is_occupied(pts)).sum()) == int((~ora.is_occupied(pts)).sum())
