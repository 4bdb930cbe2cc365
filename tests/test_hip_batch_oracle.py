"""Batched and served results against the oracle (GPU).

* SceneBatch: a sample of the batched mix that covers every material model twice, every grid edge of the mix
  (32 to 128) and 0 to 2 colliders. From identical states (each scene first runs a frame alone so that C and
  F are not trivial; the oracle then loads that state), one substep of every scene in one batched launch
  matches the oracle's substep at the single-substep bar (x, v, F, Jp rel <= 1e-4; C rel <= 1e-4 with the
  floor of helpers.c_close for bodies that move nearly rigidly; v of particles lying on a collider surface,
  at most 2%, is left out: rounding decides whether they are projected), and a whole frame's
  CFL substep count and dt equal the oracle's exactly (solver.py:176-185, 213-220); the frame itself is then
  held to the multi-step criterion.
* Open sides: every frame of a batched scene losing particles takes the oracle's substep count for the state
  it starts from, i.e. its CFL step comes from the survivors of the previous frame's removal
  (solver.py:222-240).
* The bake server: frames of requests drawn like the served workload, baked in one continuously batched
  server on the GPU, against the runner (mpm_core.bake) on the oracle backend: the same frames and emitter
  particle counts, the first frame (emission only) identical, later frames within the multi-step criterion.
"""

from __future__ import annotations

import numpy as np
import pytest
from helpers import c_close, rel_err
from oracle_backend import OracleBackend

from mpm_core.backends import register_backend
from mpm_core.backends.hip.batch import SceneBatch
from mpm_plugin_amd.workloads import mix_scene, served_request
from oracle import oracle as O

pytestmark = pytest.mark.gpu

register_backend("oracle", lambda **kw: OracleBackend(**kw))

FRAME = 1.0 / 24.0
# (model, grid edge, colliders): every model twice, every edge, 0-2 colliders
SAMPLE = [("elastic", 32, 0), ("von_mises", 48, 1), ("drucker_prager", 64, 2), ("snow", 96, 1),
          ("elastic", 128, 2), ("snow", 32, 2), ("von_mises", 128, 0), ("drucker_prager", 48, 0)]


def sample_scenes(seed: int = 21):
    out = []
    for i, (model, edge, n_col) in enumerate(SAMPLE):
        w = mix_scene(i, seed, edges=(edge,), model=model, n_colliders=n_col, n_range=(5_000, 150_000))
        out.append(w)
    return out


def oracle_of(w, state) -> O.OracleSolver:
    o = O.OracleSolver(w.domain, w.materials, w.gravity)
    if w.colliders:
        o.set_colliders(w.colliders)
    o.load_state(state)
    return o


def set_poses(w, solvers, t0: float, t1: float) -> None:
    if w.colliders:
        a, b = w.collider_transforms(t0, t1)
        for s in solvers:
            s.set_collider_transforms(a, b)


def aggregates_close(a: dict, b: dict) -> None:
    assert len(a["x"]) == len(b["x"])
    span = max(float(np.ptp(b["x"], axis=0).max()), 1e-6)
    assert np.abs(a["x"].mean(0) - b["x"].mean(0)).max() <= 0.01 * span
    for f in (np.min, np.max):
        assert np.abs(f(a["x"], 0) - f(b["x"], 0)).max() <= 0.01 * span
    ke_a, ke_b = (a["v"].astype(np.float64) ** 2).sum(), (b["v"].astype(np.float64) ** 2).sum()
    assert abs(ke_a - ke_b) <= 0.01 * ke_b + 1e-12


@pytest.fixture(scope="module")
def warmed():
    """The sample's scenes after one frame each, run alone: (workload, HIP solver, state)."""
    out = []
    for w in sample_scenes():
        s = w.make_solver("hip", "cuda:0")
        set_poses(w, [s], 0.0, FRAME)
        s.advance(FRAME)
        out.append((w, s, s.save_state()))
    print("sample: " + ", ".join(f"{w.name} ({len(st['x'])} particles)" for w, _, st in out), flush=True)
    return out


def test_batched_substep_matches_oracle(warmed):
    scenes = [s for _, s, _ in warmed]
    oracles, dts = [], []
    for w, s, st in warmed:
        s.load_state(st)
        o = oracle_of(w, st)
        dt = 0.5 * o.choose_dt(FRAME)[0]  # below the CFL step: one substep
        set_poses(w, [s, o], FRAME, FRAME + dt)
        oracles.append(o)
        dts.append(dt)
    stats = SceneBatch(scenes).advance(dts)
    for (w, s, _), o, dt, st in zip(warmed, oracles, dts, stats):
        assert st.substeps == 1 and st.dt == dt
        assert o.advance(dt)[0] == 1
        a, b = s.save_state(), o.save_state()
        # A particle that ends on a collider's surface (projected there in an earlier substep) is projected
        # again or not by the sign of a rounding-sized distance (colliders.py:140-147): its velocity is compared
        # with the others' only when it lies clear of every surface on both sides.
        clear = np.ones(len(b["x"]), bool)
        if w.colliders:
            tol = 1e-4 * w.domain.dx
            clear = (np.abs(o.collider_phi(b["x"])) > tol) & (np.abs(o.collider_phi(a["x"])) > tol)
        assert clear.mean() > 0.98
        errs = {k: rel_err(a[k], b[k]) for k in ("x", "F", "Jp")}
        errs["v"] = rel_err(a["v"][clear], b["v"][clear])
        c_ok, c_frac = c_close(a["C"], b["C"], b["v"], w.domain.dx)
        print(w.name, {k: f"{e:.1e}" for k, e in errs.items()}, f"C {c_frac:.1e} of its bound,",
              f"{int((~clear).sum())} particles on a collider surface", flush=True)
        for k, e in errs.items():
            assert e <= 1e-4, (w.name, k)
        assert c_ok, (w.name, "C")


def test_batched_frame_step_counts_match_oracle(warmed):
    scenes = [s for _, s, _ in warmed]
    oracles = []
    for w, s, st in warmed:
        s.load_state(st)
        o = oracle_of(w, st)
        set_poses(w, [s, o], FRAME, 2 * FRAME)
        oracles.append(o)
    want = [o.choose_dt(FRAME) for o in oracles]
    stats = SceneBatch(scenes).advance(FRAME)
    print("substeps per scene:", [st.substeps for st in stats], flush=True)
    for (w, s, _), o, (dt, steps), st in zip(warmed, oracles, want, stats):
        assert (st.substeps, st.dt) == (steps, dt), w.name
        o.advance(FRAME)
        aggregates_close(s.save_state(), o.save_state())


def test_open_sides_step_counts_follow_the_survivors():
    # a lower side: at an upper open side the reference's position clamp (origin + (dims - 2) dx, solver.py:48)
    # is float32(max + dx), the removal threshold itself, so `x > max + dx` (solver.py:233-234) never holds there
    walls = {"x-": False}
    w = mix_scene(5, 11, model="elastic", n_colliders=0, walls=walls, n_range=(3_000, 20_000), edges=(32, 48))
    batched = w.make_solver("hip", "cuda:0")
    plain = w.make_solver("hip", "cuda:0")
    for s in (batched, plain):
        st = s.save_state()
        fast = st["x"][:, 0] < np.median(st["x"][:, 0])
        st["v"][:] = np.float32([-0.3, 0.0, 0.0])  # the slower half stays in the domain
        st["v"][fast] = np.float32([-8.0, 0.0, 0.0])
        s.load_state(st)
    batch = SceneBatch([batched])
    after_removal = 0  # frames that started right after a frame that removed particles
    last_removed = 0
    for frame in range(8):
        if batched.particle_count == 0:
            break
        after_removal += last_removed > 0
        o = O.OracleSolver(w.domain, w.materials, w.gravity)
        o.load_state(batched.save_state())
        dt, steps = o.choose_dt(FRAME)
        got = batch.advance(FRAME)[0]
        ref = plain.advance(FRAME)
        assert (got.substeps, got.dt) == (steps, dt), frame
        assert (ref.substeps, ref.dt) == (got.substeps, got.dt), frame
        last_removed = got.removed
    assert after_removal >= 1


def test_served_bakes_match_the_oracle_runner(tmp_path):
    from test_serve import frames_of, progress, write_cache

    from mpm_core.bake import bake
    from mpm_core.scene import read_scene
    from mpm_core.serve import BakeRequest, BakeServer

    server = BakeServer(chunk=3)
    handles = []
    for i in range(4):
        scene = served_request(i, seed=4, n_range=(3_000, 15_000), edges=(32, 48), max_frames=3)
        ref = write_cache(tmp_path, f"ref_{i}", scene)
        assert bake(ref, backend="oracle").status == "done"
        out = write_cache(tmp_path, f"srv_{i}", scene)
        handles.append(server.submit(BakeRequest(read_scene(out / "scene"), cache_dir=out, name=str(i))))
    while not server.idle:
        server.step()
    server.close()
    for h in handles:
        assert h.status == "done", h.message
        a = frames_of(tmp_path / f"ref_{h.request.name}")
        b = frames_of(tmp_path / f"srv_{h.request.name}")
        assert sorted(a) == sorted(b) and len(a) > 0
        for f in sorted(a):
            for emitter in a[f]:
                (pa, va), (pb, vb) = a[f][emitter], b[f][emitter]
                assert len(pa) == len(pb)
                if len(pa) == 0:
                    continue
                if f == min(a):
                    np.testing.assert_array_equal(pa, pb)
                    continue
                aggregates_close({"x": pb, "v": vb}, {"x": pa, "v": va})
        assert progress(tmp_path / f"srv_{h.request.name}")["status"] == "done"
