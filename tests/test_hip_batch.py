"""Batched scenes (SceneBatch) against the same scenes advanced alone.

A batch advances every member by one frame per call, each with its own CFL substep count
(solver.py:213-220), through one captured launch sequence. What a scene computes must not depend on its
company: the tests below run scenes alone (``Solver.advance``, and a batch of one) and in a mix and
compare them. One frame of an elastic scene is compared particle by particle (x, v and F within 1e-4
relative, the single-substep bar; only the order of floating-point atomic additions differs). C is left out
of that comparison: for a body moving nearly rigidly it is a small difference of grid velocities (here
about 1e-6 1/s), which a frame of reordered float sums changes by its own size; its single-substep parity
is covered by the parity tests. Multi-frame runs of
every model are compared as the reference's multi-step criterion does: identical substep counts and
particle counts (mass exact), centre of mass, bounding box and kinetic energy within 1%.
"""

from __future__ import annotations

import numpy as np
import pytest

from mpm_core.backends.hip.batch import SceneBatch
from mpm_plugin_amd import _native as N
from mpm_plugin_amd.workloads import MIX_MODELS, mix_scene

pytestmark = pytest.mark.gpu

SMALL = dict(n_range=(3_000, 20_000), edges=(32, 48))


def rel_err(a, b) -> float:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def make(idx: int, seed: int = 11, **kw):
    w = mix_scene(idx, seed, **{**SMALL, **kw})
    return w, w.make_solver("hip", "cuda:0")


def set_poses(w, s, frame: int) -> None:
    if w.colliders:
        a, b = w.collider_transforms(frame * w.frame_dt, (frame + 1) * w.frame_dt)
        s.set_collider_transforms(a, b)


def assert_aggregates_close(a: dict, b: dict) -> None:
    assert len(a["x"]) == len(b["x"])
    span = max(float(np.ptp(b["x"], axis=0).max()), 1e-6)
    assert np.abs(a["x"].mean(0) - b["x"].mean(0)).max() <= 0.01 * span
    for f in (np.min, np.max):
        assert np.abs(f(a["x"], 0) - f(b["x"], 0)).max() <= 0.01 * span
    ke_a, ke_b = (a["v"].astype(np.float64) ** 2).sum(), (b["v"].astype(np.float64) ** 2).sum()
    assert abs(ke_a - ke_b) <= 0.01 * ke_b + 1e-12


def test_batch_of_one_matches_advance():
    """One elastic scene with a collider: Solver.advance vs a batch of one, frame by frame."""
    w, plain = make(0, model="elastic", n_colliders=1)
    _, batched = make(0, model="elastic", n_colliders=1)
    batch = SceneBatch([batched])
    total = 0
    for f in range(3):
        set_poses(w, plain, f)
        set_poses(w, batched, f)
        a = plain.advance(w.frame_dt)
        (b,) = batch.advance(w.frame_dt)
        assert (a.substeps, a.removed) == (b.substeps, b.removed)
        assert a.dt == pytest.approx(b.dt, rel=1e-15)
        total += b.substeps
        sa, sb = plain.save_state(), batched.save_state()
        if f == 0:
            for k in ("x", "v", "F"):
                assert rel_err(sb[k], sa[k]) <= 1e-4, k
        assert_aggregates_close(sb, sa)
    assert batch.substeps_launched == total


def test_alone_and_in_the_mix_agree():
    """Eight scenes of every model, both grid sizes, 0-2 colliders: each advanced alone through
    Solver.advance and all together in one batch, three frames."""
    specs = [dict(model=m, n_colliders=c) for m, c in zip(MIX_MODELS * 2, (0, 1, 2, 0, 1, 2, 0, 1))]
    alone = [make(i, **kw) for i, kw in enumerate(specs)]
    mixed = [make(i, **kw) for i, kw in enumerate(specs)]
    batch = SceneBatch([s for _, s in mixed])
    counts_alone, counts_mixed = [], []
    for f in range(3):
        for (w, a), (_, m) in zip(alone, mixed):
            set_poses(w, a, f)
            set_poses(w, m, f)
        counts_alone.append([a.advance(w.frame_dt).substeps for w, a in alone])
        counts_mixed.append([st.substeps for st in batch.advance([w.frame_dt for w, _ in mixed])])
        if f == 0:
            for (w, a), (_, m) in zip(alone, mixed):
                if w.materials[0].model == "elastic":
                    sa, sm = a.save_state(), m.save_state()
                    assert rel_err(sm["x"], sa["x"]) <= 1e-4
                    assert rel_err(sm["v"], sa["v"]) <= 1e-4
    assert counts_alone == counts_mixed
    assert len(set(counts_mixed[0])) > 1, "the mix should hold different substep counts"
    for (w, a), (_, m) in zip(alone, mixed):
        assert_aggregates_close(m.save_state(), a.save_state())


def test_graph_replays_match_direct_launches():
    """The captured graphs, replayed many times with host work between frames, against the same launch
    sequence issued directly."""
    specs = [dict(model="elastic", n_colliders=1), dict(model="snow", n_colliders=0),
             dict(model="drucker_prager", n_colliders=2)]
    runs = []
    for graph in (1, 0):
        pairs = [make(i, **kw) for i, kw in enumerate(specs)]
        batch = SceneBatch([s for _, s in pairs])
        batch.set_option(N.BATCH_OPT_GRAPH, graph)
        assert batch.get_option(N.BATCH_OPT_GRAPH) == graph
        for f in range(2):
            for w, s in pairs:
                set_poses(w, s, f)
            counts = [st.substeps for st in batch.advance([w.frame_dt for w, _ in pairs])]
        runs.append(([s.save_state() for _, s in pairs], counts))
    (graph_states, graph_counts), (direct_states, direct_counts) = runs
    assert graph_counts == direct_counts
    elastic_g, elastic_d = graph_states[0], direct_states[0]
    assert rel_err(elastic_g["x"], elastic_d["x"]) <= 1e-4
    for a, b in zip(graph_states, direct_states):
        assert_aggregates_close(a, b)


def test_timed_substeps_match_graph_replays():
    """With kernel timing on the batch launches its substeps directly: same substep counts and states as
    the graph replays, and every kernel class timed once per substep (the binning once per sorting
    substep)."""
    specs = [dict(model="elastic", n_colliders=1), dict(model="snow", n_colliders=0),
             dict(model="von_mises", n_colliders=2)]
    runs = []
    for timed in (False, True):
        pairs = [make(i, **kw) for i, kw in enumerate(specs)]
        batch = SceneBatch([s for _, s in pairs])
        batch.set_timing(timed)
        before = batch.substeps_launched
        for f in range(2):
            for w, s in pairs:
                set_poses(w, s, f)
            counts = [st.substeps for st in batch.advance([w.frame_dt for w, _ in pairs])]
        launched = batch.substeps_launched - before
        times = batch.kernel_times()
        runs.append(([s.save_state() for _, s in pairs], counts, launched, times))
    (g_states, g_counts, g_launched, g_times), (t_states, t_counts, t_launched, t_times) = runs
    assert g_counts == t_counts and g_launched == t_launched
    assert all(n == 0 for _, n in g_times.values())  # nothing is timed while timing is off
    print({k: (round(ms, 3), n) for k, (ms, n) in t_times.items()}, flush=True)
    for name in ("plan", "drift", "p2g", "grid_scan", "grid", "g2p", "fixup", "advance"):
        assert t_times[name][1] == t_launched and t_times[name][0] > 0, name
    assert 0 < t_times["sort"][1] < t_launched
    assert rel_err(g_states[0]["x"], t_states[0]["x"]) <= 1e-4
    for a, b in zip(g_states, t_states):
        assert_aggregates_close(a, b)


def test_members_change_between_frames():
    """Emission into a member (capacity growth), a skipped member (frame_dt 0), a member leaving, and a
    member closed while in the batch."""
    pairs = [make(i, model=m, n_colliders=0) for i, m in enumerate(("elastic", "snow", "von_mises"))]
    batch = SceneBatch([s for _, s in pairs])
    batch.advance(1 / 24)
    w0, s0 = pairs[0]
    before = s0.save_state()
    n_add = 5000
    rng = np.random.default_rng(3)
    x = rng.uniform(0.3, 0.5, (n_add, 3)).astype(np.float32)
    s0.add_particles(x, np.zeros_like(x), np.zeros(n_add, np.int32), np.ones(n_add, np.int32))
    skipped = pairs[1][1].save_state()
    stats = batch.advance([1 / 24, 0.0, 1 / 24])
    assert stats[1].substeps == 0
    np.testing.assert_array_equal(pairs[1][1].save_state()["x"], skipped["x"])
    assert s0.particle_count == len(before["x"]) + n_add
    assert np.isfinite(s0.save_state()["x"]).all()
    batch.set_scenes([pairs[0][1], pairs[2][1]])
    batch.advance(1 / 24)
    pairs[1][1].advance(1 / 24)  # the scene that left runs alone again
    pairs[2][1].close()
    with pytest.raises(ValueError):
        batch.advance(1 / 24)
    batch.set_scenes([pairs[0][1]])
    assert batch.advance(1 / 24)[0].substeps > 0


def test_membership_errors():
    pairs = [make(i, model="elastic", n_colliders=0) for i in range(2)]
    a, b = pairs[0][1], pairs[1][1]
    batch = SceneBatch([a])
    with pytest.raises(ValueError):
        batch.set_scenes([b, b])
    other = SceneBatch()
    with pytest.raises(ValueError):
        other.set_scenes([a])
    with pytest.raises(ValueError):
        batch.set_option(N.BATCH_OPT_P2G_BLOCKS, -1)
    a.set_option(N.OPT_P2G_MODE, N.P2G_ATOMIC)
    with pytest.raises(ValueError):
        batch.advance(1 / 24)
    a.set_option(N.OPT_P2G_MODE, N.P2G_BINNED)
    assert batch.advance(1 / 24)[0].substeps > 0


def test_open_walls_remove_like_advance():
    """Particles leaving through open sides are removed after the frame, as Solver.advance removes them."""
    walls = {"y-": False}
    w, plain = make(5, model="elastic", n_colliders=0, walls=walls)
    _, batched = make(5, model="elastic", n_colliders=0, walls=walls)
    v = np.tile(np.array([0.0, -6.0, 0.0], np.float32), (w.n, 1))
    for s in (plain, batched):
        st = s.save_state()
        st["v"] = v
        s.load_state(st)
    batch = SceneBatch([batched])
    removed_plain = removed_batch = 0
    for frame in range(6):
        a = plain.advance(w.frame_dt)
        b = batch.advance(w.frame_dt)[0]
        # after a removal both take the next CFL step from the survivors (solver.py:176-185, 222-240)
        assert (b.substeps, b.dt) == (a.substeps, a.dt), frame
        removed_plain += a.removed
        removed_batch += b.removed
    assert removed_plain > 0
    assert abs(removed_plain - removed_batch) <= max(2, 0.01 * removed_plain)


def test_frames_start_and_end_independently():
    """Continuous batching: every scene starts its next frame as soon as its last one ended, while the
    others are in the middle of theirs; each scene's frames match Solver.advance."""
    specs = [dict(model="elastic", n_colliders=1), dict(model="snow", n_colliders=0),
             dict(model="von_mises", n_colliders=2)]
    alone = [make(i, **kw) for i, kw in enumerate(specs)]
    mixed = [make(i, **kw) for i, kw in enumerate(specs)]
    frames = 3
    counts_alone = [[] for _ in specs]
    for i, (w, s) in enumerate(alone):
        for f in range(frames):
            set_poses(w, s, f)
            counts_alone[i].append(s.advance(w.frame_dt).substeps)
    batch = SceneBatch([s for _, s in mixed])
    next_frame = [0] * len(specs)
    counts_mixed = [[] for _ in specs]
    fdt = np.zeros(len(specs))
    for i, (w, s) in enumerate(mixed):
        set_poses(w, s, 0)
        fdt[i] = w.frame_dt
    batch.begin(fdt)
    with pytest.raises(ValueError):
        mixed[0][1].save_state()  # in the middle of a batched frame
    overlapped = False
    while any(n < frames for n in next_frame):
        done, total = batch.progress()
        _, ended = batch.run(-1)
        done2, total2 = batch.progress()
        if len(ended) < sum(1 for i in range(len(specs)) if done[i] < total[i]):
            overlapped = True
        fdt = np.zeros(len(specs))
        for i, st in ended.items():
            counts_mixed[i].append(st.substeps)
            next_frame[i] += 1
            if next_frame[i] < frames:
                w, s = mixed[i]
                set_poses(w, s, next_frame[i])
                fdt[i] = w.frame_dt
        if fdt.any():
            batch.begin(fdt)
    assert overlapped, "some frames should have ended while others were in flight"
    assert counts_mixed == counts_alone
    for (w, a), (_, m) in zip(alone, mixed):
        assert_aggregates_close(m.save_state(), a.save_state())
