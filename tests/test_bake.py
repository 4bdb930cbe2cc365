"""The bake runner (mpm_core.bake) on the CPU with a recording ballistic backend, plus two runs on the HIP
backend (marked gpu).

Covered, after the reference runner's behaviour (src/mpm_core/bake.py:56-244; its tests in
tests/test_bake.py): the frame timeline and the solver calls it makes, complete / cancelled / unstable bakes,
resume and its refusals, the cancel file, progress.json, the lock, emission warnings, object velocity,
inflow substeps, stale folders, and the command line.
"""

import json
import subprocess
import sys
from fractions import Fraction

import numpy as np
import pytest
from helpers import scale, small_scene, translate

from mpm_core.backends import create_solver, register_backend
from mpm_core.bake import (Piece, SceneChangedError, bake, first_frame_emitters, frame_pieces, main, pose_at)
from mpm_core.cache import CacheReader, exclusive_lock, read_checkpoint
from mpm_core.geometry import interpolate_transform
from mpm_core.scene import write_scene
from mpm_core.solver import AdvanceStats, ParticleData, Solver


class RecordingSolver(Solver):
    """Ballistic particles without a grid (v += g dt, x += v dt per advance) that logs every call the runner
    makes; with ``nan_on`` set, that advance (1-based) turns every position into NaN."""

    log: list = []  # the last instance's calls, for the test that made it

    def __init__(self, nan_on=None, **kwargs):
        super().__init__(**kwargs)
        self.nan_on = nan_on
        self.calls = RecordingSolver.log = []
        self.x = np.zeros((0, 3), np.float32)
        self.v = np.zeros((0, 3), np.float32)
        self.mat = np.zeros(0, np.int32)
        self.emitter = np.zeros(0, np.int32)

    def add_particles(self, positions, velocities, material_ids, emitter_ids):
        self.calls.append(("add", int(emitter_ids[0]), len(positions)))
        self.x = np.concatenate([self.x, np.asarray(positions, np.float32)])
        self.v = np.concatenate([self.v, np.asarray(velocities, np.float32)])
        self.mat = np.concatenate([self.mat, np.asarray(material_ids, np.int32)])
        self.emitter = np.concatenate([self.emitter, np.asarray(emitter_ids, np.int32)])

    def set_colliders(self, colliders):
        self.calls.append(("colliders", len(colliders)))

    def set_collider_transforms(self, start, end):
        self.calls.append(("poses", np.array(start), np.array(end)))

    def advance(self, frame_dt):
        self.calls.append(("advance", frame_dt))
        dt = np.float32(frame_dt)
        self.v = self.v + dt * np.asarray(self.gravity, np.float32)
        self.x = self.x + dt * self.v
        if sum(c[0] == "advance" for c in self.calls) == self.nan_on:
            self.x[:] = np.nan
        return AdvanceStats(1, frame_dt, 0)

    @property
    def particle_count(self):
        return len(self.x)

    def get_particles(self):
        return ParticleData(self.x.copy(), self.v.copy(), self.mat.copy(), self.emitter.copy())

    def is_occupied(self, points):
        return np.zeros(len(points), bool)

    def save_state(self):
        return {"x": self.x.copy(), "v": self.v.copy(), "mat": self.mat.copy(), "emitter": self.emitter.copy()}

    def load_state(self, state):
        self.x, self.v = np.asarray(state["x"]), np.asarray(state["v"])
        self.mat, self.emitter = np.asarray(state["mat"]), np.asarray(state["emitter"])


register_backend("recording", lambda **kw: RecordingSolver(**kw))
register_backend("recording_nan3", lambda **kw: RecordingSolver(nan_on=3, **kw))


def cache_with(tmp_path, scene):
    write_scene(scene, tmp_path / "scene")
    return tmp_path


def progress_of(cache):
    return json.loads((cache / "progress.json").read_text())


def cancel_after(frames: int):
    """A should_cancel callback that asks to stop once ``frames`` frames have been written."""
    asked = [0]

    def should_cancel():
        asked[0] += 1
        return asked[0] >= frames

    return should_cancel


# ---- the timeline ------------------------------------------------------------------------------------------

def two_inflows():
    """emitter 0: inflow over frames 3..5, 2 per frame; emitter 1: over 1..6, 3 per frame."""
    scene = small_scene(kind="inflow", emit_frames=(3, 5), frame_end=6)
    other = small_scene(kind="inflow", emit_frames=(1, 6), frame_end=6).emitters[0]
    scene.emitters[0].substeps, other.substeps = 2, 3
    scene.emitters.append(other)
    return scene


@pytest.mark.parametrize("frame,want", [
    (2, [(0, Fraction(1, 3), (1,)), (Fraction(1, 3), Fraction(2, 3), (1,)), (Fraction(2, 3), 1, (1,))]),
    # emitter 0's first frame: it emits once, at the frame itself, after emitter 1's last piece
    (3, [(0, Fraction(1, 3), (1,)), (Fraction(1, 3), Fraction(2, 3), (1,)), (Fraction(2, 3), 1, (0, 1))]),
    (4, [(0, Fraction(1, 3), (1,)), (Fraction(1, 3), Fraction(1, 2), (0,)), (Fraction(1, 2), Fraction(2, 3), (1,)),
         (Fraction(2, 3), 1, (0, 1))]),
])
def test_frame_pieces_of_two_inflows(frame, want):
    assert frame_pieces(two_inflows(), frame) == [Piece(Fraction(a), Fraction(b), e) for a, b, e in want]


def test_frame_pieces_without_emission_is_one_piece():
    scene = two_inflows()
    scene.emitters.pop()
    assert frame_pieces(scene, 6) == [Piece(Fraction(0), Fraction(1), ())]


def test_frame_pieces_tile_the_frame():
    rng = np.random.default_rng(5)
    for _ in range(50):
        scene = small_scene(kind="inflow", emit_frames=(1, 8), frame_end=8)
        scene.emitters = []
        for _ in range(int(rng.integers(1, 5))):
            start = int(rng.integers(1, 8))
            e = small_scene(kind="inflow", emit_frames=(start, int(rng.integers(start, 9))), frame_end=8).emitters[0]
            e.substeps = int(rng.integers(1, 17))
            scene.emitters.append(e)
        for frame in range(2, 9):
            pieces = frame_pieces(scene, frame)
            assert pieces[0].start == 0 and pieces[-1].end == 1
            assert all(p.start < p.end for p in pieces)
            assert all(a.end == b.start for a, b in zip(pieces, pieces[1:]))
            # every emission due in this frame happens exactly once
            due = sum((e.substeps if frame > e.frame_start else 1)
                      for e in scene.emitters if e.frame_start <= frame <= e.frame_end)
            assert sum(len(p.emitters) for p in pieces) == due


def test_first_frame_emitters():
    scene = two_inflows()  # emitter 0 starts at frame 3, emitter 1 at frame 1
    scene.emitters.append(small_scene(frame_end=6).emitters[0])  # a fill emitter
    assert first_frame_emitters(scene) == [1, 2]


def test_runner_calls_follow_the_timeline(tmp_path):
    """The solver calls of a bake are the script the timeline implies: the colliders once, the first frame's
    emissions, then per piece of every later frame the collider poses interpolated to the piece's ends, one
    advance over the piece and the emissions due at its end."""
    scene = two_inflows()
    moving = np.stack([translate(0.5, 0.5, 0.05 + 0.02 * k) for k in range(scene.frame_count)])
    scene.colliders[0].transforms = moving
    bake(cache_with(tmp_path, scene), backend="recording")
    want = [("colliders", 1)] + [("add", i) for i in first_frame_emitters(scene)]
    for frame in range(scene.frame_start + 1, scene.frame_end + 1):
        k = frame - scene.frame_start
        for p in frame_pieces(scene, frame):
            start = interpolate_transform(moving[k - 1], moving[k], float(p.start))
            want += [("poses", start, pose_at(moving, k, p.end)), ("advance", float(p.end - p.start) / scene.fps)]
            want += [("add", i) for i in p.emitters]
    got = RecordingSolver.log
    assert [c[0] for c in got] == [c[0] for c in want]
    for g, w in zip(got, want):
        if g[0] == "poses":
            np.testing.assert_array_equal(g[1][0], w[1])
            np.testing.assert_array_equal(g[2][0], w[2])
        elif g[0] == "add":
            assert g[1] == w[1] and g[2] > 0
        else:
            assert g == w
    frame_time = sum(c[1] for c in got if c[0] == "advance") / (scene.frame_count - 1)
    assert frame_time == pytest.approx(1.0 / scene.fps, rel=1e-12)


# ---- whole bakes -------------------------------------------------------------------------------------------

def test_unknown_backend_names_the_known_ones():
    with pytest.raises(KeyError, match="hip"):
        create_solver("no_such_backend", domain=None, materials=[], gravity=(0, 0, 0))


def test_complete_bake(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=5))
    result = bake(cache, backend="recording")
    assert (result.status, result.frames_written, result.message) == ("done", 5, "")
    reader = CacheReader(cache)
    assert reader.available_frames() == [1, 2, 3, 4, 5]
    assert (reader.frame_start, reader.frame_end, reader.emitter_names) == (1, 5, ["emitter_0"])
    doc = progress_of(cache)
    assert (doc["status"], doc["frame"], doc["frame_end"]) == ("done", 5, 5)
    assert doc["particles"] == len(reader.read_frame(5)["emitter_0"][0]) > 0
    assert read_checkpoint(cache)[0] == 5


def test_unstable_frame_stops_the_bake(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=6))
    result = bake(cache, backend="recording_nan3")  # the third advance produces frame 4
    assert result.status == "error" and "unstable at frame 4" in result.message
    assert CacheReader(cache).available_frames() == [1, 2, 3]
    assert (progress_of(cache)["status"], progress_of(cache)["frame"]) == ("error", 4)


def test_cancel_then_resume_equals_an_uninterrupted_bake(tmp_path):
    whole = cache_with(tmp_path / "whole", small_scene(frame_end=7))
    parts = cache_with(tmp_path / "parts", small_scene(frame_end=7))
    bake(whole, backend="recording")
    first = bake(parts, backend="recording", should_cancel=cancel_after(3))
    assert (first.status, first.frames_written) == ("cancelled", 3)
    assert progress_of(parts)["status"] == "cancelled"
    rest = bake(parts, backend="recording", resume=True)
    assert (rest.status, rest.frames_written) == ("done", 4)
    a, b = CacheReader(whole), CacheReader(parts)
    assert a.available_frames() == b.available_frames() == list(range(1, 8))
    for f in (3, 4, 7):
        np.testing.assert_array_equal(a.read_frame(f)["emitter_0"][0], b.read_frame(f)["emitter_0"][0])


def test_cancel_file_stops_after_the_frame_and_is_consumed(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=6))
    asked = [0]

    def add_on_writes_cancel_file():
        asked[0] += 1
        if asked[0] == 2:
            (cache / "cancel").touch()
        return False

    result = bake(cache, backend="recording", should_cancel=add_on_writes_cancel_file)
    # the file, written during the check after frame 2 (which the callback answers with no), is found by the
    # check after frame 3: the runner looks for the file before it asks the callback
    assert (result.status, result.frames_written) == ("cancelled", 3)
    assert not (cache / "cancel").exists()


def test_leftover_cancel_file_is_ignored(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=3))
    (cache / "cancel").touch()
    assert bake(cache, backend="recording").status == "done"


def test_progress_reads_starting_while_the_solver_is_made(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=2))
    bake(cache, backend="recording")
    assert progress_of(cache)["status"] == "done"
    seen = []

    def make(**kw):
        seen.append(progress_of(cache)["status"])
        return RecordingSolver(**kw)

    register_backend("recording_peek", make)
    bake(cache, backend="recording_peek")
    assert seen == ["starting"]


def test_unreadable_scene_is_an_error_in_progress(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=2))
    (cache / "scene" / "scene.json").write_text("{ truncated")
    with pytest.raises(ValueError):
        bake(cache, backend="recording")
    doc = progress_of(cache)
    assert doc["status"] == "error" and doc["message"].startswith("JSONDecodeError")


def test_open_emitter_mesh_is_refused_before_any_frame(tmp_path):
    scene = small_scene(kind="inflow", emit_frames=(2, 3), frame_end=4)
    scene.emitters[0].mesh.tris = scene.emitters[0].mesh.tris[2:]
    cache = cache_with(tmp_path, scene)
    with pytest.raises(ValueError, match="emitter 'emitter_0': mesh is not closed"):
        bake(cache, backend="recording")
    assert not list(cache.glob("frames/*.npz"))


@pytest.mark.parametrize("case", ["smaller_than_spacing", "partly_outside"])
def test_emission_warnings(tmp_path, case):
    scene = small_scene(frame_end=2, emitter_offset=(0.5, 0.0, 0.0) if case == "partly_outside" else (0, 0, 0))
    if case == "smaller_than_spacing":
        scene.emitters[0].transforms[:] = translate(0.5, 0.5, 0.5) @ scale(0.002)
    cache = cache_with(tmp_path, scene)
    bake(cache, backend="recording")
    warnings = progress_of(cache)["warnings"]
    pos = CacheReader(cache).read_frame(1)["emitter_0"][0]
    if case == "smaller_than_spacing":
        assert warnings == ["emitter_0: no particles, the mesh is smaller than the particle spacing (0.01562 m)"]
        assert len(pos) == 0
    else:
        assert warnings == ["emitter_0: particles outside the domain were dropped"]
        assert len(pos) > 0 and pos[:, 0].max() <= 1.0


def test_object_velocity_of_a_fill_emitter(tmp_path):
    scene = small_scene(frame_end=3)
    e = scene.emitters[0]
    e.use_object_velocity = True
    e.transforms = np.stack([translate(0.05 * k, 0.0, 0.0) @ e.transforms[0] for k in range(3)])
    cache = cache_with(tmp_path, scene)
    bake(cache, backend="recording")
    vel = CacheReader(cache).read_frame(1)["emitter_0"][1]
    # 0.05 m per frame at 24 frames/s, a forward difference on the first frame
    np.testing.assert_allclose(vel[:, 0], 1.2, atol=1e-4)
    np.testing.assert_allclose(vel[:, 1:], 0.0, atol=1e-5)


@pytest.mark.parametrize("trouble", ["changed_scene", "no_checkpoint"])
def test_resume_refusals(tmp_path, trouble):
    cache = cache_with(tmp_path, small_scene(frame_end=5))
    if trouble == "changed_scene":
        bake(cache, backend="recording", should_cancel=cancel_after(1))
        write_scene(small_scene(frame_end=5, gravity=(0, 0, -1)), cache / "scene")
        with pytest.raises(SceneChangedError, match="bake from scratch"):
            bake(cache, backend="recording", resume=True)
    else:
        with pytest.raises(FileNotFoundError, match="no checkpoint to resume from"):
            bake(cache, backend="recording", resume=True)
    assert progress_of(cache)["status"] == "error"


def test_scene_without_emitters(tmp_path):
    scene = small_scene(frame_end=3)
    scene.emitters = []
    cache = cache_with(tmp_path, scene)
    assert bake(cache, backend="recording").status == "done"
    reader = CacheReader(cache)
    assert reader.available_frames() == [1, 2, 3] and reader.read_frame(3) == {}


@pytest.mark.parametrize("resume", [False, True])
def test_surface_folder_is_cleared_by_a_fresh_bake_only(tmp_path, resume):
    cache = cache_with(tmp_path, small_scene(frame_end=4))
    bake(cache, backend="recording", should_cancel=cancel_after(1))
    marker = cache / "surface" / "meshes" / "frame_0001.npz"
    marker.parent.mkdir(parents=True)
    marker.write_bytes(b"mesh of the earlier simulation")
    bake(cache, backend="recording", resume=resume)
    assert marker.exists() == resume


def test_a_second_bake_on_the_cache_is_refused(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=2))
    with exclusive_lock(cache / "bake.lock"):
        with pytest.raises(RuntimeError, match="a bake is already running on"):
            bake(cache, backend="recording")


@pytest.mark.parametrize("substeps", [1, 6])
def test_inflow_substeps_fill_the_emitter_path(tmp_path, substeps):
    """An inflow emitter moving 0.3 m per frame: one emission per frame leaves gaps along its path, six do
    not (each emission at its own interpolated pose)."""
    scene = small_scene(kind="inflow", emit_frames=(1, 3), frame_end=3, gravity=(0, 0, 0))
    scene.emitters[0].transforms = np.stack([translate(0.1 + 0.3 * k, 0.45, 0.45) @ scale(0.1) for k in range(3)])
    scene.emitters[0].substeps = substeps
    cache = cache_with(tmp_path, scene)
    bake(cache, backend="recording")
    x = CacheReader(cache).read_frame(3)["emitter_0"][0][:, 0]
    filled, _ = np.histogram(x, bins=np.arange(0.1, 0.8, 0.02))
    assert bool((filled > 0).all()) == (substeps == 6)


def test_command_line(tmp_path, capsys):
    cache = cache_with(tmp_path, small_scene(frame_end=2))
    assert main([str(cache), "--backend", "recording"]) == 0
    out = capsys.readouterr().out.splitlines()
    assert out[-1] == "done: 2 frames written."
    assert all(json.loads(line)["status"] in ("starting", "running") for line in out[:-1])
    (cache / "scene" / "scene.json").unlink()
    assert main([str(cache), "--backend", "recording"]) == 1


# ---- on the HIP backend ----------------------------------------------------------------------------------
@pytest.mark.gpu
def test_inflow_emits_while_active(tmp_path):
    cache = cache_with(tmp_path, small_scene(kind="inflow", velocity=(0, 0, -1), emit_frames=(1, 10), frame_end=20))
    assert bake(cache).status == "done"
    reader = CacheReader(cache)
    counts = [len(reader.read_frame(f)["emitter_0"][0]) for f in range(1, 21)]
    assert counts[9] > counts[1] and counts[10:] == [counts[10]] * 10


@pytest.mark.gpu
def test_cli_on_the_hip_backend(tmp_path):
    cache = cache_with(tmp_path, small_scene(frame_end=3))
    out = subprocess.run([sys.executable, "-m", "mpm_core.bake", str(cache)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert CacheReader(cache).available_frames() == [1, 2, 3]
