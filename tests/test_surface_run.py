"""The surface runner (reference tests/test_surface_run.py; src/mpm_surface/run.py, __main__.py).

Every runner test runs twice: with the HIP surfacer on the GPU, and on the CPU with the mesher replaced by
a stub (a small closed tetrahedron at the particles' mean), which checks the runner's own logic (skipping
by key, pruning, cancel, lock, warnings) without a device.
"""

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
from helpers import small_scene, sphere_particles

from mpm_core.cache import CacheWriter, exclusive_lock, frame_path
from mpm_core.geometry import mesh_is_closed
from mpm_core.scene import write_scene
from mpm_core.solver import ParticleData
from mpm_surface import run as run_mod
from mpm_surface.meshcache import MeshCacheReader, SurfaceMesh
from mpm_surface.run import surface
from mpm_surface.settings import SurfaceSettings

DX = 1 / 32
ROOT = Path(__file__).resolve().parent.parent


def stub_mesh(positions, velocities, spacing, settings, device=None, max_bytes=None):
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    if len(p) == 0:
        return SurfaceMesh.empty()
    verts = (p.mean(0) + 0.01 * np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])).astype(np.float32)
    tris = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int32)
    return SurfaceMesh(verts, tris, np.zeros_like(verts))


@pytest.fixture(params=["stub", pytest.param("hip", marks=pytest.mark.gpu)])
def mesher(request, monkeypatch):
    if request.param == "stub":
        monkeypatch.setattr(run_mod, "surface_particles", stub_mesh)
    return request.param


def write_particle_frame(writer, frame, empty=False):
    p = sphere_particles((0.4 + 0.01 * (frame - 1), 0.5, 0.5), 0.1, DX)
    if empty:
        p = p[:0]
    n = len(p)
    writer.write_frame(frame, ParticleData(p, np.tile([0.24, 0, 0], (n, 1)).astype(np.float32),
                                           np.zeros(n, np.int32), np.zeros(n, np.int32)))


@pytest.fixture
def particle_cache(tmp_path, mesher):
    """A 4-frame particle cache of emitter_0: a sphere moving +x 1 cm per frame (no solver involved)."""
    write_scene(small_scene(frame_end=4), tmp_path / "scene")
    writer = CacheWriter(tmp_path, ["emitter_0"], 24, 1, 4, DX, "test")
    for f in range(1, 5):
        write_particle_frame(writer, f)
    return tmp_path


def progress(cache):
    return json.loads((cache / "surface" / "progress.json").read_text())


def test_meshes_all_frames(particle_cache):
    res = surface(particle_cache)
    assert res.status == "done" and res.frames_written == 4
    r = MeshCacheReader(particle_cache)
    assert r.available_frames("emitter_0") == [1, 2, 3, 4]
    assert all(mesh_is_closed(r.read("emitter_0", f).tris) for f in range(1, 5))
    doc = progress(particle_cache)
    assert doc["status"] == "done" and doc["frame"] == 4 and doc["frame_end"] == 4 and doc["particles"] > 0


def test_second_run_skips(particle_cache):
    surface(particle_cache)
    res = surface(particle_cache)
    assert (res.frames_written, res.frames_skipped) == (0, 4)


def test_changed_settings_remesh(particle_cache):
    surface(particle_cache)
    assert surface(particle_cache, settings={"emitter_0": SurfaceSettings(detail=2)}).frames_written == 4


def test_frames_subset(particle_cache):
    res = surface(particle_cache, frames=[2, 3, 9])
    assert res.frames_written == 2
    assert MeshCacheReader(particle_cache).available_frames("emitter_0") == [2, 3]


def test_disabled_emitter_removed(particle_cache):
    surface(particle_cache)
    surface(particle_cache, settings={"emitter_0": SurfaceSettings(enabled=False)})
    assert MeshCacheReader(particle_cache).emitters() == []


def test_unknown_emitter_warns(particle_cache):
    res = surface(particle_cache, settings={"nope": SurfaceSettings()})
    assert res.frames_written == 4
    assert "surface settings for unknown emitter nope ignored" in progress(particle_cache)["warnings"]


def test_partial_cache_then_rest(particle_cache):
    """Surfacing while the bake still runs: the frames written later are meshed by the next run."""
    saved = {f: frame_path(particle_cache, f).read_bytes() for f in (3, 4)}
    for f in saved:
        frame_path(particle_cache, f).unlink()
    assert surface(particle_cache).frames_written == 2
    for f, data in saved.items():
        frame_path(particle_cache, f).write_bytes(data)
    res = surface(particle_cache)
    assert (res.frames_written, res.frames_skipped) == (2, 2)


def test_empty_frame(particle_cache):
    write_particle_frame(CacheWriter(particle_cache, ["emitter_0"], 24, 1, 4, DX, "test"), 2, empty=True)
    assert surface(particle_cache).status == "done"
    assert MeshCacheReader(particle_cache).read("emitter_0", 2).tris.shape == (0, 3)


def test_cancel(particle_cache):
    res = surface(particle_cache, should_cancel=lambda: True)
    assert res.status == "cancelled" and res.frames_written == 1
    assert progress(particle_cache)["status"] == "cancelled"


def test_cancel_file(particle_cache):
    (particle_cache / "surface").mkdir(exist_ok=True)
    (particle_cache / "surface" / "cancel").write_text("")  # left over: cleared at the start
    assert surface(particle_cache).frames_written == 4


def test_meshes_without_particle_frames_are_pruned(particle_cache):
    surface(particle_cache)
    frame_path(particle_cache, 4).unlink()  # e.g. the frame range got shorter
    surface(particle_cache)
    assert MeshCacheReader(particle_cache).available_frames("emitter_0") == [1, 2, 3]


def test_awkward_emitter_names(tmp_path, mesher):
    """Emitter names are not paths, and names that differ in case get their own folders."""
    names = ["cancel", "Cream", "cream", "a/b", "progress.json"]
    write_scene(small_scene(frame_end=1), tmp_path / "scene")
    p, ids = [], []
    for i in range(len(names)):
        s = sphere_particles((0.15 + 0.17 * i, 0.5, 0.5), 0.06, DX)
        p.append(s)
        ids.append(np.full(len(s), i, np.int32))
    p = np.concatenate(p)
    CacheWriter(tmp_path, names, 24, 1, 1, DX, "test").write_frame(
        1, ParticleData(p, np.zeros_like(p), np.zeros(len(p), np.int32), np.concatenate(ids)))
    assert surface(tmp_path).status == "done"
    r = MeshCacheReader(tmp_path)
    assert r.emitters() == names
    for i, name in enumerate(names):
        assert r.read(name, 1).verts[:, 0].mean() == pytest.approx(0.15 + 0.17 * i, abs=0.01)


def test_second_concurrent_run_refused(particle_cache):
    (particle_cache / "surface").mkdir(exist_ok=True)
    with exclusive_lock(particle_cache / "surface" / "lock"):
        with pytest.raises(RuntimeError, match="already running"):
            surface(particle_cache)


def test_particle_frame_vanishing_mid_run(particle_cache):
    """A bake restarting under the surfacer: the vanished frame is skipped with a warning."""
    calls = iter(range(100))

    def bake_restarts():
        if next(calls) == 0:
            frame_path(particle_cache, 2).unlink()
        return False

    res = surface(particle_cache, should_cancel=bake_restarts)
    assert res.status == "done" and res.frames_written == 3
    assert any("frame 2" in w for w in progress(particle_cache)["warnings"])


def test_error_is_reported(particle_cache, monkeypatch):
    def broken(*args, **kwargs):
        raise ValueError("surfacing needs more GPU memory")

    monkeypatch.setattr(run_mod, "surface_particles", broken)
    with pytest.raises(ValueError):
        surface(particle_cache)
    doc = progress(particle_cache)
    assert doc["status"] == "error" and "GPU memory" in doc["message"]


@pytest.mark.gpu
def test_cli(tmp_path):
    write_scene(small_scene(frame_end=2), tmp_path / "scene")
    writer = CacheWriter(tmp_path, ["emitter_0"], 24, 1, 2, DX, "test")
    for f in (1, 2):
        write_particle_frame(writer, f)
    out = subprocess.run([sys.executable, "-m", "mpm_surface", str(tmp_path), "--frames", "1-2"],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr
    assert "done: 2 frames meshed" in out.stdout
    assert MeshCacheReader(tmp_path).available_frames("emitter_0") == [1, 2]


def test_cli_error_exit_code(tmp_path):
    out = subprocess.run([sys.executable, "-m", "mpm_surface", str(tmp_path / "missing")], capture_output=True,
                         text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 1 and out.stderr.startswith("error:")
