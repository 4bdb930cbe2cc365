"""The surfacer's runner and command line (mpm_surface.run, mpm_surface.__main__), after the reference's
src/mpm_surface/run.py and __main__.py and its tests in tests/test_surface_run.py.

The runner tests take a ``mesher``: on the CPU a stand-in that returns a small closed tetrahedron at the
particles' mean (the runner's own logic: keys, skipping, pruning, cancel, lock, warnings, errors), and marked
gpu the HIP surfacer itself. The particle caches are written directly, no solver involved: a ball moving
1 cm along +x per frame.
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
from mpm_surface import run as run_module
from mpm_surface.__main__ import frame_range, main
from mpm_surface.meshcache import MeshCacheReader, SurfaceMesh
from mpm_surface.run import particle_key, surface
from mpm_surface.settings import SurfaceSettings

DX = 1.0 / 32
REPO = Path(__file__).resolve().parent.parent
TETRA = np.int32([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])


def tetra_mesher(positions, velocities, spacing, settings, device=None, max_bytes=None):
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    if len(p) == 0:
        return SurfaceMesh.empty()
    corners = p.mean(axis=0) + np.float32(0.01) * np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    return SurfaceMesh(corners.astype(np.float32), TETRA.copy(), np.zeros((4, 3), np.float32))


@pytest.fixture(params=["tetra", pytest.param("hip", marks=pytest.mark.gpu)])
def mesher(request, monkeypatch):
    if request.param == "tetra":
        monkeypatch.setattr(run_module, "surface_particles", tetra_mesher)
    return request.param


def ball(frame: int, radius: float = 0.1) -> np.ndarray:
    return sphere_particles((0.4 + 0.01 * (frame - 1), 0.5, 0.5), radius, DX)


def put_frame(cache: Path, frame: int, parts: list[np.ndarray], names=("emitter_0",), frames=4) -> None:
    """Write particle frame ``frame`` with ``parts[i]`` as emitter i's positions (velocity 0.24 m/s along x)."""
    pos = np.concatenate(parts).astype(np.float32)
    ids = np.concatenate([np.full(len(p), i, np.int32) for i, p in enumerate(parts)])
    vel = np.zeros_like(pos)
    vel[:, 0] = 0.24
    CacheWriter(cache, list(names), 24, 1, frames, DX, "test").write_frame(
        frame, ParticleData(pos, vel, np.zeros(len(pos), np.int32), ids))


@pytest.fixture
def cache(tmp_path, mesher):
    """Four particle frames of one emitter."""
    write_scene(small_scene(frame_end=4), tmp_path / "scene")
    for f in range(1, 5):
        put_frame(tmp_path, f, [ball(f)])
    return tmp_path


def progress_of(cache: Path) -> dict:
    return json.loads((cache / "surface" / "progress.json").read_text())


def stored_frames(cache: Path) -> list[int]:
    return MeshCacheReader(cache).available_frames("emitter_0")


# ---- first runs ----------------------------------------------------------------------------------------

def test_every_frame_is_meshed(cache):
    result = surface(cache)
    assert (result.status, result.frames_written, result.frames_skipped) == ("done", 4, 0)
    assert stored_frames(cache) == [1, 2, 3, 4]
    reader = MeshCacheReader(cache)
    for f in range(1, 5):
        mesh = reader.read("emitter_0", f)
        assert mesh_is_closed(mesh.tris) and len(mesh.verts) > 0
    doc = progress_of(cache)
    assert (doc["status"], doc["frame"], doc["frame_end"]) == ("done", 4, 4)
    assert doc["particles"] == len(ball(4))


def test_mesh_keys_name_particles_and_settings(cache):
    surface(cache)
    with np.load(cache / "surface" / "meshes" / "00_emitter_0" / "frame_0002.npz") as npz:
        assert str(npz["key"]) == particle_key(ball(2), SurfaceSettings())


def test_a_subset_of_frames(cache):
    result = surface(cache, frames=[3, 2, 9])  # 9 has no particle frame
    assert result.frames_written == 2 and stored_frames(cache) == [2, 3]
    assert progress_of(cache)["frame_end"] == 3


def test_an_empty_particle_frame_gets_an_empty_mesh(cache):
    put_frame(cache, 2, [ball(2)[:0]])
    assert surface(cache).status == "done"
    assert MeshCacheReader(cache).read("emitter_0", 2).tris.shape == (0, 3)


def test_awkward_emitter_names(tmp_path, mesher):
    """Names are not paths, and names differing only in case keep separate meshes."""
    names = ["cancel", "Cream", "cream", "a/b", "progress.json"]
    write_scene(small_scene(frame_end=1), tmp_path / "scene")
    centres = [0.15 + 0.17 * i for i in range(len(names))]
    put_frame(tmp_path, 1, [sphere_particles((c, 0.5, 0.5), 0.06, DX) for c in centres], names=names, frames=1)
    assert surface(tmp_path).status == "done"
    reader = MeshCacheReader(tmp_path)
    assert reader.emitters() == names
    for name, c in zip(names, centres):
        assert reader.read(name, 1).verts[:, 0].mean() == pytest.approx(c, abs=0.01)


# ---- second runs: what is meshed again ---------------------------------------------------------------------

@pytest.mark.parametrize("change,want", [
    ("nothing", (0, 4)),
    ("settings", (4, 0)),
    ("one_frame", (1, 3)),
])
def test_a_second_run_remeshes_only_what_changed(cache, change, want):
    surface(cache)
    settings = None
    if change == "settings":
        settings = {"emitter_0": SurfaceSettings(detail=2.0)}
    elif change == "one_frame":
        put_frame(cache, 3, [ball(3, radius=0.09)])
    result = surface(cache, settings=settings)
    assert (result.frames_written, result.frames_skipped) == want


def test_frames_written_while_a_bake_runs_are_meshed_next_time(cache):
    later = {f: frame_path(cache, f).read_bytes() for f in (3, 4)}
    for f in later:
        frame_path(cache, f).unlink()
    assert surface(cache).frames_written == 2
    for f, data in later.items():
        frame_path(cache, f).write_bytes(data)
    result = surface(cache)
    assert (result.frames_written, result.frames_skipped) == (2, 2)


def test_meshes_without_a_particle_frame_are_pruned(cache):
    surface(cache)
    frame_path(cache, 4).unlink()  # a shorter frame range
    surface(cache)
    assert stored_frames(cache) == [1, 2, 3]


def test_disabling_an_emitter_removes_its_meshes(cache):
    surface(cache)
    surface(cache, settings={"emitter_0": SurfaceSettings(enabled=False)})
    assert MeshCacheReader(cache).emitters() == []


def test_settings_of_an_unknown_emitter_are_ignored_with_a_warning(cache):
    assert surface(cache, settings={"ghost": SurfaceSettings()}).frames_written == 4
    assert progress_of(cache)["warnings"] == ["surface settings for unknown emitter ghost ignored"]


# ---- stops, refusals, errors ------------------------------------------------------------------------------

def test_cancel_callback_stops_after_the_frame(cache):
    result = surface(cache, should_cancel=lambda: True)
    assert (result.status, result.frames_written) == ("cancelled", 1)
    assert progress_of(cache)["status"] == "cancelled"


def test_leftover_cancel_file_is_cleared(cache):
    (cache / "surface").mkdir(exist_ok=True)
    (cache / "surface" / "cancel").touch()
    assert surface(cache).frames_written == 4


def test_a_particle_frame_vanishing_mid_run_is_skipped(cache):
    """A bake restarting under the surfacer deletes frames; the missing one is skipped with a warning."""
    asked = [0]

    def restart_bake():
        asked[0] += 1
        if asked[0] == 1:
            frame_path(cache, 2).unlink()
        return False

    result = surface(cache, should_cancel=restart_bake)
    assert (result.status, result.frames_written) == ("done", 3)
    assert progress_of(cache)["warnings"] == [
        "particle frame 2 disappeared during surfacing (bake restarted?); skipped"]


def test_a_second_concurrent_run_is_refused(cache):
    (cache / "surface").mkdir(exist_ok=True)
    with exclusive_lock(cache / "surface" / "lock"):
        with pytest.raises(RuntimeError, match="a surfacer run is already running on"):
            surface(cache)


def test_a_mesher_error_lands_in_progress(cache, monkeypatch):
    def out_of_memory(*args, **kwargs):
        raise MemoryError("surfacing needs more device memory")

    monkeypatch.setattr(run_module, "surface_particles", out_of_memory)
    with pytest.raises(MemoryError):
        surface(cache)
    doc = progress_of(cache)
    assert (doc["status"], doc["message"]) == ("error", "MemoryError: surfacing needs more device memory")


# ---- command line --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("text,frames", [("5", [5]), ("2-4", [2, 3, 4]), ("7-7", [7])])
def test_frame_range(text, frames):
    assert frame_range(text) == frames


def test_command_line_in_process(cache, capsys):
    assert main([str(cache), "--frames", "2-3"]) == 0
    assert capsys.readouterr().out.strip() == "done: 2 frames meshed, 0 unchanged"
    assert main([str(cache / "missing")]) == 1
    assert capsys.readouterr().err.startswith("error: ")


def test_command_line_exit_status_on_error(tmp_path):
    out = subprocess.run([sys.executable, "-m", "mpm_surface", str(tmp_path / "missing")], capture_output=True,
                         text=True, timeout=300, cwd=REPO)
    assert out.returncode == 1 and out.stderr.startswith("error:")


@pytest.mark.gpu
def test_command_line_on_the_hip_surfacer(tmp_path):
    write_scene(small_scene(frame_end=2), tmp_path / "scene")
    for f in (1, 2):
        put_frame(tmp_path, f, [ball(f)], frames=2)
    out = subprocess.run([sys.executable, "-m", "mpm_surface", str(tmp_path), "--frames", "1-2"],
                         capture_output=True, text=True, timeout=300, cwd=REPO)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("done: 2 frames meshed, 0 unchanged")
    assert MeshCacheReader(tmp_path).available_frames("emitter_0") == [1, 2]
