"""Surfacing settings and the mesh cache: the file formats the Blender add-on reads and writes
(reference tests/test_surface_settings.py; src/mpm_surface/settings.py, meshcache.py)."""

import json

import numpy as np
import pytest

from mpm_surface.meshcache import MeshCacheReader, MeshCacheWriter, SurfaceMesh, folder_name
from mpm_surface.settings import SurfaceSettings, read_settings, write_settings


def test_defaults_and_roundtrip():
    s = SurfaceSettings()
    assert (s.enabled, s.detail, s.smoothing, s.iso) == (True, 1.0, 2.0, 0.35)
    assert SurfaceSettings.from_dict(s.to_dict()) == s


def test_hash_changes():
    assert SurfaceSettings(detail=2).hash() != SurfaceSettings().hash()
    assert SurfaceSettings().hash() == SurfaceSettings().hash()


@pytest.mark.parametrize("kw", [{"detail": 0}, {"smoothing": -1}, {"iso": 1.0}, {"iso": 0.0}, {"detail": float("nan")}])
def test_invalid(kw):
    with pytest.raises(ValueError):
        SurfaceSettings(**kw)


def test_settings_file_roundtrip(tmp_path):
    assert read_settings(tmp_path) == {}
    settings = {"cream": SurfaceSettings(detail=2), "sand": SurfaceSettings(enabled=False)}
    write_settings(tmp_path, settings)
    assert read_settings(tmp_path) == settings
    doc = json.loads((tmp_path / "surface_settings.json").read_text())
    assert doc["version"] == 1 and doc["emitters"]["cream"] == {"enabled": True, "detail": 2, "smoothing": 2.0,
                                                                "iso": 0.35}


def test_settings_version_checked(tmp_path):
    (tmp_path / "surface_settings.json").write_text(json.dumps({"version": 2, "emitters": {}}))
    with pytest.raises(ValueError):
        read_settings(tmp_path)


def _triangle():
    return SurfaceMesh(np.eye(3, dtype=np.float32), np.array([[0, 1, 2]], np.int32), np.zeros((3, 3), np.float32))


def test_mesh_cache_roundtrip(tmp_path):
    MeshCacheWriter(tmp_path, ["cream"]).write("cream", 3, _triangle(), "k")
    r = MeshCacheReader(tmp_path)
    assert r.emitters() == ["cream"] and r.available_frames("cream") == [3]
    m = r.read("cream", 3)
    assert np.array_equal(m.verts, _triangle().verts) and np.array_equal(m.tris, [[0, 1, 2]])
    assert m.verts.dtype == np.float32 and m.tris.dtype == np.int32 and m.vel.dtype == np.float32
    assert MeshCacheWriter(tmp_path, ["cream"]).key("cream", 3) == "k"
    assert MeshCacheWriter(tmp_path, ["cream"]).key("cream", 4) is None
    assert (tmp_path / "surface" / "meshes" / "00_cream" / "frame_0003.npz").exists()
    assert json.loads((tmp_path / "surface" / "meshes.json").read_text()) == {"version": 1,
                                                                            "emitters": {"cream": "00_cream"}}


def test_remove_emitter(tmp_path):
    w = MeshCacheWriter(tmp_path, ["cream"])
    w.write("cream", 1, _triangle(), "k")
    w.remove_emitter("cream")
    assert MeshCacheReader(tmp_path).emitters() == []


def test_stale_folders_removed(tmp_path):
    MeshCacheWriter(tmp_path, ["cream", "sand"]).write("sand", 1, _triangle(), "k")
    MeshCacheWriter(tmp_path, ["sand"])  # sand is now emitter 0: its old folder 01_sand goes
    assert not (tmp_path / "surface" / "meshes" / "01_sand").exists()


def test_folder_names_are_not_paths():
    assert folder_name(0, "a/b") == "00_a_b"
    assert folder_name(3, "..") == "03__"
    assert folder_name(1, "Cream") != folder_name(2, "cream")
    assert len(folder_name(0, "x" * 100)) == 3 + 40


def test_empty_mesh(tmp_path):
    e = SurfaceMesh.empty()
    assert e.verts.shape == (0, 3) and e.tris.shape == (0, 3) and e.vel.shape == (0, 3)
    MeshCacheWriter(tmp_path, ["cream"]).write("cream", 1, e, "k")
    assert MeshCacheReader(tmp_path).read("cream", 1).tris.shape == (0, 3)
