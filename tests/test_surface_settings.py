"""The surfacer's files that the Blender add-on reads and writes: surface_settings.json and the mesh cache
(formats of the reference's src/mpm_surface/settings.py and meshcache.py; its tests in
tests/test_surface_settings.py)."""

import hashlib
import json

import numpy as np
import pytest

from mpm_surface.meshcache import MeshCacheReader, MeshCacheWriter, SurfaceMesh, folder_name
from mpm_surface.settings import SurfaceSettings, read_settings, write_settings

# ---- settings ----------------------------------------------------------------------------------------------


def test_defaults_survive_a_dict_round_trip():
    s = SurfaceSettings()
    assert s.to_dict() == {"enabled": True, "detail": 1.0, "smoothing": 2.0, "iso": 0.35}
    assert SurfaceSettings.from_dict(s.to_dict()) == s


def test_fingerprint_is_sha1_of_sorted_json():
    """The fingerprint is part of every stored mesh key, so its text is pinned, not only its behaviour."""
    text = '{"detail": 1.0, "enabled": true, "iso": 0.35, "smoothing": 2.0}'
    assert SurfaceSettings().hash() == hashlib.sha1(text.encode()).hexdigest()
    assert SurfaceSettings(detail=2.0).hash() != SurfaceSettings().hash()
    assert SurfaceSettings(enabled=False).hash() != SurfaceSettings().hash()


@pytest.mark.parametrize("field,value,message", [
    ("detail", 0.0, "detail must be > 0"),
    ("detail", float("nan"), "detail must be > 0"),
    ("smoothing", -1.0, "smoothing must be > 0"),
    ("iso", 0.0, r"iso must be in \(0, 1\)"),
    ("iso", 1.0, r"iso must be in \(0, 1\)"),
])
def test_out_of_range_values_are_refused(field, value, message):
    with pytest.raises(ValueError, match=message):
        SurfaceSettings(**{field: value})


def test_settings_file(tmp_path):
    assert read_settings(tmp_path) == {}  # no file: every emitter takes the defaults
    chosen = {"cream": SurfaceSettings(detail=2), "sand": SurfaceSettings(enabled=False)}
    write_settings(tmp_path, chosen)
    assert json.loads((tmp_path / "surface_settings.json").read_text()) == {
        "version": 1,
        "emitters": {"cream": {"enabled": True, "detail": 2, "smoothing": 2.0, "iso": 0.35},
                     "sand": {"enabled": False, "detail": 1.0, "smoothing": 2.0, "iso": 0.35}}}
    assert read_settings(tmp_path) == chosen


def test_settings_file_of_another_version_is_refused(tmp_path):
    (tmp_path / "surface_settings.json").write_text('{"version": 2, "emitters": {}}')
    with pytest.raises(ValueError, match="unsupported surface settings version 2"):
        read_settings(tmp_path)


# ---- mesh cache --------------------------------------------------------------------------------------------

def one_triangle(shift: float = 0.0) -> SurfaceMesh:
    verts = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]) + np.float32(shift)
    return SurfaceMesh(verts, np.int32([[0, 1, 2]]), np.full((3, 3), 0.5, np.float32))


@pytest.mark.parametrize("index,name,folder", [
    (0, "cream", "00_cream"),
    (0, "a/b", "00_a_b"),
    (3, "..", "03__"),
    (7, "wet sand (2)", "07_wet_sand_2_"),
    (12, "x" * 100, "12_" + "x" * 40),
])
def test_folder_names(index, name, folder):
    assert folder_name(index, name) == folder


def test_names_differing_in_case_get_their_own_folders():
    assert folder_name(1, "Cream") != folder_name(2, "cream")


def test_mesh_files_and_manifest(tmp_path):
    writer = MeshCacheWriter(tmp_path, ["sand", "cream"])
    writer.write("cream", 3, one_triangle(), "key-a")
    assert json.loads((tmp_path / "surface" / "meshes.json").read_text()) == {
        "version": 1, "emitters": {"sand": "00_sand", "cream": "01_cream"}}
    with np.load(tmp_path / "surface" / "meshes" / "01_cream" / "frame_0003.npz") as npz:
        assert sorted(npz.files) == ["key", "tris", "vel", "verts"]
        assert (npz["verts"].dtype, npz["tris"].dtype, npz["vel"].dtype) == (np.float32, np.int32, np.float32)
        assert str(npz["key"]) == "key-a"
    reader = MeshCacheReader(tmp_path)
    # an emitter counts once its folder exists, that is once it has a mesh
    assert reader.emitters() == ["cream"] and reader.available_frames("cream") == [3]
    assert reader.available_frames("sand") == [] and reader.available_frames("nobody") == []
    mesh = reader.read("cream", 3)
    np.testing.assert_array_equal(mesh.verts, one_triangle().verts)
    np.testing.assert_array_equal(mesh.vel, one_triangle().vel)
    assert (writer.key("cream", 3), writer.key("cream", 4)) == ("key-a", None)


def test_rewriting_a_frame_replaces_it(tmp_path):
    writer = MeshCacheWriter(tmp_path, ["cream"])
    writer.write("cream", 1, one_triangle(), "old")
    writer.write("cream", 1, one_triangle(2.0), "new")
    assert writer.key("cream", 1) == "new"
    assert MeshCacheReader(tmp_path).read("cream", 1).verts.min() == 2.0
    assert not list((tmp_path / "surface" / "meshes" / "00_cream").glob("*.tmp"))


def test_removing_frames_and_emitters(tmp_path):
    writer = MeshCacheWriter(tmp_path, ["cream", "sand"])
    for frame in (1, 2):
        writer.write("cream", frame, one_triangle(), "k")
        writer.write("sand", frame, one_triangle(), "k")
    writer.remove_frame("cream", 2)
    writer.remove_frame("cream", 9)  # absent: nothing happens
    writer.remove_emitter("sand")
    reader = MeshCacheReader(tmp_path)
    assert reader.emitters() == ["cream"] and reader.available_frames("cream") == [1]


def test_folders_of_emitters_gone_from_the_cache_are_deleted(tmp_path):
    MeshCacheWriter(tmp_path, ["cream", "sand"]).write("sand", 1, one_triangle(), "k")
    MeshCacheWriter(tmp_path, ["sand"])  # sand is emitter 0 now: the folder 01_sand goes
    assert [p.name for p in (tmp_path / "surface" / "meshes").iterdir()] == []


def test_empty_mesh(tmp_path):
    empty = SurfaceMesh.empty()
    assert empty.verts.shape == empty.tris.shape == empty.vel.shape == (0, 3)
    MeshCacheWriter(tmp_path, ["cream"]).write("cream", 1, empty, "k")
    assert MeshCacheReader(tmp_path).read("cream", 1).tris.shape == (0, 3)


def test_no_manifest_means_no_meshes(tmp_path):
    reader = MeshCacheReader(tmp_path)
    assert reader.emitters() == [] and reader.available_frames("cream") == []
