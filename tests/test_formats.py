"""Formats and interface of mpm_core: the reference's CPU tests (tests/test_materials.py, test_scene.py,
test_geometry.py, test_cache.py) restated against this package."""

import json
import sys
import threading
import time
from dataclasses import replace

import numpy as np
import pytest
from helpers import icosphere, small_scene, unit_cube_mesh

from mpm_core.cache import CacheReader, CacheWriter, read_checkpoint, write_checkpoint, write_json
from mpm_core.geometry import interpolate_transform, mesh_is_closed, sample_mesh_volume, transform_points
from mpm_core.materials import (MaterialParams, delete_user_preset, list_presets, load_preset, save_user_preset)
from mpm_core.scene import Domain, read_scene, scene_hash, write_scene
from mpm_core.solver import ParticleData

PRESETS = ["clay", "gravel", "jelly", "mashed_potato", "mud", "rubber", "sand_dry", "snow_packed", "snow_powder",
           "snow_wet", "soil_dry", "sugar", "toothpaste", "whipped_cream"]


# ---- materials -------------------------------------------------------------------------------------
def test_list_presets():
    assert list_presets() == PRESETS


@pytest.mark.parametrize("preset_id", PRESETS)
def test_builtin_presets_describe_themselves(preset_id):
    p = load_preset(preset_id)
    assert p.description and not p.is_user


def test_user_presets(tmp_path):
    cream = load_preset("whipped_cream")
    firm = replace(cream.material, yield_stress=5000.0)
    saved = save_user_preset(tmp_path, " Firm cream ", firm, cream.look)
    assert saved.id == "user:firm_cream" and saved.is_user and saved.label == "Firm cream"
    assert list_presets(tmp_path)[-1] == "user:firm_cream" and "user:firm_cream" not in list_presets()
    loaded = load_preset("user:firm_cream", tmp_path)
    assert loaded.material == replace(firm, name="user:firm_cream") and loaded.look == cream.look
    save_user_preset(tmp_path, "FIRM CREAM", cream.material, cream.look)
    assert len(list(tmp_path.iterdir())) == 1 and load_preset("user:firm_cream", tmp_path).label == "FIRM CREAM"
    delete_user_preset(tmp_path, "user:firm_cream")
    assert list_presets(tmp_path) == list_presets()


def test_user_presets_sorted_by_label(tmp_path):
    look = load_preset("jelly").look
    for label in ("b", "A", "c"):
        save_user_preset(tmp_path, label, load_preset("jelly").material, look)
    assert list_presets(tmp_path)[-3:] == ["user:a", "user:b", "user:c"]


@pytest.mark.parametrize("label", ["Whipped cream", "whipped CREAM", "", "  ", "!!"])
def test_user_preset_bad_names(tmp_path, label):
    with pytest.raises(ValueError):
        save_user_preset(tmp_path, label, load_preset("jelly").material, load_preset("jelly").look)


def test_cannot_delete_builtin(tmp_path):
    with pytest.raises(ValueError):
        delete_user_preset(tmp_path, "jelly")


def test_load_cream():
    p = load_preset("whipped_cream")
    assert p.label == "Whipped cream" and p.material.model == "von_mises" and p.material.yield_stress == 2500
    assert p.look.display == "mesh"


def test_lame():
    mu, lam = MaterialParams("m", "elastic", 1.0e5, 0.3, 1000).lame()
    assert mu == pytest.approx(1.0e5 / 2.6) and lam == pytest.approx(1.0e5 * 0.3 / (1.3 * 0.4))


def test_roundtrip_dict():
    m = load_preset("sand_dry").material
    assert MaterialParams.from_dict(m.to_dict()) == m


@pytest.mark.parametrize("kw", [{"model": "goo"}, {"poisson_ratio": 0.5}, {"youngs_modulus": 0}, {"density": -1}])
def test_invalid_material(kw):
    base = dict(name="m", model="elastic", youngs_modulus=1e5, poisson_ratio=0.3, density=1000)
    with pytest.raises(ValueError):
        MaterialParams(**{**base, **kw})


def test_unknown_preset():
    with pytest.raises(KeyError):
        load_preset("custard")


# ---- scene -----------------------------------------------------------------------------------------
def test_domain_grid():
    d = Domain(min=(0, 0, 0), max=(1.0, 0.25, 0.5), resolution=64)
    assert d.dx == pytest.approx(1 / 64) and d.grid_shape == (64, 16, 32)


def test_scene_roundtrip(tmp_path):
    scene = small_scene()
    write_scene(scene, tmp_path)
    back = read_scene(tmp_path)
    assert back.domain == scene.domain and back.materials == scene.materials
    assert np.array_equal(back.emitters[0].transforms, scene.emitters[0].transforms)
    assert back.colliders[0].closed is False


def test_scene_hash_changes(tmp_path):
    write_scene(small_scene(), tmp_path)
    h = scene_hash(tmp_path)
    write_scene(small_scene(gravity=(0, 0, -5)), tmp_path)
    assert scene_hash(tmp_path) != h


def test_wrong_transform_count():
    with pytest.raises(ValueError):
        small_scene(emitter_frames=5)


def test_bad_particles_per_cell():
    with pytest.raises(ValueError):
        small_scene(particles_per_cell=6)


# ---- geometry --------------------------------------------------------------------------------------
def test_closed():
    v, t = unit_cube_mesh()
    assert mesh_is_closed(t) and not mesh_is_closed(t[:-1])


def test_cube_count():
    v, t = unit_cube_mesh()
    pts = sample_mesh_volume(v * 0.2, t, 0.01, seed=0)
    assert abs(len(pts) - 8000) <= 8000 * 0.02
    assert pts.min() >= -0.0025 and pts.max() <= 0.2025


def test_sphere_volume():
    v, t = icosphere(0.1, 4)
    t0 = time.perf_counter()
    pts = sample_mesh_volume(v, t, 0.005, seed=0)
    assert time.perf_counter() - t0 < 2.0
    assert len(pts) == pytest.approx(4 / 3 * np.pi * 0.1 ** 3 / 0.005 ** 3, rel=0.03)


def test_deterministic():
    v, t = icosphere(0.1, 3)
    assert np.array_equal(sample_mesh_volume(v, t, 0.01, 7), sample_mesh_volume(v, t, 0.01, 7))


def test_open_mesh_raises():
    v, t = unit_cube_mesh()
    with pytest.raises(ValueError, match="not closed"):
        sample_mesh_volume(v, t[:-2], 0.01, 0)


def test_scaled_mesh_volume():
    v, t = unit_cube_mesh()
    m = np.diag([0.2, 0.1, 0.4, 1.0]).astype(np.float32)
    pts = sample_mesh_volume(transform_points(m, v), t, 0.01, 0)
    assert len(pts) == pytest.approx(0.2 * 0.1 * 0.4 / 0.01 ** 3, rel=0.03)


def test_interpolate_transform():
    a = np.diag([2.0, 2.0, 2.0, 1.0])
    b = np.eye(4)
    b[:3, :3] = [[0, -2, 0], [2, 0, 0], [0, 0, 2]]
    b[:3, 3] = (1, 2, 3)
    assert np.allclose(interpolate_transform(a, b, 0), a) and np.allclose(interpolate_transform(a, b, 1), b)
    mid = interpolate_transform(a, b, 0.5)
    c = np.sqrt(0.5) * 2
    assert np.allclose(mid[:3, :3], [[c, -c, 0], [c, c, 0], [0, 0, 2]], atol=1e-6)
    assert np.allclose(mid[:3, 3], (0.5, 1, 1.5))


# ---- cache -----------------------------------------------------------------------------------------
def _particles():
    return ParticleData(np.random.rand(5, 3).astype(np.float32), np.zeros((5, 3), np.float32),
                        np.zeros(5, np.int32), np.array([0, 1, 1, 0, 1], np.int32))


def test_write_read_frame(tmp_path):
    w = CacheWriter(tmp_path, ["a", "b"], 24, 1, 10, 0.01, "hip")
    w.write_frame(3, _particles())
    r = CacheReader(tmp_path)
    assert r.available_frames() == [3] and r.emitter_names == ["a", "b"]
    f = r.read_frame(3)
    assert f["a"][0].shape == (2, 3) and f["b"][0].shape == (3, 3)


def test_no_tmp_left(tmp_path):
    CacheWriter(tmp_path, ["a", "b"], 24, 1, 10, 0.01, "hip").write_frame(1, _particles())
    assert list((tmp_path / "frames").glob("*.tmp")) == []


def test_checkpoint_roundtrip(tmp_path):
    state = {"x": np.arange(6, dtype=np.float32).reshape(2, 3), "mat": np.array([0, 1], np.int32)}
    write_checkpoint(tmp_path, 5, state, "abc")
    frame, back, h = read_checkpoint(tmp_path)
    assert frame == 5 and h == "abc"
    assert set(back) == set(state) and all(np.array_equal(back[k], state[k]) for k in state)


def test_no_checkpoint(tmp_path):
    assert read_checkpoint(tmp_path) is None


def test_concurrent_writes_to_one_path(tmp_path):
    path = tmp_path / "progress.json"
    errors = []

    def writer(n):
        try:
            for i in range(30):
                write_json(path, {"writer": n, "i": i, "pad": "x" * 20000})
        except Exception as exc:  # noqa: BLE001
            errors.append(exc)

    threads = [threading.Thread(target=writer, args=(n,)) for n in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == [] and json.loads(path.read_text())["i"] == 29
    assert list(tmp_path.glob("*.tmp")) == []


@pytest.mark.skipif(sys.platform != "win32", reason="Windows file locking")
def test_atomic_write_survives_reader_holding_file(tmp_path):
    path = tmp_path / "progress.json"
    write_json(path, {"status": "running"})
    reader = open(path, "rb")
    threading.Timer(0.05, reader.close).start()
    write_json(path, {"status": "done"})
    assert json.loads(path.read_text())["status"] == "done"
