"""The mesh cache of a bake: ``<cache>/surface/``.

Layout and formats are the reference's (src/mpm_surface/meshcache.py:1-101), which the add-on reads::

    surface/meshes.json                   {"version": 1, "emitters": {"<emitter name>": "<folder>"}}
    surface/meshes/<folder>/frame_NNNN.npz

An emitter's folder is its two-digit index in cache.json, an underscore, and its name with every run of
characters outside ``A-Za-z0-9_-`` turned into one ``_``, cut to 40 characters: names never act as paths, and
names that differ only in case get different folders. A frame file holds ``verts`` (V,3) float32 world
positions, ``tris`` (T,3) int32 triangles wound with normals out of the material, ``vel`` (V,3) float32 vertex
velocities in m/s and ``key``, the fingerprint of the particles and settings the mesh was made from.
"""

from __future__ import annotations

import json
import re
import shutil
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from mpm_core.cache import atomic_write_npz, write_json

FORMAT_VERSION = 1
_UNSAFE = re.compile(r"[^A-Za-z0-9_-]+")
_NAME_CHARS = 40


@dataclass
class SurfaceMesh:
    verts: np.ndarray
    tris: np.ndarray
    vel: np.ndarray

    @staticmethod
    def empty() -> SurfaceMesh:
        return SurfaceMesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float32))


def folder_name(index: int, name: str) -> str:
    return "{:02d}_{}".format(index, _UNSAFE.sub("_", name)[:_NAME_CHARS])


class _Layout:
    """Where an emitter's meshes live under ``<cache>/surface``."""

    def __init__(self, cache_dir: Path, folders: dict[str, str]):
        self.root = Path(cache_dir) / "surface"
        self.folders = folders

    def folder(self, emitter: str) -> Path:
        return self.root / "meshes" / self.folders[emitter]

    def path(self, emitter: str, frame: int) -> Path:
        return self.folder(emitter) / "frame_{:04d}.npz".format(frame)


class MeshCacheWriter(_Layout):
    """Lays out the folders of ``emitter_names`` (cache order), deletes folders of emitters the cache no longer
    has, and publishes the manifest (meshcache.py:42-54)."""

    def __init__(self, cache_dir: Path, emitter_names: list[str]):
        super().__init__(cache_dir, {name: folder_name(i, name) for i, name in enumerate(emitter_names)})
        meshes = self.root / "meshes"
        meshes.mkdir(parents=True, exist_ok=True)
        current = set(self.folders.values())
        for stale in (d for d in meshes.iterdir() if d.is_dir() and d.name not in current):
            shutil.rmtree(stale, ignore_errors=True)
        write_json(self.root / "meshes.json", {"version": FORMAT_VERSION, "emitters": self.folders})

    def write(self, emitter: str, frame: int, mesh: SurfaceMesh, key: str) -> None:
        path = self.path(emitter, frame)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = {"verts": (mesh.verts, np.float32), "tris": (mesh.tris, np.int32), "vel": (mesh.vel, np.float32)}
        members = {k: np.ascontiguousarray(a, dtype) for k, (a, dtype) in arrays.items()}
        members["key"] = np.array(key)
        atomic_write_npz(path, members)

    def key(self, emitter: str, frame: int) -> str | None:
        """The stored mesh's key; None when the frame has no mesh."""
        path = self.path(emitter, frame)
        if not path.is_file():
            return None
        with np.load(path) as npz:
            return str(npz["key"])

    def remove_frame(self, emitter: str, frame: int) -> None:
        self.path(emitter, frame).unlink(missing_ok=True)

    def remove_emitter(self, emitter: str) -> None:
        shutil.rmtree(self.folder(emitter), ignore_errors=True)


class MeshCacheReader(_Layout):
    """The meshes of a cache as the manifest lists them (none when there is no manifest)."""

    def __init__(self, cache_dir: Path):
        manifest = Path(cache_dir) / "surface" / "meshes.json"
        folders = json.loads(manifest.read_text(encoding="utf-8"))["emitters"] if manifest.is_file() else {}
        super().__init__(cache_dir, folders)

    def emitters(self) -> list[str]:
        """Emitters whose mesh folder exists, in cache order."""
        return [name for name in self.folders if self.folder(name).is_dir()]

    def available_frames(self, emitter: str) -> list[int]:
        if emitter not in self.folders:
            return []
        return sorted(int(p.stem[len("frame_"):]) for p in self.folder(emitter).glob("frame_*.npz"))

    def read(self, emitter: str, frame: int) -> SurfaceMesh:
        with np.load(self.path(emitter, frame)) as npz:
            return SurfaceMesh(npz["verts"], npz["tris"], npz["vel"])
