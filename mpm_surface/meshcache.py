"""The mesh cache under ``<cache>/surface/`` (reference src/mpm_surface/meshcache.py:1-101; same layout)::

    surface/meshes.json                   {"version": 1, "emitters": {"<emitter name>": "<folder>"}}
    surface/meshes/<folder>/frame_NNNN.npz

A folder is ``<two-digit emitter index>_<name with every run of characters outside [A-Za-z0-9_-] replaced
by "_", at most 40 characters>``, so names never act as paths and names that differ only in case get
different folders. Each npz holds ``verts`` (V,3) float32 world positions, ``tris`` (T,3) int32 triangles
wound with normals pointing out of the material, ``vel`` (V,3) float32 vertex velocities (m/s) and
``key``, the fingerprint of the particles and settings the mesh was made from.
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


@dataclass
class SurfaceMesh:
    verts: np.ndarray
    tris: np.ndarray
    vel: np.ndarray

    @staticmethod
    def empty() -> SurfaceMesh:
        return SurfaceMesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float32))


def folder_name(index: int, name: str) -> str:
    return "%02d_%s" % (index, re.sub(r"[^A-Za-z0-9_-]+", "_", name)[:40])


class MeshCacheWriter:
    """Lays out the folders of the given emitters (in cache order) and removes folders of emitters that are
    no longer in the cache (meshcache.py:42-54)."""

    def __init__(self, cache_dir: Path, emitter_names: list[str]):
        self.root = Path(cache_dir) / "surface"
        self.folders = {name: folder_name(i, name) for i, name in enumerate(emitter_names)}
        meshes = self.root / "meshes"
        meshes.mkdir(parents=True, exist_ok=True)
        keep = set(self.folders.values())
        for entry in meshes.iterdir():
            if entry.is_dir() and entry.name not in keep:
                shutil.rmtree(entry, ignore_errors=True)
        write_json(self.root / "meshes.json", {"version": FORMAT_VERSION, "emitters": self.folders})

    def path(self, emitter: str, frame: int) -> Path:
        return self.root / "meshes" / self.folders[emitter] / f"frame_{frame:04d}.npz"

    def write(self, emitter: str, frame: int, mesh: SurfaceMesh, key: str) -> None:
        path = self.path(emitter, frame)
        path.parent.mkdir(parents=True, exist_ok=True)
        atomic_write_npz(path, {"verts": np.ascontiguousarray(mesh.verts, np.float32),
                                "tris": np.ascontiguousarray(mesh.tris, np.int32),
                                "vel": np.ascontiguousarray(mesh.vel, np.float32),
                                "key": np.array(key)})

    def key(self, emitter: str, frame: int) -> str | None:
        """The key of the stored mesh, None when there is none."""
        path = self.path(emitter, frame)
        if not path.exists():
            return None
        with np.load(path) as data:
            return str(data["key"])

    def remove_frame(self, emitter: str, frame: int) -> None:
        self.path(emitter, frame).unlink(missing_ok=True)

    def remove_emitter(self, emitter: str) -> None:
        shutil.rmtree(self.root / "meshes" / self.folders[emitter], ignore_errors=True)


class MeshCacheReader:
    def __init__(self, cache_dir: Path):
        self.root = Path(cache_dir) / "surface"
        manifest = self.root / "meshes.json"
        self.folders: dict[str, str] = {}
        if manifest.exists():
            self.folders = json.loads(manifest.read_text(encoding="utf-8"))["emitters"]

    def emitters(self) -> list[str]:
        """The emitters that have a mesh folder, in cache order."""
        return [name for name, folder in self.folders.items() if (self.root / "meshes" / folder).is_dir()]

    def available_frames(self, emitter: str) -> list[int]:
        if emitter not in self.folders:
            return []
        folder = self.root / "meshes" / self.folders[emitter]
        return sorted(int(p.stem.split("_")[1]) for p in folder.glob("frame_*.npz"))

    def read(self, emitter: str, frame: int) -> SurfaceMesh:
        with np.load(self.root / "meshes" / self.folders[emitter] / f"frame_{frame:04d}.npz") as data:
            return SurfaceMesh(data["verts"], data["tris"], data["vel"])
