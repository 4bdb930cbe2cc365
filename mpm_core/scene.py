"""Scene folders: what the Blender add-on writes and the bake runner and the bake server read.

A scene folder is ``scene.json`` (format version 1) plus three ``.npy`` arrays per emitter and collider i:
``{emitter|collider}_{i}_verts`` (V,3) float32 local-space vertices, ``_tris`` (T,3) int32 triangles and
``_xf`` (F,4,4) float32 world matrices, one per frame from frame_start to frame_end. The format, its keys,
defaults and validation messages are the reference's (src/mpm_core/scene.py:1-12, 31-140, 148-219), so
folders written by the add-on load unchanged; the code is organised around two tables of JSON keys.
"""

from __future__ import annotations

import hashlib
import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from mpm_core.materials import MaterialParams

FORMAT_VERSION = 1
WALL_SIDES = ("x-", "x+", "y-", "y+", "z-", "z+")
EMITTER_KINDS = ("fill", "inflow")
MAX_SUBSTEPS = 16

# JSON keys of an emitter and of a collider, in the order they are written; the mesh and the per-frame
# transforms of each live in the object's .npy files
_EMITTER_KEYS = ("name", "kind", "material", "velocity", "use_object_velocity", "frame_start", "frame_end",
                 "substeps")
_COLLIDER_KEYS = ("name", "friction", "closed")
_ARRAYS = ("verts", "tris", "xf")
# scalar settings of the scene document after domain / materials / emitters / colliders
_SCENE_KEYS = ("gravity", "fps", "frame_start", "frame_end", "seed", "particles_per_cell")


def _closed_box() -> dict[str, bool]:
    return dict.fromkeys(WALL_SIDES, True)


def _floats(values) -> tuple[float, ...]:
    return tuple(map(float, values))


def _matrices(values) -> np.ndarray:
    return np.asarray(values, np.float32).reshape(-1, 4, 4)


@dataclass
class Domain:
    """The axis-aligned simulation box (scene.py:36-67); ``walls`` maps each side to closed (True) or open.
    The cell size ``dx`` is the longest side over ``resolution``."""

    min: tuple[float, float, float]
    max: tuple[float, float, float]
    resolution: int
    walls: dict[str, bool] = field(default_factory=_closed_box)
    wall_friction: float = 0.5

    def __post_init__(self):
        self.min, self.max = _floats(self.min), _floats(self.max)
        if not all(hi > lo for lo, hi in zip(self.min, self.max)):
            raise ValueError("domain max must be greater than min on every axis")
        if self.resolution < 4:
            raise ValueError("domain resolution must be at least 4")
        unknown = sorted(k for k in self.walls if k not in WALL_SIDES)
        if unknown:
            raise ValueError(f"unknown wall sides {unknown}")
        self.walls = {side: self.walls.get(side, True) for side in WALL_SIDES}

    @property
    def size(self) -> tuple[float, float, float]:
        return tuple(hi - lo for lo, hi in zip(self.min, self.max))

    @property
    def dx(self) -> float:
        return max(self.size) / self.resolution

    @property
    def grid_shape(self) -> tuple[int, int, int]:
        # cells per axis; the 1e-6 keeps a side that is an exact multiple of dx from gaining a cell
        h = self.dx
        return tuple(max(1, math.ceil(s / h - 1e-6)) for s in self.size)


@dataclass
class MeshData:
    verts: np.ndarray
    tris: np.ndarray

    def __post_init__(self):
        self.verts = np.asarray(self.verts, np.float32).reshape(-1, 3)
        self.tris = np.asarray(self.tris, np.int32).reshape(-1, 3)


@dataclass
class Emitter:
    """A mesh that emits particles: ``fill`` once at frame_start, ``inflow`` every frame in its range,
    ``substeps`` times per frame at evenly spaced times with the pose interpolated."""

    name: str
    kind: str
    material: int
    mesh: MeshData
    transforms: np.ndarray
    velocity: tuple[float, float, float] = (0.0, 0.0, 0.0)
    use_object_velocity: bool = False
    frame_start: int = 1
    frame_end: int = 1
    substeps: int = 1

    def __post_init__(self):
        if self.kind not in EMITTER_KINDS:
            raise ValueError(f"emitter {self.name!r}: kind must be one of {EMITTER_KINDS}")
        if self.substeps < 1 or self.substeps > MAX_SUBSTEPS:
            raise ValueError(f"emitter {self.name!r}: substeps must be in 1..{MAX_SUBSTEPS}")
        self.transforms = _matrices(self.transforms)
        self.velocity = _floats(self.velocity)


@dataclass
class Collider:
    name: str
    mesh: MeshData
    transforms: np.ndarray
    friction: float = 0.5
    closed: bool = True

    def __post_init__(self):
        self.transforms = _matrices(self.transforms)


@dataclass
class SceneDescription:
    domain: Domain
    materials: list[MaterialParams]
    emitters: list[Emitter]
    colliders: list[Collider]
    gravity: tuple[float, float, float] = (0.0, 0.0, -9.81)
    fps: float = 24.0
    frame_start: int = 1
    frame_end: int = 250
    seed: int = 0
    particles_per_cell: int = 8

    def __post_init__(self):
        self.gravity = _floats(self.gravity)
        if self.frame_start > self.frame_end:
            raise ValueError("frame_end must be >= frame_start")
        side = round(self.particles_per_cell ** (1.0 / 3.0))
        if side < 1 or side * side * side != self.particles_per_cell:
            raise ValueError("particles_per_cell must be a perfect cube (1, 8, 27, ...)")
        frames = self.frame_count
        for obj in self.emitters + self.colliders:
            if len(obj.transforms) != frames:
                raise ValueError(f"{obj.name!r} has {len(obj.transforms)} transforms, expected {frames} "
                                 f"(frames {self.frame_start}..{self.frame_end})")
        for e in self.emitters:
            if e.material < 0 or e.material >= len(self.materials):
                raise ValueError(f"emitter {e.name!r} references missing material {e.material}")

    @property
    def frame_count(self) -> int:
        return self.frame_end - self.frame_start + 1


# ---- folder I/O --------------------------------------------------------------------------------------

def _json_value(v):
    return list(v) if isinstance(v, tuple) else v


def _arrays_of(obj) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    return obj.mesh.verts, obj.mesh.tris, obj.transforms


def _write_objects(directory: Path, kind: str, objects, keys) -> list[dict]:
    docs = []
    for i, obj in enumerate(objects):
        for suffix, array in zip(_ARRAYS, _arrays_of(obj)):
            np.save(directory / f"{kind}_{i}_{suffix}.npy", array)
        docs.append({k: _json_value(getattr(obj, k)) for k in keys})
    return docs


def _read_objects(directory: Path, kind: str, docs: list[dict], cls) -> list:
    objects = []
    for i, doc in enumerate(docs):
        verts, tris, xf = (np.load(directory / f"{kind}_{i}_{suffix}.npy") for suffix in _ARRAYS)
        objects.append(cls(mesh=MeshData(verts, tris), transforms=xf, **doc))
    return objects


def write_scene(scene: SceneDescription, directory: Path) -> None:
    """Write the scene folder; arrays of an earlier scene in the same folder are removed first."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    for old in directory.glob("*.npy"):
        old.unlink()
    d = scene.domain
    doc = {
        "version": FORMAT_VERSION,
        "domain": {"min": list(d.min), "max": list(d.max), "resolution": d.resolution, "walls": d.walls,
                   "wall_friction": d.wall_friction},
        "materials": [m.to_dict() for m in scene.materials],
        "emitters": _write_objects(directory, "emitter", scene.emitters, _EMITTER_KEYS),
        "colliders": _write_objects(directory, "collider", scene.colliders, _COLLIDER_KEYS),
    }
    doc.update((k, _json_value(getattr(scene, k))) for k in _SCENE_KEYS)
    directory.joinpath("scene.json").write_text(json.dumps(doc, indent=2), encoding="utf-8")


def read_scene(directory: Path) -> SceneDescription:
    directory = Path(directory)
    doc = json.loads(directory.joinpath("scene.json").read_text(encoding="utf-8"))
    if doc.get("version") != FORMAT_VERSION:
        raise ValueError(f"unsupported scene format version {doc.get('version')!r}")
    dom = doc["domain"]
    return SceneDescription(
        domain=Domain(dom["min"], dom["max"], dom["resolution"], dom["walls"], dom["wall_friction"]),
        materials=[MaterialParams.from_dict(m) for m in doc["materials"]],
        emitters=_read_objects(directory, "emitter", doc["emitters"], Emitter),
        colliders=_read_objects(directory, "collider", doc["colliders"], Collider),
        **{k: doc[k] for k in _SCENE_KEYS},
    )


def scene_hash(directory: Path) -> str:
    """Identity of a scene folder for checkpoint resume: sha256 over scene.json and then every .npy file in
    sorted name order, each file's name hashed before its bytes (the reference's, scene.py:222-229, so a
    checkpoint's recorded hash means the same on either side)."""
    directory = Path(directory)
    h = hashlib.sha256()
    for path in [directory / "scene.json"] + sorted(directory.glob("*.npy")):
        h.update(path.name.encode())
        h.update(path.read_bytes())
    return h.hexdigest()
