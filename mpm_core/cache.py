"""Bake cache on disk; backend-independent, read directly by the Blender add-on.

Layout and file formats are the reference's (src/mpm_core/cache.py:1-19):

    scene/                      scene description (mpm_core.scene), written by the add-on
    cache.json                  {"version": 1, "emitters", "fps", "frame_start", "frame_end", "dx", "backend"}
    frames/frame_NNNN.npz       per emitter i: pos_i (n,3) float32, vel_i (n,3) float32
    state/checkpoint.npz        full solver state of the last finished frame + __frame, __scene_hash
    progress.json               {"status", "frame", "frame_end", "particles", "ms_per_frame", "message",
                                 "warnings"}
    cancel                      present: stop after the current frame (the runner deletes it)
    surface/                    derived meshes; a fresh bake deletes it

Every file goes to a uniquely named ``.tmp`` sibling first and is then renamed into place, so readers
never see a partial file. On Windows a reader holding the file open blocks the rename briefly; the
writer retries for about a second.
"""

from __future__ import annotations

import json
import os
import sys
import threading
import time
import uuid
from contextlib import contextmanager
from pathlib import Path

import numpy as np

from mpm_core.solver import ParticleData

FORMAT_VERSION = 1
_RENAME_TRIES = 50
_RENAME_WAIT_S = 0.02


def _atomic_write_bytes(path: Path, write) -> None:
    tmp = path.with_name(f"{path.name}.{os.getpid()}.{uuid.uuid4().hex[:8]}.tmp")
    with open(tmp, "wb") as f:
        write(f)
    for attempt in range(_RENAME_TRIES):
        try:
            os.replace(tmp, path)
            return
        except PermissionError:
            if attempt == _RENAME_TRIES - 1:
                tmp.unlink(missing_ok=True)
                raise
            time.sleep(_RENAME_WAIT_S)


def stdin_closed() -> threading.Event:
    """Event set when stdin reaches EOF, i.e. the parent that piped it has exited."""
    done = threading.Event()

    def watch():
        try:
            fd = sys.stdin.fileno()
            while os.read(fd, 4096):  # raw reads: a buffered read would block interpreter shutdown
                pass
        except (OSError, ValueError, AttributeError):
            pass
        done.set()

    threading.Thread(target=watch, daemon=True).start()
    return done


@contextmanager
def exclusive_lock(path: Path):
    """Hold an OS lock on ``path``; RuntimeError if anyone (any process or handle) holds it already."""
    f = open(path, "a+b")
    try:
        try:
            if os.name == "nt":
                import msvcrt

                f.seek(0)
                msvcrt.locking(f.fileno(), msvcrt.LK_NBLCK, 1)
            else:
                import fcntl

                fcntl.flock(f.fileno(), fcntl.LOCK_EX | fcntl.LOCK_NB)
        except OSError as exc:
            raise RuntimeError(f"{path} is locked: another process is already running here") from exc
        try:
            yield
        finally:
            if os.name == "nt":
                import msvcrt

                f.seek(0)
                msvcrt.locking(f.fileno(), msvcrt.LK_UNLCK, 1)
    finally:
        f.close()


def atomic_write_npz(path: Path, arrays: dict[str, np.ndarray]) -> None:
    _atomic_write_bytes(Path(path), lambda f: np.savez(f, **arrays))


def write_json(path: Path, doc: dict) -> None:
    _atomic_write_bytes(Path(path), lambda f: f.write(json.dumps(doc, indent=2).encode("utf-8")))


def frame_path(cache_dir: Path, frame: int) -> Path:
    return Path(cache_dir) / "frames" / f"frame_{frame:04d}.npz"


def frame_arrays(emitter_count: int, particles: ParticleData) -> dict[str, np.ndarray]:
    """The npz members of one frame: pos_i / vel_i per emitter, in particle order."""
    out = {}
    for i in range(emitter_count):
        sel = particles.emitter_ids == i
        out[f"pos_{i}"] = np.ascontiguousarray(particles.positions[sel], np.float32)
        out[f"vel_{i}"] = np.ascontiguousarray(particles.velocities[sel], np.float32)
    return out


class CacheWriter:
    def __init__(self, cache_dir: Path, emitter_names: list[str], fps: float, frame_start: int, frame_end: int,
                 dx: float, backend: str):
        self.cache_dir = Path(cache_dir)
        self.emitter_names = list(emitter_names)
        (self.cache_dir / "frames").mkdir(parents=True, exist_ok=True)
        write_json(self.cache_dir / "cache.json", {
            "version": FORMAT_VERSION, "emitters": self.emitter_names, "fps": fps, "frame_start": frame_start,
            "frame_end": frame_end, "dx": dx, "backend": backend})

    def write_frame(self, frame: int, particles: ParticleData) -> None:
        atomic_write_npz(frame_path(self.cache_dir, frame), frame_arrays(len(self.emitter_names), particles))


class CacheReader:
    def __init__(self, cache_dir: Path):
        self.cache_dir = Path(cache_dir)
        doc = json.loads((self.cache_dir / "cache.json").read_text(encoding="utf-8"))
        if doc.get("version") != FORMAT_VERSION:
            raise ValueError(f"unsupported cache format version {doc.get('version')!r}")
        self.emitter_names: list[str] = doc["emitters"]
        self.fps: float = doc["fps"]
        self.frame_start: int = doc["frame_start"]
        self.frame_end: int = doc["frame_end"]
        self.dx: float = doc["dx"]

    def available_frames(self) -> list[int]:
        return sorted(int(p.stem.split("_")[1]) for p in (self.cache_dir / "frames").glob("frame_*.npz"))

    def read_frame(self, frame: int) -> dict[str, tuple[np.ndarray, np.ndarray]]:
        with np.load(frame_path(self.cache_dir, frame)) as data:
            return {name: (data[f"pos_{i}"], data[f"vel_{i}"]) for i, name in enumerate(self.emitter_names)}


def _checkpoint_path(cache_dir: Path) -> Path:
    return Path(cache_dir) / "state" / "checkpoint.npz"


def write_checkpoint(cache_dir: Path, frame: int, state: dict[str, np.ndarray], scene_hash: str) -> None:
    path = _checkpoint_path(cache_dir)
    path.parent.mkdir(parents=True, exist_ok=True)
    atomic_write_npz(path, {**state, "__frame": np.array(frame), "__scene_hash": np.array(scene_hash)})


def read_checkpoint(cache_dir: Path) -> tuple[int, dict[str, np.ndarray], str] | None:
    path = _checkpoint_path(cache_dir)
    if not path.exists():
        return None
    with np.load(path) as data:
        state = {k: data[k] for k in data.files if not k.startswith("__")}
        return int(data["__frame"]), state, str(data["__scene_hash"])
