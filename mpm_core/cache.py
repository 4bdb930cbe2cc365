"""The bake cache folder: what the runner writes and the Blender add-on and the surfacer read.

Files and formats are an interface and follow the reference (src/mpm_core/cache.py:1-19, 121-177):
``cache.json`` {"version": 1, "emitters", "fps", "frame_start", "frame_end", "dx", "backend"};
``frames/frame_NNNN.npz`` with ``pos_i`` / ``vel_i`` (float32, particle order) per emitter i;
``state/checkpoint.npz``, the solver state of the last finished frame plus ``__frame`` and
``__scene_hash``; ``progress.json``; a ``cancel`` file the runner removes when it stops; ``surface/``.

Readers may open any file while a bake runs, so nothing is written in place: every file is produced under
a temporary name in its own folder and renamed over the target, which is atomic. On Windows a reader that
holds the target open makes the rename fail for a moment; the rename is retried for about a second.
"""

from __future__ import annotations

import json
import os
import sys
import tempfile
import threading
import time
from contextlib import contextmanager
from pathlib import Path

import numpy as np

from mpm_core.solver import ParticleData

FORMAT_VERSION = 1
RENAME_PATIENCE_S = 1.0  # how long a rename over a file a reader holds open is retried


def _replace_patiently(src: str, dst: Path) -> None:
    deadline = time.monotonic() + RENAME_PATIENCE_S
    while True:
        try:
            os.replace(src, dst)
            return
        except PermissionError:
            if time.monotonic() >= deadline:
                os.unlink(src)
                raise
            time.sleep(0.02)


def _publish(path: Path, fill) -> None:
    """Write a file by calling ``fill(binary file)`` on a fresh temporary sibling, then rename it over
    ``path``. The temporary name is unique to this writer, so concurrent writers never share one."""
    path = Path(path)
    fd, tmp = tempfile.mkstemp(prefix=f"{path.name}.", suffix=".tmp", dir=path.parent)
    try:
        with os.fdopen(fd, "wb") as f:
            fill(f)
    except BaseException:
        os.unlink(tmp)
        raise
    _replace_patiently(tmp, path)


def atomic_write_npz(path: Path, arrays: dict[str, np.ndarray]) -> None:
    _publish(path, lambda f: np.savez(f, **arrays))


def write_json(path: Path, doc: dict) -> None:
    text = json.dumps(doc, indent=2).encode("utf-8")
    _publish(path, lambda f: f.write(text))


# ---- process plumbing used by the command-line runners ---------------------------------------------

def stdin_closed() -> threading.Event:
    """An event that is set once standard input reaches end of file, i.e. once the process that started
    this one with a pipe on stdin has gone (the add-on's way to stop an orphaned bake). The watcher reads
    the raw descriptor: a thread parked in a buffered ``sys.stdin`` read blocks interpreter shutdown."""
    gone = threading.Event()

    def drain() -> None:
        try:
            fd = sys.stdin.fileno()
            while os.read(fd, 1 << 12):
                pass
        except (OSError, ValueError, AttributeError):
            pass
        gone.set()

    threading.Thread(target=drain, name="stdin-watch", daemon=True).start()
    return gone


def _try_lock(f) -> None:
    if os.name == "nt":
        import msvcrt

        f.seek(0)
        msvcrt.locking(f.fileno(), msvcrt.LK_NBLCK, 1)
    else:
        import fcntl

        fcntl.flock(f.fileno(), fcntl.LOCK_EX | fcntl.LOCK_NB)


def _unlock(f) -> None:
    if os.name == "nt":  # flock locks go with the descriptor; msvcrt byte locks must be released
        import msvcrt

        f.seek(0)
        msvcrt.locking(f.fileno(), msvcrt.LK_UNLCK, 1)


@contextmanager
def exclusive_lock(path: Path):
    """Hold an operating-system lock on ``path`` (dropped with the process if it dies). RuntimeError when
    another holder, in this process or another, has it already."""
    with open(path, "a+b") as f:
        try:
            _try_lock(f)
        except OSError as exc:
            raise RuntimeError(f"{path} is locked: another process is already running here") from exc
        try:
            yield
        finally:
            _unlock(f)


# ---- frames ------------------------------------------------------------------------------------------

def frame_path(cache_dir: Path, frame: int) -> Path:
    return Path(cache_dir, "frames", "frame_%04d.npz" % frame)


def frame_arrays(emitter_count: int, particles: ParticleData) -> dict[str, np.ndarray]:
    """One frame's npz members: positions and velocities of each emitter's particles, in particle order.
    A stable sort by emitter id groups them in one pass instead of one mask per emitter."""
    ids = np.asarray(particles.emitter_ids)
    order = np.argsort(ids, kind="stable")
    bounds = np.searchsorted(ids[order], np.arange(emitter_count + 1))
    pos = np.asarray(particles.positions, np.float32)[order]
    vel = np.asarray(particles.velocities, np.float32)[order]
    out: dict[str, np.ndarray] = {}
    for i in range(emitter_count):
        part = slice(bounds[i], bounds[i + 1])
        out[f"pos_{i}"] = np.ascontiguousarray(pos[part])
        out[f"vel_{i}"] = np.ascontiguousarray(vel[part])
    return out


class CacheWriter:
    """Writes ``cache.json`` on creation (the frames folder made if missing), then frames as they finish."""

    def __init__(self, cache_dir: Path, emitter_names: list[str], fps: float, frame_start: int, frame_end: int,
                 dx: float, backend: str):
        self.cache_dir = Path(cache_dir)
        self.emitter_names = list(emitter_names)
        self.cache_dir.joinpath("frames").mkdir(parents=True, exist_ok=True)
        header = dict(version=FORMAT_VERSION, emitters=self.emitter_names, fps=fps, frame_start=frame_start,
                      frame_end=frame_end, dx=dx, backend=backend)
        write_json(self.cache_dir / "cache.json", header)

    def write_frame(self, frame: int, particles: ParticleData) -> None:
        atomic_write_npz(frame_path(self.cache_dir, frame), frame_arrays(len(self.emitter_names), particles))


class CacheReader:
    """The header of a cache and its frames."""

    def __init__(self, cache_dir: Path):
        self.cache_dir = Path(cache_dir)
        header = json.loads(self.cache_dir.joinpath("cache.json").read_text(encoding="utf-8"))
        if header.get("version") != FORMAT_VERSION:
            raise ValueError(f"unsupported cache format version {header.get('version')!r}")
        self.emitter_names: list[str] = list(header["emitters"])
        self.fps: float = header["fps"]
        self.frame_start: int = header["frame_start"]
        self.frame_end: int = header["frame_end"]
        self.dx: float = header["dx"]

    def available_frames(self) -> list[int]:
        names = (p.name for p in self.cache_dir.joinpath("frames").glob("frame_*.npz"))
        return sorted(int(name[len("frame_"):-len(".npz")]) for name in names)

    def read_frame(self, frame: int) -> dict[str, tuple[np.ndarray, np.ndarray]]:
        with np.load(frame_path(self.cache_dir, frame)) as npz:
            return {name: (npz["pos_%d" % i], npz["vel_%d" % i]) for i, name in enumerate(self.emitter_names)}


# ---- checkpoint ------------------------------------------------------------------------------------

CHECKPOINT = Path("state", "checkpoint.npz")


def write_checkpoint(cache_dir: Path, frame: int, state: dict[str, np.ndarray], scene_hash: str) -> None:
    path = Path(cache_dir) / CHECKPOINT
    path.parent.mkdir(parents=True, exist_ok=True)
    members = dict(state)
    members["__frame"] = np.array(frame)
    members["__scene_hash"] = np.array(scene_hash)
    atomic_write_npz(path, members)


def read_checkpoint(cache_dir: Path) -> tuple[int, dict[str, np.ndarray], str] | None:
    """(frame, state, scene hash) of the last checkpoint, or None when there is none."""
    path = Path(cache_dir) / CHECKPOINT
    if not path.is_file():
        return None
    with np.load(path) as npz:
        meta = {k for k in npz.files if k.startswith("__")}
        return int(npz["__frame"]), {k: npz[k] for k in npz.files if k not in meta}, str(npz["__scene_hash"])
