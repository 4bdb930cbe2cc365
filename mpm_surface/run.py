"""The surface runner: particle cache in, mesh cache out; a re-run re-meshes only what changed
(reference src/mpm_surface/run.py:1-130).

Progress goes to ``<cache>/surface/progress.json`` (the bake runner's schema); creating
``<cache>/surface/cancel`` stops the run after the current frame. One run per cache at a time
(``<cache>/surface/lock``): a second raises RuntimeError. A fresh bake deletes ``surface/``, so stop the
surfacer before starting one.
"""

from __future__ import annotations

import hashlib
import json
import time
from collections.abc import Callable
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from mpm_core.cache import CacheReader, exclusive_lock, write_json
from mpm_surface.field import surface_particles
from mpm_surface.meshcache import MeshCacheReader, MeshCacheWriter
from mpm_surface.settings import SurfaceSettings, read_settings


@dataclass
class SurfaceResult:
    status: str  # "done" | "cancelled" | "error"
    frames_written: int
    frames_skipped: int
    message: str = ""


def particle_key(positions: np.ndarray, settings: SurfaceSettings) -> str:
    """What a mesh was made from: the positions' bytes and the settings (run.py:33-34)."""
    return hashlib.sha1(np.ascontiguousarray(positions).tobytes()).hexdigest() + ":" + settings.hash()


def surface(cache_dir: Path, *, settings: dict[str, SurfaceSettings] | None = None, frames: list[int] | None = None,
            device: str | None = None, should_cancel: Callable[[], bool] | None = None) -> SurfaceResult:
    """Mesh every particle frame of the cache (or those in ``frames``) for every enabled emitter
    (run.py:37-50). ``settings`` replaces the settings file when given."""
    cache_dir = Path(cache_dir)
    surf_dir = cache_dir / "surface"
    surf_dir.mkdir(parents=True, exist_ok=True)
    try:
        lock = exclusive_lock(surf_dir / "lock")
        lock.__enter__()
    except RuntimeError as exc:
        raise RuntimeError(f"a surfacer run is already running on {cache_dir}") from exc
    try:
        return _run(cache_dir, surf_dir, settings, frames, device, should_cancel)
    finally:
        lock.__exit__(None, None, None)


def _run(cache_dir: Path, surf_dir: Path, settings, frames, device, should_cancel) -> SurfaceResult:
    """run.py:53-130."""
    cancel_file = surf_dir / "cancel"
    cancel_file.unlink(missing_ok=True)  # left over from an earlier run
    doc = {"status": "starting", "frame": None, "frame_end": None, "particles": 0, "ms_per_frame": None,
           "message": "", "warnings": []}

    def update(**fields):
        doc.update(fields)
        write_json(surf_dir / "progress.json", doc)

    def cancelled() -> bool:
        if cancel_file.exists():
            cancel_file.unlink()
            return True
        return bool(should_cancel and should_cancel())

    update()
    written = skipped = 0
    try:
        reader = CacheReader(cache_dir)
        scene = json.loads((cache_dir / "scene" / "scene.json").read_text(encoding="utf-8"))
        spacing = reader.dx / round(scene["particles_per_cell"] ** (1 / 3))
        settings = read_settings(cache_dir) if settings is None else settings
        doc["warnings"].extend(f"surface settings for unknown emitter {name} ignored"
                               for name in settings if name not in reader.emitter_names)
        resolved = {name: settings.get(name, SurfaceSettings()) for name in reader.emitter_names}
        writer = MeshCacheWriter(cache_dir, reader.emitter_names)
        active = []
        for name, s in resolved.items():
            if s.enabled:
                active.append(name)
            else:
                writer.remove_emitter(name)  # the add-on shows the particles instead

        # meshes whose particle frame is gone (a shorter frame range, an interrupted re-bake)
        particle_frames = set(reader.available_frames())
        mesh_reader = MeshCacheReader(cache_dir)
        for name in active:
            for frame in set(mesh_reader.available_frames(name)) - particle_frames:
                writer.remove_frame(name, frame)

        todo = reader.available_frames()
        if frames is not None:
            wanted = set(frames)
            todo = [f for f in todo if f in wanted]
        doc["frame_end"] = todo[-1] if todo else None
        timings = []
        for index, frame in enumerate(todo):
            t0 = time.perf_counter()
            try:
                data = reader.read_frame(frame)
            except FileNotFoundError:
                doc["warnings"].append(f"particle frame {frame} disappeared during surfacing (bake restarted?); "
                                       f"skipped")
                continue
            meshed = False
            for name in active:
                positions, velocities = data[name]
                key = particle_key(positions, resolved[name])
                if writer.key(name, frame) == key:
                    continue
                writer.write(name, frame, surface_particles(positions, velocities, spacing, resolved[name], device),
                             key)
                meshed = True
            if meshed:
                written += 1
                timings.append((time.perf_counter() - t0) * 1000)
            else:
                skipped += 1
            update(status="running", frame=frame, particles=int(sum(len(data[n][0]) for n in active)),
                   ms_per_frame=round(float(np.mean(timings)), 1) if timings else None)
            if index < len(todo) - 1 and cancelled():
                update(status="cancelled")
                return SurfaceResult("cancelled", written, skipped)
        update(status="done")
        return SurfaceResult("done", written, skipped)
    except Exception as exc:
        update(status="error", message=f"{type(exc).__name__}: {exc}")
        raise
