"""The surfacer's runner: a particle cache in, a mesh cache out; a second run re-meshes only what changed.

Behaviour follows the reference runner (src/mpm_surface/run.py:1-130): progress in
``<cache>/surface/progress.json`` (the bake runner's keys), ``<cache>/surface/cancel`` stops the run after the
current frame, one run per cache at a time (``<cache>/surface/lock``; a second raises RuntimeError). A mesh is
remade only when its key, the SHA-1 of its particle positions plus the settings fingerprint, changed. A fresh
bake deletes ``surface/``, so the surfacer should be stopped before one starts.
"""

from __future__ import annotations

import hashlib
import json
import time
from collections.abc import Callable
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from mpm_core.cache import CacheReader
from mpm_core.jobs import FolderJob
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
    """What a mesh is made from: SHA-1 of the positions' bytes, a colon, the settings fingerprint."""
    digest = hashlib.sha1(np.ascontiguousarray(positions).tobytes()).hexdigest()
    return f"{digest}:{settings.hash()}"


@dataclass
class _Plan:
    """What one run does: the particle cache, the lattice spacing, per-emitter settings, the emitters to
    surface and the frames to visit."""

    reader: CacheReader
    spacing: float
    settings: dict[str, SurfaceSettings]
    active: list[str]
    frames: list[int]


def _plan(cache_dir: Path, given: dict[str, SurfaceSettings] | None, frames: list[int] | None, progress) -> _Plan:
    reader = CacheReader(cache_dir)
    scene = json.loads((cache_dir / "scene" / "scene.json").read_text(encoding="utf-8"))
    spacing = reader.dx / round(scene["particles_per_cell"] ** (1.0 / 3.0))
    chosen = read_settings(cache_dir) if given is None else given
    for name in chosen:
        if name not in reader.emitter_names:
            progress.doc["warnings"].append(f"surface settings for unknown emitter {name} ignored")
    settings = {name: chosen.get(name, SurfaceSettings()) for name in reader.emitter_names}
    todo = reader.available_frames()
    if frames is not None:
        todo = sorted(set(todo).intersection(frames))
    return _Plan(reader, spacing, settings, [n for n in reader.emitter_names if settings[n].enabled], todo)


def _tidy(cache_dir: Path, plan: _Plan) -> MeshCacheWriter:
    """Lay out the mesh cache: drop the meshes of disabled emitters (the add-on then shows their particles) and
    meshes whose particle frame is gone (a shorter frame range, an interrupted re-bake)."""
    writer = MeshCacheWriter(cache_dir, plan.reader.emitter_names)
    for name in plan.reader.emitter_names:
        if name not in plan.active:
            writer.remove_emitter(name)
    present = set(plan.reader.available_frames())
    stored = MeshCacheReader(cache_dir)
    for name in plan.active:
        for frame in set(stored.available_frames(name)) - present:
            writer.remove_frame(name, frame)
    return writer


def surface(cache_dir: Path, *, settings: dict[str, SurfaceSettings] | None = None, frames: list[int] | None = None,
            device: str | None = None, should_cancel: Callable[[], bool] | None = None) -> SurfaceResult:
    """Mesh every particle frame of the cache, or those in ``frames``, for every enabled emitter; ``settings``
    replaces the settings file when given (run.py:37-50)."""
    cache_dir = Path(cache_dir)
    surf_dir = cache_dir / "surface"
    surf_dir.mkdir(parents=True, exist_ok=True)
    busy = f"a surfacer run is already running on {cache_dir}"
    with FolderJob(surf_dir, "lock", busy, should_cancel) as job, job.recording_errors():
        progress = job.progress
        plan = _plan(cache_dir, settings, frames, progress)
        writer = _tidy(cache_dir, plan)
        progress.doc["frame_end"] = plan.frames[-1] if plan.frames else None
        written = skipped = 0
        ms = []
        for n, frame in enumerate(plan.frames, start=1):
            t0 = time.perf_counter()
            try:
                particles = plan.reader.read_frame(frame)
            except FileNotFoundError:
                progress.doc["warnings"].append(
                    f"particle frame {frame} disappeared during surfacing (bake restarted?); skipped")
                continue
            changed = 0
            for name in plan.active:
                positions, velocities = particles[name]
                key = particle_key(positions, plan.settings[name])
                if writer.key(name, frame) != key:
                    mesh = surface_particles(positions, velocities, plan.spacing, plan.settings[name], device)
                    writer.write(name, frame, mesh, key)
                    changed += 1
            if changed:
                written += 1
                ms.append(1000.0 * (time.perf_counter() - t0))
            else:
                skipped += 1
            progress.update(status="running", frame=frame,
                            particles=int(sum(len(particles[name][0]) for name in plan.active)),
                            ms_per_frame=round(float(np.mean(ms)), 1) if ms else None)
            if n < len(plan.frames) and job.cancel_requested():
                progress.update(status="cancelled")
                return SurfaceResult("cancelled", written, skipped)
        progress.update(status="done")
        return SurfaceResult("done", written, skipped)
