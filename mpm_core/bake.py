"""Bake runner: a scene folder in, a cache folder out.

    python -m mpm_core.bake CACHE_DIR [--backend hip] [--device cuda:0] [--resume] [--watch-stdin]

Reads the scene in ``CACHE_DIR/scene`` and writes frames, checkpoints and ``progress.json`` into CACHE_DIR in the
formats of mpm_core.cache. Exit status 0 when the bake is done or cancelled, 1 on error.

A bake's frames depend on what the runner does between solver calls, so that is the reference runner's
(src/mpm_core/bake.py:56-244):

* the timeline: the first frame only emits; every later frame is advanced from the previous one in pieces
  that end where inflow emitters emit (``frame_pieces``), the colliders' poses interpolated over each piece;
* emission: jittered lattice points of the emitter's mesh at its interpolated pose, clipped to the domain,
  inflow points dropped where material already is, velocities from the emitter (plus the object's own
  motion when asked);
* the stops: NaN/inf in the particles, a ``cancel`` file or callback, errors into progress.json;
* resume from the last checkpoint when the scene hash matches.

The same timeline and emission drive mpm_core.serve, which runs many bakes at once. The default backend is
the HIP solver.
"""

from __future__ import annotations

import argparse
import shutil
import sys
import time
from collections.abc import Callable
from dataclasses import dataclass
from fractions import Fraction
from pathlib import Path

import numpy as np

from mpm_core.backends import create_solver
from mpm_core.cache import CacheWriter, read_checkpoint, write_checkpoint
from mpm_core.geometry import interpolate_transform, mesh_is_closed, sample_mesh_volume, transform_points
from mpm_core.jobs import FolderJob, ProgressFile, command_line
from mpm_core.scene import Emitter, SceneDescription, read_scene, scene_hash
from mpm_core.solver import ColliderGeometry, Solver

DEFAULT_BACKEND = "hip"
FRESH_SUBFOLDERS = ("frames", "state", "surface")  # removed when a bake starts from scratch


@dataclass
class BakeResult:
    status: str  # "done" | "cancelled" | "error"
    frames_written: int
    message: str = ""


class SceneChangedError(RuntimeError):
    pass


def require_closed_emitters(scene: SceneDescription) -> None:
    for e in scene.emitters:
        if not mesh_is_closed(e.mesh.tris):
            raise ValueError(f"emitter {e.name!r}: mesh is not closed (every edge must be shared by exactly "
                             f"two triangles; in Blender, try Merge by Distance)")


# ---- the timeline ----------------------------------------------------------------------------------------

@dataclass(frozen=True)
class Piece:
    """One ``Solver.advance`` of a frame: from fraction ``start`` of the frame to ``end``, after which the
    inflow emitters in ``emitters`` emit (in emitter order)."""

    start: Fraction
    end: Fraction
    emitters: tuple[int, ...] = ()


def _inflow_on(e: Emitter, frame: int) -> bool:
    return e.kind == "inflow" and e.frame_start <= frame <= e.frame_end


def first_frame_emitters(scene: SceneDescription) -> list[int]:
    """Emitters that emit on the scene's first frame, at the frame itself: every fill emitter and every
    inflow emitter whose range holds that frame (bake.py:81-90)."""
    return [i for i, e in enumerate(scene.emitters) if e.kind == "fill" or _inflow_on(e, scene.frame_start)]


def frame_pieces(scene: SceneDescription, frame: int) -> list[Piece]:
    """The pieces that advance ``frame - 1`` to ``frame`` (a frame after the first). An inflow emitter with m
    substeps emits at j/m, j = 1..m, in every frame of its range but its first, where it emits once, at the
    frame itself (j = m). Pieces are cut at every such time; the last ends at 1, with or without emissions
    (bake.py:93-113, 175-190)."""
    due: dict[Fraction, list[int]] = {Fraction(1): []}
    for i, e in enumerate(scene.emitters):
        if _inflow_on(e, frame):
            for j in range(1 if frame > e.frame_start else e.substeps, e.substeps + 1):
                due.setdefault(Fraction(j, e.substeps), []).append(i)
    ends = sorted(due)
    return [Piece(start, end, tuple(due[end])) for start, end in zip([Fraction(0)] + ends[:-1], ends)]


def pose_at(transforms: np.ndarray, k: int, s: Fraction) -> np.ndarray:
    """World matrix a fraction ``s`` of the way from frame index k-1 to k; on frame index 0, that frame's."""
    return interpolate_transform(transforms[k - 1], transforms[k], float(s)) if k > 0 else transforms[k]


def collider_transforms(scene: SceneDescription, frame: int, s0: Fraction, s1: Fraction):
    """(start, end) matrices of every collider for the piece from ``frame - 1 + s0`` to ``frame - 1 + s1``."""
    k = frame - scene.frame_start
    return tuple(np.stack([pose_at(c.transforms, k, s) for c in scene.colliders]) for s in (s0, s1))


# ---- emission --------------------------------------------------------------------------------------------

def emission_seed(scene: SceneDescription, emitter_index: int, frame: int, s: Fraction):
    """Seed of an emission's jitter: ``seed + i`` for a fill emitter (it emits once); for an inflow emitter
    ``[seed, frame, i, s.numerator, s.denominator]``, one stream per emission (bake.py:63-64)."""
    if scene.emitters[emitter_index].kind == "fill":
        return scene.seed + emitter_index
    return [scene.seed, frame, emitter_index, s.numerator, s.denominator]


def emitter_points(scene: SceneDescription, emitter_index: int, frame: int, s: Fraction,
                   progress: ProgressFile | None = None, sampler=sample_mesh_volume) -> np.ndarray:
    """Points of one emission at ``frame - 1 + s``: the emitter's mesh at its interpolated pose sampled on a
    jittered lattice of spacing dx / cbrt(particles per cell), points outside the domain dropped, each
    condition reported once as a warning (bake.py:56-74). ``sampler`` has the signature and results of
    mpm_core.geometry.sample_mesh_volume (the HIP solver offers one that runs on the GPU)."""
    e = scene.emitters[emitter_index]
    spacing = scene.domain.dx / round(scene.particles_per_cell ** (1.0 / 3.0))
    world = transform_points(pose_at(e.transforms, frame - scene.frame_start, s), e.mesh.verts)
    pts = sampler(world, e.mesh.tris, spacing, emission_seed(scene, emitter_index, frame, s))
    note = progress.warn if progress is not None else (lambda message: None)
    if len(pts) == 0:
        note(f"{e.name}: no particles, the mesh is smaller than the particle spacing ({spacing:.4g} m)")
        return pts
    inside = np.logical_and(pts >= np.asarray(scene.domain.min), pts <= np.asarray(scene.domain.max)).all(axis=1)
    if not inside.all():
        note(f"{e.name}: particles outside the domain were dropped")
        pts = pts[inside]
    return pts


def emit(scene: SceneDescription, solver: Solver, emitter_index: int, frame: int, s: Fraction,
         progress: ProgressFile) -> int:
    """One emission into ``solver`` at ``frame - 1 + s`` (the frame itself on the scene's first frame);
    returns the number of particles added."""
    e = scene.emitters[emitter_index]
    pts = emitter_points(scene, emitter_index, frame, s, progress,
                         getattr(solver, "sample_mesh_volume", sample_mesh_volume))
    if e.kind == "inflow" and len(pts) and solver.particle_count:
        pts = pts[~solver.is_occupied(pts)]  # no new material where material already is
    n = len(pts)
    if n == 0:
        return 0
    vel = np.repeat(np.asarray(e.velocity, np.float32)[None], n, axis=0)
    k = frame - scene.frame_start
    if e.use_object_velocity and len(e.transforms) > 1:
        # the object's motion over the frame at the points' local positions; a forward difference on frame 0
        a, b = (k - 1, k) if k > 0 else (0, 1)
        local = transform_points(np.linalg.inv(pose_at(e.transforms, k, s)), pts)
        vel += (transform_points(e.transforms[b], local) - transform_points(e.transforms[a], local)) * scene.fps
    solver.add_particles(pts, vel, np.full(n, e.material, np.int32), np.full(n, emitter_index, np.int32))
    return n


# ---- the runner ------------------------------------------------------------------------------------------

class _Runner:
    """One bake of one cache folder, inside its FolderJob."""

    def __init__(self, job: FolderJob, backend: str, device: str | None):
        self.job = job
        self.dir = job.folder
        self.progress = job.progress
        self.backend = backend
        self.device = device
        self.written = 0

    def _open(self, resume: bool) -> int:
        """Read the scene, make the solver and the writer; returns the first frame to produce."""
        scene_dir = self.dir / "scene"
        self.scene = scene = read_scene(scene_dir)
        self.hash = scene_hash(scene_dir)
        self.progress.doc["frame_end"] = scene.frame_end
        require_closed_emitters(scene)
        self.solver = create_solver(self.backend, domain=scene.domain, materials=scene.materials,
                                    gravity=scene.gravity, particles_per_cell=scene.particles_per_cell,
                                    device=self.device)
        self.solver.set_colliders([ColliderGeometry(c.mesh, c.closed, c.friction) for c in scene.colliders])
        self.writer = CacheWriter(self.dir, [e.name for e in scene.emitters], scene.fps, scene.frame_start,
                                  scene.frame_end, scene.domain.dx, self.backend)
        if not resume:
            for sub in FRESH_SUBFOLDERS:  # including the previous simulation's surface meshes
                shutil.rmtree(self.dir / sub, ignore_errors=True)
            (self.dir / "frames").mkdir()
            return scene.frame_start
        saved = read_checkpoint(self.dir)
        if saved is None:
            raise FileNotFoundError(f"no checkpoint to resume from in {self.dir}")
        last, state, saved_hash = saved
        if saved_hash != self.hash:
            raise SceneChangedError("the scene changed since this bake was started; bake from scratch")
        self.solver.load_state(state)
        return last + 1

    def _produce(self, frame: int) -> None:
        scene, solver = self.scene, self.solver
        if frame == scene.frame_start:
            for i in first_frame_emitters(scene):
                emit(scene, solver, i, frame, Fraction(1), self.progress)
            return
        for piece in frame_pieces(scene, frame):
            if scene.colliders:
                solver.set_collider_transforms(*collider_transforms(scene, frame, piece.start, piece.end))
            solver.advance(float(piece.end - piece.start) / scene.fps)
            for i in piece.emitters:
                emit(scene, solver, i, frame, piece.end, self.progress)

    def run(self, resume: bool) -> BakeResult:
        with self.job.recording_errors():
            first = self._open(resume)
            ms = []
            for frame in range(first, self.scene.frame_end + 1):
                t0 = time.perf_counter()
                self._produce(frame)
                particles = self.solver.get_particles()
                if not (np.isfinite(particles.positions).all() and np.isfinite(particles.velocities).all()):
                    message = f"simulation became unstable at frame {frame} (NaN/inf in particle positions)"
                    self.progress.update(status="error", frame=frame, message=message)
                    return BakeResult("error", self.written, message)
                self.writer.write_frame(frame, particles)
                write_checkpoint(self.dir, frame, self.solver.save_state(), self.hash)
                self.written += 1
                ms.append(1000.0 * (time.perf_counter() - t0))
                self.progress.update(status="running", frame=frame, particles=len(particles.positions),
                                     ms_per_frame=round(float(np.mean(ms)), 1))
                if frame < self.scene.frame_end and self.job.cancel_requested():
                    self.progress.update(status="cancelled")
                    return BakeResult("cancelled", self.written)
            self.progress.update(status="done")
            return BakeResult("done", self.written)


def bake(cache_dir: Path, *, backend: str = DEFAULT_BACKEND, device: str | None = None, resume: bool = False,
         should_cancel: Callable[[], bool] | None = None) -> BakeResult:
    """Bake ``cache_dir``. One bake at a time per cache folder: RuntimeError while another holds it.
    ``should_cancel`` is asked after every frame but the last, when no ``cancel`` file is there."""
    cache_dir = Path(cache_dir)
    with FolderJob(cache_dir, "bake.lock", f"a bake is already running on {cache_dir}", should_cancel) as job:
        return _Runner(job, backend, device).run(resume)


def main(argv: list[str] | None = None) -> int:
    def options(ap: argparse.ArgumentParser) -> None:
        ap.add_argument("--backend", default=DEFAULT_BACKEND)
        ap.add_argument("--resume", action="store_true")

    def run(args, parent_gone):
        progress_file = args.cache_dir / "progress.json"
        seen = {"mtime": None}

        def poll() -> bool:
            """Echo progress.json as one line whenever it changed; cancel once the parent has gone."""
            if progress_file.exists():
                mtime = progress_file.stat().st_mtime
                if mtime != seen["mtime"]:
                    seen["mtime"] = mtime
                    print(" ".join(progress_file.read_text(encoding="utf-8").splitlines()), flush=True)
            return parent_gone is not None and parent_gone()

        return bake(args.cache_dir, backend=args.backend, device=args.device, resume=args.resume,
                    should_cancel=poll)

    return command_line("python -m mpm_core.bake", __doc__, argv, options, run,
                        lambda r: f"{r.status}: {r.frames_written} frames written. {r.message}".strip())


if __name__ == "__main__":
    sys.exit(main())
