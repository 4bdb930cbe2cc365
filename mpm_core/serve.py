"""Bake server: many bakes in flight on one GPU, continuously batched.

The reference bakes one scene per runner process (src/mpm_core/bake.py:116-208). A BakeServer keeps many
bakes in flight on one GPU instead:

* requests are admitted in arrival order while their estimated HBM footprint fits the budget and a slot
  is free; each admitted bake keeps its particle and grid state resident on the device (one HipSolver
  context) until it ends;
* every admitted bake belongs to one SceneBatch. A bake's frame is split into the pieces the runner
  passes to ``Solver.advance`` (the whole frame, or the stretch up to the next inflow emission,
  bake.py:100-113 and 175-190); a piece is started with ``SceneBatch.begin`` and runs, substep by
  substep, together with every other bake's piece in flight; pieces end independently, so a bake whose
  piece ended is serviced (emission, frame output, collider poses, next piece) while the others keep
  running. Each bake's sequence of advance, emission and pose calls is exactly the runner's, so its
  result is the runner's up to the order of floating-point atomic additions;
* finished frames are streamed into the bake's cache in the runner's format (mpm_core.cache): frames,
  checkpoints, progress; writer threads do the file work.

The server reports per-frame latency (time between a bake's consecutive finished frames; for the first
frame, from admission) at p50/p99, the queueing delay (arrival to admission), and the aggregate
particle-substeps/s. With ``batched=False`` and
``max_active=1`` it is the serial runner (one bake at a time, ``Solver.advance``), the baseline.

Latency target. Bakes in one batch share the device, so a bake's frame latency grows with the work of every
other bake in flight. With ``latency_target_ms`` set, the number of bakes admitted at once is capped and the
cap follows the measured latency: after every ``LATENCY_WINDOW`` finished frames the p99 of the last
``LATENCY_WINDOW`` frame latencies is compared with the target; above it the cap drops to
``LATENCY_BACKOFF`` of the bakes in flight (at least one), below ``LATENCY_HEADROOM`` of it the cap grows by
an eighth (at least one). The cap starts at ``LATENCY_START_CAP``. Requests beyond the cap wait in the queue
(their wait is reported apart); bakes already admitted are never stopped, so after a drop the cap bounds
p99 once the bakes admitted under the higher cap have finished.
"""

from __future__ import annotations

import shutil
import time
from collections import deque
from collections.abc import Callable, Sequence
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from fractions import Fraction
from pathlib import Path

import numpy as np

from mpm_core.bake import (Piece, collider_transforms, emit, first_frame_emitters, frame_pieces,
                            require_closed_emitters)
from mpm_core.cache import CacheWriter, write_checkpoint
from mpm_core.jobs import ProgressFile
from mpm_core.scene import SceneDescription, scene_hash
from mpm_core.solver import ColliderGeometry, ParticleData

# device bytes per particle slot (two particle sets of 32 float + 2 int planes, emitter ids, 5 sort
# arrays, the FP64 work list, the drift-routing flag), per grid node (two float4 arrays, cell offsets),
# per brick (flags, occupied list, slot map) and per occupied brick (the staged 6^3 core of its P2G window and the
# drift-routing counters and lists of its 64 cells); a scene holds min(bricks, particle slots) of those
BYTES_PER_SLOT = 2 * 34 * 4 + 4 + 6 * 4 + 1
BYTES_PER_NODE = 2 * 16 + 4
BYTES_PER_BRICK = 4 + 1 + 4 + 4
BYTES_PER_OCCUPIED_BRICK = 216 * 16 + 64 * 4 + 64 * 4 * 4
GROWTH = 1.5  # capacity grows by half when particles are added
LATENCY_WINDOW = 32     # finished frames between two adjustments of the admission cap
LATENCY_BACKOFF = 0.7   # the cap after a window whose p99 exceeded the target, as a share of bakes in flight
LATENCY_HEADROOM = 0.5  # the cap grows after a window whose p99 stayed below this share of the target
LATENCY_START_CAP = 8   # bakes admitted at once before the first latency window


def mesh_volume(verts: np.ndarray, tris: np.ndarray) -> float:
    """Enclosed volume of a closed triangle mesh (divergence theorem)."""
    v = np.asarray(verts, np.float64)
    a, b, c = v[tris[:, 0]], v[tris[:, 1]], v[tris[:, 2]]
    return abs(float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0)


def hbm_footprint(scene: SceneDescription) -> int:
    """Device bytes a bake of ``scene`` is expected to hold at its peak: grid arrays plus particle slots for
    every particle its emitters can add (fill: the mesh volume at the scene's particle density; inflow:
    that per emission, for every emission, at most twice the domain's fill)."""
    dims = [s + 6 for s in scene.domain.grid_shape]
    bricks = int(np.prod([(d + 3) // 4 for d in dims]))
    grid = bricks * 64 * BYTES_PER_NODE + bricks * BYTES_PER_BRICK
    density = scene.particles_per_cell / scene.domain.dx ** 3
    particles = 0.0
    for e in scene.emitters:
        scale = float(np.max([abs(np.linalg.det(np.asarray(m, np.float64)[:3, :3])) for m in e.transforms]))
        per = mesh_volume(e.mesh.verts, e.mesh.tris) * scale * density
        if e.kind == "inflow":
            per *= e.substeps * max(0, e.frame_end - e.frame_start + 1)
        particles += per
    cells = float(np.prod(scene.domain.grid_shape))
    particles = min(particles, 2.0 * cells * scene.particles_per_cell)
    slots = GROWTH * particles + 256
    return int(grid + BYTES_PER_SLOT * slots + BYTES_PER_OCCUPIED_BRICK * min(bricks, slots))


@dataclass
class BakeRequest:
    scene: SceneDescription
    cache_dir: Path | None = None  # frames, checkpoints, progress go here (None: only to on_frame)
    name: str = ""
    on_frame: Callable[[int, ParticleData], None] | None = None  # called on a writer thread
    footprint: int | None = None  # HBM bytes to reserve (default: hbm_footprint(scene))


@dataclass
class BakeHandle:
    """What the server reports about one request."""

    request: BakeRequest
    status: str = "queued"  # queued | running | done | cancelled | error
    frames_written: int = 0
    message: str = ""
    arrival: float = 0.0
    admitted: float | None = None
    finished: float | None = None
    frame_latency: list[float] = field(default_factory=list)  # seconds, per finished frame
    particle_substeps: int = 0
    cancel_requested: bool = False

    @property
    def done(self) -> bool:
        return self.status in ("done", "cancelled", "error")


class _Bake:
    """An admitted request: its solver and where it is in its frame loop."""

    def __init__(self, handle: BakeHandle, solver, writer, progress, hash_: str, footprint: int):
        self.handle = handle
        self.scene: SceneDescription = handle.request.scene
        self.solver = solver
        self.writer = writer
        self.progress = progress
        self.hash = hash_
        self.footprint = footprint
        self.frame = self.scene.frame_start  # the frame being produced
        self.pieces: list[Piece] = []
        self.piece = 0
        self.in_piece = False  # a piece was started in the batch and has not ended
        self.last_output = 0.0
        self.last_write: Future | None = None
        self.count_at_piece = 0
        self.timings: list[float] = []
        self.frame_t0 = 0.0


class BakeServer:
    def __init__(self, device: str = "cuda:0", hbm_budget: float = 250e9, max_active: int = 1024,
                 batched: bool = True, chunk: int = 4, checkpoint_every: int = 1, writer_threads: int = 4,
                 backend: str = "hip", latency_target_ms: float | None = None):
        from mpm_core.backends.hip.batch import SceneBatch
        from mpm_core.backends.hip.solver import device_ordinal

        self.device = device
        self.hbm_budget = float(hbm_budget)
        self.max_active = int(max_active)
        self.latency_target = None if latency_target_ms is None else float(latency_target_ms) / 1000.0
        # admission cap: all of max_active, or with a latency target a small start that then follows it
        self.cap = self.max_active if latency_target_ms is None else min(self.max_active, LATENCY_START_CAP)
        self.cap_log: list[tuple[float, int]] = [(time.perf_counter(), self.cap)]
        self.recent: deque[float] = deque(maxlen=LATENCY_WINDOW)
        self.since_adjust = 0
        self.batched = bool(batched)
        self.chunk = int(chunk)
        self.checkpoint_every = max(1, int(checkpoint_every))
        self.backend = backend
        self.queue: deque[BakeHandle] = deque()
        self.active: list[_Bake] = []
        self.finished: list[BakeHandle] = []
        self.reserved = 0
        self.pool = ThreadPoolExecutor(max_workers=max(1, writer_threads))
        self.batch = SceneBatch(device=device_ordinal(device)) if self.batched else None
        self.t_first: float | None = None
        self.t_last: float | None = None
        self.particle_substeps = 0
        self.launches = 0
        self.run_seconds = 0.0  # host time spent inside batch runs (device work and its synchronisation)
        self.work_log: list[tuple[float, int]] = []  # (time, particle-substeps so far) after each step

    # ---- requests -------------------------------------------------------------------------------------
    def submit(self, request: BakeRequest, arrival: float | None = None) -> BakeHandle:
        h = BakeHandle(request, arrival=time.perf_counter() if arrival is None else arrival)
        if self.t_first is None:
            self.t_first = h.arrival
        self.queue.append(h)
        return h

    def cancel(self, handle: BakeHandle) -> None:
        """Stop after the current frame (the runner's cancel); a queued request is dropped."""
        handle.cancel_requested = True
        if handle in self.queue:
            self.queue.remove(handle)
            handle.status = "cancelled"
            handle.finished = time.perf_counter()
            self.finished.append(handle)

    @property
    def idle(self) -> bool:
        return not self.queue and not self.active

    # ---- the loop -------------------------------------------------------------------------------------
    def run(self, arrivals: Sequence[tuple[float, BakeRequest]] = (), stop_after: float | None = None) -> list[BakeHandle]:
        """Submit each request at its offset (seconds from now) and serve until everything has ended, or
        until `stop_after` seconds have passed (what is still queued or in flight is then left as it is)."""
        t0 = time.perf_counter()
        pending = deque(sorted(arrivals, key=lambda a: a[0]))
        handles = []
        while pending or not self.idle:
            now = time.perf_counter()
            if stop_after is not None and now - t0 >= stop_after:
                break
            while pending and t0 + pending[0][0] <= now:
                off, req = pending.popleft()
                handles.append(self.submit(req, arrival=t0 + off))
            if self.idle and pending:
                time.sleep(max(0.0, t0 + pending[0][0] - time.perf_counter()))
                continue
            self.step()
        self.drain()
        return handles

    def step(self) -> None:
        """Admit what fits, then advance the bakes in flight: up to `chunk` substeps of every piece in flight
        (batched), or one piece of each bake in turn (serial)."""
        self._admit()
        if self.batched:
            if any(b.in_piece for b in self.active):
                t = time.perf_counter()
                launched, ended = self.batch.run(self.chunk)
                self.run_seconds += time.perf_counter() - t
                self.launches += launched
                for i in sorted(ended):
                    self._piece_ended(self.active[i], ended[i].substeps)
        else:
            for b in self.active:
                if b.handle.done:
                    continue
                try:
                    st = b.solver.advance(self._prepare_piece(b))
                except Exception as exc:  # noqa: BLE001
                    self._fail_bake(b, exc)
                    continue
                self._piece_ended(b, st.substeps)
        self._retire()
        self._admit()
        self._begin_pieces()
        self.work_log.append((time.perf_counter(), self.particle_substeps))

    def drain(self) -> None:
        """Wait for every file write in flight."""
        for b in self.active:
            self._wait_writes(b)

    def close(self) -> None:
        self.drain()
        if self.batch is not None:
            self.batch.set_scenes([])
        for b in self.active:
            _release(b.solver)
        self.active = []
        if self.batch is not None:
            self.batch.close()
        self.pool.shutdown(wait=True)

    # ---- admission ------------------------------------------------------------------------------------
    def _admit(self) -> None:
        added = False
        while self.queue and len(self.active) < min(self.max_active, self.cap):
            h = self.queue[0]
            fp = h.request.footprint if h.request.footprint is not None else hbm_footprint(h.request.scene)
            if self.reserved + fp > self.hbm_budget and self.active:
                break  # wait (in arrival order) until enough memory is free
            self.queue.popleft()
            b = self._start(h, fp)
            if b is None:
                h.finished = time.perf_counter()
                self.finished.append(h)
                continue
            self.active.append(b)
            self.reserved += fp
            added = True
        if added and self.batched:
            self.batch.set_scenes([b.solver for b in self.active])

    def _start(self, h: BakeHandle, footprint: int) -> _Bake | None:
        """Create the bake's solver and produce its first frame (emission only, as the runner does)."""
        from mpm_core.backends import create_solver

        req, scene = h.request, h.request.scene
        progress = ProgressFile(None)
        try:
            require_closed_emitters(scene)
            writer = None
            hash_ = ""
            if req.cache_dir is not None:
                cache_dir = Path(req.cache_dir)
                cache_dir.mkdir(parents=True, exist_ok=True)
                progress = ProgressFile(cache_dir / "progress.json")
                progress.doc["frame_end"] = scene.frame_end
                progress.update()
                if (cache_dir / "scene").exists():
                    hash_ = scene_hash(cache_dir / "scene")
                for sub in ("frames", "state", "surface"):
                    shutil.rmtree(cache_dir / sub, ignore_errors=True)
                (cache_dir / "frames").mkdir()
                writer = CacheWriter(cache_dir, [e.name for e in scene.emitters], scene.fps, scene.frame_start,
                                     scene.frame_end, scene.domain.dx, self.backend)
            solver = create_solver(self.backend, domain=scene.domain, materials=scene.materials,
                                   gravity=scene.gravity, particles_per_cell=scene.particles_per_cell,
                                   device=self.device)
            solver.set_colliders([ColliderGeometry(c.mesh, c.closed, c.friction) for c in scene.colliders])
        except Exception as exc:  # noqa: BLE001 - one bad request must not stop the server
            self._fail(h, progress, exc)
            return None
        h.status = "running"
        h.admitted = time.perf_counter()
        b = _Bake(h, solver, writer, progress, hash_, footprint)
        b.frame_t0 = h.admitted
        b.last_output = h.admitted
        try:
            for i in first_frame_emitters(scene):
                emit(scene, solver, i, scene.frame_start, Fraction(1), progress)
            self._frame_out(b)
            self._next_frame(b)
        except Exception as exc:  # noqa: BLE001
            self._fail_bake(b, exc)
        return b

    # ---- frames and pieces ----------------------------------------------------------------------------
    def _prepare_piece(self, b: _Bake) -> float:
        """Collider poses of the bake's next piece (bake.py:184-186); returns the piece's length."""
        piece = b.pieces[b.piece]
        if b.scene.colliders:
            b.solver.set_collider_transforms(*collider_transforms(b.scene, b.frame, piece.start, piece.end))
        b.count_at_piece = b.solver.particle_count
        return float(piece.end - piece.start) / b.scene.fps

    def _next_frame(self, b: _Bake) -> None:
        b.frame += 1
        b.frame_t0 = time.perf_counter()
        if b.frame > b.scene.frame_end:
            b.pieces = []
            b.handle.status = "done"
            self._chain(b, lambda prog=b.progress: prog.update(status="done"))
            return
        b.pieces = frame_pieces(b.scene, b.frame)
        b.piece = 0

    def _begin_pieces(self) -> None:
        """Start the next piece of every bake that has none in flight."""
        if not self.batched:
            return
        fdt = np.zeros(len(self.active))
        any_ = False
        for i, b in enumerate(self.active):
            if b.in_piece or b.handle.done or not b.pieces:
                continue
            try:
                fdt[i] = self._prepare_piece(b)
            except Exception as exc:  # noqa: BLE001
                self._fail_bake(b, exc)
                continue
            b.in_piece = True
            any_ = True
        if any_:
            self.batch.begin(fdt)

    def _piece_ended(self, b: _Bake, substeps: int) -> None:
        h = b.handle
        b.in_piece = False
        try:
            self.particle_substeps += b.count_at_piece * substeps
            h.particle_substeps += b.count_at_piece * substeps
            piece = b.pieces[b.piece]
            for i in piece.emitters:
                emit(b.scene, b.solver, i, b.frame, piece.end, b.progress)
            b.piece += 1
            if piece.end == 1:
                self._frame_out(b)
                if b.frame < b.scene.frame_end and self._cancelled(b):
                    h.status = "cancelled"
                    self._chain(b, lambda prog=b.progress: prog.update(status="cancelled"))
                    return
                self._next_frame(b)
        except Exception as exc:  # noqa: BLE001
            self._fail_bake(b, exc)

    def _cancelled(self, b: _Bake) -> bool:
        if b.handle.cancel_requested:
            return True
        cd = b.handle.request.cache_dir
        if cd is not None and (Path(cd) / "cancel").exists():
            (Path(cd) / "cancel").unlink()
            return True
        return False

    def _frame_out(self, b: _Bake) -> None:
        """The runner's end of frame (bake.py:191-207): NaN stop, frame file, checkpoint, progress."""
        h, frame = b.handle, b.frame
        particles = b.solver.get_particles()
        if not (np.isfinite(particles.positions).all() and np.isfinite(particles.velocities).all()):
            raise FloatingPointError(f"simulation became unstable at frame {frame} (NaN/inf in particle positions)")
        now = time.perf_counter()
        h.frame_latency.append(now - b.last_output)
        self._track_latency(now - b.last_output)
        b.last_output = now
        b.timings.append((now - b.frame_t0) * 1000.0)
        self.t_last = now
        last = frame == b.scene.frame_end
        state = None
        if b.writer is not None and ((frame - b.scene.frame_start) % self.checkpoint_every == 0 or last):
            state = b.solver.save_state()
        h.frames_written += 1
        ms = round(float(np.mean(b.timings)), 1)

        def write(frame=frame, particles=particles, state=state, ms=ms):
            if b.writer is not None:
                b.writer.write_frame(frame, particles)
                if state is not None:
                    write_checkpoint(b.writer.cache_dir, frame, state, b.hash)
            if h.request.on_frame is not None:
                h.request.on_frame(frame, particles)
            b.progress.update(status="running", frame=frame, particles=len(particles.positions), ms_per_frame=ms)

        self._chain(b, write)

    def _track_latency(self, seconds: float) -> None:
        """Adjust the admission cap to the latency target (see the module notes)."""
        if self.latency_target is None:
            return
        self.recent.append(seconds)
        self.since_adjust += 1
        if self.since_adjust < LATENCY_WINDOW:
            return
        self.since_adjust = 0
        p99 = float(np.percentile(np.asarray(self.recent), 99))
        if p99 > self.latency_target:
            cap = max(1, int(LATENCY_BACKOFF * len(self.active)))
        elif p99 < LATENCY_HEADROOM * self.latency_target:
            cap = min(self.max_active, self.cap + max(1, self.cap // 8))
        else:
            return
        if cap != self.cap:
            self.cap = cap
            self.cap_log.append((time.perf_counter(), cap))

    def _chain(self, b: _Bake, fn) -> None:
        """Run fn on a writer thread after the bake's previous write (its files stay in order)."""
        prev = b.last_write

        def task():
            if prev is not None:
                prev.result()
            fn()

        b.last_write = self.pool.submit(task)

    def _wait_writes(self, b: _Bake) -> None:
        if b.last_write is None:
            return
        try:
            b.last_write.result()
        except Exception as exc:  # noqa: BLE001 - a failed file write fails the bake
            b.handle.status = "error"
            b.handle.message = f"{type(exc).__name__}: {exc}"

    def _fail_bake(self, b: _Bake, exc: Exception) -> None:
        """An admitted bake failed: its error status goes after its pending file writes."""
        h = b.handle
        h.status = "error"
        h.message = f"{type(exc).__name__}: {exc}" if not isinstance(exc, FloatingPointError) else str(exc)
        self._chain(b, lambda prog=b.progress, msg=h.message, frame=b.frame:
                    prog.update(status="error", frame=frame, message=msg))

    def _fail(self, h: BakeHandle, progress, exc: Exception) -> None:
        h.status = "error"
        h.message = f"{type(exc).__name__}: {exc}"
        if progress is not None:
            try:
                progress.update(status="error", message=h.message)
            except OSError:
                pass

    # ---- retirement -----------------------------------------------------------------------------------
    def _retire(self) -> None:
        leaving = [b for b in self.active if b.handle.done]
        if not leaving:
            return
        self.active = [b for b in self.active if not b.handle.done]
        if self.batched:
            self.batch.set_scenes([b.solver for b in self.active])
        for b in leaving:
            self._wait_writes(b)
            b.handle.finished = time.perf_counter()
            self.finished.append(b.handle)
            self.reserved -= b.footprint
            _release(b.solver)

    # ---- report ---------------------------------------------------------------------------------------
    def work_between(self, t0: float, t1: float) -> int:
        """Particle-substeps of the pieces that ended between t0 and t1."""
        before = [w for t, w in self.work_log if t <= t0]
        upto = [w for t, w in self.work_log if t <= t1]
        return (upto[-1] if upto else 0) - (before[-1] if before else 0)

    def report(self, handles: Sequence[BakeHandle] | None = None) -> dict:
        """Latency percentiles over finished frames (first frames apart: they include queueing) and the
        aggregate particle-substeps/s from the first arrival to the last frame."""
        hs = list(handles) if handles is not None else list(self.finished)
        first = [h.frame_latency[0] for h in hs if h.frame_latency]
        waits = [h.admitted - h.arrival for h in hs if h.admitted is not None]
        later = [x for h in hs for x in h.frame_latency[1:]]
        span = (self.t_last - self.t_first) if self.t_first is not None and self.t_last is not None else 0.0

        def pct(v, q):
            return float(np.percentile(v, q)) * 1000.0 if v else None

        return {
            "bakes": len(hs),
            "done": sum(h.status == "done" for h in hs),
            "errors": sum(h.status == "error" for h in hs),
            "frames": sum(h.frames_written for h in hs),
            "frame_latency_ms": {"p50": pct(later, 50), "p99": pct(later, 99),
                                 "mean": float(np.mean(later)) * 1000.0 if later else None},
            "first_frame_ms": {"p50": pct(first, 50), "p99": pct(first, 99)},
            "queue_wait_ms": {"p50": pct(waits, 50), "p99": pct(waits, 99)},
            "particle_substeps": int(self.particle_substeps),
            "seconds": span,
            "particle_substeps_per_s": self.particle_substeps / span if span > 0 else None,
            "substep_launches": self.launches,
            "seconds_in_batch_runs": round(self.run_seconds, 3),
            "latency_target_ms": None if self.latency_target is None else self.latency_target * 1000.0,
            "admission_cap": {"last": self.cap, "min": min(c for _, c in self.cap_log),
                              "max": max(c for _, c in self.cap_log), "changes": len(self.cap_log) - 1},
        }


def _release(solver) -> None:
    """Free the solver's device state now (Solver has no close; the HIP one does)."""
    close = getattr(solver, "close", None)
    if close is not None:
        close()

