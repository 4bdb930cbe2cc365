"""SceneBatch: many HipSolver scenes advanced one frame each through one batched launch sequence.

The reference bakes one scene per runner process (src/mpm_core/bake.py:116-208) and has no batching; a
batch is this backend's addition. Every member keeps behaving as its own ``Solver`` between frames (poses,
state reads, emission, checkpoints) and ``SceneBatch.advance`` does, per scene, what ``Solver.advance``
does (src/mpm_core/backends/warp/solver.py:213-240: CFL step count from the scene's own speed maximum,
its substeps, escaped-particle removal), only for all scenes at once: each substep of the batch is one
replay of a captured HIP graph in which every still-running scene takes part.

``begin`` and ``run`` expose the same engine with frames that start and end independently (continuous
batching): a scene whose frame ended can be serviced (frame written, emission, poses) and start its next
frame while the others are still in theirs.
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Sequence

import numpy as np

from mpm_core.solver import AdvanceStats
from mpm_plugin_amd import _native as N

from .solver import HipSolver


class SceneBatch:
    """A set of ``HipSolver`` scenes on one GPU, advanced together."""

    def __init__(self, scenes: Sequence[HipSolver] = (), device: int | None = None):
        self._lib = N.lib()
        if device is None:
            device = scenes[0].ordinal if scenes else 0
        self.ordinal = int(device)
        handle = C.c_void_p()
        N.check(self._lib.mpm_batch_create(self.ordinal, C.byref(handle)))
        self._b = handle
        self._scenes: list[HipSolver] = []
        if scenes:
            self.set_scenes(scenes)

    def close(self) -> None:
        if getattr(self, "_b", None):
            self._lib.mpm_batch_destroy(self._b)
            self._b = None
            self._scenes = []

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def __len__(self) -> int:
        return len(self._scenes)

    @property
    def scenes(self) -> list[HipSolver]:
        return list(self._scenes)

    def set_scenes(self, scenes: Sequence[HipSolver]) -> None:
        """Replace the members; their order is the order of ``advance``'s arguments and results."""
        for s in scenes:
            if not isinstance(s, HipSolver) or s._ctx is None:
                raise ValueError("batch members must be open HipSolver instances")
        arr = (C.c_void_p * max(1, len(scenes)))(*[s._ctx.value for s in scenes])
        N.check(self._lib.mpm_batch_set_scenes(self._b, arr, len(scenes)))
        self._scenes = list(scenes)

    def advance(self, frame_dt: float | Sequence[float]) -> list[AdvanceStats]:
        """One frame of every member (frame_dt per member, or one value for all; <= 0 skips a member)."""
        n = len(self._scenes)
        self._check_members()
        fdt = np.broadcast_to(np.asarray(frame_dt, np.float64), (n,)).copy()
        stats = (N.AdvanceStats * max(1, n))()
        N.check(self._lib.mpm_batch_advance(self._b, fdt.ctypes.data_as(C.POINTER(C.c_double)), stats))
        return [AdvanceStats(int(stats[i].substeps), float(stats[i].dt), int(stats[i].removed)) for i in range(n)]

    def begin(self, frame_dt: float | Sequence[float]) -> None:
        """Start a frame of every member with frame_dt > 0 (it must not be in a frame)."""
        n = len(self._scenes)
        self._check_members()
        fdt = np.broadcast_to(np.asarray(frame_dt, np.float64), (n,)).copy()
        N.check(self._lib.mpm_batch_begin(self._b, fdt.ctypes.data_as(C.POINTER(C.c_double))))

    def run(self, max_launches: int = -1) -> tuple[int, dict[int, AdvanceStats]]:
        """Replay substeps (max_launches > 0: that many; 0: until every frame in flight ended; < 0: until the
        first one ended). Returns (substep launches, {member index: stats of its frame that ended})."""
        n = len(self._scenes)
        self._check_members()
        launched = C.c_int64()
        finished = np.zeros(max(1, n), np.int32)
        stats = (N.AdvanceStats * max(1, n))()
        N.check(self._lib.mpm_batch_run(self._b, int(max_launches), C.byref(launched), N.iptr(finished), stats))
        done = {i: AdvanceStats(int(stats[i].substeps), float(stats[i].dt), int(stats[i].removed))
                for i in range(n) if finished[i]}
        return int(launched.value), done

    def progress(self) -> tuple[np.ndarray, np.ndarray]:
        """(substeps done, substeps of the frame in flight) per member."""
        n = len(self._scenes)
        done, total = np.zeros(max(1, n), np.int32), np.zeros(max(1, n), np.int32)
        N.check(self._lib.mpm_batch_progress(self._b, N.iptr(done), N.iptr(total)))
        return done[:n], total[:n]

    def _check_members(self) -> None:
        if any(s._ctx is None for s in self._scenes) or self._lib.mpm_batch_scene_count(self._b) != len(self._scenes):
            raise ValueError("a batch member was closed: call set_scenes with the open ones")

    def set_option(self, option: int, value: int) -> None:
        N.check(self._lib.mpm_batch_set_option(self._b, int(option), int(value)))

    def get_option(self, option: int) -> int:
        return int(self._lib.mpm_batch_get_option(self._b, int(option)))

    def set_timing(self, enabled: bool) -> None:
        """Bracket every kernel class of the batched substeps with HIP events (substeps are then launched
        directly instead of replayed from the captured graphs)."""
        N.check(self._lib.mpm_batch_set_timing(self._b, 1 if enabled else 0))

    def reset_timing(self) -> None:
        N.check(self._lib.mpm_batch_reset_timing(self._b))

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        """(device milliseconds, launches) per kernel class since the last reset."""
        out = {}
        for k, name in enumerate(N.BATCH_KERNEL_NAMES):
            ms, count = C.c_double(), C.c_int64()
            N.check(self._lib.mpm_batch_get_kernel_time(self._b, k, C.byref(ms), C.byref(count)))
            out[name] = (ms.value, count.value)
        return out

    @property
    def substeps_launched(self) -> int:
        return self.get_option(N.BATCH_OPT_SUBSTEPS_LAUNCHED)
