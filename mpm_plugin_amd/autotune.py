"""Autotuner for the batched engine's launch shapes.

The batched substep (libmpm_hip.so, batch.hpp) runs persistent grids whose sizes, the G2P tile (particles
per workgroup) and the re-sort cadence are launch parameters that do not change any result beyond the
order of floating-point atomic additions (the sort cadence changes only which particles share a
summation order). Their best values depend on the mix of scenes (how many, how large, how many run at
once), so they are measured on the batch itself: coordinate descent over one knob at a time, each
candidate scored by the particle-substeps per second of a few substep launches of a frame in flight
(candidates of one knob interleaved, best of `repeats`). The scenes advance while the tuner measures, as
they would in production; the frame in flight is completed afterwards, so the caller sees whole frames
only.

Results are cached per device and workload signature (scene count and particle total, rounded to powers
of two) in a JSON file, ``$MPM_TUNE_CACHE`` or ``~/.cache/mpm_plugin_amd/tune.json``; a cache that cannot
be written is skipped silently.
"""

from __future__ import annotations

import json
import math
import os
import time
from collections.abc import Callable, Sequence

from mpm_plugin_amd import _native as N

# knob -> candidate values (the first value of each list is the library default)
BATCH_SPACE: dict[int, list[int]] = {
    N.BATCH_OPT_G2P_TILE: [256, 128, 64],
    N.BATCH_OPT_P2G_WAVES: [1, 2],
    N.BATCH_OPT_G2P_BLOCKS: [2048, 1024, 4096, 8192],
    N.BATCH_OPT_P2G_BLOCKS: [8192, 4096, 16384],
    N.BATCH_OPT_GRID_BLOCKS: [2048, 512, 1024, 4096],
    N.BATCH_OPT_SORT_INTERVAL: [16, 24, 8],
}
KNOB_NAMES = {N.BATCH_OPT_G2P_TILE: "g2p_tile", N.BATCH_OPT_G2P_BLOCKS: "g2p_blocks",
              N.BATCH_OPT_P2G_BLOCKS: "p2g_blocks", N.BATCH_OPT_GRID_BLOCKS: "grid_blocks",
              N.BATCH_OPT_SORT_INTERVAL: "sort_interval", N.BATCH_OPT_P2G_WAVES: "p2g_waves"}


def coordinate_descent(space: dict[int, Sequence[int]], measure: Callable[..., float],
                       repeats: int = 2) -> tuple[dict[int, int], float]:
    """Maximise measure(config, knob) one knob at a time, starting from each knob's first value. The
    candidates of one knob are measured in interleaved rounds (v1 v2 v3 v1 v2 v3 ...) and each keeps its
    best, so a measure that drifts with time (scenes finishing their frames) treats them alike."""
    best = {k: v[0] for k, v in space.items()}
    best_score = -math.inf
    for knob, values in space.items():
        scores = {v: -math.inf for v in values}
        for _ in range(repeats):
            for v in values:
                trial = dict(best)
                trial[knob] = v
                scores[v] = max(scores[v], measure(trial, knob))
        pick = max(values, key=lambda v: scores[v])
        best[knob] = pick
        best_score = scores[pick]
    return best, best_score


def signature(batch) -> str:
    n = len(batch)
    particles = sum(s.particle_count for s in batch.scenes)
    return f"scenes~2^{round(math.log2(max(n, 1)))}_particles~2^{round(math.log2(max(particles, 1)))}"


def _device_name() -> str:
    try:
        with open("/sys/class/kfd/kfd/topology/nodes/1/name") as f:
            return f.read().strip() or "gpu"
    except OSError:
        return "gpu"


def _cache_path() -> str:
    return os.environ.get("MPM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mpm_plugin_amd",
                                                         "tune.json"))


def load_cached(key: str) -> dict[int, int] | None:
    try:
        with open(_cache_path()) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    entry = doc.get(key)
    if not entry:
        return None
    names = {v: k for k, v in KNOB_NAMES.items()}
    return {names[k]: int(v) for k, v in entry.items() if k in names}


def store_cached(key: str, config: dict[int, int]) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        try:
            with open(path) as f:
                doc = json.load(f)
        except (OSError, ValueError):
            doc = {}
        doc[key] = {KNOB_NAMES[k]: int(v) for k, v in config.items()}
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(doc, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass


def apply(batch, config: dict[int, int]) -> None:
    for knob, value in config.items():
        batch.set_option(knob, value)


def tune_batch(batch, frame_dt, launches: int = 16, repeats: int = 2, space: dict[int, Sequence[int]] | None = None,
               use_cache: bool = True) -> dict:
    """Pick the batch's launch shapes by measurement; applies and returns them.

    Every member must be between frames. One frame of every member is started and run to its end;
    candidate shapes are timed on its substeps (a frame shorter than the measurements needs is followed
    by further frames of the same length). A sort-cadence candidate is timed over two cadences of the
    longest candidate, so that every candidate's window holds its sorts (a window without one would favour
    the longest cadence). A window's rate also depends on where in the frame it falls (early launches
    carry every scene, late ones only the scenes with the most substeps), so the result replaces the
    library's defaults only if it is faster over a whole frame."""
    space = dict(space or BATCH_SPACE)
    key = f"{_device_name()}|{signature(batch)}"
    if use_cache:
        cached = load_cached(key)
        if cached:
            apply(batch, cached)
            return {"config": {KNOB_NAMES[k]: v for k, v in cached.items()}, "cached": True, "key": key}
    n = len(batch)
    counts = [s.particle_count for s in batch.scenes]
    in_frame = [False]

    sort_values = space.get(N.BATCH_OPT_SORT_INTERVAL, [0])
    long_window = max(launches, 2 * max(sort_values))

    def measure(config: dict[int, int], knob: int | None = None) -> float:
        """Particle-substeps per second over the substep launches of one window with this config."""
        window = long_window if knob in (N.BATCH_OPT_SORT_INTERVAL, None) else launches
        apply(batch, config)
        if not in_frame[0]:
            batch.begin(frame_dt)
            in_frame[0] = True
        batch.run(1)  # (re)captures the graphs outside the timed window
        done0, total = batch.progress()
        if all(done0[i] >= total[i] for i in range(n)):
            batch.begin(frame_dt)
            done0, _ = batch.progress()
        t0 = time.perf_counter()
        batch.run(window)
        dt = time.perf_counter() - t0
        done1, total = batch.progress()
        work = sum(counts[i] * (int(done1[i]) - int(done0[i])) for i in range(n))
        if all(done1[i] >= total[i] for i in range(n)):
            in_frame[0] = False
        return work / dt if dt > 0 else 0.0

    def whole_frame(config: dict[int, int]) -> float:
        """Particle-substeps per second over one whole frame of every member with this config."""
        apply(batch, config)
        batch.begin(frame_dt)
        batch.run(1)  # (re)captures the graphs outside the timed window
        done0, _ = batch.progress()
        t0 = time.perf_counter()
        batch.run(0)
        dt = time.perf_counter() - t0
        done1, _ = batch.progress()
        work = sum(counts[i] * (int(done1[i]) - int(done0[i])) for i in range(n))
        return work / dt if dt > 0 else 0.0

    t_start = time.perf_counter()
    config, score = coordinate_descent(space, measure, repeats)
    default = {k: v[0] for k, v in space.items()}
    if config != default:
        if in_frame[0]:
            batch.run(0)
            in_frame[0] = False
        rate_default = whole_frame(default)
        rate_tuned = whole_frame(config)
        if rate_tuned < 1.02 * rate_default:
            config, score = default, rate_default
        else:
            score = rate_tuned
    apply(batch, config)
    if in_frame[0]:
        batch.run(0)  # finish the frame in flight
    if use_cache:
        store_cached(key, config)
    return {"config": {KNOB_NAMES[k]: v for k, v in config.items()}, "cached": False, "key": key,
            "particle_substeps_per_s": score, "tuning_seconds": time.perf_counter() - t_start}
