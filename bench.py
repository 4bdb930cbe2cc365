"""Benchmark: particle-substeps/s of the HIP MLS-MPM substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload multi_material|elastic_block]

A step is one ``Solver.advance(dt)`` of the workload, which is exactly one substep (dt below the CFL
step): collider pose upload, CFL speed reduction, P2G, grid update, collider contact, G2P, particle
projection. Unless --steps is given, the timed steps run to the end of the workload's configured
substeps (2000 for the multi-material mix, whose collider reaches the sand column after about 580), after
the warm-up and the profile pass. With --steps K there are two windows of K steps: right after the profile
pass, and from substep 800 (collider in contact), the steps between them run untimed; the value is the
particle-substeps of both windows over their time. The value's run brackets no launch with events (a
dispatch that records events holds the next one back: about 11 us per substep with G2P bracketed); P2G's
and G2P's mean launch times come from a second run of the same substeps with both bracketed. Inputs are
resident in HBM before the timed region; the timed region is bracketed by device synchronisation.
Per-kernel device time comes from HIP events recorded by the library on its own stream around the
launches it brackets.

A second measurement (skipped with --no-batch) runs the batched engine on the 256-scene mix
(mpm_plugin_amd.workloads.scene_mix: grid edges 32-128, 5k-400k particles per scene, four material models,
0-2 colliders each): the same frames advanced one scene at a time (``Solver.advance`` per scene, in turn,
each scene's frame finished before the next starts: ``plain``), with the same calls not waiting for each
other (each scene's stream may overlap the next scenes' on the GPU: ``plain_overlapped``), and through one
``SceneBatch``; every count is a sum of particles x substeps actually run. ``speedup`` is batched over
plain.

A third measurement (skipped with --no-serve) runs the bake server (mpm_core.serve.BakeServer) on
requests arriving as a Poisson process at 20 per second, each drawn like the scene mix with a
lognormal frame count (median 48) and 20-200 substeps per frame, admitted against a 250 GB HBM budget:
once continuously batched and once as serial plain runs (one bake at a time, Solver.advance), with the
same arrival times. Every bake streams its frames into a cache folder in the runner's format (frame files,
checkpoint and progress every frame) on a scratch file system (tmpfs when it has room). Reported:
per-frame latency p50/p99 and aggregate particle-substeps/s.

A fourth measurement (skipped with --no-surface), taken right after the timed runs, meshes their final
frame with the surfacer (mpm_surface.field.surface_particles, default settings: field cell = particle spacing dx/2,
kernel radius 4 cells, iso 0.35), one mesh per emitter as the surface runner does: wall time per frame
with host arrays in and host meshes out, and the device time of the upload, splat, normalisation and
marching cubes.

The printed JSON line carries:
  roofline      for the kernel with the largest device time: algorithmic bytes per launch (bytes per
                particle x particles + bytes per active grid node x active nodes; P2G 88 + 16, G2P 148 + 12)
                over its mean launch time, against 8 TB/s HBM
  cpu_baseline  the CPU oracle (a restatement of the reference Warp kernels in C++, OpenMP, FP32) timed
                on a bounded number of substeps of the same workload on every usable host core, the median
                of three runs; with it, the
                surfacing object gets cpu_splat, the oracle's numpy splat timed on a bounded sample of the
                surfaced frame
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
PROFILE_STEPS = 32  # steps with every kernel class timed, before the timed steps (two sort intervals)

# Algorithmic bytes per particle and per active node (each array touched once). P2G: what it reads, x and v
# and the 64-byte record {C, Kirchhoff stress, mass} in place of the reference's C, F, mat and p_mass (it
# never touches F, mat or Jp), and the node's momentum and mass it produces. G2P: the reference's data flow
# (kernels.py:183-263), which is less than this build moves (it also writes the stress and mass, 28 B more).
P2G_BYTES_PARTICLE = 88  # read x12 v12 record64
P2G_BYTES_SNOW = 0
G2P_BYTES_PARTICLE = 148  # read x12 F36 mat4, write x12 v12 C36 F36
G2P_BYTES_SNOW = 8  # Jp read + write
P2G_BYTES_NODE = 16  # write momentum 12 + mass 4
G2P_BYTES_NODE = 12  # read velocity
CONTACT_SUBSTEP = 800  # the multi-material mix: the collider has been in the sand column for ~200 substeps


CU_COUNT = 256
CLOCK_HZ = 2.4e9
PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_multi_material_128.md")


def measured_counters(workload: str) -> dict:
    """HBM traffic and VALU instruction counts per launch of P2G and G2P on the 1M multi-material workload,
    from the committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration given there); empty for any other workload or without the summary."""
    out = {}
    if not workload.startswith("multi_material") or not os.path.exists(PMC_SUMMARY):
        return out
    with open(PMC_SUMMARY) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch":
                out[("traffic", parts[1])] = int(parts[2])
            elif len(parts) == 3 and parts[0] == "valu_instructions_per_launch":
                out[("valu", parts[1])] = int(parts[2])
    return out


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the rest of the workload's configured run, so every phase of "
                         "it, collider contact included, is in the measurement)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="multi_material", choices=["multi_material", "elastic_block"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-collider", action="store_true")
    ap.add_argument("--p2g-mode", default="binned", choices=["binned", "atomic", "fixed"],
                    help="P2G of the single-scene run: the library's default binned form, the reference's "
                         "per-particle global atomics, or fixed point")
    ap.add_argument("--precision", default="fallback", choices=["fp32", "fallback", "fp64"],
                    help="SVD/return-map precision of the single-scene run (library default: fallback)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched scene-mix measurement")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-frames", type=int, default=2, help="timed frames of the scene mix")
    ap.add_argument("--no-tune", action="store_true", help="batched scene mix with default launch shapes")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-requests measurement")
    ap.add_argument("--no-surface", action="store_true", help="skip the surfacing measurement")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="request arrivals per second")
    return ap.parse_args(argv)


def build_workload(name: str, with_collider: bool):
    from mpm_plugin_amd import workloads

    if name == "elastic_block":
        return workloads.elastic_block()
    return workloads.multi_material(with_collider=with_collider)


def active_nodes(solver) -> int:
    _, gm = solver.get_grid(1)
    return int((gm > 0).sum())


def usable_cores() -> tuple[int, str]:
    """Host cores this process may run on: its CPU affinity, capped by a cgroup CPU quota if one is set."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    note = f"{cores} in the affinity mask"
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            q = max(1, int(int(quota) // int(period)))
            if q < cores:
                cores, note = q, f"{note}, cgroup quota {q}"
    except (OSError, ValueError):
        pass
    return cores, note


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seconds: float, runs: int = 3) -> dict:
    """The FP32 CPU oracle on the same workload from its initial state, on every usable host core: the median
    particle-substeps/s of ``runs`` runs of about ``seconds / runs`` each."""
    from oracle import oracle as O

    threads, note = usable_cores()
    rates, counts = [], []
    for _ in range(runs):
        ora = O.OracleSolver(w.domain, w.materials, w.gravity, threads=threads)
        ora.add_particles(w.positions, w.velocities, w.material_ids, w.emitter_ids)
        if w.colliders:
            ora.set_colliders(w.colliders)
        t_sim, steps, elapsed = 0.0, 0, 0.0
        t0 = time.perf_counter()
        while elapsed < seconds / runs and steps < w.substeps:
            if w.colliders:
                ora.set_collider_transforms(*w.collider_transforms(t_sim, t_sim + w.dt))
            ora.advance(w.dt)
            t_sim += w.dt
            steps += 1
            elapsed = time.perf_counter() - t0
        rates.append(w.n * steps / elapsed)
        counts.append(steps)
    median = float(np.median(rates))
    return {"value": median, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "runs": [round(r, 1) for r in rates],
            "sample": f"median of {runs} runs of {min(counts)}-{max(counts)} substeps each from the initial state of "
                      f"the same {w.n}-particle workload (FP32 C++ restatement of the reference Warp kernels, "
                      f"OpenMP, {threads} threads: {note}; {cpu_model()})"}


def cpu_surface_baseline(particles, spacing: float, seconds: float = 5.0) -> dict:
    """Time the CPU oracle's splat (numpy, the reference's field.py:35-63 restated) on a bounded sample of the
    surfaced frame: the first particles of emitter 0, as many as about ``seconds`` of work allows. The
    reference has no CPU surfacer (Warp), so this is the CPU path beside the device splat, not a target."""
    from mpm_surface.field import field_grid
    from mpm_surface.settings import SurfaceSettings
    from oracle import surface as OS

    x = particles.positions[particles.emitter_ids == particles.emitter_ids.min()]
    v = particles.velocities[particles.emitter_ids == particles.emitter_ids.min()]
    origin, dims, cell, h = field_grid(x, spacing, SurfaceSettings())
    done, elapsed, chunk = 0, 0.0, 256
    t0 = time.perf_counter()
    while elapsed < seconds and done < len(x):
        OS.splat(x[done:done + chunk], v[done:done + chunk], origin.astype(np.float32), cell, h, dims)
        done = min(done + chunk, len(x))
        elapsed = time.perf_counter() - t0
    return {"value": done / elapsed, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"splat of {done} of emitter 0's {len(x)} particles into its {dims[0]}x{dims[1]}x{dims[2]} "
                      f"field (numpy restatement, one thread), {elapsed:.1f} s"}


def run_frames(pairs, frames: int, first_frame: int, advance_all) -> tuple[float, int]:
    """Advance every scene `frames` frames; returns (seconds, particle-substeps)."""
    work = 0
    t0 = time.perf_counter()
    for f in range(first_frame, first_frame + frames):
        for w, s in pairs:
            if w.colliders:
                s.set_collider_transforms(*w.collider_transforms(f * w.frame_dt, (f + 1) * w.frame_dt))
        counts = [s.particle_count for _, s in pairs]
        stats = advance_all(pairs)
        work += sum(n * st.substeps for n, st in zip(counts, stats))
        log(f"  frame {f}: substeps per scene {min(st.substeps for st in stats)}-"
            f"{max(st.substeps for st in stats)}, {time.perf_counter() - t0:.2f} s")
    return time.perf_counter() - t0, work


def batch_bench(n_scenes: int, frames: int, tune: bool = True) -> dict:
    """The scene mix advanced one scene at a time and through one SceneBatch, same frames."""
    from mpm_core.backends.hip.batch import SceneBatch
    from mpm_plugin_amd.workloads import scene_mix

    mix = scene_mix(n_scenes)
    total = sum(w.n for w in mix)
    log(f"scene mix: {len(mix)} scenes, {total} particles")
    result = {"workload": f"scene_mix_{n_scenes}", "scenes": len(mix), "particles": int(total),
              "frames": frames, "frame_dt": mix[0].frame_dt}

    def plain(pairs):
        # one scene at a time: each scene's frame ends before the next scene starts
        out = []
        for w, s in pairs:
            out.append(s.advance(w.frame_dt))
            s.synchronize()
        return out

    def overlapped(pairs):
        # the same calls in turn without waiting: each scene's substeps run on its own stream and the
        # GPU may overlap them with the next scenes'
        return [s.advance(w.frame_dt) for w, s in pairs]

    pairs = [(w, w.make_solver(device="cuda:0")) for w in mix]
    run_frames(pairs, 1, 0, plain)  # warm-up frame
    secs, work = run_frames(pairs, frames, 1, plain)
    result["plain"] = {"value": work / secs, "seconds": round(secs, 3), "particle_substeps": int(work)}
    log(f"plain: {work / secs:.3e} particle-substeps/s ({secs:.2f} s)")
    secs, work = run_frames(pairs, frames, 1 + frames, overlapped)
    for _, s in pairs:
        s.synchronize()
    result["plain_overlapped"] = {"value": work / secs, "seconds": round(secs, 3), "particle_substeps": int(work)}
    log(f"plain, scenes overlapping on their streams: {work / secs:.3e} particle-substeps/s ({secs:.2f} s)")
    for _, s in pairs:
        s.close()
    del pairs

    pairs = [(w, w.make_solver(device="cuda:0")) for w in mix]
    batch = SceneBatch([s for _, s in pairs])

    def batched(ps):
        return batch.advance([w.frame_dt for w, _ in ps])

    run_frames(pairs, 1, 0, batched)
    if tune:
        from mpm_plugin_amd.autotune import tune_batch

        tuned = tune_batch(batch, [w.frame_dt for w, _ in pairs], use_cache=False)
        result["tuned"] = tuned
        log(f"autotuned launch shapes: {tuned['config']} ({tuned['tuning_seconds']:.1f} s)")
        first = 2
    else:
        first = 1
    launched = batch.substeps_launched
    secs, work = run_frames(pairs, frames, first, batched)
    result["batched"] = {"value": work / secs, "seconds": round(secs, 3), "particle_substeps": int(work),
                         "substep_launches": int(batch.substeps_launched - launched)}
    log(f"batched: {work / secs:.3e} particle-substeps/s ({secs:.2f} s)")
    result["speedup"] = result["batched"]["value"] / result["plain"]["value"]
    result["speedup_vs_overlapped"] = result["batched"]["value"] / result["plain_overlapped"]["value"]
    # one more frame with every kernel class bracketed by the library's HIP events (launched directly
    # rather than from the graphs while timed): where a batched substep's device time goes
    batch.reset_timing()
    batch.set_timing(True)
    launched = batch.substeps_launched
    t_secs, _ = run_frames(pairs, 1, first + frames, batched)
    batch.set_timing(False)
    per = max(1, batch.substeps_launched - launched)
    kernels = {k: {"ms": round(ms, 3), "launches": n, "us_per_substep": round(1000.0 * ms / per, 2)}
               for k, (ms, n) in batch.kernel_times().items()}
    result["batched_kernels"] = {"substep_launches": per, "wall_s": round(t_secs, 3), "kernels": kernels}
    log("batched kernel classes, us per substep launch: "
        + ", ".join(f"{k}={v['us_per_substep']}" for k, v in kernels.items()))
    batch.close()
    for _, s in pairs:
        s.close()
    return result


def cache_root() -> tuple[str, str]:
    """A scratch folder for the served bakes' caches: tmpfs when it has room, else the system temp folder."""
    import shutil
    import tempfile

    for base in ("/dev/shm", tempfile.gettempdir()):
        try:
            if os.path.isdir(base) and shutil.disk_usage(base).free > 24e9:
                return tempfile.mkdtemp(prefix="mpm_served_", dir=base), base
        except OSError:
            continue
    return tempfile.mkdtemp(prefix="mpm_served_"), tempfile.gettempdir()


SERVE_LATENCY_TARGET_MS = 250.0  # the p99 frame latency the latency-bounded server holds itself to


def serve_bench(n_requests: int, rate: float) -> dict:
    """Poisson-arriving bake requests served three ways: serial plain runs (one bake at a time), the
    continuously batched server, and the batched server holding p99 frame latency to a target by capping how
    many bakes it admits at once. Each bake streams its frames into its own cache folder in the runner's
    format (frame files per emitter, the checkpoint and progress file every frame, mpm_core.cache) on a
    scratch file system, so a frame's latency includes its download and file writes. All three get the same
    arrivals and are stopped at the same horizon, twice the arrival window (no server drains its backlog
    within it): throughput and latencies describe the same stretch of time, latency percentiles cover the
    frames finished in it."""
    import shutil

    from mpm_core.serve import BakeRequest, BakeServer
    from mpm_plugin_amd.workloads import served_requests

    reqs = served_requests(n_requests, rate)
    window = reqs[-1][0]
    horizon = 2.0 * window
    root, fs = cache_root()
    out = {"requests": n_requests, "rate_per_s": rate, "hbm_budget_gb": 250, "window_s": round(window, 3),
           "horizon_s": round(horizon, 3), "latency_target_ms": SERVE_LATENCY_TARGET_MS,
           "frames_requested": int(sum(s.frame_end - s.frame_start + 1 for _, s in reqs)),
           "cache_files": f"every frame and checkpoint written under {fs}"}
    legs = (("serial", dict(batched=False, max_active=1)),
            ("batched", dict(batched=True)),
            ("batched_latency_bound", dict(batched=True, latency_target_ms=SERVE_LATENCY_TARGET_MS)))
    try:
        for name, kw in legs:
            server = BakeServer(hbm_budget=250e9, **kw)
            t0 = time.perf_counter()
            handles = server.run([(off, BakeRequest(scene, name=f"r{i}",
                                                    cache_dir=os.path.join(root, name, f"r{i}")))
                                  for i, (off, scene) in enumerate(reqs)], stop_after=horizon)
            rep = server.report(handles)
            rep["horizon_particle_substeps_per_s"] = server.work_between(t0, t0 + horizon) / horizon
            server.close()
            rep["wall_s"] = round(time.perf_counter() - t0, 3)
            out[name] = rep
            shutil.rmtree(os.path.join(root, name), ignore_errors=True)
            lat = rep["frame_latency_ms"]
            log(f"served ({name}): {rep['horizon_particle_substeps_per_s']:.3e} particle-substeps/s over "
                f"{horizon:.1f} s; {rep['frames']} frames, latency p50 {lat['p50'] or 0:.1f} ms p99 "
                f"{lat['p99'] or 0:.1f} ms; {rep['done']}/{n_requests} bakes done")
    finally:
        shutil.rmtree(root, ignore_errors=True)
    serial = out["serial"]["horizon_particle_substeps_per_s"]
    out["throughput_speedup"] = out["batched"]["horizon_particle_substeps_per_s"] / serial
    out["throughput_speedup_latency_bound"] = out["batched_latency_bound"]["horizon_particle_substeps_per_s"] / serial
    return out


def surface_bench(particles, spacing: float, reps: int = 5) -> dict:
    """The surfacer (mpm_surface.field.surface_particles, default settings) on every emitter of a frame of
    particles, as the surface runner meshes one frame: host arrays in, host meshes out. Reported: wall time
    per frame (best of ``reps``), and the device time of its phases from the surfacer's own HIP events."""
    from mpm_surface.field import field_grid, surface_particles, surfacer
    from mpm_surface.settings import SurfaceSettings

    settings = SurfaceSettings()
    sets = [(particles.positions[particles.emitter_ids == e], particles.velocities[particles.emitter_ids == e])
            for e in np.unique(particles.emitter_ids)]
    for x, v in sets:  # warm-up: buffers grown to the largest emitter
        surface_particles(x, v, spacing, settings, "cuda:0")
    walls, device = [], None
    for _ in range(reps):
        phases = dict.fromkeys(("upload", "splat", "normalize", "marching_cubes"), 0.0)
        verts = tris = 0
        t0 = time.perf_counter()
        for x, v in sets:
            m = surface_particles(x, v, spacing, settings, "cuda:0")
            verts, tris = verts + len(m.verts), tris + len(m.tris)
            for k, ms in surfacer("cuda:0").last_times.items():
                phases[k] += ms
        walls.append(time.perf_counter() - t0)
        if device is None or sum(phases.values()) < sum(device.values()):
            device = phases
    cells = sum(int(np.prod(field_grid(x, spacing, settings)[1])) for x, _ in sets)
    wall = min(walls)
    out = {"particles": int(len(particles.positions)), "emitters": len(sets), "field_nodes": cells,
           "verts": verts, "tris": tris, "settings": settings.to_dict(), "spacing": spacing,
           "ms_per_frame": round(wall * 1000, 2), "particles_per_s": len(particles.positions) / wall,
           "device_ms": {k: round(v, 3) for k, v in device.items()}}
    log(f"surfacing: {out['particles']} particles in {len(sets)} emitters -> {verts} verts, {tris} tris; "
        f"{out['ms_per_frame']} ms per frame (device: {out['device_ms']})")
    return out


def main(argv=None) -> int:
    args = parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (independent scenes shard "
              f"across GPUs without communication, so the per-GPU figure is the one to multiply)",
              file=sys.stderr)
        return 2
    w = build_workload(args.workload, with_collider=not args.no_collider)
    log(f"workload {w.name}: {w.n} particles, grid {w.domain.grid_shape}, dt {w.dt}")
    from mpm_plugin_amd import _native as N

    def make_run():
        """A fresh solver of the workload and its step function (one advance of dt with the pose update)."""
        solver = w.make_solver(device="cuda:0")
        solver.set_option(N.OPT_PRECISION, {"fp32": N.PRECISION_FP32, "fallback": N.PRECISION_FALLBACK,
                                            "fp64": N.PRECISION_FP64}[args.precision])
        solver.set_option(N.OPT_P2G_MODE, {"binned": N.P2G_BINNED, "atomic": N.P2G_ATOMIC,
                                           "fixed": N.P2G_FIXED}[args.p2g_mode])
        t_sim = 0.0

        def step():
            nonlocal t_sim
            if w.colliders:
                solver.set_collider_transforms(*w.collider_transforms(t_sim, t_sim + w.dt))
            st = solver.advance(w.dt)
            if st.substeps != 1:
                raise RuntimeError(f"advance({w.dt}) took {st.substeps} substeps; the workload assumes one")
            t_sim += w.dt

        return solver, step

    first = args.warmup + PROFILE_STEPS
    if args.steps is None:
        windows = [(first, w.substeps)]  # the rest of the configured run, every phase of it
    else:
        windows = [(first, first + args.steps)]
        if w.colliders and CONTACT_SUBSTEP >= first + args.steps and CONTACT_SUBSTEP + args.steps <= w.substeps:
            windows.append((CONTACT_SUBSTEP, CONTACT_SUBSTEP + args.steps))  # in collider contact as well

    def timed_run(kernels: tuple, profile: bool):
        """A fresh run: warm-up, [profile pass], then the windows (untimed steps between them) with only the
        launches of `kernels` (none: no launch) bracketed by the library's events. Returns (solver, seconds
        per window, kernel times, profile, active nodes per window)."""
        solver, step = make_run()
        for _ in range(args.warmup):
            step()
        solver.synchronize()
        prof = None
        if profile:
            # per-class kernel times over two sort intervals, every class bracketed with events
            solver.reset_timing()
            solver.set_timing(True)
            for _ in range(PROFILE_STEPS):
                step()
            solver.set_timing(False)
            prof = solver.kernel_times()
        else:
            for _ in range(PROFILE_STEPS):
                step()
        k = first
        secs = []
        nodes = []
        solver.reset_timing()
        for a, b in windows:
            for _ in range(k, a):
                step()
            solver.set_timing(bool(kernels), phases=list(kernels))
            solver.synchronize()
            t0 = time.perf_counter()
            for _ in range(a, b):
                step()
            solver.synchronize()
            secs.append(time.perf_counter() - t0)
            solver.set_timing(False)
            nodes.append(active_nodes(solver))
            k = b
        return solver, secs, solver.kernel_times(), prof, nodes

    solver, secs, _, times, nodes = timed_run((), profile=True)
    # grid nodes with mass at the end of each window, weighted by the window's steps (the mean per launch)
    n_active = int(round(sum(m * (b - a) for m, (a, b) in zip(nodes, windows)) / sum(b - a for a, b in windows)))
    n_snow = int(sum((w.material_ids == i).sum() for i, m in enumerate(w.materials) if m.model == "snow"))
    log(f"profile pass ({PROFILE_STEPS} steps), kernel ms: "
        + ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in times.items()))
    elapsed = sum(secs)
    total_steps = sum(b - a for a, b in windows)
    n = solver.particle_count
    solver.close()
    # P2G's and G2P's launch times over the same substeps, from a second run with both bracketed
    solver, secs_bracketed, kt, _, _ = timed_run(("g2p", "p2g"), profile=False)
    timed = {"g2p": kt["g2p"], "p2g": kt["p2g"]}
    final = solver.get_particles() if not args.no_surface else None
    log(f"timed windows {windows}: " + ", ".join(f"{s:.3f} s" for s in secs)
        + f"; a second run of the same substeps with P2G and G2P bracketed: "
        + ", ".join(f"{s:.3f} s" for s in secs_bracketed)
        + f", g2p {timed['g2p'][0]:.2f} ms / {timed['g2p'][1]} launches, p2g {timed['p2g'][0]:.2f} ms / "
          f"{timed['p2g'][1]} launches; active nodes {n_active}")
    bytes_per_launch = {
        "p2g": P2G_BYTES_PARTICLE * n + P2G_BYTES_SNOW * n_snow + P2G_BYTES_NODE * n_active,
        "g2p": G2P_BYTES_PARTICLE * n + G2P_BYTES_SNOW * n_snow + G2P_BYTES_NODE * n_active,
    }
    pmc = measured_counters(w.name)

    def roofline_of(kernel: str) -> dict:
        ms, launches = timed[kernel]
        mean_s = ms / 1000.0 / max(launches, 1)
        achieved = bytes_per_launch[kernel] / mean_s / 1e9
        lib_name = {"p2g": "k_p2g_split" if args.p2g_mode == "binned" else "k_p2g_atomic", "g2p": "k_g2p"}[kernel]
        out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc.get(("traffic", lib_name)),
               "kernel": kernel, "mean_launch_us": round(mean_s * 1e6, 2),
               "algorithmic_bytes": int(bytes_per_launch[kernel])}
        valu = pmc.get(("valu", lib_name))
        if valu:
            # wave instructions over what the CUs issue in the launch (one per CU per cycle)
            out["valu_busy_frac"] = round(valu / (mean_s * CU_COUNT * CLOCK_HZ), 3)
        return out

    dominant = max(("p2g", "g2p"), key=lambda k: timed[k][0] / max(timed[k][1], 1))
    roofline = roofline_of(dominant)
    roofline["other_kernel"] = roofline_of("g2p" if dominant == "p2g" else "p2g")
    kernels = {k: {"ms_total": round(v[0], 3), "launches": v[1]} for k, v in times.items() if v[1]}
    kernels["profile_steps"] = PROFILE_STEPS

    # surfacing right after the run, in the state the surface runner meets in its own process (after the
    # batched legs' hundreds of contexts the same host-to-device copies measured three times slower)
    surfaced = None
    if final is not None:
        surfaced = surface_bench(final, w.domain.dx / 2)

    baseline = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        baseline = cpu_baseline(w, args.cpu_seconds)
        log(f"cpu baseline {baseline['value']:.3e} particle-substeps/s")
        if surfaced is not None:
            surfaced["cpu_splat"] = cpu_surface_baseline(final, w.domain.dx / 2)
            surfaced["device_splat_particles_per_s"] = surfaced["particles"] / (surfaced["device_ms"]["splat"] / 1e3)
            log(f"cpu surfacing baseline {surfaced['cpu_splat']['value']:.3e} particles/s splatted "
                f"(device: {surfaced['device_splat_particles_per_s']:.3e})")

    batched = None
    solver.close()
    if not args.no_batch:
        batched = batch_bench(args.batch_scenes, args.batch_frames, tune=not args.no_tune)
    served = None
    if not args.no_serve:
        served = serve_bench(args.serve_requests, args.serve_rate)

    value = n * total_steps / elapsed
    line = {
        "metric": "particle-substeps/s at 1M pts/128^3",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": total_steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1000.0 / total_steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "particles": n, "grid": list(w.domain.grid_shape), "dt": w.dt,
                   "materials": [m.model for m in w.materials], "colliders": len(w.colliders),
                   "substeps_per_step": 1, "svd_precision": args.precision, "p2g_mode": args.p2g_mode,
                   "timed_substeps": [list(ab) for ab in windows],
                   "window_values": [round(n * (b - a) / t, 1) for (a, b), t in zip(windows, secs)],
                   "value_with_p2g_g2p_bracketed": round(n * total_steps / sum(secs_bracketed), 1),
                   "active_nodes": n_active},
        "roofline": roofline,
        "cpu_baseline": baseline,
        "kernels": kernels,
        "batched": batched,
        "served": served,
        "surface": surfaced,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
