"""The bake server: admission, pieces, frame output, cancel and errors against the bake runner.

Runner semantics are checked on the CPU with the ballistic fake backend (tests/test_bake.py), where the
server in serial mode must write byte-identical caches. The batched server on the GPU is compared with
the runner on the HIP backend: same frames and emitter particle counts, the first frame identical, and
every later frame within the multi-step criterion (centre of mass, bounding box and kinetic energy within
1%; the runs differ only in the order of floating-point atomic additions).
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from helpers import small_scene
from test_bake import RecordingSolver  # noqa: F401 - registers the "recording" backends

from mpm_core.bake import bake
from mpm_core.cache import CacheReader
from mpm_core.scene import read_scene
from mpm_core.serve import BakeRequest, BakeServer, hbm_footprint, mesh_volume
from mpm_core.scene import write_scene


def scenes():
    return {
        "fill": small_scene(frame_end=5),
        "inflow": small_scene(frame_end=6, kind="inflow", emit_frames=(2, 4), velocity=(0.0, 0.0, -1.0)),
        "moving": small_scene(frame_end=4, velocity=(1.0, 0.0, 0.0)),
    }


def write_cache(tmp_path, name, scene):
    d = tmp_path / name
    write_scene(scene, d / "scene")
    return d


def frames_of(cache_dir):
    r = CacheReader(cache_dir)
    return {f: r.read_frame(f) for f in r.available_frames()}


def progress(cache_dir):
    return json.loads((cache_dir / "progress.json").read_text())


def test_mesh_volume_and_footprint():
    from helpers import unit_cube_mesh

    v, t = unit_cube_mesh()
    assert mesh_volume(v, t) == pytest.approx(1.0)
    assert mesh_volume(v * 2, t) == pytest.approx(8.0)
    a = hbm_footprint(small_scene(frame_end=3))
    b = hbm_footprint(small_scene(frame_end=3, resolution=64))
    assert 0 < a < b


def test_serial_server_writes_the_runners_cache(tmp_path):
    """Three bakes served together (pieces interleaved) write exactly what three runner bakes write."""
    server = BakeServer(batched=False, backend="recording", max_active=3)
    handles = []
    for name, scene in scenes().items():
        ref = write_cache(tmp_path, "ref_" + name, scene)
        assert bake(ref, backend="recording").status == "done"
        out = write_cache(tmp_path, "srv_" + name, scene)
        handles.append(server.submit(BakeRequest(read_scene(out / "scene"), cache_dir=out, name=name)))
    while not server.idle:
        server.step()
    server.close()
    for h in handles:
        assert h.status == "done", h.message
        ref, out = tmp_path / ("ref_" + h.request.name), tmp_path / ("srv_" + h.request.name)
        a, b = frames_of(ref), frames_of(out)
        assert sorted(a) == sorted(b)
        for f in a:
            for emitter in a[f]:
                for k in range(2):
                    np.testing.assert_array_equal(a[f][emitter][k], b[f][emitter][k])
        assert progress(out)["status"] == "done"
        assert json.loads((ref / "cache.json").read_text()) == json.loads((out / "cache.json").read_text())
        with np.load(ref / "state" / "checkpoint.npz") as x, np.load(out / "state" / "checkpoint.npz") as y:
            assert int(x["__frame"]) == int(y["__frame"])
            np.testing.assert_array_equal(x["x"], y["x"])


def test_error_and_cancel_do_not_stop_the_others(tmp_path):
    server = BakeServer(batched=False, backend="recording_nan3", max_active=4)
    bad = write_cache(tmp_path, "bad", small_scene(frame_end=8))
    h_bad = server.submit(BakeRequest(read_scene(bad / "scene"), cache_dir=bad))
    server2 = BakeServer(batched=False, backend="recording", max_active=4)
    keep = write_cache(tmp_path, "keep", small_scene(frame_end=4))
    stop = write_cache(tmp_path, "stop", small_scene(frame_end=8))
    h_keep = server2.submit(BakeRequest(read_scene(keep / "scene"), cache_dir=keep))
    h_stop = server2.submit(BakeRequest(read_scene(stop / "scene"), cache_dir=stop))
    while not server.idle:
        server.step()
    assert h_bad.status == "error" and "unstable" in h_bad.message
    assert progress(bad)["status"] == "error"
    server2.step()  # admits both and runs the first piece
    (stop / "cancel").write_text("")
    while not server2.idle:
        server2.step()
    server.close()
    server2.close()
    assert h_keep.status == "done" and h_keep.frames_written == 4
    assert h_stop.status == "cancelled" and 1 < h_stop.frames_written < 8
    assert progress(stop)["status"] == "cancelled"
    assert not (stop / "cancel").exists()


def test_admission_waits_for_memory():
    """A request that does not fit the budget waits until a running bake has ended."""
    server = BakeServer(batched=False, backend="recording", max_active=8, hbm_budget=100)
    a = server.submit(BakeRequest(small_scene(frame_end=3), footprint=60))
    b = server.submit(BakeRequest(small_scene(frame_end=3), footprint=60))
    c = server.submit(BakeRequest(small_scene(frame_end=2), footprint=30))
    while not server.idle:
        server.step()
    server.close()
    assert all(h.status == "done" for h in (a, b, c))
    assert b.admitted >= a.finished  # in arrival order: c waited behind b
    assert c.admitted >= a.finished
    rep = server.report()
    assert rep["bakes"] == 3 and rep["frames"] == 3 + 3 + 2


def test_arrivals_and_report():
    server = BakeServer(batched=False, backend="recording", max_active=2)
    reqs = [(0.01 * i, BakeRequest(small_scene(frame_end=3))) for i in range(4)]
    handles = server.run(reqs)
    server.close()
    assert [h.status for h in handles] == ["done"] * 4
    rep = server.report()
    assert rep["frames"] == 12 and rep["frame_latency_ms"]["p99"] >= rep["frame_latency_ms"]["p50"] >= 0
    assert all(h.arrival <= h.admitted <= h.finished for h in handles)


@pytest.mark.gpu
def test_batched_server_matches_the_runner(tmp_path):
    """The GPU server with four bakes in flight (one with inflow pieces) against runner bakes."""
    server = BakeServer(chunk=3)
    handles = []
    for name, scene in [*scenes().items(), ("fill2", small_scene(frame_end=5, resolution=48))]:
        ref = write_cache(tmp_path, "ref_" + name, scene)
        assert bake(ref, backend="hip").status == "done"
        out = write_cache(tmp_path, "srv_" + name, scene)
        handles.append(server.submit(BakeRequest(read_scene(out / "scene"), cache_dir=out, name=name)))
    while not server.idle:
        server.step()
    server.close()
    assert server.launches > 0
    for h in handles:
        assert h.status == "done", h.message
        a = frames_of(tmp_path / ("ref_" + h.request.name))
        b = frames_of(tmp_path / ("srv_" + h.request.name))
        assert sorted(a) == sorted(b)
        for f in sorted(a):
            for emitter in a[f]:
                (pa, va), (pb, vb) = a[f][emitter], b[f][emitter]
                assert len(pa) == len(pb)
                if len(pa) == 0:
                    continue
                if f == min(a):
                    np.testing.assert_array_equal(pa, pb)
                    continue
                span = max(float(np.ptp(pa, axis=0).max()), 1e-6)
                assert np.abs(pa.mean(0) - pb.mean(0)).max() <= 0.01 * span
                assert np.abs(pa.min(0) - pb.min(0)).max() <= 0.01 * span
                assert np.abs(pa.max(0) - pb.max(0)).max() <= 0.01 * span
                ke_a, ke_b = (va.astype(np.float64) ** 2).sum(), (vb.astype(np.float64) ** 2).sum()
                assert abs(ke_a - ke_b) <= 0.01 * ke_b + 1e-9
        assert progress(tmp_path / ("srv_" + h.request.name))["status"] == "done"


def test_latency_target_moves_the_admission_cap():
    """The cap starts small, drops to LATENCY_BACKOFF of the bakes in flight after a window whose p99 is over
    the target, and grows by an eighth (at least one) after a window well under it."""
    from mpm_core import serve

    server = BakeServer(batched=False, backend="recording", max_active=64, latency_target_ms=100.0)
    assert server.cap == serve.LATENCY_START_CAP
    server.active = [object()] * 10  # ten bakes in flight
    for _ in range(serve.LATENCY_WINDOW - 1):
        server._track_latency(0.25)
    assert server.cap == serve.LATENCY_START_CAP  # no decision before a full window
    server._track_latency(0.25)
    assert server.cap == int(serve.LATENCY_BACKOFF * 10)
    for _ in range(serve.LATENCY_WINDOW):
        server._track_latency(0.08)  # between the headroom and the target: the cap holds
    assert server.cap == int(serve.LATENCY_BACKOFF * 10)
    for _ in range(serve.LATENCY_WINDOW):
        server._track_latency(0.01)
    assert server.cap == int(serve.LATENCY_BACKOFF * 10) + 1
    server.active = []
    report = server.report([])
    assert report["latency_target_ms"] == 100.0 and report["admission_cap"]["changes"] == 2
    server.close()


def test_admission_cap_holds_requests_in_the_queue(tmp_path):
    server = BakeServer(batched=False, backend="recording", max_active=64, latency_target_ms=1e6)
    server.cap = 2
    handles = [server.submit(BakeRequest(small_scene(frame_end=3), cache_dir=write_cache(tmp_path, f"c{i}",
                                                                                     small_scene(frame_end=3))))
               for i in range(5)]
    server._admit()
    assert len(server.active) == 2 and len(server.queue) == 3
    while not server.idle:
        server.step()
    server.close()
    assert all(h.status == "done" for h in handles)
