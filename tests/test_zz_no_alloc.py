"""mpm_is_occupied_noalloc (no device allocation, fixed per-GPU volume buffer) against mpm_is_occupied (GPU).

Both are called through the C-ABI on the same context and query points; the no-alloc entry point reads and
writes device buffers that the test allocates with the HIP runtime the library is linked against. Two calls
in a row, the second with fewer points after the particles moved, must each equal the allocating entry point
bit for bit (an integer result); so must a call on a second, smaller grid after the first grid's call, which
leaves its volumes in the shared buffer.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
from helpers import add, box_particles, make_solver, sphere_particles

from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain
from mpm_plugin_amd import _native as N

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
JELLY = MaterialParams("jelly", "elastic", 5.0e4, 0.3, 1000.0)


def hip_runtime() -> C.CDLL:
    N.lib()  # loads libmpm_hip.so and with it the runtime it links, so this finds the same library
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (L.hipMalloc, L.hipFree, L.hipMemcpy):
        f.restype = C.c_int
    return L


class DeviceArray:
    def __init__(self, hip, nbytes: int):
        self.hip, self.nbytes, self.ptr = hip, nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(nbytes, 1)) == 0

    def upload(self, a: np.ndarray) -> None:
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, a: np.ndarray) -> np.ndarray:
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self) -> None:
        self.hip.hipFree(self.ptr)


def queries(lo, hi, step, rng) -> np.ndarray:
    axes = [np.arange(a, b, step) for a, b in zip(lo, hi)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    return (g + rng.uniform(-0.3, 0.3, g.shape) * step).astype(np.float32)


def both(solver, pts: np.ndarray, d_pts: DeviceArray, d_out: DeviceArray):
    L = N.lib()
    pts = np.ascontiguousarray(pts, np.float32)
    want = np.zeros(len(pts), np.uint8)
    N.check(L.mpm_is_occupied(solver._ctx, len(pts), N.fptr(pts), N.u8ptr(want)))
    d_pts.upload(pts)
    N.check(L.mpm_is_occupied_noalloc(solver._ctx, len(pts), C.cast(d_pts.ptr, C.POINTER(C.c_float)),
                                      C.cast(d_out.ptr, C.POINTER(C.c_uint8))))
    N.check(L.mpm_synchronize(solver._ctx))
    got = d_out.download(np.zeros(len(pts), np.uint8))
    return got, want


def test_noalloc_matches_allocating_entry_point_twice_and_across_grids():
    hip = hip_runtime()
    rng = np.random.default_rng(7)
    dom = Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 64)
    s = make_solver("cuda:0", dom, [JELLY])
    add(s, sphere_particles((0.5, 0.5, 0.45), 0.2, dom.dx, seed=1), velocity=(0.5, -0.3, 0.2))
    add(s, box_particles((0.15, 0.15, 0.05), (0.4, 0.35, 0.3), dom.dx, seed=2))
    first = queries((0.05, 0.05, 0.0), (0.95, 0.95, 0.8), 1 / 96, rng)  # ~600k points
    d_pts = DeviceArray(hip, first.nbytes)
    d_out = DeviceArray(hip, len(first))
    try:
        got, want = both(s, first, d_pts, d_out)
        print(f"first call: {len(first)} points, {int(want.sum())} occupied", flush=True)
        assert 0 < want.sum() < len(first)
        assert np.array_equal(got, want)
        for _ in range(5):
            s.advance(1 / 240)  # the particles move; the buffer still holds the first call's volumes
        second = queries((0.2, 0.2, 0.1), (0.8, 0.8, 0.7), 1 / 40, rng)
        assert len(second) < len(first)
        got, want = both(s, second, d_pts, d_out)
        print(f"second call: {len(second)} points, {int(want.sum())} occupied", flush=True)
        assert 0 < want.sum() < len(second)
        assert np.array_equal(got, want)
        # a smaller grid on the same GPU reuses the front of the buffer the 64^3 grid filled
        small = Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 24)
        t = make_solver("cuda:0", small, [JELLY])
        add(t, box_particles((0.3, 0.3, 0.3), (0.6, 0.6, 0.6), small.dx, seed=3))
        third = queries((0.1, 0.1, 0.1), (0.9, 0.9, 0.9), 1 / 30, rng)
        got, want = both(t, third, d_pts, d_out)
        print(f"smaller grid: {len(third)} points, {int(want.sum())} occupied", flush=True)
        assert 0 < want.sum() < len(third)
        assert np.array_equal(got, want)
    finally:
        d_pts.free()
        d_out.free()


def test_noalloc_edge_cases():
    hip = hip_runtime()
    L = N.lib()
    dom = Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 32)
    s = make_solver("cuda:0", dom, [JELLY])
    d = DeviceArray(hip, 1024)
    try:
        fp, up = C.cast(d.ptr, C.POINTER(C.c_float)), C.cast(d.ptr, C.POINTER(C.c_uint8))
        # no particles: nothing is occupied
        pts = np.array([[0.5, 0.5, 0.5], [0.1, 0.2, 0.3]], np.float32)
        got, want = both(s, pts, d, DeviceArray(hip, 16))
        assert np.array_equal(got, want) and not want.any()
        N.check(L.mpm_is_occupied_noalloc(s._ctx, 0, None, None))  # empty query: nothing to do
        with pytest.raises(ValueError):
            N.check(L.mpm_is_occupied_noalloc(s._ctx, 4, None, up))
        with pytest.raises(ValueError):
            N.check(L.mpm_is_occupied_noalloc(s._ctx, -1, fp, up))
    finally:
        d.free()
    # 256^3: (262 -> 264)^3 = 18.4 M nodes, beyond the fixed buffer's 8.4 M
    big = make_solver("cuda:0", Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 256), [JELLY])
    d = DeviceArray(hip, 64)
    try:
        with pytest.raises(ValueError, match="fixed volume buffer"):
            N.check(L.mpm_is_occupied_noalloc(big._ctx, 1, C.cast(d.ptr, C.POINTER(C.c_float)),
                                              C.cast(d.ptr, C.POINTER(C.c_uint8))))
    finally:
        d.free()
