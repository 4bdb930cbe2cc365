"""Surfacing of one particle set on the GPU: density splat -> marching cubes -> vertex velocities
(reference src/mpm_surface/field.py:1-142).

Density is an SPH-style sum of a cubic B-spline kernel of support radius ``h = 2 * spacing * smoothing``,
normalised so that a uniform particle lattice at ``spacing`` reads 1; the surface is its ``settings.iso``
level set. Velocity is the kernel-weighted mean of the particles' velocities, sampled trilinearly at the
vertices. The device work runs in libmpm_hip.so (mpm_surface_build); this module places the field grid,
computes the rest density and the memory budget, and turns grid units into world positions, as the
reference's host code does.
"""

from __future__ import annotations

import atexit
import functools
import math
import threading

import numpy as np

from mpm_core.backends.hip.solver import device_ordinal
from mpm_plugin_amd import _native as N
from mpm_surface.meshcache import SurfaceMesh
from mpm_surface.settings import SurfaceSettings

# device bytes per field node: density (4) and velocity (12), crossed-edge flags and vertex ids (12 + 12),
# triangle counts and offsets per cell (4 + 4), and 8 for the mesh and the scans' partial sums
BYTES_PER_CELL = 56


def cubic_kernel(q: np.ndarray) -> np.ndarray:
    """The unnormalised cubic B-spline of support [0, 1) (field.py:24-32)."""
    q = np.asarray(q, np.float64)
    return np.where(q < 0.5, 1.0 - 6.0 * q**2 + 6.0 * q**3, np.where(q < 1.0, 2.0 * (1.0 - q) ** 3, 0.0))


@functools.lru_cache(maxsize=64)
def rest_density(spacing: float, h: float) -> float:
    """The kernel sum at a point of a uniform particle lattice of this spacing, averaged over the eight
    offsets {0, 1/2}^3 spacing of the point within a lattice cell (field.py:88-97)."""
    n = math.ceil(h / spacing) + 1
    r = np.arange(-n, n + 1, dtype=np.float64)
    lattice = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * spacing
    offsets = np.array(list(np.ndindex(2, 2, 2)), np.float64) * 0.5 * spacing
    return float(np.mean([cubic_kernel(np.linalg.norm(lattice - o, axis=1) / h).sum() for o in offsets]))


class Surfacer:
    """A device's surfacing buffers (mpm_surface_create), kept between calls so that meshing frame after
    frame reallocates only when a frame needs more. Calls are serialised."""

    def __init__(self, device: int = 0):
        self.device = device
        self._lib = N.lib()
        self._lock = threading.Lock()
        handle = N.C.c_void_p()
        N.check(self._lib.mpm_surface_create(device, N.C.byref(handle)))
        self._handle = handle
        self.last_times: dict[str, float] | None = None

    def close(self) -> None:
        if self._handle:
            self._lib.mpm_surface_destroy(self._handle)
            self._handle = N.C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, desc: N.SurfaceDesc, x: np.ndarray, v: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Vertices (grid units), triangles and vertex velocities of the iso surface of x's density."""
        x, v = _rows(x, v)
        with self._lock:
            nv, nt = N.C.c_int64(), N.C.c_int64()
            N.check(self._lib.mpm_surface_build(self._handle, N.C.byref(desc), len(x), N.fptr(x), N.fptr(v),
                                                N.C.byref(nv), N.C.byref(nt)))
            return self._fetch(nv.value, nt.value)

    def mesh(self, x: np.ndarray, v: np.ndarray, describe):
        """Upload the particles, let ``describe(lo, hi)`` place the field from their bounds (computed on the
        device; it returns the description and anything else to hand back, or raises), and mesh them.
        Returns (vertices in grid units, triangles, vertex velocities, what describe handed back)."""
        x, v = _rows(x, v)
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        with self._lock:
            N.check(self._lib.mpm_surface_upload(self._handle, len(x), N.fptr(x), N.fptr(v), N.fptr(lo),
                                                 N.fptr(hi)))
            desc, extra = describe(lo, hi)
            nv, nt = N.C.c_int64(), N.C.c_int64()
            N.check(self._lib.mpm_surface_build_uploaded(self._handle, N.C.byref(desc), N.C.byref(nv),
                                                         N.C.byref(nt)))
            return self._fetch(nv.value, nt.value) + (extra,)

    def _fetch(self, nv: int, nt: int):
        verts = np.empty((nv, 3), np.float32)
        vel = np.empty((nv, 3), np.float32)
        tris = np.empty((nt, 3), np.int32)
        N.check(self._lib.mpm_surface_fetch(self._handle, N.fptr(verts), N.iptr(tris), N.fptr(vel)))
        ms = np.zeros(4, np.float32)
        N.check(self._lib.mpm_surface_times(self._handle, N.fptr(ms)))
        self.last_times = dict(zip(("upload", "splat", "normalize", "marching_cubes"), map(float, ms)))
        return verts, tris, vel

    def field(self, dims) -> tuple[np.ndarray, np.ndarray]:
        """The last build's normalised density (D0,D1,D2) and velocity (D0,D1,D2,3)."""
        density = np.empty(tuple(dims), np.float32)
        velocity = np.empty(tuple(dims) + (3,), np.float32)
        with self._lock:
            N.check(self._lib.mpm_surface_get_field(self._handle, N.fptr(density), N.fptr(velocity)))
        return density, velocity


def _rows(x: np.ndarray, v: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    if len(v) != len(x):
        raise ValueError("positions and velocities differ in length")
    return x, v


_surfacers: dict[int, Surfacer] = {}
_surfacers_lock = threading.Lock()


def surfacer(device: str | int | None = None) -> Surfacer:
    """The process's surfacer of a device (created on first use)."""
    ordinal = device if isinstance(device, int) else device_ordinal(device)
    with _surfacers_lock:
        if ordinal not in _surfacers:
            _surfacers[ordinal] = Surfacer(ordinal)
        return _surfacers[ordinal]


@atexit.register
def _close_surfacers() -> None:
    """Free every device's buffers while the HIP runtime is still up."""
    with _surfacers_lock:
        for s in _surfacers.values():
            s.close()
        _surfacers.clear()


def free_device_memory(device: int) -> int:
    free, total = N.C.c_int64(), N.C.c_int64()
    N.check(N.lib().mpm_device_memory(device, N.C.byref(free), N.C.byref(total)))
    return free.value


def grid_from_bounds(lo: np.ndarray, hi: np.ndarray, spacing: float, settings: SurfaceSettings):
    """The field's origin (float64), node counts, cell and kernel radius from the particles' float32 bounds
    (field.py:112-118): the bounds padded by h + 2 cells, so that the density falls below iso before the
    grid's edge and every mesh closes."""
    cell = spacing / settings.detail
    h = 2.0 * spacing * settings.smoothing
    pad = h + 2.0 * cell
    origin = np.asarray(lo, np.float32).astype(np.float64) - pad
    dims = tuple(int(math.ceil(e / cell)) + 1 for e in np.asarray(hi, np.float32) + pad - origin)
    return origin, dims, cell, h


def field_grid(positions: np.ndarray, spacing: float, settings: SurfaceSettings):
    """grid_from_bounds of the particles' bounds computed on the host."""
    # per column: numpy's axis-0 reduction over (N, 3) rows is ten times slower, with the same result
    lo = np.array([positions[:, a].min() for a in range(3)], np.float32)
    hi = np.array([positions[:, a].max() for a in range(3)], np.float32)
    return grid_from_bounds(lo, hi, spacing, settings)


def surface_particles(positions: np.ndarray, velocities: np.ndarray, spacing: float, settings: SurfaceSettings,
                      device: str | None = None, max_bytes: int | None = None) -> SurfaceMesh:
    """The closed surface of a particle set (field.py:100-142). max_bytes defaults to 70% of the device's
    free memory; a field that needs more raises ValueError."""
    positions = np.asarray(positions, np.float32).reshape(-1, 3)
    if len(positions) == 0:
        return SurfaceMesh.empty()
    ordinal = device_ordinal(device)
    if max_bytes is None:
        max_bytes = int(0.7 * (free_device_memory(ordinal) or 8 * 1024**3))

    def describe(lo, hi):
        """The field for the particles' bounds (from the device), within the memory budget."""
        origin, dims, cell, h = grid_from_bounds(lo, hi, spacing, settings)
        cells = dims[0] * dims[1] * dims[2]
        needed = cells * BYTES_PER_CELL
        if needed > max_bytes:
            raise ValueError(f"surfacing needs a {dims[0]}x{dims[1]}x{dims[2]} grid ({cells:,} cells, about "
                             f"{needed / 1e9:.1f} GB of GPU memory), over the budget of {max_bytes / 1e9:.1f} GB; "
                             f"lower the detail or split the emitter")
        desc = N.SurfaceDesc()
        desc.dims[:] = dims
        desc.origin[:] = [float(o) for o in origin]
        desc.cell = cell
        desc.h = h
        desc.inv_rest = 1.0 / rest_density(spacing, h)
        desc.iso = settings.iso
        return desc, (origin, cell)

    g_verts, tris, vel, (origin, cell) = surfacer(ordinal).mesh(positions, velocities, describe)
    if len(g_verts) == 0:
        return SurfaceMesh.empty()
    verts = (origin + g_verts.astype(np.float64) * cell).astype(np.float32)
    return SurfaceMesh(verts, tris, vel)
