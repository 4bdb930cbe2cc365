"""HipSolver: the MI355X backend behind the unchanged ``mpm_core.solver.Solver`` interface.

Host side of the reference's WarpSolver (src/mpm_core/backends/warp/solver.py:26-271) over the C-ABI of
libmpm_hip.so: the derived constants are computed here in FP64 exactly as the reference does
(solver.py:37-71, 164-172) and handed to the library, which owns all device state and runs every kernel.
"""

from __future__ import annotations

import ctypes as C
import math
import re
from collections.abc import Sequence

import numpy as np

from mpm_core.geometry import mesh_is_closed
from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain
from mpm_core.solver import AdvanceStats, ColliderGeometry, ParticleData, Solver
from mpm_plugin_amd import _native as N

PAD = 3  # padding cells around the domain so stencils never leave the grid (solver.py:18)
CFL = 0.4
COLLIDER_THICKNESS = 2.0  # open-mesh colliders are solid this many cells behind their surface
MODEL_IDS = {"elastic": 0, "von_mises": 1, "drucker_prager": 2, "snow": 3}
_STATE_KEYS = ("x", "v", "C", "F", "Jp", "mat", "emitter")


def device_ordinal(device: str | None) -> int:
    """``None``, ``"cuda"``, ``"hip"``, ``"cuda:N"`` or ``"hip:N"`` -> GPU ordinal."""
    if device is None:
        return 0
    m = re.fullmatch(r"(?:cuda|hip|gpu)(?::(\d+))?", str(device).strip())
    if not m:
        raise ValueError(f"unsupported device {device!r}: the HIP backend runs on 'cuda:N' / 'hip:N'")
    return int(m.group(1) or 0)


def dp_alpha(friction_angle_deg: float) -> float:
    """Drucker-Prager alpha from the friction angle (solver.py:263-265)."""
    s = math.sin(math.radians(friction_angle_deg))
    return math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


def domain_desc(domain: Domain, gravity, particles_per_cell: int) -> N.DomainDesc:
    """Derived domain constants (solver.py:37-54), FP64 on the host, FP32 at the boundary."""
    dx = domain.dx
    dims = tuple(s + 2 * PAD for s in domain.grid_shape)
    origin = np.asarray(domain.min, np.float64) - PAD * dx
    d = N.DomainDesc()
    d.dims[:] = dims
    d.origin[:] = origin.tolist()
    d.dx, d.inv_dx, d.p_vol = dx, 1.0 / dx, dx ** 3 / particles_per_cell
    d.gravity[:] = [float(g) for g in gravity]
    # last node index at/below each lo wall and first at/above each hi wall
    d.wall_lo_node[:] = [int(math.floor((lo - o) / dx + 1e-3)) for lo, o in zip(domain.min, origin)]
    d.wall_hi_node[:] = [int(math.ceil((hi - o) / dx - 1e-3)) for hi, o in zip(domain.max, origin)]
    d.lo_on[:] = [int(domain.walls[f"{a}-"]) for a in "xyz"]
    d.hi_on[:] = [int(domain.walls[f"{a}+"]) for a in "xyz"]
    d.wall_lo[:] = list(domain.min)
    d.wall_hi[:] = list(domain.max)
    d.clamp_lo[:] = (origin + dx).tolist()
    d.clamp_hi[:] = (origin + (np.asarray(dims) - 2) * dx).tolist()
    d.wall_friction = domain.wall_friction
    d.cell_vol = dx ** 3
    # removal thresholds compare float32 positions with float32(min - dx) (solver.py:230-234)
    d.remove_lo[:] = [-math.inf if domain.walls[f"{a}-"] else float(np.float32(lo - dx))
                      for a, lo in zip("xyz", domain.min)]
    d.remove_hi[:] = [math.inf if domain.walls[f"{a}+"] else float(np.float32(hi + dx))
                      for a, hi in zip("xyz", domain.max)]
    d.dx_f64 = dx
    return d


def material_table(materials: Sequence[MaterialParams], domain: Domain, particles_per_cell: int):
    """Per-material constants (solver.py:56-71); p_mass = float32(density) * float32(p_vol)."""
    p_vol32 = np.float32(domain.dx ** 3 / particles_per_cell)
    arr = (N.MaterialDesc * max(1, len(materials)))()
    for i, m in enumerate(materials):
        mu, lam = m.lame()
        e = arr[i]
        e.model = MODEL_IDS[m.model]
        e.mu, e.lam = mu, lam
        e.yield_stress, e.alpha = m.yield_stress, dp_alpha(m.friction_angle_deg)
        e.hardening, e.max_hardening = m.hardening, m.max_hardening
        e.crit_c, e.crit_s = m.critical_compression, m.critical_stretch
        e.p_mass = float(np.float32(m.density) * p_vol32)
        e.wave_speed = math.sqrt((lam + 2 * mu) * (m.max_hardening if m.model == "snow" else 1.0) / m.density)
    return arr


class HipSolver(Solver):
    """Dense-domain MLS-MPM on one AMD GPU through libmpm_hip.so."""

    def __init__(self, domain: Domain, materials: Sequence[MaterialParams], gravity: tuple[float, float, float],
                 particles_per_cell: int = 8, device: str | None = None):
        super().__init__(domain, materials, gravity, particles_per_cell, device)
        for m in self.materials:
            if m.model not in MODEL_IDS:
                raise NotImplementedError(f"material model {m.model!r} is not supported by the HIP backend")
        self._lib = N.lib()
        self.ordinal = device_ordinal(device)
        self.device = f"hip:{self.ordinal}"
        self.dx = domain.dx
        self._domain_desc = domain_desc(domain, self.gravity, particles_per_cell)
        self._materials_desc = material_table(self.materials, domain, particles_per_cell)
        self.dims = tuple(self._domain_desc.dims)
        self._n_colliders = 0
        handle = C.c_void_p()
        N.check(self._lib.mpm_create(self.ordinal, C.byref(self._domain_desc), self._materials_desc,
                                     len(self.materials), C.byref(handle)))
        self._ctx = handle

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.mpm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    # -- particles -------------------------------------------------------------------------------
    def add_particles(self, positions, velocities, material_ids, emitter_ids) -> None:
        n = len(positions)
        if n == 0:
            return
        x = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(n, 3))
        v = np.ascontiguousarray(np.asarray(velocities, np.float32).reshape(n, 3))
        mat = np.ascontiguousarray(np.asarray(material_ids, np.int32).reshape(n))
        em = np.ascontiguousarray(np.asarray(emitter_ids, np.int32).reshape(n))
        if mat.min() < 0 or mat.max() >= len(self.materials):
            raise ValueError("material id out of range")
        N.check(self._lib.mpm_add_particles(self._ctx, n, N.fptr(x), N.fptr(v), N.iptr(mat), N.iptr(em)))

    @property
    def particle_count(self) -> int:
        return int(self._lib.mpm_particle_count(self._ctx))

    def _download(self, keys: Sequence[str]) -> dict[str, np.ndarray]:
        n = self.particle_count
        shapes = {"x": ((n, 3), np.float32), "v": ((n, 3), np.float32), "C": ((n, 3, 3), np.float32),
                  "F": ((n, 3, 3), np.float32), "Jp": ((n,), np.float32), "mat": ((n,), np.int32),
                  "emitter": ((n,), np.int32)}
        out = {k: np.empty(*shapes[k]) for k in keys}
        ptrs = [(N.iptr(out[k]) if k in ("mat", "emitter") else N.fptr(out[k])) if k in out else None
                for k in _STATE_KEYS]
        N.check(self._lib.mpm_get_state(self._ctx, *ptrs))
        return out

    def get_particles(self) -> ParticleData:
        s = self._download(("x", "v", "mat", "emitter"))
        return ParticleData(s["x"], s["v"], s["mat"], s["emitter"])

    def save_state(self) -> dict[str, np.ndarray]:
        return self._download(_STATE_KEYS)

    def load_state(self, state: dict[str, np.ndarray]) -> None:
        n = len(state["x"])
        arrs = {
            "x": np.ascontiguousarray(np.asarray(state["x"], np.float32).reshape(n, 3)),
            "v": np.ascontiguousarray(np.asarray(state["v"], np.float32).reshape(n, 3)),
            "C": np.ascontiguousarray(np.asarray(state["C"], np.float32).reshape(n, 3, 3)),
            "F": np.ascontiguousarray(np.asarray(state["F"], np.float32).reshape(n, 3, 3)),
            "Jp": np.ascontiguousarray(np.asarray(state["Jp"], np.float32).reshape(n)),
            "mat": np.ascontiguousarray(np.asarray(state["mat"], np.int32).reshape(n)),
            "emitter": np.ascontiguousarray(np.asarray(state["emitter"], np.int32).reshape(n)),
        }
        N.check(self._lib.mpm_load_state(self._ctx, n, N.fptr(arrs["x"]), N.fptr(arrs["v"]), N.fptr(arrs["C"]),
                                         N.fptr(arrs["F"]), N.fptr(arrs["Jp"]), N.iptr(arrs["mat"]),
                                         N.iptr(arrs["emitter"])))

    # -- colliders -------------------------------------------------------------------------------
    def set_colliders(self, colliders: Sequence[ColliderGeometry]) -> None:
        cs = list(colliders)
        nv = np.array([len(c.mesh.verts) for c in cs], np.int32)
        nt = np.array([len(c.mesh.tris) for c in cs], np.int32)
        verts = np.ascontiguousarray(np.concatenate([c.mesh.verts for c in cs]) if cs else np.zeros((0, 3)),
                                     np.float32)
        tris = np.ascontiguousarray(np.concatenate([c.mesh.tris for c in cs]) if cs else np.zeros((0, 3)), np.int32)
        closed = np.array([int(c.closed) for c in cs], np.int32)
        friction = np.array([c.friction for c in cs], np.float32)
        N.check(self._lib.mpm_set_colliders(self._ctx, len(cs), N.iptr(nv), N.iptr(nt), N.fptr(verts), N.iptr(tris),
                                            N.iptr(closed), N.fptr(friction)))
        self._n_colliders = len(cs)
        identity = np.tile(np.eye(4, dtype=np.float32), (len(cs), 1, 1))
        self.set_collider_transforms(identity, identity)

    def set_collider_transforms(self, start: np.ndarray, end: np.ndarray) -> None:
        start = np.ascontiguousarray(np.asarray(start, np.float32).reshape(-1, 4, 4))
        end = np.ascontiguousarray(np.asarray(end, np.float32).reshape(-1, 4, 4))
        if len(start) != self._n_colliders or len(end) != self._n_colliders:
            raise ValueError(f"expected {self._n_colliders} collider transforms")
        # decomposed by the library in FP64 (geometry.decompose_transform's polar split, solver.py:258-260)
        N.check(self._lib.mpm_set_collider_transforms(self._ctx, self._n_colliders, N.fptr(start), N.fptr(end)))

    # -- stepping --------------------------------------------------------------------------------
    def advance(self, frame_dt: float) -> AdvanceStats:
        st = N.AdvanceStats()
        N.check(self._lib.mpm_advance(self._ctx, float(frame_dt), C.byref(st)))
        return AdvanceStats(int(st.substeps), float(st.dt), int(st.removed))

    def substep(self, dt: float, s: float = 1.0, frame_dt: float | None = None) -> None:
        """One fixed-dt substep at frame fraction ``s`` (not part of the Solver interface)."""
        N.check(self._lib.mpm_substep(self._ctx, float(dt), float(s), float(frame_dt if frame_dt else dt)))

    # -- queries ---------------------------------------------------------------------------------
    def is_occupied(self, points: np.ndarray) -> np.ndarray:
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        if len(pts) == 0:
            return np.zeros(0, bool)
        out = np.zeros(len(pts), np.uint8)
        N.check(self._lib.mpm_is_occupied(self._ctx, len(pts), N.fptr(pts), N.u8ptr(out)))
        return out.astype(bool)

    def sample_mesh_volume(self, verts: np.ndarray, tris: np.ndarray, spacing: float, seed) -> np.ndarray:
        """``mpm_core.geometry.sample_mesh_volume`` with the lattice test on the GPU (geometry.py:89-155).

        The inside lattice points come from the library in the reference's order with its FP64 arithmetic;
        the jitter is the reference's single ``default_rng(seed).uniform`` draw over them, so the result is
        the same array. Not part of the Solver interface: the runner uses it when a backend offers it.
        """
        if not mesh_is_closed(tris):
            raise ValueError("emitter mesh is not closed (every edge must be shared by exactly two triangles)")
        v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        count = C.c_int64()
        N.check(self._lib.mpm_lattice_inside(self._ctx, N.dptr(v), len(v), N.iptr(t), len(t), float(spacing), 0,
                                             None, C.byref(count)))
        if count.value == 0:
            return np.zeros((0, 3), np.float32)
        pts = np.empty((count.value, 3), np.float64)
        N.check(self._lib.mpm_lattice_inside(self._ctx, N.dptr(v), len(v), N.iptr(t), len(t), float(spacing),
                                             count.value, N.dptr(pts), C.byref(count)))
        pts += np.random.default_rng(seed).uniform(-0.25, 0.25, pts.shape) * spacing
        return pts.astype(np.float32)

    # -- phase hooks and measurement (parity tests, benchmarks) ----------------------------------
    def debug_p2g(self, dt: float) -> tuple[np.ndarray, np.ndarray]:
        N.check(self._lib.mpm_debug_p2g(self._ctx, float(dt)))
        return self.get_grid(0)

    def debug_grid_update(self, dt: float, s: float = 1.0, frame_dt: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
        N.check(self._lib.mpm_debug_grid_update(self._ctx, float(dt), float(s), float(frame_dt)))
        return self.get_grid(1)

    def debug_g2p(self, dt: float, s: float = 1.0, frame_dt: float = 1.0) -> None:
        N.check(self._lib.mpm_debug_g2p(self._ctx, float(dt), float(s), float(frame_dt)))

    def get_grid(self, which: int) -> tuple[np.ndarray, np.ndarray]:
        gv = np.zeros(self.dims + (3,), np.float32)
        gm = np.zeros(self.dims, np.float32)
        N.check(self._lib.mpm_get_grid(self._ctx, which, N.fptr(gv), N.fptr(gm)))
        return gv, gm

    def set_grid(self, grid_v: np.ndarray, grid_m: np.ndarray) -> None:
        gv = np.ascontiguousarray(grid_v, np.float32).reshape(self.dims + (3,))
        gm = np.ascontiguousarray(grid_m, np.float32).reshape(self.dims)
        N.check(self._lib.mpm_set_grid(self._ctx, N.fptr(gv), N.fptr(gm)))

    def binning(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = self.particle_count
        nb = C.c_int64()
        N.check(self._lib.mpm_get_binning(self._ctx, None, None, None, C.byref(nb)))
        keys = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        offsets = np.zeros(nb.value + 1, np.int32)
        N.check(self._lib.mpm_get_binning(self._ctx, N.iptr(keys), N.iptr(perm), N.iptr(offsets), C.byref(nb)))
        return keys, perm, offsets

    def debug_sort(self) -> None:
        """Bin the particles by brick now (stable radix sort + reorder), as the next substep would."""
        N.check(self._lib.mpm_debug_sort(self._ctx))

    def set_option(self, option: int, value: int) -> None:
        """Tuning knobs of the library (mpm_plugin_amd._native.OPT_*)."""
        N.check(self._lib.mpm_set_option(self._ctx, int(option), int(value)))

    def get_option(self, option: int) -> int:
        return int(self._lib.mpm_get_option(self._ctx, int(option)))

    def set_timing(self, enabled: bool, phases=None) -> None:
        """Bracket kernel launches with HIP events: every kernel class, or only those named in ``phases``
        (names of ``N.KERNEL_NAMES``)."""
        mask = 0
        if enabled:
            names = N.KERNEL_NAMES if phases is None else phases
            for name in names:
                mask |= 1 << N.KERNEL_NAMES.index(name)
        N.check(self._lib.mpm_set_timing(self._ctx, mask))

    def reset_timing(self) -> None:
        N.check(self._lib.mpm_reset_timing(self._ctx))

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        out = {}
        for k, name in enumerate(N.KERNEL_NAMES):
            ms, count = C.c_double(), C.c_int64()
            N.check(self._lib.mpm_get_kernel_time(self._ctx, k, C.byref(ms), C.byref(count)))
            out[name] = (ms.value, count.value)
        return out

    def synchronize(self) -> None:
        N.check(self._lib.mpm_synchronize(self._ctx))
