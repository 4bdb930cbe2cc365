"""ctypes wrapper of the CPU oracle (cpu_ref.cpp). TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the CPU-baseline leg of bench.py import this module, as the checker
or the timed CPU baseline; the HIP product never does.

The derived constants are restated here from the reference's host code independently of the product:
solver.py:37-54 (dx, inv_dx, p_vol, dims with PAD = 3, origin = min - 3 dx, wall flags, clamp box, wall
node indices with the +-1e-3 guards), solver.py:56-71 (material table, DP alpha :263-265, P-wave speed
with worst-case snow hardening :60-62), solver.py:105 (p_mass = float32(density) * float32(p_vol)),
solver.py:176-185 (CFL dt and step count), solver.py:222-240 (escaped-particle removal) and
solver.py:164-172 (collider pose decomposition and AABB margin).
"""

from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

PAD = 3
CFL = 0.4
MODEL_IDS = {"elastic": 0, "von_mises": 1, "drucker_prager": 2, "snow": 3}
_HERE = os.path.dirname(os.path.abspath(__file__))
# MPM_ORACLE_LIB names another build of cpu_ref.cpp (the mutation tests load deliberately broken copies)
_LIB_PATH = os.environ.get("MPM_ORACLE_LIB") or os.path.join(_HERE, "libmpm_oracle.so")


class OracleDomain(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_float * 3), ("dx", C.c_float), ("inv_dx", C.c_float),
                ("p_vol", C.c_float), ("gravity", C.c_float * 3), ("wall_lo_node", C.c_int32 * 3),
                ("wall_hi_node", C.c_int32 * 3), ("lo_on", C.c_int32 * 3), ("hi_on", C.c_int32 * 3),
                ("wall_lo", C.c_float * 3), ("wall_hi", C.c_float * 3), ("clamp_lo", C.c_float * 3),
                ("clamp_hi", C.c_float * 3), ("wall_friction", C.c_float), ("cell_vol", C.c_float)]


class OracleMaterial(C.Structure):
    _fields_ = [("model", C.c_int32), ("mu", C.c_float), ("lam", C.c_float), ("yield_stress", C.c_float),
                ("alpha", C.c_float), ("hardening", C.c_float), ("max_hardening", C.c_float),
                ("crit_c", C.c_float), ("crit_s", C.c_float), ("p_mass", C.c_float)]


class OracleCollider(C.Structure):
    _fields_ = [("verts", C.POINTER(C.c_float)), ("tris", C.POINTER(C.c_int32)), ("n_verts", C.c_int32),
                ("n_tris", C.c_int32), ("closed", C.c_int32), ("friction", C.c_float),
                ("aabb_lo", C.c_float * 3), ("aabb_hi", C.c_float * 3), ("margin", C.c_float),
                ("t0", C.c_float * 3), ("t1", C.c_float * 3), ("q0", C.c_float * 4), ("q1", C.c_float * 4),
                ("S0", C.c_float * 9), ("S1", C.c_float * 9)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle/build.sh)")
        L = C.CDLL(_LIB_PATH)
        fp, dp, ip, u8 = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        D, M, K = C.POINTER(OracleDomain), C.POINTER(OracleMaterial), C.POINTER(OracleCollider)
        i64, i32, f64 = C.c_int64, C.c_int32, C.c_double
        sig = {
            "oracle_svd3_f64": ([dp, dp, dp, dp], None),
            "oracle_svd3_f32": ([fp, fp, fp, fp], None),
            "oracle_tau_f64": ([i32, f64, f64, f64, f64, f64, dp, dp], None),
            "oracle_project_f64": ([i32, f64, f64, f64, f64, f64, f64, dp, dp], None),
            "oracle_bspline_w_f64": ([f64, i32], f64),
            "oracle_p2g_f32": ([D, M, i64, fp, fp, fp, fp, fp, ip, f64, fp, fp, i32], None),
            "oracle_p2g_f64": ([D, M, i64, fp, fp, fp, fp, fp, ip, f64, dp, dp, i32], None),
            "oracle_grid_update_f32": ([D, f64, fp, fp, i32], None),
            "oracle_grid_update_f64": ([D, f64, dp, dp, i32], None),
            "oracle_grid_colliders_f32": ([D, K, i32, f64, f64, fp, fp, i32], None),
            "oracle_g2p_f32": ([D, M, i64, fp, fp, fp, fp, fp, ip, fp, f64, i32], None),
            "oracle_g2p_f64": ([D, M, i64, fp, fp, fp, fp, fp, ip, dp, f64, i32], None),
            "oracle_project_particles_f32": ([D, K, i32, i64, fp, fp, f64, f64, i32], None),
            "oracle_collider_sdf_f64": ([K, f64, dp, f64, dp, dp], None),
            "oracle_collider_velocity_f64": ([K, f64, dp, f64, dp], None),
            "oracle_collider_phi_f32": ([D, K, i32, i64, fp, f64, fp, i32], None),
            "oracle_substep_f32": ([D, M, K, i32, i64, fp, fp, fp, fp, fp, ip, f64, f64, f64, fp, fp, i32], None),
            "oracle_substep_f64": ([D, M, K, i32, i64, fp, fp, fp, fp, fp, ip, f64, f64, f64, dp, dp, i32], None),
            "oracle_max_speed": ([i64, fp], C.c_float),
            "oracle_is_occupied": ([D, i64, fp, i64, fp, u8], None),
            "oracle_max_threads": ([], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# --------------------------------------------------------------------------------------------------
# host-side constants (restated from solver.py)

def make_domain(domain, gravity, particles_per_cell: int = 8) -> OracleDomain:
    dx = domain.dx
    dims = tuple(s + 2 * PAD for s in domain.grid_shape)
    origin = np.asarray(domain.min, np.float64) - PAD * dx
    d = OracleDomain()
    d.dims[:] = dims
    d.origin[:] = [float(o) for o in origin]
    d.dx = dx
    d.inv_dx = 1.0 / dx
    d.p_vol = dx ** 3 / particles_per_cell
    d.gravity[:] = [float(g) for g in gravity]
    d.wall_lo_node[:] = [int(math.floor((lo - o) / dx + 1e-3)) for lo, o in zip(domain.min, origin)]
    d.wall_hi_node[:] = [int(math.ceil((hi - o) / dx - 1e-3)) for hi, o in zip(domain.max, origin)]
    d.lo_on[:] = [int(domain.walls[f"{a}-"]) for a in "xyz"]
    d.hi_on[:] = [int(domain.walls[f"{a}+"]) for a in "xyz"]
    d.wall_lo[:] = list(domain.min)
    d.wall_hi[:] = list(domain.max)
    d.clamp_lo[:] = [float(v) for v in origin + dx]
    d.clamp_hi[:] = [float(v) for v in origin + (np.asarray(dims) - 2) * dx]
    d.wall_friction = domain.wall_friction
    d.cell_vol = dx ** 3
    return d


def dp_alpha(friction_angle_deg: float) -> float:
    s = math.sin(math.radians(friction_angle_deg))
    return math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


def make_materials(materials, domain, particles_per_cell: int = 8):
    p_vol32 = np.float32(domain.dx ** 3 / particles_per_cell)
    arr = (OracleMaterial * max(1, len(materials)))()
    for i, m in enumerate(materials):
        mu, lam = m.lame()
        arr[i].model = MODEL_IDS[m.model]
        arr[i].mu, arr[i].lam = mu, lam
        arr[i].yield_stress = m.yield_stress
        arr[i].alpha = dp_alpha(m.friction_angle_deg)
        arr[i].hardening, arr[i].max_hardening = m.hardening, m.max_hardening
        arr[i].crit_c, arr[i].crit_s = m.critical_compression, m.critical_stretch
        arr[i].p_mass = float(np.float32(m.density) * p_vol32)
    return arr


def wave_speeds(materials) -> np.ndarray:
    out = []
    for m in materials:
        mu, lam = m.lame()
        out.append(math.sqrt((lam + 2 * mu) * (m.max_hardening if m.model == "snow" else 1.0) / m.density))
    return np.array(out, np.float64)


def choose_dt(frame_dt: float, vmax: float, dx: float, wave_speed_present: float) -> tuple[float, int]:
    if not math.isfinite(vmax):
        vmax = 0.0
    dt = min(CFL * dx / (wave_speed_present + vmax), frame_dt)
    steps = math.ceil(frame_dt / dt - 1e-9)
    return frame_dt / steps, steps


def escaped_mask(x: np.ndarray, domain) -> np.ndarray:
    out = np.zeros(len(x), bool)
    for a, ax in enumerate("xyz"):
        if not domain.walls[f"{ax}-"]:
            out |= x[:, a] < domain.min[a] - domain.dx
        if not domain.walls[f"{ax}+"]:
            out |= x[:, a] > domain.max[a] + domain.dx
    return out


def _quat_from_rotation(R: np.ndarray) -> np.ndarray:
    """Rotation matrix -> unit quaternion (x, y, z, w) by the largest-diagonal (Shepperd) method."""
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = 2.0 * math.sqrt(1.0 + tr)
        q = [(R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s]
    else:
        i = int(np.argmax([R[0, 0], R[1, 1], R[2, 2]]))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * math.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
        q = [0.0, 0.0, 0.0, (R[k, j] - R[j, k]) / s]
        q[i] = 0.25 * s
        q[j] = (R[j, i] + R[i, j]) / s
        q[k] = (R[k, i] + R[i, k]) / s
    q = np.asarray(q)
    return q / np.linalg.norm(q)


def decompose(m: np.ndarray):
    """(t, q xyzw, S) with M = [R S | t], R proper (geometry.py:30-39), computed independently."""
    A = np.asarray(m, np.float64)[:3, :3]
    U, s, Vt = np.linalg.svd(A)
    if np.linalg.det(U @ Vt) < 0:
        U[:, 2] = -U[:, 2]
        s[2] = -s[2]
    return np.asarray(m, np.float64)[:3, 3], _quat_from_rotation(U @ Vt), Vt.T @ np.diag(s) @ Vt


class Colliders:
    """Oracle collider set: meshes plus start/end poses, kept alive for the ctypes pointers."""

    def __init__(self, geometries, dx: float):
        self.geoms = list(geometries)
        self.dx = dx
        self.arr = (OracleCollider * max(1, len(self.geoms)))()
        self._keep = []
        for i, g in enumerate(self.geoms):
            v = np.ascontiguousarray(g.mesh.verts, np.float32)
            t = np.ascontiguousarray(g.mesh.tris, np.int32)
            self._keep += [v, t]
            c = self.arr[i]
            c.verts, c.tris = _ptr(v, C.c_float), _ptr(t, C.c_int32)
            c.n_verts, c.n_tris = len(v), len(t)
            c.closed, c.friction = int(g.closed), g.friction
            c.aabb_lo[:] = [float(a) for a in v.min(0)]
            c.aabb_hi[:] = [float(a) for a in v.max(0)]
        eye = np.tile(np.eye(4, dtype=np.float32), (len(self.geoms), 1, 1))
        self.set_transforms(eye, eye)

    def set_transforms(self, start, end) -> None:
        start = np.asarray(start, np.float32).reshape(-1, 4, 4)
        end = np.asarray(end, np.float32).reshape(-1, 4, 4)
        for i, (a, b) in enumerate(zip(start, end)):
            (t0, q0, S0), (t1, q1, S1) = decompose(a), decompose(b)
            if np.dot(q0, q1) < 0:
                q1 = -q1
            c = self.arr[i]
            c.t0[:], c.t1[:] = [float(v) for v in t0], [float(v) for v in t1]
            c.q0[:], c.q1[:] = [float(v) for v in q0], [float(v) for v in q1]
            c.S0[:], c.S1[:] = [float(v) for v in S0.ravel()], [float(v) for v in S1.ravel()]
            inv_scale = max(1.0 / max(np.linalg.svd(m[:3, :3], compute_uv=False).min(), 1e-12) for m in (a, b))
            c.margin = 3.0 * self.dx * inv_scale

    def __len__(self):
        return len(self.geoms)


# --------------------------------------------------------------------------------------------------
# whole-state oracle solver (the reference's WarpSolver data flow, solver.py:26-255, on the CPU)

@dataclass
class State:
    x: np.ndarray
    v: np.ndarray
    C: np.ndarray
    F: np.ndarray
    Jp: np.ndarray
    mat: np.ndarray
    emitter: np.ndarray

    @classmethod
    def from_dict(cls, s: dict) -> State:
        return cls(*(np.array(s[k], t, copy=True, order="C") for k, t in
                     (("x", np.float32), ("v", np.float32), ("C", np.float32), ("F", np.float32),
                      ("Jp", np.float32), ("mat", np.int32), ("emitter", np.int32))))

    def to_dict(self) -> dict:
        return {k: getattr(self, k).copy() for k in ("x", "v", "C", "F", "Jp", "mat", "emitter")}


class OracleSolver:
    """CPU oracle with the reference's solver semantics. ``precision`` is "f32" or "f64"."""

    def __init__(self, domain, materials, gravity, particles_per_cell: int = 8, precision: str = "f32",
                 threads: int | None = None):
        self.domain = domain
        self.materials = list(materials)
        self.ppc = particles_per_cell
        self.d = make_domain(domain, gravity, particles_per_cell)
        self.m = make_materials(self.materials, domain, particles_per_cell)
        self.speeds = wave_speeds(self.materials)
        self.precision = precision
        self.threads = threads or lib().oracle_max_threads()
        self.colliders = Colliders([], domain.dx)
        self.frame_dt = 1.0
        dims = tuple(self.d.dims)
        ft = np.float32 if precision == "f32" else np.float64
        self.gv = np.zeros(dims + (3,), ft)
        self.gm = np.zeros(dims, ft)
        self.s = State(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3, 3), np.float32),
                       np.zeros((0, 3, 3), np.float32), np.zeros(0, np.float32), np.zeros(0, np.int32),
                       np.zeros(0, np.int32))

    @property
    def n(self) -> int:
        return len(self.s.x)

    def add_particles(self, x, v, mat, emitter) -> None:
        n = len(x)
        if n == 0:
            return
        new = State(np.asarray(x, np.float32).reshape(n, 3), np.asarray(v, np.float32).reshape(n, 3),
                    np.zeros((n, 3, 3), np.float32), np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)),
                    np.ones(n, np.float32), np.asarray(mat, np.int32), np.asarray(emitter, np.int32))
        self.s = State(*(np.ascontiguousarray(np.concatenate([getattr(self.s, k), getattr(new, k)]))
                         for k in ("x", "v", "C", "F", "Jp", "mat", "emitter")))

    def load_state(self, st: dict) -> None:
        self.s = State.from_dict(st)

    def save_state(self) -> dict:
        return self.s.to_dict()

    def set_colliders(self, geoms) -> None:
        self.colliders = Colliders(geoms, self.domain.dx)

    def set_collider_transforms(self, start, end) -> None:
        self.colliders.set_transforms(start, end)

    def _args(self):
        s = self.s
        return (_ptr(s.x, C.c_float), _ptr(s.v, C.c_float), _ptr(s.C, C.c_float), _ptr(s.F, C.c_float),
                _ptr(s.Jp, C.c_float))

    def substep(self, dt: float, s_frac: float = 1.0) -> None:
        L = lib()
        ct = C.c_float if self.precision == "f32" else C.c_double
        fn = L.oracle_substep_f32 if self.precision == "f32" else L.oracle_substep_f64
        fn(C.byref(self.d), self.m, self.colliders.arr, len(self.colliders), self.n, *self._args(),
           _ptr(self.s.mat, C.c_int32), dt, s_frac, 1.0 / self.frame_dt, _ptr(self.gv, ct), _ptr(self.gm, ct),
           self.threads)

    def choose_dt(self, frame_dt: float) -> tuple[float, int]:
        vmax = float(lib().oracle_max_speed(self.n, _ptr(self.s.v, C.c_float)))
        present = np.unique(self.s.mat)
        return choose_dt(frame_dt, vmax, self.domain.dx, float(self.speeds[present].max()))

    def advance(self, frame_dt: float):
        if self.n == 0:
            return 0, frame_dt, 0
        self.frame_dt = frame_dt
        dt, steps = self.choose_dt(frame_dt)
        for k in range(steps):
            self.substep(dt, (k + 1) / steps)
        out = escaped_mask(self.s.x, self.domain)
        removed = int(out.sum())
        if removed:
            keep = ~out
            self.s = State(*(np.ascontiguousarray(getattr(self.s, k)[keep])
                             for k in ("x", "v", "C", "F", "Jp", "mat", "emitter")))
        return steps, dt, removed

    def is_occupied(self, pts) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.uint8)
        if len(pts):
            lib().oracle_is_occupied(C.byref(self.d), self.n, _ptr(self.s.x, C.c_float), len(pts),
                                     _ptr(pts, C.c_float), _ptr(out, C.c_uint8))
        return out.astype(bool)

    # phase hooks (one kernel each) ------------------------------------------------------------
    def p2g(self, dt: float):
        ct = C.c_float if self.precision == "f32" else C.c_double
        self.gv[...] = 0
        self.gm[...] = 0
        fn = lib().oracle_p2g_f32 if self.precision == "f32" else lib().oracle_p2g_f64
        s = self.s
        fn(C.byref(self.d), self.m, self.n, _ptr(s.x, C.c_float), _ptr(s.v, C.c_float), _ptr(s.C, C.c_float),
           _ptr(s.F, C.c_float), _ptr(s.Jp, C.c_float), _ptr(s.mat, C.c_int32), dt, _ptr(self.gv, ct),
           _ptr(self.gm, ct), self.threads)
        return self.gv.copy(), self.gm.copy()

    def grid_update(self, dt: float, s_frac: float = 1.0):
        ct = C.c_float if self.precision == "f32" else C.c_double
        fn = lib().oracle_grid_update_f32 if self.precision == "f32" else lib().oracle_grid_update_f64
        fn(C.byref(self.d), dt, _ptr(self.gv, ct), _ptr(self.gm, ct), self.threads)
        if len(self.colliders):
            assert self.precision == "f32"
            lib().oracle_grid_colliders_f32(C.byref(self.d), self.colliders.arr, len(self.colliders), s_frac,
                                            1.0 / self.frame_dt, _ptr(self.gv, C.c_float),
                                            _ptr(self.gm, C.c_float), self.threads)
        return self.gv.copy()

    def g2p(self, dt: float, s_frac: float = 1.0, grid_v: np.ndarray | None = None):
        ct = C.c_float if self.precision == "f32" else C.c_double
        if grid_v is not None:
            self.gv[...] = grid_v
        fn = lib().oracle_g2p_f32 if self.precision == "f32" else lib().oracle_g2p_f64
        fn(C.byref(self.d), self.m, self.n, *self._args(), _ptr(self.s.mat, C.c_int32), _ptr(self.gv, ct), dt,
           self.threads)
        if len(self.colliders):
            self.project_particles(s_frac)

    def collider_phi(self, x, s_frac: float = 1.0) -> np.ndarray:
        """Smallest signed distance of each point to the colliders at frame fraction s_frac."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        out = np.full(len(x), 1e30, np.float32)
        if len(x) and len(self.colliders):
            lib().oracle_collider_phi_f32(C.byref(self.d), self.colliders.arr, len(self.colliders), len(x),
                                          _ptr(x, C.c_float), s_frac, _ptr(out, C.c_float), self.threads)
        return out

    def project_particles(self, s_frac: float = 1.0) -> None:
        """project_particles alone (colliders.py:120-149) on the current state."""
        lib().oracle_project_particles_f32(C.byref(self.d), self.colliders.arr, len(self.colliders), self.n,
                                           _ptr(self.s.x, C.c_float), _ptr(self.s.v, C.c_float), s_frac,
                                           1.0 / self.frame_dt, self.threads)


def svd3(F: np.ndarray, precision: str = "f64"):
    if precision == "f64":
        F = np.ascontiguousarray(F, np.float64)
        U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
        lib().oracle_svd3_f64(_ptr(F, C.c_double), _ptr(U, C.c_double), _ptr(s, C.c_double), _ptr(V, C.c_double))
    else:
        F = np.ascontiguousarray(F, np.float32)
        U, s, V = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        lib().oracle_svd3_f32(_ptr(F, C.c_float), _ptr(U, C.c_float), _ptr(s, C.c_float), _ptr(V, C.c_float))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def tau(model: str, F, mu, lam, hardening=0.0, max_hardening=5.0, Jp=1.0) -> np.ndarray:
    F = np.ascontiguousarray(F, np.float64)
    out = np.zeros(9)
    lib().oracle_tau_f64(MODEL_IDS[model], mu, lam, hardening, max_hardening, Jp, _ptr(F, C.c_double),
                         _ptr(out, C.c_double))
    return out.reshape(3, 3)


def project(model: str, sig, mu=1.0, lam=1.0, yield_stress=0.0, alpha=0.0, crit_c=0.0, crit_s=0.0) -> np.ndarray:
    sig = np.ascontiguousarray(sig, np.float64)
    out = np.zeros(3)
    lib().oracle_project_f64(MODEL_IDS[model], mu, lam, yield_stress, alpha, crit_c, crit_s,
                             _ptr(sig, C.c_double), _ptr(out, C.c_double))
    return out


def bspline_w(fx: float, i: int) -> float:
    return lib().oracle_bspline_w_f64(fx, i)


def collider_sdf(geom, xf_start, xf_end, s: float, p, thickness: float, dx: float):
    cs = Colliders([geom], dx)
    cs.set_transforms(np.asarray(xf_start)[None], np.asarray(xf_end)[None])
    p = np.ascontiguousarray(p, np.float64)
    n, phi = np.zeros(3), np.zeros(1)
    lib().oracle_collider_sdf_f64(cs.arr, s, _ptr(p, C.c_double), thickness, _ptr(n, C.c_double),
                                  _ptr(phi, C.c_double))
    return n, float(phi[0])


def collider_velocity(geom, xf_start, xf_end, s: float, p, inv_frame_dt: float, dx: float):
    cs = Colliders([geom], dx)
    cs.set_transforms(np.asarray(xf_start)[None], np.asarray(xf_end)[None])
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros(3)
    lib().oracle_collider_velocity_f64(cs.arr, s, _ptr(p, C.c_double), inv_frame_dt, _ptr(out, C.c_double))
    return out
