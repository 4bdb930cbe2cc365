"""Shared test helpers: matrices, meshes, particle bodies, solvers, a small scene and error measures.

Meshes come from the workload module (mpm_plugin_amd.workloads), so tests and benchmarks build bodies and
colliders from the same geometry.
"""

from __future__ import annotations

import numpy as np

from mpm_core.materials import load_preset
from mpm_core.scene import WALL_SIDES, Collider, Domain, Emitter, MeshData, SceneDescription
from mpm_plugin_amd import workloads

ALL_WALLS = dict.fromkeys(WALL_SIDES, True)


# ---- 4x4 matrices (float32, column vectors) -------------------------------------------------------------

def translate(x, y, z) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    return m


def scale(x, y=None, z=None) -> np.ndarray:
    return np.diag([x, x if y is None else y, x if z is None else z, 1.0]).astype(np.float32)


def _rotation(axis: int, degrees: float) -> np.ndarray:
    """Right-handed rotation about coordinate axis ``axis`` by ``degrees``."""
    c, s = np.cos(np.radians(degrees)), np.sin(np.radians(degrees))
    i, j = [(1, 2), (2, 0), (0, 1)][axis]  # the plane the rotation turns, in its positive sense
    m = np.eye(4, dtype=np.float32)
    m[i, i] = m[j, j] = c
    m[i, j], m[j, i] = -s, s
    return m


def rot_x(degrees: float) -> np.ndarray:
    return _rotation(0, degrees)


def rot_y(degrees: float) -> np.ndarray:
    return _rotation(1, degrees)


# ---- meshes -------------------------------------------------------------------------------------------

def unit_cube_mesh(centered: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """Closed unit cube wound outward: [0,1]^3, or [-0.5,0.5]^3 when ``centered``."""
    v, t = workloads.box_mesh()
    return (v - np.float32(0.5) if centered else v), t


def plane_mesh(size: float) -> tuple[np.ndarray, np.ndarray]:
    """An open square of side ``size`` in the xy plane, centred on the origin, normal +z."""
    corners = np.array([(-1, -1), (1, -1), (1, 1), (-1, 1)], np.float32) * np.float32(size / 2)
    verts = np.concatenate([corners, np.zeros((4, 1), np.float32)], axis=1)
    return verts, np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def icosphere(radius: float, subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    v, t = workloads.icosphere_f64(subdivisions)
    return (v * radius).astype(np.float32), t


# ---- particle bodies and solvers ---------------------------------------------------------------------------

def box_particles(lo, hi, dx: float, seed: int = 0) -> np.ndarray:
    """The box [lo, hi] sampled at 8 particles per cell (spacing dx/2)."""
    return workloads.box_points(lo, hi, dx / 2, seed)


def sphere_particles(center, radius: float, dx: float, seed: int = 0) -> np.ndarray:
    """A ball sampled at 8 particles per cell through a 4-times subdivided icosphere."""
    from mpm_core.geometry import sample_mesh_volume

    v, t = icosphere(radius, 4)
    return sample_mesh_volume(v + np.asarray(center, np.float32), t, dx / 2, seed)


def make_solver(device, domain, materials, gravity=(0.0, 0.0, -9.81), backend: str = "hip"):
    from mpm_core.backends import create_solver

    return create_solver(backend, domain=domain, materials=list(materials), gravity=gravity, device=device)


def add(solver, positions, material_id: int = 0, emitter_id: int = 0, velocity=(0.0, 0.0, 0.0)) -> None:
    n = len(positions)
    vel = np.broadcast_to(np.asarray(velocity, np.float32), (n, 3)).copy()
    solver.add_particles(positions, vel, np.full(n, material_id, np.int32), np.full(n, emitter_id, np.int32))


def solver_with_box(device, domain, material, lo, hi):
    s = make_solver(device, domain, [material])
    add(s, box_particles(lo, hi, domain.dx))
    return s


def solver_with_cube(device, domain, lo, hi):
    """A jelly box (the built-in preset)."""
    return solver_with_box(device, domain, load_preset("jelly").material, lo, hi)


def solver_with_sphere(device, domain, material, center, radius):
    s = make_solver(device, domain, [material])
    add(s, sphere_particles(center, radius, domain.dx))
    return s


def solver_with_presets(device, domain, boxes):
    """``boxes``: [(preset id, lo, hi)]; box i gets material i and emitter i."""
    s = make_solver(device, domain, [load_preset(p).material for p, _, _ in boxes])
    for i, (_, lo, hi) in enumerate(boxes):
        add(s, box_particles(lo, hi, domain.dx, seed=i), material_id=i, emitter_id=i)
    return s


def run(solver, frames: int, fps: float = 24.0) -> None:
    for _ in range(frames):
        solver.advance(1.0 / fps)


# ---- a small scene -------------------------------------------------------------------------------------

def small_scene(frame_end: int = 10, gravity=(0.0, 0.0, -9.81), particles_per_cell: int = 8,
                emitter_frames: int | None = None, emitter_offset=(0.0, 0.0, 0.0), kind: str = "fill",
                velocity=(0.0, 0.0, 0.0), emit_frames: tuple[int, int] | None = None,
                resolution: int = 32, preset: str = "jelly") -> SceneDescription:
    """Frames 1..frame_end in the unit-cube domain: one 0.2 m cube emitter ``emitter_0`` with its corner at
    (0.4, 0.4, 0.1) + ``emitter_offset`` (``emitter_frames`` transforms, default one per frame; it emits over
    ``emit_frames``, default frame 1) and one open 0.8 m square collider ``floor`` at z = 0.05."""
    frames = frame_end if emitter_frames is None else emitter_frames
    corner = np.add((0.4, 0.4, 0.1), emitter_offset)
    first, last = emit_frames if emit_frames is not None else (1, 1)
    emitter = Emitter(name="emitter_0", kind=kind, material=0, mesh=MeshData(*unit_cube_mesh()),
                      transforms=np.stack([translate(*corner) @ scale(0.2)] * frames), velocity=tuple(velocity),
                      frame_start=first, frame_end=last)
    floor = Collider(name="floor", mesh=MeshData(*plane_mesh(0.8)),
                     transforms=np.stack([translate(0.5, 0.5, 0.05)] * frame_end), friction=0.5, closed=False)
    return SceneDescription(domain=Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), resolution),
                            materials=[load_preset(preset).material], emitters=[emitter], colliders=[floor],
                            gravity=tuple(gravity), frame_start=1, frame_end=frame_end,
                            particles_per_cell=particles_per_cell)


# ---- error measures ------------------------------------------------------------------------------------

def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """Norm-wise relative error ||a - b|| / ||b|| in FP64 (||a|| when b vanishes)."""
    a64, b64 = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale_b = np.linalg.norm(b64)
    return float(np.linalg.norm(a64 - b64) / scale_b) if scale_b > 0 else float(np.linalg.norm(a64))


def c_close(a_C: np.ndarray, b_C: np.ndarray, b_v: np.ndarray, dx: float, rel: float = 1e-4) -> tuple[bool, float]:
    """APIC C against a reference: ||dC|| <= rel ||C|| + 1e-6 (4/dx) ||v||.

    C = 4/dx^2 sum_i w_i v_i (x_i - x_p)^T is a difference of grid velocities; for a body moving nearly
    rigidly it is rounding, and its size only means something against the velocity-gradient scale 4/dx |v|.
    FP32 grid velocities put about 1e-7 of that scale into C, so the floor is 1e-6 of it; where C is large the
    relative bar decides. Returns (ok, error as a fraction of the bound)."""
    d = float(np.linalg.norm(np.asarray(a_C, np.float64) - np.asarray(b_C, np.float64)))
    bound = rel * float(np.linalg.norm(np.asarray(b_C, np.float64))) + \
        1e-6 * 4.0 / dx * float(np.linalg.norm(np.asarray(b_v, np.float64)))
    return d <= bound, d / bound if bound > 0 else float(d > 0)
