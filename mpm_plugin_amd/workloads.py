"""Synthetic workloads: the benchmark and parity scenes, built through the public interface.

* ``elastic_block``: unit-cube domain at 64^3 (dx = 1/64), walls with friction 1e6 (sticky), a 0.2 m
  elastic cube centred at (0.5, 0.6, 0.5) sampled at 8 particles per cell (26^3 = 17,576 particles,
  jitter +-0.25 of the particle spacing), E = 1e4, nu = 0.3, rho = 1000, gravity (0, -9.8, 0),
  dt = 1e-4 for 1000 substeps.
* ``multi_material``: unit cube at 128^3, gravity (0, 0, -9.81), walls on; a von Mises sphere
  (r 0.15 at (0.22, 0.5, 0.30), E 5e3, nu 0.45, yield 200, rho 1000), a Drucker-Prager column
  ([0.4, 0.6]^2 x [0, 0.72], E 3.5e4, nu 0.3, 30 deg, rho 1600) and a snow sphere (r 0.16 at
  (0.80, 0.5, 0.35), E 1.4e5, nu 0.2, xi 10, theta_c 2.5e-2, theta_s 7.5e-3, h_max 5, rho 400) thrown at
  (-3, 0, 0) m/s: about 1.0 M particles at 8 per cell; one closed icosphere collider (2562 vertices,
  r 0.1, friction 0.5) moving from (0.5, 0.5, 0.88) to (0.5, 0.5, 0.70) over 0.1 s; dt = 5e-5 for 2000
  substeps (one substep per advance: the CFL step is about 6.6e-5).
* ``scene_mix``: many independent scenes for the batched engine. Grid edge uniform in
  {32, 48, 64, 96, 128} over the unit cube; particle count log-uniform in [5k, 400k], at most
  0.35 * 8 * edge^3 so the body fits at 8 particles per cell; one body (sphere or box) of one material
  drawn uniformly from the four models, its parameters jittered by a uniform factor in [0.8, 1.2] around
  the values above (elastic: the elastic block's); random velocity up to 3 m/s; 0-2 icosphere colliders
  (162 vertices) on straight paths; frames of 1/24 s. Bodies are jittered lattices (spacing dx/2, jitter
  +-0.25 of the spacing), generated directly rather than through mesh sampling for speed.
* ``served_requests``: bake requests for the server, Poisson arrivals (20 per second by default), each a
  scene description drawn like ``scene_mix`` (the body a fill emitter, sampled by the runner) with a
  lognormal frame count (median 48, sigma 0.5, clamped to [1, 240]) and a frame rate chosen so that the
  first frame takes a substep count drawn uniformly from [20, 200] at the scene's CFL step (rather than
  fixing dt = (1/24)/k, which could undercut the CFL step; the runner's Solver.advance chooses substeps
  by CFL either way).

Particles are sampled with mpm_core.geometry.sample_mesh_volume, as the runner emits them.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from mpm_core.geometry import sample_mesh_volume, transform_points
from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain, MeshData
from mpm_core.solver import ColliderGeometry


def _translate(x, y, z) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    return m


def _scale(x, y=None, z=None) -> np.ndarray:
    return np.diag([x, x if y is None else y, x if z is None else z, 1.0]).astype(np.float32)


def _wound_outward(verts: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """Triangles of a convex mesh, each reordered where needed so that its normal (b - a) x (c - a) points
    away from the centroid of the vertices."""
    v = np.asarray(verts, np.float64)
    a, b, c = (v[tris[:, k]] for k in range(3))
    away = np.einsum("ij,ij->i", np.cross(b - a, c - a), (a + b + c) / 3.0 - v.mean(axis=0))
    out = np.array(tris, np.int32)
    out[away < 0] = out[away < 0][:, ::-1]
    return out


def box_mesh() -> tuple[np.ndarray, np.ndarray]:
    """Unit cube [0,1]^3: corner i at (i & 1, i >> 1 & 1, i >> 2 & 1), two triangles per face, wound outward."""
    corners = np.array([[(i >> axis) & 1 for axis in range(3)] for i in range(8)], np.float32)
    tris = []
    for axis in range(3):
        u, w = (a for a in range(3) if a != axis)
        for side in (0, 1):
            quad = [(side << axis) | (i << u) | (j << w) for i, j in ((0, 0), (1, 0), (1, 1), (0, 1))]
            tris += [quad[:3], [quad[0], quad[2], quad[3]]]
    return corners, _wound_outward(corners, np.array(tris))


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere, vertices in float32 (see icosphere_f64)."""
    v, t = icosphere_f64(subdivisions)
    return v.astype(np.float32), t


def icosphere_f64(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: 10 * 4^s + 2 vertices, 20 * 4^s triangles. The icosahedron's vertices are the cyclic
    permutations of (0, +-1, +-golden ratio), its faces the triples of mutually nearest vertices; every
    subdivision splits each triangle into four at its edge midpoints, pushed out to the sphere."""
    g = (1.0 + 5.0 ** 0.5) / 2.0
    base = [np.array(p) for s in (-1.0, 1.0) for t in (-g, g) for p in ((0.0, s, t), (s, t, 0.0), (t, 0.0, s))]
    verts = [p / np.linalg.norm(p) for p in base]
    dist = np.linalg.norm(np.subtract.outer(np.array(verts), np.array(verts)).diagonal(axis1=1, axis2=3), axis=2)
    near = np.isclose(dist, dist[dist > 0].min())
    faces = _wound_outward(np.array(verts), np.array(
        [(i, j, k) for i in range(12) for j in range(i + 1, 12) for k in range(j + 1, 12)
         if near[i, j] and near[j, k] and near[i, k]])).tolist()
    for _ in range(subdivisions):
        midpoint: dict[frozenset, int] = {}

        def split(a: int, b: int) -> int:
            edge = frozenset((a, b))
            if edge not in midpoint:
                m = verts[a] + verts[b]
                midpoint[edge] = len(verts)
                verts.append(m / np.linalg.norm(m))
            return midpoint[edge]

        finer = []
        for a, b, c in faces:
            ab, bc, ca = split(a, b), split(b, c), split(c, a)
            finer += [[a, ab, ca], [ab, b, bc], [ca, bc, c], [ab, bc, ca]]
        faces = finer
    return np.array(verts), np.array(faces, np.int32)


def box_points(lo, hi, spacing: float, seed) -> np.ndarray:
    v, t = box_mesh()
    m = _translate(*lo) @ _scale(*np.subtract(hi, lo))
    return sample_mesh_volume(transform_points(m, v), t, spacing, seed)


def sphere_points(center, radius: float, spacing: float, seed, subdivisions: int = 4) -> np.ndarray:
    v, t = icosphere(subdivisions)
    return sample_mesh_volume(v * np.float32(radius) + np.asarray(center, np.float32), t, spacing, seed)


@dataclass
class Workload:
    """One scene: what to pass to create_solver, the particles to add, and the stepping."""

    name: str
    domain: Domain
    materials: list[MaterialParams]
    gravity: tuple[float, float, float]
    positions: np.ndarray
    velocities: np.ndarray
    material_ids: np.ndarray
    emitter_ids: np.ndarray
    dt: float
    substeps: int
    colliders: list[ColliderGeometry] = field(default_factory=list)
    collider_path: list[tuple[np.ndarray, np.ndarray]] = field(default_factory=list)  # (start, end) per collider
    duration: float = 0.0
    collider_scale: list[float] = field(default_factory=list)  # uniform scale per collider (default 0.1)
    frame_dt: float = 1.0 / 24.0

    @property
    def n(self) -> int:
        return len(self.positions)

    def collider_transforms(self, t0: float, t1: float) -> tuple[np.ndarray, np.ndarray]:
        """(C,4,4) start/end world matrices at times t0 and t1 of a linear translation path: translate(p)
        @ scale(s), written out (the product of those float32 matrices is exactly this one)."""
        n = min(len(self.collider_path), len(self.colliders))
        out = np.zeros((2, n, 4, 4), np.float32)
        for c in range(n):
            a, b = self.collider_path[c]
            scale = self.collider_scale[c] if c < len(self.collider_scale) else 0.1
            for k, t in enumerate((t0, t1)):
                u = min(max(t / self.duration, 0.0), 1.0) if self.duration > 0 else 0.0
                m = out[k, c]
                m[0, 0] = m[1, 1] = m[2, 2] = scale
                m[3, 3] = 1.0
                m[:3, 3] = a + u * (b - a)
        return out[0], out[1]

    def make_solver(self, backend: str = "hip", device: str | None = None):
        from mpm_core.backends import create_solver

        s = create_solver(backend, domain=self.domain, materials=self.materials, gravity=self.gravity,
                          device=device)
        s.add_particles(self.positions, self.velocities, self.material_ids, self.emitter_ids)
        if self.colliders:
            s.set_colliders(self.colliders)
        return s


def elastic_block(seed: int = 0) -> Workload:
    domain = Domain((0, 0, 0), (1, 1, 1), 64, wall_friction=1.0e6)
    mat = MaterialParams("elastic_block", "elastic", 1.0e4, 0.3, 1000.0)
    x = box_points((0.4, 0.5, 0.4), (0.6, 0.7, 0.6), domain.dx / 2, seed)
    n = len(x)
    return Workload("elastic_block_64", domain, [mat], (0.0, -9.8, 0.0), x, np.zeros_like(x),
                    np.zeros(n, np.int32), np.zeros(n, np.int32), dt=1.0e-4, substeps=1000)


def multi_material(seed: int = 1, with_collider: bool = True) -> Workload:
    domain = Domain((0, 0, 0), (1, 1, 1), 128)
    mats = [
        MaterialParams("plastic", "von_mises", 5.0e3, 0.45, 1000.0, yield_stress=200.0),
        MaterialParams("sand", "drucker_prager", 3.5e4, 0.3, 1600.0, friction_angle_deg=30.0),
        MaterialParams("snow", "snow", 1.4e5, 0.2, 400.0, hardening=10.0, critical_compression=2.5e-2,
                       critical_stretch=7.5e-3, max_hardening=5.0),
    ]
    h = domain.dx / 2
    bodies = [
        sphere_points((0.22, 0.5, 0.30), 0.15, h, seed + 0),
        box_points((0.4, 0.4, 0.0), (0.6, 0.6, 0.72), h, seed + 1),
        sphere_points((0.80, 0.5, 0.35), 0.16, h, seed + 2),
    ]
    vels = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (-3.0, 0.0, 0.0)]
    x = np.concatenate(bodies)
    v = np.concatenate([np.tile(np.asarray(u, np.float32), (len(b), 1)) for b, u in zip(bodies, vels)])
    ids = np.concatenate([np.full(len(b), i, np.int32) for i, b in enumerate(bodies)])
    w = Workload("multi_material_128", domain, mats, (0.0, 0.0, -9.81), x, v, ids, ids.copy(), dt=5.0e-5,
                 substeps=2000, duration=0.1)
    if with_collider:
        verts, tris = icosphere(4)
        w.colliders = [ColliderGeometry(MeshData(verts, tris), closed=True, friction=0.5)]
        w.collider_path = [(np.array([0.5, 0.5, 0.88]), np.array([0.5, 0.5, 0.70]))]
    return w


MIX_EDGES = (32, 48, 64, 96, 128)
MIX_MODELS = ("elastic", "von_mises", "drucker_prager", "snow")


def _jittered_material(model: str, rng: np.random.Generator) -> MaterialParams:
    def j(v: float) -> float:
        return float(v * rng.uniform(0.8, 1.2))

    if model == "elastic":
        return MaterialParams("elastic", "elastic", j(1.0e4), min(j(0.3), 0.45), j(1000.0))
    if model == "von_mises":
        return MaterialParams("plastic", "von_mises", j(5.0e3), min(j(0.45), 0.49), j(1000.0), yield_stress=j(200.0))
    if model == "drucker_prager":
        return MaterialParams("sand", "drucker_prager", j(3.5e4), min(j(0.3), 0.45), j(1600.0),
                              friction_angle_deg=j(30.0))
    return MaterialParams("snow", "snow", j(1.4e5), min(j(0.2), 0.45), j(400.0), hardening=j(10.0),
                          critical_compression=j(2.5e-2), critical_stretch=j(7.5e-3), max_hardening=j(5.0))


def lattice_body(shape: str, center, size: float, spacing: float, rng: np.random.Generator) -> np.ndarray:
    """Jittered lattice points (jitter +-0.25 spacing) inside a sphere (size = radius) or box (size = side)."""
    half = size if shape == "sphere" else size / 2
    axes = [np.arange(c - half + spacing / 2, c + half, spacing) for c in center]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    if shape == "sphere":
        g = g[np.sum((g - np.asarray(center)) ** 2, axis=1) <= half * half]
    g = g + rng.uniform(-0.25 * spacing, 0.25 * spacing, g.shape)
    return g.astype(np.float32)


def mix_scene(index: int, seed: int = 3, n_range=(5_000, 400_000), edges=MIX_EDGES, model: str | None = None,
              n_colliders: int | None = None, walls: dict | None = None) -> Workload:
    """Scene `index` of the batched mix (deterministic in (seed, index)); `model`, `n_colliders` and
    `walls` pin what is otherwise drawn (all walls closed by default)."""
    rng = np.random.default_rng([seed, index])
    edge = int(edges[rng.integers(len(edges))])
    n_target = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
    n_target = min(n_target, 0.35 * 8 * edge ** 3)
    drawn = MIX_MODELS[rng.integers(len(MIX_MODELS))]
    model = model or drawn
    mat = _jittered_material(model, rng)
    domain = Domain((0, 0, 0), (1, 1, 1), edge, walls=dict(walls or {}))
    dx = domain.dx
    volume = n_target / (8.0 / dx ** 3)
    shape = "sphere" if rng.random() < 0.5 else "box"
    size = (3.0 * volume / (4.0 * np.pi)) ** (1.0 / 3.0) if shape == "sphere" else volume ** (1.0 / 3.0)
    half = size if shape == "sphere" else size / 2
    lo, hi = half + 2 * dx, 1.0 - half - 2 * dx
    center = [float(rng.uniform(lo, hi)) if hi > lo else 0.5 for _ in range(3)]
    x = lattice_body(shape, center, size, dx / 2, rng)
    direction = rng.normal(size=3)
    vel = (direction / np.linalg.norm(direction) * rng.uniform(0.0, 3.0)).astype(np.float32)
    n = len(x)
    w = Workload(f"mix_{index}_{model}_{edge}", domain, [mat], (0.0, -9.81, 0.0), x, np.tile(vel, (n, 1)),
                 np.zeros(n, np.int32), np.zeros(n, np.int32), dt=0.0, substeps=0, duration=1.0)
    n_col = int(rng.integers(3))
    n_col = n_col if n_colliders is None else n_colliders
    if n_col:
        verts, tris = icosphere(2)
        for _ in range(n_col):
            w.colliders.append(ColliderGeometry(MeshData(verts, tris), closed=True, friction=float(rng.uniform(0, 1))))
            w.collider_path.append((rng.uniform(0.15, 0.85, 3), rng.uniform(0.15, 0.85, 3)))
            w.collider_scale.append(float(rng.uniform(0.06, 0.15)))
    return w


def scene_mix(n_scenes: int = 256, seed: int = 3, **kw) -> list[Workload]:
    """The batched mix: ``n_scenes`` scenes of :func:`mix_scene` (about 25 M particles for 256)."""
    return [mix_scene(i, seed, **kw) for i in range(n_scenes)]


def _wave_speed(m: MaterialParams) -> float:
    mu, lam = m.lame()
    return float(np.sqrt((lam + 2.0 * mu) / m.density))


def served_request(index: int, seed: int = 4, n_range=(5_000, 400_000), edges=MIX_EDGES,
                   median_frames: float = 48.0, max_frames: int = 240):
    """Scene description of request `index` of the served workload (deterministic in (seed, index))."""
    from mpm_core.scene import Collider, Emitter, SceneDescription

    w = mix_scene(index, seed, n_range=n_range, edges=edges)
    rng = np.random.default_rng([seed, index, 1])
    frames = int(np.clip(round(median_frames * np.exp(rng.normal(0.0, 0.5))), 1, max_frames))
    target = float(rng.uniform(20, 200))
    speed = float(np.linalg.norm(w.velocities[0]))
    cfl_dt = 0.4 * w.domain.dx / (_wave_speed(w.materials[0]) + speed)
    fps = 1.0 / (target * cfl_dt)
    # the body as a fill emitter mesh of the same extent as the lattice body
    lo, hi = w.positions.min(0), w.positions.max(0)
    centre, ext = (lo + hi) / 2, (hi - lo)
    if rng.random() < 0.5:
        verts, tris = icosphere(3)
        xf = _translate(*centre) @ _scale(*(ext / 2))
    else:
        verts, tris = box_mesh()
        xf = _translate(*(centre - ext / 2)) @ _scale(*ext)
    emitter = Emitter(name="body", kind="fill", material=0, mesh=MeshData(verts, tris),
                      transforms=np.repeat(xf[None], frames, 0), velocity=tuple(float(v) for v in w.velocities[0]))
    colliders = []
    for c, (geom, (a, b)) in enumerate(zip(w.colliders, w.collider_path)):
        sc = w.collider_scale[c]
        xfs = np.stack([_translate(*(a + (b - a) * (k / max(frames - 1, 1)))) @ _scale(sc) for k in range(frames)])
        colliders.append(Collider(name=f"collider_{c}", mesh=geom.mesh, transforms=xfs, friction=geom.friction,
                                  closed=geom.closed))
    return SceneDescription(domain=w.domain, materials=w.materials, emitters=[emitter], colliders=colliders,
                            gravity=w.gravity, fps=fps, frame_start=1, frame_end=frames, seed=seed + index)


def served_requests(n_requests: int, rate: float = 20.0, seed: int = 4, **kw):
    """[(arrival offset in seconds, scene description)] of a Poisson arrival process at `rate` per second."""
    rng = np.random.default_rng([seed, 2])
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    return [(float(t[i]), served_request(i, seed, **kw)) for i in range(n_requests)]
