"""The CPU oracle behind the Solver interface (test infrastructure): runs the bake runner and the
reference's behaviour tests on the CPU for comparison with the HIP backend."""

from __future__ import annotations

import numpy as np

from mpm_core.solver import AdvanceStats, ParticleData, Solver
from oracle import oracle as O


class OracleBackend(Solver):
    def __init__(self, domain, materials, gravity, particles_per_cell: int = 8, device=None, precision="f32"):
        super().__init__(domain, materials, gravity, particles_per_cell, device)
        self.o = O.OracleSolver(domain, self.materials, self.gravity, particles_per_cell, precision=precision)

    def add_particles(self, positions, velocities, material_ids, emitter_ids):
        mat = np.asarray(material_ids, np.int32)
        if len(mat) and (mat.min() < 0 or mat.max() >= len(self.materials)):
            raise ValueError("material id out of range")
        self.o.add_particles(positions, velocities, mat, emitter_ids)

    def set_colliders(self, colliders):
        self.o.set_colliders(colliders)
        self._n_col = len(colliders)

    def set_collider_transforms(self, start, end):
        start = np.asarray(start, np.float32).reshape(-1, 4, 4)
        if len(start) != getattr(self, "_n_col", 0):
            raise ValueError("collider transform count")
        self.o.set_collider_transforms(start, end)

    def advance(self, frame_dt):
        steps, dt, removed = self.o.advance(frame_dt)
        return AdvanceStats(steps, dt, removed)

    @property
    def particle_count(self):
        return self.o.n

    def get_particles(self):
        s = self.o.s
        return ParticleData(s.x.copy(), s.v.copy(), s.mat.copy(), s.emitter.copy())

    def is_occupied(self, points):
        return self.o.is_occupied(points)

    def save_state(self):
        return self.o.save_state()

    def load_state(self, state):
        self.o.load_state(state)
