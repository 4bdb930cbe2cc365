"""Closed forms that pin the oracle, written from the mathematics rather than from any kernel's code.

Stresses come from spectral decompositions (numpy.linalg.eigh) instead of a 3x3 SVD; return maps are
stated as the geometric facts they produce; contact rules as the velocities a rigid plane imposes.
"""

from __future__ import annotations

import math

import numpy as np


def rotation(rng) -> np.ndarray:
    """Uniformly random proper rotation from a random unit quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def polar_rotation(F: np.ndarray) -> np.ndarray:
    """R of F = R S with S symmetric positive definite (det F > 0)."""
    lam, Q = np.linalg.eigh(F.T @ F)
    return F @ (Q @ np.diag(1.0 / np.sqrt(lam)) @ Q.T)


def corotated_tau(F: np.ndarray, mu: float, lam: float) -> np.ndarray:
    """Kirchhoff stress of the fixed-corotated model: 2 mu (F - R) F^T + lam J (J - 1) I."""
    J = float(np.linalg.det(F))
    return 2.0 * mu * (F - polar_rotation(F)) @ F.T + lam * J * (J - 1.0) * np.eye(3)


def hencky_tau(F: np.ndarray, mu: float, lam: float) -> np.ndarray:
    """Kirchhoff stress of StVK in Hencky strain: eps = log of the left stretch, tau = 2 mu eps + lam tr(eps) I."""
    b, Q = np.linalg.eigh(F @ F.T)
    eps = np.log(np.maximum(np.sqrt(b), 1.0e-4))
    return Q @ np.diag(2.0 * mu * eps + lam * eps.sum()) @ Q.T


def snow_tau(F: np.ndarray, mu: float, lam: float, xi: float, max_h: float, jp: float) -> np.ndarray:
    """Stomakhin et al. 2013: the corotated stress with both Lame parameters scaled by min(e^{xi (1 - Jp)}, max)."""
    h = min(math.exp(xi * (1.0 - jp)), max_h)
    return corotated_tau(F, mu * h, lam * h)


def material_tau(model: str, F: np.ndarray, mu: float, lam: float, xi: float = 0.0, max_h: float = 1.0,
                 jp: float = 1.0) -> np.ndarray:
    if model in ("von_mises", "drucker_prager"):
        return hencky_tau(F, mu, lam)
    if model == "snow":
        return snow_tau(F, mu, lam, xi, max_h, jp)
    return corotated_tau(F, mu, lam)


def f32(a) -> float:
    """The float32 rounding of a value, as a Python float."""
    return float(np.float32(a))
