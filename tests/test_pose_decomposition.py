"""The library's FP64 pose decomposition (mpm_decompose_transform, used by set_collider_transforms) against
the Python restatement of the reference's (mpm_core.geometry.decompose_transform, geometry.py:30-49,
itself pinned by vectors recorded from the reference), on rotations times stretches, uniform and mirrored
scales. Host code only: runs without a GPU."""

import ctypes as C

import numpy as np
import pytest

from mpm_core.geometry import decompose_transform_sv

N = pytest.importorskip("mpm_plugin_amd._native")


def _rotation(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _library(m32):
    t, q, S, smin = (C.c_double * 3)(), (C.c_double * 4)(), (C.c_double * 9)(), C.c_double()
    N.check(N.lib().mpm_decompose_transform(N.fptr(m32), t, q, S, C.byref(smin)))
    return np.array(t), np.array(q), np.array(S).reshape(3, 3), smin.value


def _cases(rng, n):
    for i in range(n):
        R = _rotation(rng.normal(size=4))
        kind = i % 4
        if kind == 0:
            S = np.diag(rng.uniform(0.05, 3.0, 3))
        elif kind == 1:
            S = np.eye(3) * rng.uniform(0.05, 3.0)  # uniform scale: equal singular values
        elif kind == 2:
            Q = _rotation(rng.normal(size=4))
            S = Q @ np.diag(rng.uniform(0.05, 3.0, 3)) @ Q.T
        else:
            S = np.diag(rng.uniform(0.05, 3.0, 3) * np.array([1.0, 1.0, -1.0]))  # mirrored
        m = np.eye(4)
        m[:3, :3] = R @ S if i % 8 else S
        m[:3, 3] = rng.normal(size=3)
        yield np.ascontiguousarray(m.astype(np.float32))


def test_decomposition_matches_restatement():
    rng = np.random.default_rng(7)
    for m in _cases(rng, 800):
        t, q, S, smin = _library(m)
        rt, rq, rS, rsmin = decompose_transform_sv(m)
        assert np.array_equal(t, rt)
        np.testing.assert_allclose(q, rq, rtol=0, atol=1e-10)
        np.testing.assert_allclose(S, rS, rtol=0, atol=1e-12 * np.abs(rS).max())
        assert abs(smin - rsmin) <= 1e-12 * rsmin
        # the decomposition reproduces the matrix: R(q) S = A
        np.testing.assert_allclose(_rotation(q) @ S, m[:3, :3].astype(np.float64), rtol=0, atol=1e-12)


def test_identity_and_pure_translation_are_exact():
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (0.25, -1.5, 3.0)
    t, q, S, smin = _library(m)
    assert np.array_equal(t, [0.25, -1.5, 3.0])
    assert np.array_equal(q, [0.0, 0.0, 0.0, 1.0])
    assert np.array_equal(S, np.eye(3))
    assert smin == 1.0
