"""Drift routing: P2G between two sorts, when particles no longer sit in the cell they were sorted into.

Particles are binned by cell every `sort_interval` substeps. In between, a particle whose stencil moved to
another cell is handed to the incoming list of that cell's P2G lane (up to four per cell per substep; the
rest are scattered with atomics by the lane that holds them). Three states are compared after a substep
taken with particles that moved about 0.4 cells since the last sort (a large share drifted, many cells
over their list capacity):
  routed   sort interval 8: P2G with drift routing
  resorted sort interval 1: the same substep after a fresh sort (no drift)
  oracle   the FP32 CPU oracle's substep from the same state
Routed and resorted differ only in the order of float additions (bars a few rounding steps wide); both
meet the one-substep bar (1e-4) against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest
from test_precision import ELASTIC, RES, gpu_solver, phase_b, rel

from mpm_core.scene import Domain
from mpm_plugin_amd import _native as N

pytestmark = pytest.mark.gpu

DT = 2.0e-4
PRIOR = 3  # substeps after the sort and before the compared one


def moving_state(n=60000, seed=3):
    st = phase_b(n, seed=seed)
    rng = np.random.default_rng(seed + 100)
    # |v| about 4 m/s: 0.1 cell per substep at dx = 1/128
    st["v"] = (rng.normal(size=st["v"].shape) * 2.0 + np.array([3.0, -2.0, 1.0])).astype(np.float32)
    return st


def cells(x):
    return np.floor(np.asarray(x, np.float64) * RES - 0.5).astype(np.int64)


@pytest.fixture(scope="module")
def runs():
    st = moving_state()
    out = {}
    for name, interval in (("routed", 8), ("resorted", 1)):
        s = gpu_solver(ELASTIC, N.PRECISION_FALLBACK, N.P2G_BINNED)
        s.set_option(N.OPT_SORT_INTERVAL, 8)
        s.load_state(st)
        for _ in range(PRIOR):
            s.substep(DT, 1.0, DT)
        before = s.save_state()
        s.set_option(N.OPT_SORT_INTERVAL, interval)
        s.substep(DT, 1.0, DT)
        out[name] = (before, s.save_state())
    out["initial"] = st
    return out


def test_the_state_has_drifted(runs):
    """The compared substep starts with many particles outside their sorted cell, and with cells that
    receive more drifted particles than their list holds."""
    before, _ = runs["routed"]
    c0, c1 = cells(runs["initial"]["x"]), cells(before["x"])
    drifted = np.any(c0 != c1, axis=1)
    assert drifted.mean() > 0.2
    _, per_cell = np.unique(c1[drifted], axis=0, return_counts=True)
    assert per_cell.max() > 4


def test_routed_matches_resorted(runs):
    (b_r, a_r), (b_s, a_s) = runs["routed"], runs["resorted"]
    # the same runs up to the compared substep (which of a full cell's drifted particles are routed
    # follows the order of atomics, so the sums can differ in rounding)
    assert rel(b_r["x"], b_s["x"]) <= 1e-7 and rel(b_r["v"], b_s["v"]) <= 1e-6
    assert rel(a_r["x"], a_s["x"]) <= 1e-6
    assert rel(a_r["F"], a_s["F"]) <= 1e-6
    assert rel(a_r["v"], a_s["v"]) <= 1e-5
    assert rel(a_r["C"], a_s["C"]) <= 1e-4


def test_routed_substep_matches_the_oracle(runs):
    from oracle import oracle as O

    before, after = runs["routed"]
    ora = O.OracleSolver(Domain((0, 0, 0), (1, 1, 1), RES), [ELASTIC], (0, 0, -9.81), precision="f32")
    ora.load_state(before)
    ora.substep(DT)
    ref = ora.save_state()
    for k in ("x", "v", "C", "F"):
        assert rel(after[k], ref[k]) <= 1e-4, k


@pytest.mark.parametrize("mode", ["binned", "fixed"])
@pytest.mark.parametrize("split", [2, 4])
def test_waves_per_brick_match_one(mode, split):
    """P2G with 2 or 4 waves per brick (wave h of n takes every n-th particle of a cell, windows summed)
    against one wave per brick, on the drifted state: the grids differ only in the order of float
    additions."""
    st = moving_state()
    p2g_mode = N.P2G_BINNED if mode == "binned" else N.P2G_FIXED
    grids, states = {}, {}
    for waves in (1, split):
        s = gpu_solver(ELASTIC, N.PRECISION_FALLBACK, p2g_mode)
        s.set_option(N.OPT_SORT_INTERVAL, 8)
        s.set_option(N.OPT_P2G_WAVES, waves)
        assert s.get_option(N.OPT_P2G_WAVES) == waves
        s.load_state(st)
        for _ in range(PRIOR + 1):
            s.substep(DT, 1.0, DT)
        states[waves] = s.save_state()
        grids[waves] = s.get_grid(1)  # the last substep's grid velocity and mass
    (v1, m1), (v2, m2) = grids[1], grids[split]
    assert rel(m2, m1) <= 1e-6 and rel(v2, v1) <= 1e-6
    for k in ("x", "v", "F"):
        assert rel(states[split][k], states[1][k]) <= 1e-5, k
    assert rel(states[split]["C"], states[1]["C"]) <= 1e-4


def test_atomic_flush_matches_staged_windows():
    """Brick windows added into the grid with global atomics (option P2G_FLUSH) against the staged windows
    the grid update sums, on the drifted state: the grids differ only in the order of float additions."""
    st = moving_state()
    out = {}
    for flush in (N.P2G_FLUSH_STAGED, N.P2G_FLUSH_ATOMIC):
        s = gpu_solver(ELASTIC, N.PRECISION_FALLBACK, N.P2G_BINNED)
        s.set_option(N.OPT_SORT_INTERVAL, 8)
        s.set_option(N.OPT_P2G_FLUSH, flush)
        assert s.get_option(N.OPT_P2G_FLUSH) == flush
        s.load_state(st)
        for _ in range(PRIOR + 1):
            s.substep(DT, 1.0, DT)
        out[flush] = (s.save_state(), s.get_grid(1), s.debug_p2g(DT))
    (a, (va, ma), (pa, qa)), (b, (vb, mb), (pb, qb)) = out[N.P2G_FLUSH_STAGED], out[N.P2G_FLUSH_ATOMIC]
    assert rel(mb, ma) <= 1e-6 and rel(vb, va) <= 1e-6
    assert rel(qb, qa) <= 1e-6 and rel(pb, pa) <= 1e-6  # the P2G download resolves either form
    for k in ("x", "v", "F"):
        assert rel(b[k], a[k]) <= 1e-5, k
