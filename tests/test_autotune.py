"""The launch-shape autotuner: the search on the CPU, the measurement on a real batch on the GPU."""

from __future__ import annotations

import json

import numpy as np
import pytest

from mpm_plugin_amd import _native as N
from mpm_plugin_amd import autotune as T


def test_coordinate_descent_finds_the_separable_optimum():
    space = {0: [1, 2, 3, 4], 1: [10, 20, 30]}
    calls = []

    def score(c, knob=None):
        calls.append(dict(c))
        return -((c[0] - 3) ** 2) - (c[1] - 20) ** 2

    best, s = T.coordinate_descent(space, score, repeats=2)
    assert best == {0: 3, 1: 20} and s == 0
    assert len(calls) == 2 * (4 + 3)


def test_interleaving_cancels_a_drifting_measure():
    """Scores that fall over time (scenes finishing their frames) still pick the truly faster value."""
    space = {0: [1, 2]}
    t = [0]

    def score(c, knob=None):
        t[0] += 1
        return (1.1 if c[0] == 2 else 1.0) * 100.0 / t[0]

    best, _ = T.coordinate_descent(space, score, repeats=1)
    assert best[0] == 1  # measured first, so favoured by the drift at repeats=1 ...
    t[0] = 0
    best, _ = T.coordinate_descent({0: [2, 1]}, score, repeats=3)
    assert best[0] == 2  # ... and with the faster value first it is kept


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("MPM_TUNE_CACHE", str(tmp_path / "tune.json"))
    cfg = {N.BATCH_OPT_G2P_TILE: 128, N.BATCH_OPT_SORT_INTERVAL: 16}
    T.store_cached("dev|sig", cfg)
    assert T.load_cached("dev|sig") == cfg
    assert json.loads((tmp_path / "tune.json").read_text())["dev|sig"] == {"g2p_tile": 128, "sort_interval": 16}
    assert T.load_cached("other") is None


@pytest.mark.gpu
def test_tune_batch_applies_a_config_from_the_space(tmp_path, monkeypatch):
    from mpm_core.backends.hip.batch import SceneBatch
    from mpm_plugin_amd.workloads import mix_scene

    monkeypatch.setenv("MPM_TUNE_CACHE", str(tmp_path / "tune.json"))
    ws = [mix_scene(i, 5, n_range=(3_000, 20_000), edges=(32, 48)) for i in range(4)]
    solvers = [w.make_solver("hip", "cuda:0") for w in ws]
    batch = SceneBatch(solvers)
    space = {N.BATCH_OPT_G2P_TILE: [256, 64], N.BATCH_OPT_G2P_BLOCKS: [2048, 512],
             N.BATCH_OPT_SORT_INTERVAL: [8, 4]}
    out = T.tune_batch(batch, [w.frame_dt for w in ws], launches=3, repeats=1, space=space)
    assert not out["cached"] and out["particle_substeps_per_s"] > 0
    for knob, values in space.items():
        assert batch.get_option(knob) in values
        assert out["config"][T.KNOB_NAMES[knob]] == batch.get_option(knob)
    done, total = batch.progress()
    assert np.all(done == total)  # between frames again
    assert all(np.isfinite(s.save_state()["x"]).all() for s in solvers)
    again = T.tune_batch(batch, [w.frame_dt for w in ws], space=space)
    assert again["cached"] and again["config"] == out["config"]
