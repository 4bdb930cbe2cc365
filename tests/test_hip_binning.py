"""Cell binning on the GPU: keys, permutation and cell offsets are bit-exact against a CPU stable sort.

The key of a particle is its stencil base node floor(x / dx - 0.5) (src/mpm_core/backends/warp/kernels.py:25-27)
numbered brick-major (brick * 64 + cell within the 4x4x4-node brick), computed here in float32 exactly as on
the device.
"""

import numpy as np
import pytest
from test_hip_parity import pair, seeded_state

pytestmark = pytest.mark.gpu


def cpu_keys(x: np.ndarray, dims, origin, inv_dx) -> np.ndarray:
    xp = (x.astype(np.float32) - np.asarray(origin, np.float32)) * np.float32(inv_dx)
    base = np.floor(xp - np.float32(0.5)).astype(np.int64)
    base = np.clip(base, 0, np.asarray(dims) - 3)
    bd = (np.asarray(dims) + 3) // 4
    b = base >> 2
    brick = (b[:, 0] * bd[1] + b[:, 1]) * bd[2] + b[:, 2]
    cell = (base[:, 0] & 3) * 16 + (base[:, 1] & 3) * 4 + (base[:, 2] & 3)
    return (brick * 64 + cell).astype(np.int32)


@pytest.mark.parametrize("res,n,seed", [(32, 3000, 0), (64, 40000, 1), (128, 200000, 2)])
def test_binning_bit_exact(res, n, seed):
    state = seeded_state(n, res, seed)
    rng = np.random.default_rng(seed)
    # scatter some particles over the whole domain so many bricks and both key bytes are exercised
    k = len(state["x"]) // 4
    state["x"][:k] = rng.uniform(0.05, 0.95, size=(k, 3)).astype(np.float32)
    hip, _ = pair(state, res=res)
    hip.debug_sort()
    keys, perm, offsets = hip.binning()
    d = hip._domain_desc
    expect = cpu_keys(state["x"], tuple(d.dims), tuple(d.origin), d.inv_dx)
    assert np.array_equal(keys, expect)
    assert np.array_equal(perm, np.argsort(expect, kind="stable").astype(np.int32))
    counts = np.bincount(expect, minlength=len(offsets) - 1)
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(counts)]).astype(np.int32))


def test_binning_keeps_insertion_order_for_output():
    state = seeded_state(20000, 64, 5)
    hip, _ = pair(state)
    hip.debug_sort()
    back = hip.save_state()
    for k in ("x", "v", "C", "F", "Jp", "mat", "emitter"):
        assert np.array_equal(back[k], state[k]), k


def test_empty_solver_sorts_and_steps():
    from mpm_core.backends import create_solver
    from mpm_core.materials import load_preset
    from mpm_core.scene import Domain

    s = create_solver("hip", domain=Domain((0, 0, 0), (1, 1, 1), 16), materials=[load_preset("jelly").material],
                      gravity=(0, 0, -9.81), device="cuda:0")
    s.debug_sort()
    assert s.advance(1 / 24).substeps == 0
    assert s.particle_count == 0
