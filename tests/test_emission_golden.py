"""mpm_core.geometry against vectors recorded from the reference's own sampler and transform helpers
(tests/golden/make_emission_vectors.py): emission counts and jittered positions bit for bit."""

import hashlib
import os

import numpy as np
import pytest

from mpm_core.geometry import decompose_transform, interpolate_transform, sample_mesh_volume, transform_points

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "emission_vectors.npz"))
CASES = sorted({k.split("/")[0] for k in GOLDEN.files if "/count" in k})


@pytest.mark.parametrize("name", CASES)
def test_sample_mesh_volume_bit_exact(name):
    g = {k.split("/", 1)[1]: GOLDEN[k] for k in GOLDEN.files if k.startswith(name + "/")}
    seed = g["seed"].tolist() if bool(g["seed_is_list"]) else int(g["seed"][0])
    pts = sample_mesh_volume(transform_points(g["xf"], g["verts"]), g["tris"], float(g["spacing"]), seed)
    assert len(pts) == int(g["count"])
    assert np.array_equal(pts[:8], g["head"]) and np.array_equal(pts[-8:], g["tail"])
    assert hashlib.sha256(np.ascontiguousarray(pts, np.float32).tobytes()).hexdigest() == str(g["sha256"])


@pytest.mark.parametrize("i", range(4))
def test_decompose_transform_matches_reference(i):
    t, q, S = decompose_transform(GOLDEN[f"decompose{i}/m"])
    np.testing.assert_allclose(t, GOLDEN[f"decompose{i}/t"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(q, GOLDEN[f"decompose{i}/q"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(S, GOLDEN[f"decompose{i}/S"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("i", range(3))
def test_interpolate_transform_matches_reference(i):
    out = interpolate_transform(GOLDEN[f"interp{i}/a"], GOLDEN[f"interp{i}/b"], float(GOLDEN[f"interp{i}/s"]))
    np.testing.assert_array_equal(out, GOLDEN[f"interp{i}/out"])
