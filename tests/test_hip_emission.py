"""Emission on the GPU: the HIP solver's sample_mesh_volume (lattice test in FP64 on the device, the
reference's jitter draw on the host) returns the same array, bit for bit, as the host function the runner
uses (mpm_core.geometry.sample_mesh_volume, a restatement of the reference's geometry.py:89-155 pinned to the
reference's own output by tests/test_emission_golden.py): emission counts and positions are part of a bake's
result."""

import numpy as np
import pytest
from helpers import icosphere, rot_x, rot_y, scale, translate, unit_cube_mesh

from mpm_core.geometry import sample_mesh_volume, transform_points
from mpm_core.materials import MaterialParams
from mpm_core.scene import Domain

pytestmark = pytest.mark.gpu

JELLY = MaterialParams("jelly", "elastic", 1.0e4, 0.3, 1000.0)


@pytest.fixture(scope="module")
def solver():
    from mpm_core.backends import create_solver

    return create_solver("hip", domain=Domain((0, 0, 0), (1, 1, 1), 32), materials=[JELLY], gravity=(0, 0, -9.81),
                         device="cuda:0")


def two_spheres():
    """Two overlapping spheres as one mesh: columns through both cross the surface four times."""
    v1, t1 = icosphere(0.12, 3)
    v2, t2 = icosphere(0.09, 3)
    verts = np.concatenate([v1 + [0.5, 0.5, 0.35], v2 + [0.53, 0.48, 0.52]])
    return verts, np.concatenate([t1, t2 + len(v1)])


CASES = {
    "sphere": (lambda: (icosphere(0.15, 3)[0] + [0.5, 0.5, 0.5], icosphere(0.15, 3)[1]), 1.0 / 128, 5),
    "sphere_fine": (lambda: (icosphere(0.1, 4)[0] + [0.31, 0.62, 0.44], icosphere(0.1, 4)[1]), 1.0 / 256, 11),
    "box_rotated": (lambda: (transform_points(translate(0.5, 0.4, 0.3) @ rot_x(17) @ rot_y(31) @ scale(0.3, 0.2, 0.1),
                                              unit_cube_mesh(centered=True)[0]), unit_cube_mesh(centered=True)[1]),
                    1.0 / 200, [3, 7, 1, 1, 2]),
    "two_spheres": (two_spheres, 1.0 / 160, 2),
    "tiny": (lambda: (icosphere(0.002, 2)[0] + [0.5, 0.5, 0.5], icosphere(0.002, 2)[1]), 1.0 / 64, 0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_lattice_matches_host(solver, case):
    make, spacing, seed = CASES[case]
    verts, tris = make()
    want = sample_mesh_volume(verts, tris, spacing, seed)
    got = solver.sample_mesh_volume(verts, tris, spacing, seed)
    assert got.dtype == np.float32 and got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    if case != "tiny":
        assert len(got) > 1000


def test_open_mesh_is_rejected(solver):
    verts, tris = unit_cube_mesh()
    with pytest.raises(ValueError):
        solver.sample_mesh_volume(verts, tris[:-1], 0.05, 0)


def test_bad_meshes_are_errors(solver):
    verts, tris = icosphere(0.1, 2)
    bad = verts.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ValueError):
        solver.sample_mesh_volume(bad + 0.5, tris, 0.01, 0)
    with pytest.raises(ValueError):
        solver.sample_mesh_volume(verts + 0.5, tris, 0.0, 0)
