"""Writes tests/golden/emission_vectors.npz from the reference's own volume sampler.

sample_mesh_volume (src/mpm_core/geometry.py:89-155) decides emission counts and jittered positions; its
exact output (lattice order, ray parity, one jitter draw from numpy's PCG64) is not fixed by any closed
form, so its input/output vectors are taken from the reference itself. The reference module is loaded by
file path in the build container only; the fixture holds inputs (meshes, spacing, seed) and outputs
(count, SHA-256 of the float32 bytes, a few points). decompose_transform (geometry.py:30-39) and
interpolate_transform (geometry.py:71-86) vectors are recorded alongside.

    python tests/golden/make_emission_vectors.py /root/reference
"""

from __future__ import annotations

import hashlib
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import icosphere, rot_x, rot_y, scale, translate, unit_cube_mesh  # noqa: E402


def load_reference_geometry(root: str):
    path = os.path.join(root, "src", "mpm_core", "geometry.py")
    spec = importlib.util.spec_from_file_location("reference_geometry", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cases():
    cube_v, cube_t = unit_cube_mesh()
    out = []
    # the benchmark's elastic block: 0.2 m cube at (0.4, 0.5, 0.4), spacing 1/128, seed 0
    out.append(("block_64", cube_v, cube_t, translate(0.4, 0.5, 0.4) @ scale(0.2), 1.0 / 128, 0))
    out.append(("cube_tilted", cube_v, cube_t, translate(0.3, 0.2, 0.1) @ rot_x(17) @ rot_y(33) @ scale(0.15, 0.1, 0.2),
                1.0 / 100, 7))
    for sub, r, c, h, seed in ((2, 0.1, (0.5, 0.5, 0.5), 0.01, 3), (3, 0.12, (0.4, 0.6, 0.3), 1.0 / 128, [5, 2, 1, 1, 1]),
                               (4, 0.15, (0.22, 0.5, 0.30), 1.0 / 256, 1)):
        v, t = icosphere(1.0, sub)
        out.append((f"sphere_s{sub}", v, t, translate(*c) @ scale(r), h, seed))
    return out


def main(root: str) -> None:
    g = load_reference_geometry(root)
    rec = {}
    for name, v, t, xf, spacing, seed in cases():
        world = g.transform_points(xf, v)
        pts = g.sample_mesh_volume(world, t, spacing, seed)
        rec[f"{name}/verts"] = v
        rec[f"{name}/tris"] = t
        rec[f"{name}/xf"] = xf
        rec[f"{name}/spacing"] = np.array(spacing)
        rec[f"{name}/seed"] = np.array(seed if isinstance(seed, list) else [seed], np.int64)
        rec[f"{name}/seed_is_list"] = np.array(isinstance(seed, list))
        rec[f"{name}/count"] = np.array(len(pts))
        rec[f"{name}/sha256"] = np.array(hashlib.sha256(np.ascontiguousarray(pts, np.float32).tobytes()).hexdigest())
        rec[f"{name}/head"] = pts[:8]
        rec[f"{name}/tail"] = pts[-8:]
        print(name, len(pts))
    mats = [np.diag([2.0, 2.0, 2.0, 1.0]), translate(1, 2, 3) @ rot_x(40) @ scale(0.3, 0.5, 2.0),
            translate(-1, 0.5, 0) @ rot_y(170) @ scale(1.0, 1.0, -1.0), rot_x(179.9) @ rot_y(-95)]
    for i, m in enumerate(mats):
        t_, q_, S_ = g.decompose_transform(np.asarray(m, np.float32))
        rec[f"decompose{i}/m"] = np.asarray(m, np.float32)
        rec[f"decompose{i}/t"], rec[f"decompose{i}/q"], rec[f"decompose{i}/S"] = t_, q_, S_
    for i, (a, b, s) in enumerate(((mats[0], mats[1], 0.3), (mats[1], mats[2], 0.5), (mats[3], mats[1], 0.9))):
        rec[f"interp{i}/a"], rec[f"interp{i}/b"], rec[f"interp{i}/s"] = (np.asarray(a, np.float32),
                                                                         np.asarray(b, np.float32), np.array(s))
        rec[f"interp{i}/out"] = g.interpolate_transform(np.asarray(a, np.float32), np.asarray(b, np.float32), s)
    np.savez_compressed(os.path.join(HERE, "emission_vectors.npz"), **rec)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
