"""Mutation check of the oracle's pins (CPU): plausible misreadings of the reference, planted one at a time in
a copy of oracle/cpu_ref.cpp, must each fail the pin suite (tests/test_oracle.py, tests/test_oracle_pins.py).

Every mutant is built with g++ into a temporary folder and the pin suite runs in a child process that loads it
through MPM_ORACLE_LIB. An unmutated copy built the same way must pass, so a failure is the mutation's doing.
The mistakes planted here live only in those temporary copies of the oracle; nothing on the GPU is involved.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCE = os.path.join(ROOT, "oracle", "cpu_ref.cpp")
PIN_FILES = [os.path.join(ROOT, "tests", "test_oracle.py"), os.path.join(ROOT, "tests", "test_oracle_pins.py")]

# name -> (text of the oracle, misreading); each text occurs exactly once in cpu_ref.cpp
MUTATIONS = {
    # P2G, kernels.py:126-140
    "p2g_stress_factor_3": ("const T k = -dt * p_vol * T(4) * inv_dx * inv_dx;",
                            "const T k = -dt * p_vol * T(3) * inv_dx * inv_dx;"),
    "p2g_stress_sign": ("const T k = -dt * p_vol * T(4) * inv_dx * inv_dx;",
                        "const T k = dt * p_vol * T(4) * inv_dx * inv_dx;"),
    "p2g_affine_without_mass": ("+ pm * T(C[9 * p + i])", "+ T(C[9 * p + i])"),
    "p2g_hardening_sign": ("std::exp(T(m.hardening) * (T(1) - T(Jp[p])))",
                           "std::exp(T(m.hardening) * (T(Jp[p]) - T(1)))"),
    # stresses, kernels.py:30-60
    "corotated_transposed": ("mmul(FmR, mtrans(F))", "mmul(mtrans(F), FmR)"),
    "hencky_without_2mu": ("d[i] = T(2) * mu * e[i] + lam * tr;", "d[i] = mu * e[i] + lam * tr;"),
    # G2P, kernels.py:224-263
    "g2p_affine_factor_3": ("const T wc = T(4) * inv_dx * inv_dx * w;", "const T wc = T(3) * inv_dx * inv_dx * w;"),
    "f_update_transposed": ("mmul(G, load_m<T>(F + 9 * p))", "mmul(load_m<T>(F + 9 * p), G)"),
    "f_update_sign": ("G.a[i / 3][i % 3] += dt * nC.a[i / 3][i % 3];", "G.a[i / 3][i % 3] -= dt * nC.a[i / 3][i % 3];"),
    "snow_jp_ratio_inverted": ("(sig[0] * sig[1] * sig[2]) / (ns[0] * ns[1] * ns[2])",
                               "(ns[0] * ns[1] * ns[2]) / (sig[0] * sig[1] * sig[2])"),
    # grid contact, colliders.py:96-117
    "collider_band_2dx": ("band = dx;", "band = T(2) * dx;"),
    "in_band_approach_dropped": ("vel[a] = vcol[a] + rt[a] + vn * n[a];", "vel[a] = vcol[a] + rt[a];"),
    "inside_keeps_approach": ("if (phi < T(0)) {\n            for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a];",
                              "if (phi < T(0)) {\n            for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a] + vn * n[a];"),
    "friction_sign": ("T(1) - T(cs[c].friction) * (-vn) / rl", "T(1) - T(cs[c].friction) * vn / rl"),
    # particle projection, colliders.py:137-147
    "projection_wrong_way": ("pos[a] -= phi * nrm[a];", "pos[a] += phi * nrm[a];"),
    "projection_removes_leaving": ("if (vn < T(0))\n          for (int a = 0; a < 3; ++a) vel[a] -= vn * nrm[a];",
                                   "if (vn > T(0))\n          for (int a = 0; a < 3; ++a) vel[a] -= vn * nrm[a];"),
}


def _build(src_text: str, out_dir: str) -> str:
    os.makedirs(out_dir, exist_ok=True)
    src = os.path.join(out_dir, "cpu_ref.cpp")
    with open(src, "w") as f:
        f.write(src_text)
    lib = os.path.join(out_dir, "libmpm_oracle.so")
    subprocess.run(["g++", "-O1", "-fopenmp", "-shared", "-fPIC", "-std=c++17", "-o", lib, src], check=True,
                   capture_output=True)
    return lib


def _run_pins(lib: str) -> subprocess.CompletedProcess:
    env = dict(os.environ, MPM_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-m", "pytest", *PIN_FILES, "-x", "-q", "-p", "no:cacheprovider",
                           "-p", "no:xdist", "-m", "not gpu"], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=600)


@pytest.fixture(scope="module")
def outcomes(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    base = tmp_path_factory.mktemp("oracle_mutants")
    original = open(SOURCE).read()
    jobs = {"unmutated": original}
    for name, (text, wrong) in MUTATIONS.items():
        assert original.count(text) == 1, f"{name}: the oracle text to mutate is not unique or is gone"
        jobs[name] = original.replace(text, wrong)

    def one(item):
        name, text = item
        return name, _run_pins(_build(text, os.path.join(base, name)))

    with ThreadPoolExecutor(max_workers=4) as pool:
        return dict(pool.map(one, jobs.items()))


def test_unmutated_copy_passes(outcomes):
    r = outcomes["unmutated"]
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutation_fails_the_pins(outcomes, name):
    r = outcomes[name]
    # exit status 1: tests ran and at least one failed (not a collection or usage error)
    assert r.returncode == 1, f"{name} survived the pin suite (exit {r.returncode}):\n{r.stdout[-2000:]}"
