"""CPU check of the tiled force kernel's index arithmetic on real cell
lists (bounds of every staged window and LDS access, coverage of every
neighbour row, one block per bead)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import radius_for, walk

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("tp") / "tile_plan_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "tile_plan_check.cpp")],
                   check=True)
    return str(exe)


@pytest.mark.parametrize("N,cvf,csl_h,het_frac,seed", [
    (1, 0.2, 2.0, 1.0, 1), (1000, 0.2, 2.0, 0.5, 2), (30386, 0.2, 2.0, 0.5, 3),
    (131072, 0.3, 2.5, 0.3, 4), (50000, 0.05, 2.0, 0.5, 5),
])
def test_tile_plan_bounds_and_coverage(checker, tmp_path, N, cvf, csl_h, het_frac, seed):
    R_n = radius_for(N, cvf) if N > 1 else 5.0
    r = walk(N, R_n, seed).astype(np.float32)
    het = (np.random.default_rng(seed).random(N) < het_frac).astype(np.uint8)
    pf, hf = tmp_path / "r.f32", tmp_path / "het.u8"
    r.tofile(pf)
    het.tofile(hf)
    d_m = float(np.float32(R_n))
    out = subprocess.run([checker, str(pf), str(N), "1.25", str(csl_h), repr(d_m), str(hf)],
                         capture_output=True, text=True)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "fails=0" in out.stdout
