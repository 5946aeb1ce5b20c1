"""The reference's command-line driver (solve_example) built against the C++
facade: MatrixMarket in, x.txt out, #META lines, exit codes."""
import os
import subprocess

import numpy as np
import pytest

from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bootcmatchg_amd", "solve_example")
SETTINGS = os.path.join(ROOT, "tests", "golden", "settings_c0.txt")


def write_mm_symmetric(path, A):
    row, col, val = A
    n = len(row) - 1
    r = np.repeat(np.arange(n), np.diff(row))
    low = col <= r
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% lower triangle\n")
        f.write("%d %d %d\n" % (n, n, int(low.sum())))
        for i, j, v in zip(r[low], col[low], val[low]):
            f.write("%d %d %.17g\n" % (i + 1, j + 1, v))


def test_solve_example_matches_oracle(tmp_path):
    A = ops.lognormal3d(9, 4.0, seed=4)
    n = len(A[0]) - 1
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    p = subprocess.run([EXE, mm, SETTINGS], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "#META sol;num_iteration int" in p.stdout
    assert "#META time;solve_time float" in p.stdout
    x = np.loadtxt(tmp_path / "x.txt")
    rc, xo, res = orc.solve(A, np.ones(n), orc.params_from_file(SETTINGS)[1])
    assert rc == 0
    its = int(p.stdout.split("#META sol;num_iteration int")[1].split()[0])
    assert abs(its - res.iterations) <= 2
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_solve_example_bad_settings_exits_like_the_reference(tmp_path):
    A = ops.poisson3d(4)
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    bad = tmp_path / "s"
    bad.write_text("7\n" * 14)
    p = subprocess.run([EXE, mm, str(bad)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 255
    p = subprocess.run([EXE, mm, str(tmp_path / "missing")], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 255 and "Setting file not found!" in p.stdout
