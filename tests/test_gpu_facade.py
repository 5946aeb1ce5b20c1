"""The command-line solver (bootcmatchg_amd/bcmg_solve) over the C++ facade:
MatrixMarket and MTX2 in (src/matrix/matrixIO.cu:4-289), x.txt out
("%.14e", src/matrix/vector.cu:444-466), the #META lines of solve()
(src/SOLVER/solver.cu:106-118) and of every hierarchy
(src/AMG/matchingAggregation.cu:412-414) with the level table
(src/AMG/AMG.cu:241-251), and the reference's exit codes."""
import os
import re
import subprocess

import numpy as np
import pytest

from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bootcmatchg_amd", "bcmg_solve")
GOLDEN = os.path.join(ROOT, "tests", "golden")
SETTINGS = os.path.join(GOLDEN, "settings_c0.txt")


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


def write_mtx2_shuffled(path, A, seed=0):
    """MTX2: n, nnz, the 1-based rows, the 1-based columns, the values; the
    entries in a random order (the reader sorts each row)."""
    row, col, val = A
    n = len(row) - 1
    r = np.repeat(np.arange(n), np.diff(row))
    perm = np.random.default_rng(seed).permutation(len(col))
    with open(path, "w") as f:
        f.write("%d %d\n" % (n, len(col)))
        f.write("\n".join(str(v) for v in r[perm] + 1) + "\n")
        f.write("\n".join(str(v) for v in col[perm] + 1) + "\n")
        f.write("\n".join("%.17g" % v for v in val[perm]) + "\n")


def write_settings(path, rtol, base=SETTINGS):
    lines = open(base).read().splitlines()
    lines[13] = "%g" % rtol
    open(path, "w").write("\n".join(lines) + "\n")


def run(args, cwd):
    print("bcmg_solve", " ".join(args[1:]), flush=True)
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)


def meta(out, key):
    return [float(v) for v in re.findall(r"#META %s (?:int|float) (\S+)" % re.escape(key), out)]


def test_cli_defaults_match_the_oracle(tmp_path):
    """No options: the reference driver's behaviour (b = 1, x.txt)."""
    A = ops.lognormal3d(9, 4.0, seed=4)
    n = len(A[0]) - 1
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    p = run([EXE, mm, SETTINGS], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    x = np.loadtxt(tmp_path / "x.txt")
    rc, xo, res = orc.solve(A, np.ones(n), orc.params_from_file(SETTINGS)[1])
    assert rc == 0
    its = int(meta(p.stdout, "sol;num_iteration")[0])
    assert abs(its - res.iterations) <= 2
    if its == res.iterations:
        # x.txt holds 15 significant digits
        assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    else:
        r = np.ones(n) - ops.to_scipy(A) @ x
        assert np.linalg.norm(r) <= 1e-6 * (1 + 1e-6) * np.sqrt(n)


def test_cli_prints_every_hierarchy_like_the_reference(tmp_path):
    """Three hierarchies (bootstrap): one level table, cmplx lines and
    agg;level_number / avg_coarse_ratio / OpCmplx per hierarchy, sizes and
    complexities equal to the oracle's hierarchies."""
    A = ops.lognormal3d(12, 4.0, seed=1)
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    cfg = os.path.join(GOLDEN, "settings_c2.txt")
    p = run([EXE, "-b", "rand:0", "-o", "-", mm, cfg], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    out = p.stdout
    O = orc.Boot(A, orc.params_from_file(cfg)[1])
    nh, _ = O.info()
    assert meta(out, "agg;hierarchy_levels_num") == [nh]
    assert len(meta(out, "agg;level_number")) == nh
    assert out.count("TOTAL_MUL_TIME: ") == nh
    assert out.count("Current cmplx for V-cycle: ") == nh
    assert "Number of hierarchy: %d" % nh in out
    tables = re.findall(r"^A(\d+) n: (\d+) nnz: (\d+) avg_nnz: \S+$", out, flags=re.M)
    expect = []
    for k in range(nh):
        sizes = [(len(O.level_A(k, l)[0]) - 1, len(O.level_A(k, l)[1]))
                 for l in range(O.levels(k))]
        expect += [(str(l), str(a), str(b)) for l, (a, b) in enumerate(sizes)]
        nnz = [b for _, b in sizes]
        ratio = np.mean([sizes[l - 1][0] / sizes[l][0] for l in range(1, len(sizes))])
        assert meta(out, "agg;level_number")[k] == len(sizes) - 1
        assert abs(meta(out, "agg;OpCmplx")[k] - sum(nnz) / nnz[0]) <= 1e-6 * sum(nnz) / nnz[0]
        assert abs(meta(out, "agg;avg_coarse_ratio")[k] - ratio) <= 1e-6 * ratio
        assert "%d] %d" % (k, sizes[-1][0]) in out
    assert tables == expect


def test_cli_mtx2_random_rhs_fp32(tmp_path):
    """MTX2 input in shuffled entry order, a seeded right-hand side and the
    FP32 preconditioner: the true residual meets rtol and the solution agrees
    with the FP64 oracle's."""
    A = ops.lognormal3d(10, 4.0, seed=2)
    n = len(A[0]) - 1
    f = str(tmp_path / "a.mtx2")
    write_mtx2_shuffled(f, A, seed=5)
    cfg = str(tmp_path / "s")
    write_settings(cfg, 1e-10)
    p = run([EXE, "-f", "mtx2", "-b", "rand:7", "-p", "fp32", "-o", "sol.txt", f, cfg],
            tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    x = np.loadtxt(tmp_path / "sol.txt")
    b = ops.uniform_pm1(7, n)
    r = b - ops.to_scipy(A) @ x
    assert np.linalg.norm(r) <= 1e-10 * (1 + 1e-6) * np.linalg.norm(b)
    rc, xo, res = orc.solve(A, b, orc.params_from_file(cfg)[1])
    assert rc == 0
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)


def test_cli_rhs_file_and_repeat(tmp_path):
    """A right-hand side read from a file, three solves on one composite:
    every solve reports the same iteration count, the result equals the
    oracle's."""
    A = ops.poisson3d(8)
    n = len(A[0]) - 1
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    b = ops.uniform_pm1(11, n)
    np.savetxt(tmp_path / "b.txt", b, fmt="%.17g")
    p = run([EXE, "-b", "b.txt", "-r", "3", mm, SETTINGS], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    its = re.findall(r"^solve \d: \S+ ms, (\d+) iterations$", p.stdout, flags=re.M)
    assert len(its) == 3 and len(set(its)) == 1
    x = np.loadtxt(tmp_path / "x.txt")
    rc, xo, res = orc.solve(A, b, orc.params_from_file(SETTINGS)[1])
    assert rc == 0 and abs(int(its[0]) - res.iterations) <= 2
    if int(its[0]) == res.iterations:
        assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)


def test_cli_exits_like_the_reference(tmp_path):
    A = ops.poisson3d(4)
    mm = str(tmp_path / "a.mtx")
    write_mm_symmetric(mm, A)
    bad = tmp_path / "s"
    bad.write_text("7\n" * 14)
    p = run([EXE, mm, str(bad)], tmp_path)
    assert p.returncode == 255
    p = run([EXE, mm, str(tmp_path / "missing")], tmp_path)
    assert p.returncode == 255 and "Setting file not found!" in p.stdout
    p = run([EXE, mm], tmp_path)
    assert p.returncode == 1
    # zero right-hand side: the reference exits with 1 (src/SOLVER/FCG.cu:459-462)
    z = tmp_path / "zero.txt"
    z.write_text("0\n" * (len(A[0]) - 1))
    p = run([EXE, "-b", str(z), mm, SETTINGS], tmp_path)
    assert p.returncode == 1
