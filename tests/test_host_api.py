"""Host-side checks that need no GPU: the settings parser (product and
oracle agree, error cases), the exported C-ABI, the facade header."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import bootcmatchg_amd as bc
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

FIELDS = [f for f, _ in bc.Params._fields_]


def write_settings(path, values, comments=True):
    with open(path, "w") as f:
        for v in values:
            f.write(("%s    %% comment after the value\n" % v) if comments else ("%s\n" % v))


def params_equal(p, q):
    return all(getattr(p, f) == getattr(q, f) for f in FIELDS)


@pytest.mark.parametrize("fname", ["settings_c0.txt", "settings_c2.txt", "settings_c3.txt"])
def test_settings_files_parse_identically(fname):
    path = os.path.join(GOLDEN, fname)
    p = bc.params_from_file(path)
    rc, q = orc.params_from_file(path)
    assert rc == 0 and params_equal(p, q)
    assert p.matchtype == 3


def test_settings_values_and_layout(tmp_path):
    vals = ["2", "3", "0.5", "2", "0", "39", "1", "0", "4", "16", "4", "3", "2000", "1.e-8"]
    path = str(tmp_path / "s")
    write_settings(path, vals)
    p = bc.params_from_file(path)
    assert (p.solver_type, p.max_hrc, p.conv_ratio, p.aggrsweeps, p.aggrtype, p.max_levels,
            p.cycle_type, p.coarse_solver, p.relax_type, p.relaxnumber_coarse,
            p.prerelax_sweeps, p.postrelax_sweeps, p.itnlim, p.rtol) == \
        (2, 3, 0.5, 2, 0, 39, 1, 0, 4, 16, 4, 3, 2000, 1e-8)


@pytest.mark.parametrize("line,value", [(0, "3"), (4, "2"), (6, "3"), (7, "1"), (8, "5")])
def test_settings_out_of_range(tmp_path, line, value):
    vals = ["0", "1", "0.8", "2", "0", "39", "0", "4", "4", "16", "4", "4", "2000", "1e-6"]
    vals[line] = value
    path = str(tmp_path / "s")
    write_settings(path, vals, comments=False)
    with pytest.raises(bc.BcmgError) as e:
        bc.params_from_file(path)
    assert e.value.code == bc.ESETTINGS
    assert orc.params_from_file(path)[0] == -2


def test_settings_missing_and_short(tmp_path):
    with pytest.raises(bc.BcmgError) as e:
        bc.params_from_file(str(tmp_path / "nope"))
    assert e.value.code == bc.ESETTINGS
    path = str(tmp_path / "short")
    write_settings(path, ["0", "1", "0.8"])
    with pytest.raises(bc.BcmgError):
        bc.params_from_file(path)
    assert orc.params_from_file(path)[0] == -2


def test_smoothed_aggregation_is_parsed(tmp_path):
    vals = ["0", "1", "0.8", "2", "1", "39", "0", "4", "4", "16", "4", "4", "2000", "1e-6"]
    path = str(tmp_path / "s")
    write_settings(path, vals)
    p = bc.params_from_file(path)
    assert p.aggrtype == 1


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bcmg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(bcmg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = bc.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", bc.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    assert set(names) <= exported


def test_no_device_is_reported_not_faked():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(bc.BcmgError) as e:
        bc.Context(0)
    assert e.value.code == bc.ENODEV


REF_EXAMPLE = "/root/reference/src/solve_example.cu"


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE),
                    reason="the reference tree is not present on this machine")
def test_reference_driver_compiles_and_links_against_the_facade(tmp_path):
    """The reference's own command-line driver, read where it lies in the
    reference tree, compiles unchanged against include/BootCMatchG.h and links
    against libbcmg.so (built in the temporary directory, not run: that needs
    a GPU)."""
    exe = str(tmp_path / "solve_example")
    # fed on stdin, so that its #include "BootCMatchG.h" finds this header and
    # not the CUDA one beside it
    with open(REF_EXAMPLE) as src:
        subprocess.run(["g++", "-std=c++17", "-x", "c++", "-I", os.path.join(ROOT, "include"),
                        "-o", exe, "-", "-x", "none", "-L", os.path.dirname(bc.LIB_PATH),
                        "-lbcmg"], stdin=src, check=True)
    assert os.path.exists(exe)


def test_headers_and_cli_compile():
    """bcmg.h is plain C; the command-line solver compiles with warnings as errors."""
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-x", "c",
                    os.path.join(ROOT, "include", "bcmg.h")], check=True)
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "bootcmatchg_amd", "csrc", "bcmg_cli.cpp")], check=True)


def test_operator_generators_are_deterministic():
    from bootcmatchg_amd import operators as ops
    a = ops.lognormal3d(5, 4.0, seed=1)
    b = ops.lognormal3d(5, 4.0, seed=1)
    c = ops.lognormal3d(5, 4.0, seed=2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[2], c[2])


def test_extension_entry_points_validate_arguments():
    """The extensions (precision, multi-column FCG, serving loop) reject bad
    handles and arguments with BCMG_EINVAL before touching a device."""
    L = bc.lib()
    vp = ctypes.c_void_p
    i = ctypes.c_int()
    ll = ctypes.c_longlong()
    assert L.bcmg_set_precision(vp(), 1) == bc.EINVAL
    assert L.bcmg_fcg_restarts(vp(), ctypes.byref(i)) == bc.EINVAL
    assert L.bcmg_fcg_multi(vp(), vp(), 1, vp(), vp(), vp(), vp(), vp()) == bc.EINVAL
    assert L.bcmg_server_create(vp(), ctypes.c_size_t(1 << 30), 8, ctypes.byref(vp())) == \
        bc.EINVAL
    assert L.bcmg_server_add_operator(vp(), vp(), ctypes.byref(i)) == bc.EINVAL
    assert L.bcmg_server_submit(vp(), 0, vp(), vp(), ctypes.byref(ll)) == bc.EINVAL
    assert L.bcmg_server_step(vp(), ctypes.byref(i)) == bc.EINVAL
    assert L.bcmg_server_result(vp(), ctypes.c_longlong(1), None, None, None, None, None) == \
        bc.EINVAL
    assert L.bcmg_server_completed(vp(), None, 0, ctypes.byref(i)) == bc.EINVAL
    assert L.bcmg_server_info(vp(), None, None, None, None) == bc.EINVAL
    assert L.bcmg_server_destroy(vp()) == bc.OK
    # context on a caller's stream: the null stream is refused
    assert L.bcmg_ctx_create_on_stream(0, vp(), ctypes.byref(vp())) == bc.EINVAL
    # smoothed aggregation's product and prolongator accessors
    assert L.bcmg_spgemm(vp(), 1, 1, 1, vp(), vp(), vp(), vp(), vp(), vp(),
                         ctypes.c_longlong(0), ctypes.byref(vp())) == bc.EINVAL
    assert L.bcmg_level_Pinfo(vp(), 0, 0, None, None, None) == bc.EINVAL
    assert L.bcmg_level_Pcsr(vp(), 0, 0, None, None, None) == bc.EINVAL


READER_SRC = r'''
#include "BootCMatchG.h"
int main(int argc, char **argv)
{
    CSR *A = readMatrixFromFile(argv[1], std::atoi(argv[2]), false);
    std::printf("%d %d %d\n", A->n, A->m, A->nnz);
    for (int i = 0; i <= A->n; i++) std::printf("%d\n", A->row[i]);
    for (int k = 0; k < A->nnz; k++) std::printf("%d %.17g\n", A->col[k], A->val[k]);
    CSRm::free(A);
    return 0;
}
'''


@pytest.fixture(scope="module")
def reader_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("reader")
    (d / "r.cpp").write_text(READER_SRC)
    exe = str(d / "r")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(d / "r.cpp"),
                    "-o", exe, "-L", os.path.dirname(bc.LIB_PATH), "-lbcmg",
                    "-Wl,-rpath," + os.path.dirname(bc.LIB_PATH)], check=True)
    return exe


def read_with(exe, path, m_type):
    out = subprocess.run([exe, path, str(m_type)], capture_output=True, text=True, check=True,
                         timeout=60).stdout.split("\n")
    n, m, nnz = map(int, out[0].split())
    row = np.array([int(v) for v in out[1:n + 2]], dtype=np.int32)
    ent = [l.split() for l in out[n + 2:n + 2 + nnz]]
    col = np.array([int(c) for c, _ in ent], dtype=np.int32)
    val = np.array([float(v) for _, v in ent])
    return n, m, row, col, val


def test_matrix_readers_give_the_sorted_csr(reader_exe, tmp_path):
    """readMatrixFromFile on the host (no device): a symmetric MatrixMarket
    file (lower triangle), a general one in shuffled order and an MTX2 file in
    shuffled order all give the operator's CSR (columns ascending)."""
    from bootcmatchg_amd import operators as ops
    A = ops.lognormal3d(5, 4.0, seed=3)
    row, col, val = A
    n = len(row) - 1
    r = np.repeat(np.arange(n), np.diff(row))
    low = col <= r
    p1 = str(tmp_path / "sym.mtx")
    with open(p1, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n%c\n%c\n")
        f.write("%d %d %d\n" % (n, n, int(low.sum())))
        for i, j, v in zip(r[low], col[low], val[low]):
            f.write("%d %d %.17g\n" % (i + 1, j + 1, v))
    perm = np.random.default_rng(0).permutation(len(col))
    p2 = str(tmp_path / "gen.mtx")
    with open(p2, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write("%d %d %d\n" % (n, n, len(col)))
        for k in perm:
            f.write("%d %d %.17g\n" % (r[k] + 1, col[k] + 1, val[k]))
    p3 = str(tmp_path / "a.mtx2")
    with open(p3, "w") as f:
        f.write("%d %d\n" % (n, len(col)))
        f.write(" ".join(str(v) for v in r[perm] + 1) + "\n")
        f.write(" ".join(str(v) for v in col[perm] + 1) + "\n")
        f.write(" ".join("%.17g" % v for v in val[perm]) + "\n")
    for path, m_type in ((p1, 0), (p2, 0), (p3, 1)):
        nn, m, rr, cc, vv = read_with(reader_exe, path, m_type)
        assert (nn, m) == (n, n), path
        assert np.array_equal(rr, row) and np.array_equal(cc, col), path
        assert np.array_equal(vv, val), path


def test_matrix_reader_errors(reader_exe, tmp_path):
    """A matrix type other than 0/1 exits with the reference's message and
    status 1; an MTX2 index out of range and a short file exit with 255."""
    p = str(tmp_path / "x")
    open(p, "w").write("2 3\n1 2 3\n1 2 2\n1 1 1\n")
    q = subprocess.run([reader_exe, p, "2"], capture_output=True, text=True, timeout=60)
    assert q.returncode == 1 and "-F/--inputype" in q.stdout
    q = subprocess.run([reader_exe, p, "1"], capture_output=True, text=True, timeout=60)
    assert q.returncode == 255
    open(p, "w").write("2 3\n1 2 2\n1 2 2\n1 1\n")
    q = subprocess.run([reader_exe, p, "1"], capture_output=True, text=True, timeout=60)
    assert q.returncode == 255
