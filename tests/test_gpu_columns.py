"""Column forms of streamed operators: the stencil pattern (one byte per row
over a sorted table of at most 8 column offsets), the 16-bit offsets and the
int32 columns give the same bits in every row kernel and in whole solves.

The pattern is kept for square operators whose FP64 stream exceeds 128 MB
and whose column offsets take at most 8 distinct values; the 16-bit copy
for those above the non-temporal size (192 MB) whose offsets fit 16 bits. A/B runs
restrict one matrix handle to a form (bcmg_csr_set_columns); composites are
set up after the restriction, so their captured launches use it.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import SETTINGS, bits

pytestmark = pytest.mark.gpu

NARROW_NNZ = (192 << 20) // 12
RTOL_ROW = 1e-13


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


def csr_of(S):
    S = S.tocsr()
    S.sort_indices()
    return (S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64))


def row_ops(ctx, dA, n, seed):
    """SpMV, residual and three Jacobi/L1 sweeps on seeded vectors; the results."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    d = rng.uniform(7, 8, n)
    dx, df, dd = dev(ctx, x), dev(ctx, f), dev(ctx, d)
    y = bc.DeviceArray(ctx, n)
    out = []
    bc.spmv(ctx, dA, dx, y)
    out.append(y.numpy())
    bc.residual(ctx, dA, dx, df, y)
    out.append(y.numpy())
    u, u2 = dev(ctx, x), bc.DeviceArray(ctx, n)
    bc.relax(ctx, dA, dd, df, u, u2, 3)
    out.append(u.numpy())
    for a in (dx, df, dd, y, u, u2):
        a.free()
    return out, x


def same_bits(a, b):
    return all(np.array_equal(bits(p), bits(q)) for p, q in zip(a, b))


def test_pattern_forms_bitexact(ctx):
    """3D 7-point Poisson 140^3 (7 offsets, all within 16 bits): the pattern,
    the 16-bit offsets and the int32 columns give the same bits in SpMV,
    residual and sweeps, within the oracle's row bound."""
    A = ops.poisson3d(140)
    n = len(A[0]) - 1
    assert len(A[1]) > NARROW_NNZ
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == 0
    res = {}
    for form in (0, 2, 4):
        dA.set_columns(form)
        assert dA.column_bytes == form
        res[form], x = row_ops(ctx, dA, n, 21)
    assert same_bits(res[0], res[2]) and same_bits(res[0], res[4])
    absmv = ops.to_scipy((A[0], A[1], np.abs(A[2]))) @ np.abs(x)
    assert np.all(np.abs(res[0][0] - orc.spmv(A, x)) <= RTOL_ROW * absmv + 1e-300)
    dA.free()


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_pattern_solve_bitexact(ctx, prec):
    """Whole setup + FCG solves on 140^3 with the drop-in settings: every
    column form gives the same iterations and solution bits (FP64 and the
    FP32 hierarchy, whose level-0 sweeps stream FP32 values)."""
    A = ops.poisson3d(140)
    n = len(A[0]) - 1
    b = ops.uniform_pm1(0, n)
    dA = bc.Csr(ctx, A)
    db = dev(ctx, b)
    out = {}
    for form in (0, 2, 4):
        dA.set_columns(form)
        B = bc.Boot(ctx, dA, bc.make_params(**dict(SETTINGS["V_L1_2sweeps"])))
        if prec != bc.PREC_FP64:
            B.set_precision(prec)
        dx = bc.DeviceArray(ctx, n)
        it, res = B.fcg(db, dx)
        out[form] = (it, res, dx.numpy())
        print("form %d: %d iterations, residual %.3e" % (form, it, res), flush=True)
        B.free()
        dx.free()
    for form in (2, 4):
        assert out[form][0] == out[0][0] and out[form][1] == out[0][1]
        assert np.array_equal(bits(out[form][2]), bits(out[0][2]))
    S = ops.to_scipy(A)
    assert np.linalg.norm(b - S @ out[0][2]) <= 1e-8 * (1 + 1e-6) * np.linalg.norm(b)
    db.free()
    dA.free()


def test_pattern_without_16bit_offsets(ctx):
    """3D 7-point Poisson 190^3: the plane offset 190^2 = 36100 needs more
    than 16 bits, so there is no 16-bit copy, but 7 offsets make a pattern;
    it gives the int32 kernels' bits."""
    A = ops.poisson3d(190)
    n = len(A[0]) - 1
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == 0
    with pytest.raises(bc.BcmgError) as e:
        dA.set_columns(2)
    assert e.value.code == bc.EUNSUP
    p, _ = row_ops(ctx, dA, n, 22)
    dA.set_columns(4)
    assert dA.column_bytes == 4
    w, _ = row_ops(ctx, dA, n, 22)
    assert same_bits(p, w)
    dA.free()


def test_no_pattern_with_nine_offsets(ctx):
    """A 2D 9-point operator on 1800^2 (9 offsets): the 16-bit copy only."""
    N = 1800
    T = sp.diags([np.ones(N - 1), np.ones(N), np.ones(N - 1)], [-1, 0, 1])
    S = 10.0 * sp.eye(N * N) - sp.kron(T, T)
    A = csr_of(S)
    assert len(A[1]) > NARROW_NNZ
    assert len(np.unique(A[1] - np.repeat(np.arange(N * N, dtype=np.int32), np.diff(A[0])))) == 9
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == 2
    with pytest.raises(bc.BcmgError) as e:
        dA.set_columns(0)
    assert e.value.code == bc.EUNSUP
    dA.free()


def test_pattern_anisotropic_2d_with_boundary_rows(ctx):
    """2D 5-point anisotropic 2048^2 (5 offsets; boundary rows miss some):
    pattern and int32 columns agree bit for bit."""
    A = ops.aniso2d(2048, 0.01)
    n = len(A[0]) - 1
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == 0
    p, _ = row_ops(ctx, dA, n, 23)
    dA.set_columns(4)
    w, _ = row_ops(ctx, dA, n, 23)
    assert same_bits(p, w)
    dA.free()
