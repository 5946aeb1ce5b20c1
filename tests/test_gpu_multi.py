"""The multi-column solve (bcmg_fcg_multi): several right-hand sides share
every launch; column j must equal the single-column solve of b_j bit for bit
(same iterations, same residual, same x), in every precision, composite type
and cycle, whatever the other columns hold (zero right-hand sides, columns
that converge at different iterations, more than 32 columns)."""
import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops

from common import OPERATORS, SETTINGS, bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


def rhs_set(n, k, zero_at=None):
    cols = []
    for j in range(k):
        b = ops.uniform_pm1(100 + j, n)
        if j % 3 == 1:
            b = b * 1e-3          # converges in fewer iterations than its neighbours
        if zero_at is not None and j == zero_at:
            b = np.zeros(n)
        cols.append(b)
    return cols


def check_multi_vs_single(ctx, B, bs):
    n = len(bs[0])
    dbs = [dev(ctx, b) for b in bs]
    xs = [bc.DeviceArray(ctx, n) for _ in bs]
    it, res, st = B.fcg_multi(dbs, xs)
    restarts_multi = B.restarts()
    for j, b in enumerate(bs):
        if not np.any(b):
            assert st[j] == bc.EZERORHS and it[j] == 0
            assert not np.any(xs[j].numpy())
            continue
        assert st[j] == bc.OK
        x1 = bc.DeviceArray(ctx, n)
        it1, res1 = B.fcg(dbs[j], x1)
        assert (it[j], res[j]) == (it1, res1), (j, it[j], it1, res[j], res1)
        assert np.array_equal(bits(xs[j].numpy()), bits(x1.numpy())), j
    return it, restarts_multi


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32, bc.PREC_BF16])
@pytest.mark.parametrize("name,setting", [
    ("poisson3d_17", "V_L1_2sweeps"),
    ("aniso2d_48", "W_L1_3sweeps"),
    ("poisson2d_33", "H_jacobi_1sweep"),
    ("lognormal3d_14", "boot3_mult"),
    ("lognormal3d_10", "boot3_sym"),
    ("poisson3d_12", "boot2_add_mixed"),
])
def test_multi_equals_single(ctx, name, setting, prec):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS[setting]))
    B.set_precision(prec)
    it, _ = check_multi_vs_single(ctx, B, rhs_set(n, 6, zero_at=4))
    assert len(set(int(v) for v in it if v)) >= 1


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
@pytest.mark.parametrize("k", [5, 20, 32])
def test_multi_column_counts(ctx, k, prec):
    """An operator above the one-workgroup relaxation's 4096 rows, so that
    the cycle's last sweep writes FCG's FP64 block itself: 5, 20 and 32
    columns give column groups that divide a 256-lane block and ones that
    do not (the results staged through LDS, or stored lane by lane), with
    columns stopping at different iterations."""
    A = OPERATORS["poisson3d_17"]()
    n = len(A[0]) - 1
    assert n > 4096
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    B.set_precision(prec)
    check_multi_vs_single(ctx, B, rhs_set(n, k))


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_multi_equals_single_narrow_columns(ctx, prec):
    """An operator whose FP64 stream exceeds 192 MB: the single-column
    solve streams its stencil pattern on level 0 (sweeps, residual, the fused
    FCG SpMV), the multi-column kernels its 32-bit columns; the columns still
    agree bit for bit."""
    A = ops.poisson3d(140)
    n = len(A[0]) - 1
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == 0
    B = bc.Boot(ctx, dA, bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    B.set_precision(prec)
    check_multi_vs_single(ctx, B, rhs_set(n, 3))


def test_multi_more_than_32_columns_and_graphs(ctx):
    A = OPERATORS["poisson3d_12"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    bs = rhs_set(n, 35)
    check_multi_vs_single(ctx, B, bs)
    # host launches instead of graph replay: same bits
    dbs = [dev(ctx, b) for b in bs[:5]]
    xg = [bc.DeviceArray(ctx, n) for _ in dbs]
    xe = [bc.DeviceArray(ctx, n) for _ in dbs]
    itg, resg, _ = B.fcg_multi(dbs, xg)
    bc.use_graphs(ctx, False)
    ite, rese, _ = B.fcg_multi(dbs, xe)
    bc.use_graphs(ctx, True)
    assert np.array_equal(itg, ite) and np.array_equal(resg, rese)
    for a, b in zip(xg, xe):
        assert np.array_equal(bits(a.numpy()), bits(b.numpy()))


def test_multi_long_rows(ctx):
    """A fine operator whose 256-row blocks exceed the LDS tile (120 entries
    per row): the group-order SpMV and the unfused products are mirrored."""
    n = 300
    M = sp.random(n, n, density=0.4, random_state=3, format="csr")
    M = (M + M.T + sp.eye(n) * n).tocsr()
    M.sort_indices()
    A = (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    check_multi_vs_single(ctx, B, rhs_set(n, 4))


def test_multi_restarts_match_single(ctx):
    """True-residual restarts (rtol below what FP64 reaches) happen per column
    exactly as in the single-column solve."""
    A = ops.lognormal3d(32, 25.0, seed=7)
    n = len(A[0]) - 1
    prm = bc.make_params(**dict(SETTINGS["V_L1_2sweeps"], rtol=1e-15, itnlim=60))
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    B.set_precision(bc.PREC_FP32)
    _, restarts = check_multi_vs_single(ctx, B, rhs_set(n, 3))
    assert restarts >= 1


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_multi_streamed_blocks(ctx, prec):
    """32 columns on 3D Poisson 100^3: the FCG column blocks exceed the
    non-temporal size (192 MB), so the FCG SpMV and update stream them
    non-temporally, and in FP64 so do the level-0 interleaved blocks (256 MB;
    128 MB in FP32), so the row kernels and the last sweep stream too; every
    column keeps its single solve's iterations and bits."""
    A = ops.poisson3d(100)
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    B.set_precision(prec)
    check_multi_vs_single(ctx, B, rhs_set(n, 32))
