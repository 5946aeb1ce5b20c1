"""Degenerate operators through the whole path, GPU against the oracle.

- 2x2, a diagonal-only operator (no edge to match: every aggregate is a
  singleton), two disconnected blocks, and an operator already below the
  coarse-size limit: hierarchies bit-exact, FCG iterations within +-2,
  solutions within 1e-12 of the oracle's when the counts agree, also through
  the multi-column path;
- 1x1: the first FCG step is exact, so the loop pass FCG v3 always makes
  (src/SOLVER/FCG.cu:498-565) divides 0 by 0 (the reference returns NaN);
  both sides report the breakdown instead.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import SETTINGS, bits, check_fcg_vs_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def csr(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))


CASES = {
    "n2": lambda: csr(np.array([[2.0, -1.0], [-1.0, 2.0]])),
    "diagonal_only": lambda: csr(sp.diags(np.arange(1, 1001, dtype=float))),
    "two_blocks": lambda: csr(sp.block_diag([ops.to_scipy(ops.poisson3d(8)),
                                              ops.to_scipy(ops.poisson2d(20))])),
    "below_coarse_limit": lambda: ops.poisson3d(4),
}


def compare(B, O, k):
    assert B.levels(k) == O.levels(k)
    for l in range(B.levels(k)):
        g, o = B.level_A(k, l), O.level_A(k, l)
        assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]), l
        assert np.array_equal(bits(g[2]), bits(o[2])), l
        if l + 1 < B.levels(k):
            ag, pg = B.level_P(k, l)
            ao, po = O.level_P(k, l)
            assert np.array_equal(ag, ao) and np.array_equal(bits(pg), bits(po)), l


@pytest.mark.parametrize("setting", ["V_L1_2sweeps", "H_jacobi_1sweep", "boot3_mult"])
@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_operators(ctx, name, setting):
    A = CASES[name]()
    n = len(A[0]) - 1
    prm = SETTINGS[setting]
    print("edge", name, setting, n, flush=True)
    O = orc.Boot(A, orc.params(**prm))
    assert O.rc == 0
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    assert B.info()[0] == O.info()[0]
    compare(B, O, 0)
    b = ops.uniform_pm1(1, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dx = bc.DeviceArray(ctx, n)
    db = bc.DeviceArray.from_numpy(ctx, b)
    it, _ = B.fcg(db, dx)
    check_fcg_vs_oracle(dx.numpy(), it, xo, res_o.iterations, A, b, prm["rtol"], name)
    # the same operator through the multi-column path: bit-identical columns
    xs = [bc.DeviceArray(ctx, n) for _ in range(2)]
    its, _, st = B.fcg_multi([db, db], xs)
    assert list(its) == [it, it] and list(st) == [0, 0]
    for xm in xs:
        assert np.array_equal(bits(xm.numpy()), bits(dx.numpy()))


def test_one_by_one_reports_breakdown(ctx):
    A = csr(np.array([[4.0]]))
    O = orc.Boot(A, orc.params(**SETTINGS["V_L1_2sweeps"]))
    rc, _, _, _ = O.fcg(np.array([1.0]))
    assert rc == -5                                   # ORC_EBREAKDOWN
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    with pytest.raises(bc.BcmgError) as e:
        B.fcg(bc.DeviceArray.from_numpy(ctx, np.array([1.0])), bc.DeviceArray(ctx, 1))
    assert e.value.code == bc.EBREAKDOWN
