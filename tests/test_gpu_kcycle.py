"""The K-cycle (cycle_type 3: two inner flexible CG steps on each level whose
next level is not the coarsest, src/SOLVER/FCG.cu:620-723) on the GPU against
the oracle. The hierarchy is bit-exact as for the other cycles; one
preconditioner application agrees with the oracle's to 1e-10 (different
summation orders in the inner products); FCG iterations within +-2."""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import OPERATORS, SETTINGS, bits

pytestmark = pytest.mark.gpu
K = SETTINGS["K_L1_1sweep_agg"]


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


@pytest.mark.parametrize("name", ["poisson3d_17", "lognormal3d_14", "aniso2d_48"])
def test_kcycle_precond_and_fcg(ctx, name):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    O = orc.Boot(A, orc.params(**K))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**K))
    assert B.levels(0) == O.levels(0) and B.levels(0) >= 4
    r = ops.uniform_pm1(2, n)
    z = dev(ctx, np.zeros(n))
    B.precond(dev(ctx, r), z)
    zo = O.precond(r, np.zeros(n))
    assert np.linalg.norm(z.numpy() - zo) <= 1e-10 * np.linalg.norm(zo)
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dx = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(dev(ctx, b), dx)
    assert abs(it - res_o.iterations) <= 2, (it, res_o.iterations)
    x = dx.numpy()
    assert np.linalg.norm(b - ops.to_scipy(A) @ x) <= 10 * K["rtol"] * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)


def test_kcycle_reduced_precision_and_graphs(ctx):
    A = OPERATORS["lognormal3d_14"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**K))
    b = dev(ctx, ops.uniform_pm1(3, n))
    x64 = bc.DeviceArray(ctx, n)
    it64, _ = B.fcg(b, x64)
    B.set_precision(bc.PREC_FP32)
    bc.use_graphs(ctx, False)
    xe = bc.DeviceArray(ctx, n)
    ite, re_ = B.fcg(b, xe)
    bc.use_graphs(ctx, True)
    xg = bc.DeviceArray(ctx, n)
    itg, rg = B.fcg(b, xg)
    assert (itg, rg) == (ite, re_)
    assert np.array_equal(bits(xg.numpy()), bits(xe.numpy()))
    assert abs(ite - it64) <= 2
    bh = b.numpy()
    assert np.linalg.norm(bh - ops.to_scipy(A) @ xe.numpy()) <= K["rtol"] * np.linalg.norm(bh) * 1.000001


def test_kcycle_multi_column_is_unsupported(ctx):
    A = OPERATORS["poisson3d_12"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**K))
    with pytest.raises(bc.BcmgError) as e:
        B.fcg_multi([dev(ctx, np.ones(n))], [bc.DeviceArray(ctx, n)])
    assert e.value.code == bc.EUNSUP
