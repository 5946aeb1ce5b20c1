"""The K-cycle (cycle_type 3: two inner flexible CG steps on each level whose
next level is not the coarsest, src/SOLVER/FCG.cu:620-723) on the GPU against
the oracle. The hierarchy is bit-exact as for the other cycles; one
preconditioner application agrees with the oracle's to 1e-10 (different
summation orders in the inner products); FCG iterations within +-2. The
multi-column and serving forms give every column the single solve's bits.
"""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import OPERATORS, SETTINGS, bits, check_fcg_vs_oracle

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
    check_fcg_vs_oracle(dx.numpy(), it, xo, res_o.iterations, A, b, K["rtol"], "K " + name)


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


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_kcycle_multi_column_equals_single(ctx, prec):
    """Several right-hand sides through one K-cycle composite: every column
    bit-identical to its single solve (per-column inner iterations, dots in
    the single path's order), also with more columns than one block holds."""
    A = OPERATORS["lognormal3d_14"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**K))
    B.set_precision(prec)
    k = 5
    bs = [dev(ctx, ops.uniform_pm1(40 + j, n)) for j in range(k)]
    xs = [bc.DeviceArray(ctx, n) for _ in range(k)]
    its, res, st = B.fcg_multi(bs, xs)
    assert list(st) == [0] * k
    for j in range(k):
        x1 = bc.DeviceArray(ctx, n)
        it1, r1 = B.fcg(bs[j], x1)
        assert its[j] == it1 and res[j] == r1, (j, its[j], it1)
        assert np.array_equal(bits(xs[j].numpy()), bits(x1.numpy()))


def test_kcycle_serving_alone_equals_mix(ctx):
    """A K-cycle composite served next to a V-cycle one, requests arriving
    over several steps: each result equals its single solve."""
    Ak = OPERATORS["poisson3d_17"]()
    Av = OPERATORS["aniso2d_48"]()
    Bk = bc.Boot(ctx, bc.Csr(ctx, Ak), bc.make_params(**K))
    Bv = bc.Boot(ctx, bc.Csr(ctx, Av), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    S = bc.Server(ctx, 8 << 30, 4)
    ok, ov = S.add(Bk), S.add(Bv)
    reqs = []
    for q in range(6):
        B, op, A = (Bk, ok, Ak) if q % 2 == 0 else (Bv, ov, Av)
        n = len(A[0]) - 1
        reqs.append([B, op, dev(ctx, ops.uniform_pm1(60 + q, n)), bc.DeviceArray(ctx, n), n,
                     None, None])
    step = 0
    while any(r[6] is None for r in reqs):
        for q, r in enumerate(reqs):
            if r[5] is None and q <= step:
                r[5] = S.submit(r[1], r[2], r[3])
        S.step()
        for r in reqs:
            if r[5] is not None and r[6] is None:
                r[6] = S.result(r[5])
        step += 1
        assert step < 20000
    S.close()
    for B, _, b, x, n, _, res in reqs:
        x1 = bc.DeviceArray(ctx, n)
        it1, r1 = B.fcg(b, x1)
        assert res["status"] == 0 and res["iterations"] == it1 and res["residual"] == r1
        assert np.array_equal(bits(x.numpy()), bits(x1.numpy()))
