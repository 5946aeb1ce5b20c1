"""The bench workloads at their full sizes, checked through properties that
do not depend on size (the oracle is run where it finishes in seconds):

- 3D Poisson 160^3, FP64 drop-in (settings 0,1,0.8,2,0,39,0,4,4,16,4,4,2000,
  1e-8, b uniform[-1,1] seed 0): true relative residual <= rtol; level-0
  aggregates cover every row once, with at most 2^aggrsweeps members and
  sum of p^2 = 1 over each aggregate; every level's operator symmetric to
  1e-14; every level (operator and prolongator) bit-exact against the
  oracle's, and FCG against the oracle's (iterations within +-2, solution
  within 1e-12 when the counts agree);
- 3D Poisson 32^3, b = 1, rtol 1e-6 (the case the reference runs on a CPU):
  the same comparison with the oracle;
- 2D anisotropic 2048^2 (eps 0.01), aggrsweeps 3, W-cycle, 32 right-hand
  sides (seeds 0..31) with an FP32 hierarchy: every column at rtol in the
  true residual, and columns equal to their single solves bit for bit;
- cell-coefficient diffusion 128^3 (exp(g), g ~ N(0, 4)), bootstrap of 3
  hierarchies, against the oracle's FP64 solve (about 100 s on one core):
  the FP64 hierarchies of the first two bootstrap steps bit-exact, the FP64
  GPU solve to the oracle's tolerance; FP32 and BF16 hierarchies at rtol in
  the true residual, within 1e-5 of the oracle's FP64 solution (both stop at
  rtol 1e-8), iterations within +25% + 2 of the oracle's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import bits, check_fcg_vs_oracle

pytestmark = pytest.mark.gpu

DROPIN = dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, aggrtype=0,
              max_levels=39, cycle_type=0, coarse_solver=4, relax_type=4,
              relaxnumber_coarse=16, prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000,
              rtol=1e-8)


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def rel_res(S, b, x):
    return np.linalg.norm(b - S @ x) / np.linalg.norm(b)


def to_sp(t):
    row, col, val = t
    n = len(row) - 1
    return sp.csr_matrix((val, col, row), shape=(n, n))


def test_poisson160_dropin(ctx):
    A = ops.poisson3d(160)
    n = len(A[0]) - 1
    S = ops.to_scipy(A)
    b = ops.uniform_pm1(0, n)
    print("poisson 160^3: n", n, flush=True)
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**DROPIN))
    dx = bc.DeviceArray(ctx, n)
    it, resid = B.fcg(bc.DeviceArray.from_numpy(ctx, b), dx)
    x = dx.numpy()
    assert rel_res(S, b, x) <= DROPIN["rtol"] * 1.01
    # aggregation invariants of level 0
    agg, p = B.level_P(0, 0)
    nc = B.level_size(0, 1)[0]
    assert agg.min() == 0 and agg.max() == nc - 1
    counts = np.bincount(agg, minlength=nc)
    assert counts.min() >= 1 and counts.max() <= 2 ** DROPIN["aggrsweeps"]
    p2 = np.bincount(agg, weights=p * p, minlength=nc)
    assert np.allclose(p2, 1.0, rtol=0, atol=1e-14)
    for l in range(B.levels(0)):
        Al = to_sp(B.level_A(0, l))
        d = abs(Al - Al.T).max()
        assert d <= 1e-14 * abs(Al).max(), (l, d)
    # the oracle at this size (about 15 s on one core)
    print("oracle solve", flush=True)
    O = orc.Boot(A, orc.params(**DROPIN))
    compare_levels(B, O, 0)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    check_fcg_vs_oracle(x, it, xo, res_o.iterations, A, b, DROPIN["rtol"], "poisson 160^3")


def compare_levels(B, O, k):
    """Every level's operator and prolongator, bit for bit."""
    assert B.levels(k) == O.levels(k)
    for l in range(B.levels(k)):
        g, o = B.level_A(k, l), O.level_A(k, l)
        assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]), (k, l)
        assert np.array_equal(bits(g[2]), bits(o[2])), (k, l)
        if l + 1 < B.levels(k):
            (ag, pg), (ao, po) = B.level_P(k, l), O.level_P(k, l)
            assert np.array_equal(ag, ao) and np.array_equal(bits(pg), bits(po)), (k, l)
    print("hierarchy %d: %d levels bit-exact" % (k, B.levels(k)), flush=True)


def test_poisson32_reference_cpu_case(ctx):
    """The reference's CPU-runnable configuration: 32^3, b = 1, rtol 1e-6."""
    A = ops.poisson3d(32)
    n = len(A[0]) - 1
    b = np.ones(n)
    prm = dict(DROPIN, rtol=1e-6)
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    O = orc.Boot(A, orc.params(**prm))
    compare_levels(B, O, 0)
    dx = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(bc.DeviceArray.from_numpy(ctx, b), dx)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    check_fcg_vs_oracle(dx.numpy(), it, xo, res_o.iterations, A, b, 1e-6, "poisson 32^3 b=1")


def test_aniso2048_32_rhs(ctx):
    A = ops.aniso2d(2048, 0.01)
    n = len(A[0]) - 1
    S = ops.to_scipy(A)
    prm = bc.make_params(**dict(DROPIN, aggrsweeps=3, cycle_type=2))
    print("aniso 2048^2: n", n, flush=True)
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    B.set_precision(bc.PREC_FP32)
    bh = [ops.uniform_pm1(j, n) for j in range(32)]
    bs = [bc.DeviceArray.from_numpy(ctx, b) for b in bh]
    xs = [bc.DeviceArray(ctx, n) for _ in range(32)]
    its, res, st = B.fcg_multi(bs, xs)
    assert list(st) == [0] * 32
    for j in range(32):
        assert rel_res(S, bh[j], xs[j].numpy()) <= DROPIN["rtol"] * (1 + 1e-6), j
    for j in (0, 17, 31):
        x1 = bc.DeviceArray(ctx, n)
        it1, r1 = B.fcg(bs[j], x1)
        assert its[j] == it1 and res[j] == r1
        assert np.array_equal(bits(xs[j].numpy()), bits(x1.numpy()))


def test_lognormal128_bootstrap_precision(ctx):
    A = ops.lognormal3d(128, 4.0, seed=0)
    n = len(A[0]) - 1
    S = ops.to_scipy(A)
    b = ops.uniform_pm1(0, n)
    settings = dict(DROPIN, max_hrc=3, conv_ratio=0.5)
    prm = bc.make_params(**settings)
    print("lognormal 128^3: n", n, flush=True)
    dA = bc.Csr(ctx, A)
    db = bc.DeviceArray.from_numpy(ctx, b)
    out = {}
    for prec in (bc.PREC_FP64, bc.PREC_FP32, bc.PREC_BF16):
        B = bc.Boot(ctx, dA, prm)
        B.set_precision(prec)
        dx = bc.DeviceArray(ctx, n)
        it, _ = B.fcg(db, dx)
        x = dx.numpy()
        assert rel_res(S, b, x) <= DROPIN["rtol"] * (1 + 1e-6), prec
        out[prec] = (it, x)
        if prec == bc.PREC_FP64:
            B64 = B
        else:
            B.free()
    print("oracle: bootstrap of 3 hierarchies and FCG (about 100 s)", flush=True)
    O = orc.Boot(A, orc.params(**settings))
    assert O.info()[0] == B64.info()[0] == 3
    for k in range(2):                  # built from w = 1: bit-exact
        compare_levels(B64, O, k)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    it64, x64 = out[bc.PREC_FP64]
    check_fcg_vs_oracle(x64, it64, xo, res_o.iterations, A, b, DROPIN["rtol"],
                        "lognormal 128^3 FP64")
    for prec in (bc.PREC_FP32, bc.PREC_BF16):
        it, x = out[prec]
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        print("lognormal 128^3 precision %d: %d iterations (oracle %d), relative error "
              "against the FP64 oracle %.2e" % (prec, it, res_o.iterations, err), flush=True)
        assert err <= 1e-5, prec
        assert it <= 1.25 * res_o.iterations + 2, (prec, it, res_o.iterations)
