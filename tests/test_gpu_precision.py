"""The reduced-precision preconditioner (FP32 hierarchy, or bfloat16 operator
values on levels >= 1) inside FP64 FCG, against the FP64 oracle.

The hierarchy is built in FP64 (bit-exact with the oracle, checked in
test_gpu_parity) and converted, so what is compared here is the effect of
the precision on the solve:
- one preconditioner application vs the oracle's FP64 application:
  relative 2-norm difference <= PREC_TOL (FP32: ~170 u32, u32 = 2^-24;
  BF16: ~13 u_bf16, u_bf16 = 2^-8);
- the FCG solution: true relative residual ||b - A x|| / ||b|| <= rtol (the
  FP64 true-residual check enforces it), and relative error against the
  oracle's FP64 solution <= SOL_TOL = 1e-5 (both solves stop at rtol 1e-8
  on operators with condition numbers below ~1e3);
- FCG iterations: FP32 within +-2 of the oracle's FP64 count, BF16 within
  +25% + 2 (its coarse-level values carry 8-bit mantissas).
"""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import OPERATORS, SETTINGS, bits

pytestmark = pytest.mark.gpu

PREC_TOL = {bc.PREC_FP32: 1e-5, bc.PREC_BF16: 5e-2}
SOL_TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


@pytest.mark.parametrize("prec", [bc.PREC_FP32, bc.PREC_BF16])
@pytest.mark.parametrize("name,setting", [
    ("poisson3d_17", "V_L1_2sweeps"),
    ("lognormal3d_14", "boot3_mult"),
    ("aniso2d_48", "W_L1_3sweeps"),
])
def test_precond_reduced_precision(ctx, name, setting, prec):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm = SETTINGS[setting]
    O = orc.Boot(A, orc.params(**prm))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    B.set_precision(prec)
    r = ops.uniform_pm1(4, n)
    z = dev(ctx, np.zeros(n))
    B.precond(dev(ctx, r), z)
    zo = O.precond(r, np.zeros(n))
    err = np.linalg.norm(z.numpy() - zo) / np.linalg.norm(zo)
    print("%s %s prec=%d precond rel err %.3e" % (name, setting, prec, err))
    assert err <= PREC_TOL[prec], err


@pytest.mark.parametrize("prec", [bc.PREC_FP32, bc.PREC_BF16])
@pytest.mark.parametrize("name,setting", [
    ("poisson3d_17", "V_L1_2sweeps"),
    ("lognormal3d_14", "boot3_mult"),
    ("aniso2d_48", "W_L1_3sweeps"),
    ("poisson2d_33", "H_jacobi_1sweep"),
])
def test_fcg_reduced_precision(ctx, name, setting, prec):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm = SETTINGS[setting]
    O = orc.Boot(A, orc.params(**prm))
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    B.set_precision(prec)
    dx = bc.DeviceArray(ctx, n)
    it, resid = B.fcg(dev(ctx, b), dx)
    x = dx.numpy()
    S = ops.to_scipy(A)
    rtol = prm["rtol"]
    true_rel = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("%s %s prec=%d it=%d (fp64 oracle %d) restarts=%d true_rel=%.3e err=%.3e"
          % (name, setting, prec, it, res_o.iterations, B.restarts(), true_rel, err))
    assert true_rel <= rtol * (1 + 1e-6)
    assert err <= SOL_TOL
    if prec == bc.PREC_FP32:
        assert abs(it - res_o.iterations) <= 2, (it, res_o.iterations)
    else:
        assert it <= 1.25 * res_o.iterations + 2, (it, res_o.iterations)


def test_precision_switch_back_is_drop_in(ctx):
    """After FP32 and BF16 solves, BCMG_PREC_FP64 gives the drop-in's bits."""
    A = OPERATORS["lognormal3d_14"]()
    n = len(A[0]) - 1
    prm = bc.make_params(**SETTINGS["V_L1_2sweeps"])
    b = dev(ctx, ops.uniform_pm1(5, n))
    B0 = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    x0 = bc.DeviceArray(ctx, n)
    it0, r0 = B0.fcg(b, x0)
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    x = bc.DeviceArray(ctx, n)
    for p in (bc.PREC_FP32, bc.PREC_BF16, bc.PREC_FP64):
        B.set_precision(p)
        it, r = B.fcg(b, x)
    assert it == it0 and r == r0
    assert np.array_equal(bits(x.numpy()), bits(x0.numpy()))
    assert B.restarts() == 0
    with pytest.raises(bc.BcmgError) as e:
        B.set_precision(7)
    assert e.value.code == bc.EINVAL


def test_reduced_precision_graph_replay(ctx):
    """Captured FCG iterations replay the FP32 cycle with the same bits as
    host launches."""
    A = OPERATORS["poisson3d_17"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    B.set_precision(bc.PREC_FP32)
    b = dev(ctx, ops.uniform_pm1(6, n))
    bc.use_graphs(ctx, False)
    xe = bc.DeviceArray(ctx, n)
    ite, re_ = B.fcg(b, xe)
    bc.use_graphs(ctx, True)
    for _ in range(2):
        xg = bc.DeviceArray(ctx, n)
        itg, rg = B.fcg(b, xg)
        assert itg == ite and rg == re_
        assert np.array_equal(bits(xg.numpy()), bits(xe.numpy()))


def test_true_residual_restart(ctx):
    """rtol below what FP64 reaches on a high-contrast operator (cell
    coefficients exp(g), g ~ N(0, 25)): the recursive residual passes the
    test, the true residual does not, so FCG restarts from b - A x until
    itnlim. The iteration count stays bounded by itnlim + 1 (the reference's
    count) and the reported residual is the true one."""
    A = ops.lognormal3d(32, 25.0, seed=7)
    n = len(A[0]) - 1
    itnlim = 120
    prm = bc.make_params(**dict(SETTINGS["V_L1_2sweeps"], rtol=1e-15, itnlim=itnlim))
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    B.set_precision(bc.PREC_FP32)
    b = ops.uniform_pm1(0, n)
    dx = bc.DeviceArray(ctx, n)
    it, resid = B.fcg(dev(ctx, b), dx)
    x = dx.numpy()
    true_abs = np.linalg.norm(b - ops.to_scipy(A) @ x)
    print("restarts=%d it=%d resid=%.3e numpy true=%.3e" % (B.restarts(), it, resid, true_abs))
    assert B.restarts() >= 1
    assert it <= itnlim + 1
    assert abs(resid - true_abs) <= 1e-2 * true_abs
