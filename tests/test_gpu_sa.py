"""Smoothed aggregation (aggrtype 1) on the GPU against the oracle.

Setup results are bit-exact: the general sparse product (pattern, column
order and every value: each entry sums its products in the order of the left
row, then the right row), omega, the smoothed prolongators, R = P^T and every
coarse operator of hierarchies built from w = 1. The solve is compared with
the tolerances of the unsmoothed path: one preconditioner application within
1e-12 (relative 2-norm; the kernels fuse multiply-adds, the oracle does not),
FCG iterations within +-2 and the solution within 1e-12 of the oracle's when
the counts agree (else the true residual at rtol). Multi-column solves equal
single solves bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import OPERATORS, SETTINGS, bits, check_fcg_vs_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


def triple(M):
    M = sp.csr_matrix(M)
    return (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))


def rnd(rng, n, m, d):
    M = sp.random(n, m, density=d, random_state=rng, format="csr")
    M.data = rng.uniform(-1, 1, M.nnz)
    return M


def spgemm_cases():
    rng = np.random.default_rng(0)
    cases = {
        "short_rows": (rnd(rng, 300, 200, 0.03), rnd(rng, 200, 150, 0.05), 150),
        "long_rows": (rnd(rng, 50, 400, 0.5), rnd(rng, 400, 100, 0.3), 100),
    }
    # a right row longer than a chunk of a long row
    Rb = sp.csr_matrix(np.vstack([rng.uniform(-1, 1, (1, 3000)),
                                  rnd(rng, 5, 3000, 0.2).toarray()]))
    Lb = sp.csr_matrix(np.array([[1.0, 0.5, 0, 0, 0, 0], [0, 0, 1, 1, 1, 1],
                                 [2.0, 0, 0, 0, 0, 0]]))
    cases["chunk_spanning"] = (Lb, Rb, 3000)
    # long rows interleaved with short ones; rows of exactly 1024 products
    Rm = rnd(rng, 600, 900, 0.02)
    Lm = rnd(rng, 400, 600, 0.01).tolil()
    for i in range(0, 400, 37):
        Lm[i, :] = rng.uniform(-1, 1, 600) * (rng.uniform(0, 1, 600) < 0.5)
    cases["mixed"] = (Lm.tocsr(), Rm, 900)
    # long rows touching more columns than one LDS sort holds
    cases["wide_rows"] = (rnd(rng, 6, 300, 0.9), rnd(rng, 300, 5000, 0.4), 5000)
    cases["at_capacity"] = (sp.csr_matrix(np.ones((700, 2))),
                            sp.csr_matrix(rng.uniform(-1, 1, (2, 512))), 512)
    cases["empty_rows"] = (sp.csr_matrix(np.array([[0.0] * 4, [1, 1, 0, 0], [0] * 4])),
                           sp.csr_matrix(np.hstack([rng.uniform(-1, 1, (4, 5)),
                                                    np.zeros((4, 3))])), 8)
    return cases


CASES = spgemm_cases()


@pytest.mark.parametrize("name", list(CASES))
def test_spgemm_bitexact(ctx, name):
    L, R, m = CASES[name]
    Lt, Rt = triple(L), triple(R)
    print("spgemm case", name, L.shape, R.shape, flush=True)
    C = bc.spgemm(ctx, Lt, Rt, m)
    row, col, val = C.download()
    orow, ocol, oval = orc.spgemm(Lt, Rt, m)
    assert np.array_equal(row, orow)
    assert np.array_equal(col, ocol)
    assert np.array_equal(bits(val), bits(oval))


@pytest.mark.parametrize("name,per_pass", [("mixed", 5000), ("short_rows", 700),
                                           ("chunk_spanning", 1)])
def test_spgemm_row_batches(ctx, name, per_pass):
    """Rows taken in batches of a few thousand slots (long rows alone) give
    the same bits as one pass."""
    L, R, m = CASES[name]
    Lt, Rt = triple(L), triple(R)
    print("spgemm batches", name, per_pass, flush=True)
    row, col, val = bc.spgemm(ctx, Lt, Rt, m, slots_per_pass=per_pass).download()
    orow, ocol, oval = orc.spgemm(Lt, Rt, m)
    assert np.array_equal(row, orow)
    assert np.array_equal(col, ocol)
    assert np.array_equal(bits(val), bits(oval))


def test_spgemm_summation_order_and_duplicates(ctx):
    """(0 + 1e16) + 1 = 1e16: only the stated order gives these bits; a right
    row may repeat a column (entries in storage order)."""
    Lt = (np.array([0, 3], np.int32), np.array([0, 1, 2], np.int32), np.ones(3))
    Rt = (np.array([0, 1, 2, 3], np.int32), np.zeros(3, np.int32),
          np.array([1e16, 1.0, -1e16]))
    _, col, val = bc.spgemm(ctx, Lt, Rt, 1).download()
    assert col.tolist() == [0] and val[0] == 0.0
    Lr = (np.array([0, 1], np.int32), np.array([0], np.int32), np.array([1.0]))
    Rr = (np.array([0, 3], np.int32), np.zeros(3, np.int32), np.array([1e16, -1e16, 1.0]))
    _, col, val = bc.spgemm(ctx, Lr, Rr, 1).download()
    assert col.tolist() == [0] and val[0] == 1.0


def test_spgemm_rejects_bad_indices(ctx):
    Lt = (np.array([0, 1], np.int32), np.array([3], np.int32), np.ones(1))
    Rt = (np.array([0, 1, 1], np.int32), np.array([0], np.int32), np.ones(1))
    with pytest.raises(bc.BcmgError) as e:
        bc.spgemm(ctx, Lt, Rt, 1)
    assert e.value.code == bc.EINVAL


SA = dict(aggrtype=1)


def compare_sa_hierarchy(B, O, k):
    L = B.levels(k)
    assert L == O.levels(k)
    for l in range(L):
        Ag, Ao = B.level_A(k, l), O.level_A(k, l)
        assert np.array_equal(Ag[0], Ao[0]) and np.array_equal(Ag[1], Ao[1]), (k, l)
        assert np.array_equal(bits(Ag[2]), bits(Ao[2])), (k, l)
        Dg, Mg = B.level_DM(k, l)
        Do, Mo = O.level_DM(k, l)
        assert (Dg is None) == (Do is None) and (Mg is None) == (Mo is None)
        if Do is not None:
            assert np.array_equal(bits(Dg), bits(Do)), (k, l)
        if Mo is not None:
            assert np.array_equal(bits(Mg), bits(Mo)), (k, l)
        if l + 1 < L:
            ag, pg = B.level_P(k, l)
            ao, po = O.level_P(k, l)
            assert np.array_equal(ag, ao) and np.array_equal(bits(pg), bits(po)), (k, l)
            rg, cg, vg, sg, og = B.level_Pcsr(k, l)
            ro, co, vo, so, oo = O.level_Pcsr(k, l)
            assert sg and so and og == oo, (k, l, og, oo)
            assert np.array_equal(rg, ro) and np.array_equal(cg, co), (k, l)
            assert np.array_equal(bits(vg), bits(vo)), (k, l)


@pytest.mark.parametrize("name,setting", [
    ("poisson3d_17", "V_L1_2sweeps"),
    ("aniso2d_48", "W_L1_3sweeps"),
    ("lognormal3d_14", "V_L1_2sweeps"),
    ("poisson2d_33", "H_jacobi_1sweep"),
    ("lognormal3d_14", "K_L1_1sweep_agg"),
])
def test_sa_hierarchy_and_fcg(ctx, name, setting):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm = dict(SETTINGS[setting], **SA)
    print("sa", name, setting, flush=True)
    O = orc.Boot(A, orc.params(**prm))
    assert O.rc == 0
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    compare_sa_hierarchy(B, O, 0)
    r = ops.uniform_pm1(2, n)
    z = dev(ctx, np.zeros(n))
    B.precond(dev(ctx, r), z)
    zo = O.precond(r, np.zeros(n))
    assert np.linalg.norm(z.numpy() - zo) <= 1e-12 * np.linalg.norm(zo)
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dx = bc.DeviceArray(ctx, n)
    it, resid = B.fcg(dev(ctx, b), dx)
    x = dx.numpy()
    check_fcg_vs_oracle(x, it, xo, res_o.iterations, A, b, prm["rtol"], "sa " + name)


@pytest.mark.parametrize("name,setting", [
    ("poisson3d_12", "boot3_mult"),
    ("lognormal3d_10", "boot3_sym"),
])
def test_sa_bootstrap_composites(ctx, name, setting):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm = dict(SETTINGS[setting], **SA)
    O = orc.Boot(A, orc.params(**prm))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    nh_o, er_o = O.info()
    nh, er = B.info()
    assert nh == nh_o
    print("sa %s %s: estimated ratio %.17g (oracle %.17g)" % (name, setting, er, er_o))
    assert abs(er - er_o) <= 1e-12 * abs(er_o), (er, er_o)
    for k in range(min(nh, 2)):
        compare_sa_hierarchy(B, O, k)
    b = ops.uniform_pm1(1, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dx = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(dev(ctx, b), dx)
    check_fcg_vs_oracle(dx.numpy(), it, xo, res_o.iterations, A, b, prm["rtol"],
                        "sa " + name + " " + setting)


def test_sa_larger_operator_long_coarse_rows(ctx):
    """A 3-D operator big enough for coarse rows whose products exceed one
    packed workgroup (the long-row path of the product)."""
    A = ops.poisson3d(40)
    prm = dict(SETTINGS["V_L1_2sweeps"], **SA)
    O = orc.Boot(A, orc.params(**prm))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    compare_sa_hierarchy(B, O, 0)


@pytest.mark.parametrize("prec", [bc.PREC_FP32, bc.PREC_BF16])
def test_sa_reduced_precision(ctx, prec):
    A = OPERATORS["lognormal3d_14"]()
    n = len(A[0]) - 1
    prm = dict(SETTINGS["V_L1_2sweeps"], **SA)
    O = orc.Boot(A, orc.params(**prm))
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**prm))
    B.set_precision(prec)
    dx = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(dev(ctx, b), dx)
    x = dx.numpy()
    assert np.linalg.norm(b - ops.to_scipy(A) @ x) <= prm["rtol"] * (1 + 1e-6) * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)
    assert it <= 1.25 * res_o.iterations + 2, (it, res_o.iterations)


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_sa_multi_equals_single(ctx, prec):
    A = OPERATORS["poisson3d_17"]()
    n = len(A[0]) - 1
    prm = bc.make_params(**dict(SETTINGS["V_L1_2sweeps"], **SA))
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    B.set_precision(prec)
    bs = [dev(ctx, ops.uniform_pm1(20 + j, n)) for j in range(5)]
    xs = [bc.DeviceArray(ctx, n) for _ in range(5)]
    its, res, st = B.fcg_multi(bs, xs)
    for j in range(5):
        x1 = bc.DeviceArray(ctx, n)
        it1, r1 = B.fcg(bs[j], x1)
        assert its[j] == it1 and res[j] == r1, (j, its[j], it1)
        assert np.array_equal(bits(xs[j].numpy()), bits(x1.numpy()))


def test_sa_solve_and_settings_file(ctx, tmp_path):
    A = ops.poisson3d(14)
    n = len(A[0]) - 1
    f = tmp_path / "s.txt"
    f.write_text("\n".join(["0", "1", "0.8", "2", "1", "39", "0", "4", "4", "16", "4", "4",
                            "2000", "1e-8"]) + "\n")
    prm = bc.params_from_file(str(f))
    assert prm.aggrtype == 1
    x, info = bc.solve(A, prm, np.ones(n))
    rc, xo, res = orc.solve(A, np.ones(n), orc.params_from_file(str(f))[1])
    assert rc == 0
    check_fcg_vs_oracle(x, info["iterations"], xo, res.iterations, A, np.ones(n), 1e-8,
                        "sa solve() poisson 14^3")
