"""Parity of the gfx950 library (through the C-ABI) against the serial oracle.

Integer and setup results (matching weights, matchings, aggregates, coarse
patterns and values, prolongators, smoother divisors) must be bit-exact.
Solve-phase floating point is compared with stated tolerances: the kernels
sum rows in a lane-strided order with fma, the oracle sequentially.
FCG iteration counts must agree within +-2.
"""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import OPERATORS, SETTINGS, bits, check_fcg_vs_oracle

pytestmark = pytest.mark.gpu

# |y_gpu - y_oracle| <= RTOL_ROW * (|A| |x|)_i : two orders of summation
RTOL_ROW = 1e-13


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return bc.DeviceArray.from_numpy(ctx, a)


def absmv(A, x):
    return ops.to_scipy((A[0], A[1], np.abs(A[2]))) @ np.abs(x)


@pytest.mark.parametrize("name", ["poisson3d_17", "aniso2d_48", "lognormal3d_14"])
def test_spmv_residual_relax(ctx, name):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    dA = bc.Csr(ctx, A)
    dx, df, dy = dev(ctx, x), dev(ctx, f), bc.DeviceArray(ctx, n)
    bound = RTOL_ROW * (absmv(A, x) + np.abs(f)) + 1e-300
    bc.spmv(ctx, dA, dx, dy)
    assert np.all(np.abs(dy.numpy() - orc.spmv(A, x)) <= bound)
    bc.residual(ctx, dA, dx, df, dy)
    assert np.all(np.abs(dy.numpy() - (f - orc.spmv(A, x))) <= bound)
    S = ops.to_scipy(A)
    d = np.abs(S).sum(axis=1).A1
    u = dev(ctx, x)
    u2 = bc.DeviceArray(ctx, n)
    bc.relax(ctx, dA, dev(ctx, d), df, u, u2, 4)
    uo = orc.relax(A, d, f, x, 4)
    assert np.max(np.abs(u.numpy() - uo)) <= 1e-12 * max(1.0, np.max(np.abs(uo)))


def test_spmv_long_rows(ctx):
    """Rows longer than a wave (64 lanes) and empty-ish rows: a dense-ish
    random symmetric matrix with a diagonal in every row."""
    rng = np.random.default_rng(5)
    n = 300
    import scipy.sparse as sp
    M = sp.random(n, n, density=0.4, random_state=3, format="csr")
    M = M + M.T + sp.eye(n) * n
    M = M.tocsr()
    M.sort_indices()
    A = (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))
    x = rng.uniform(-1, 1, n)
    dA = bc.Csr(ctx, A)
    dy = bc.DeviceArray(ctx, n)
    bc.spmv(ctx, dA, dev(ctx, x), dy)
    bound = RTOL_ROW * absmv(A, x) + 1e-300
    assert np.all(np.abs(dy.numpy() - orc.spmv(A, x)) <= bound)


def _with_far_zeros(A):
    """A plus explicit zeros at (0, n-1) and (n-1, 0): the same sums up to
    +-0 terms, but offsets no 16-bit column can hold."""
    import scipy.sparse as sp
    n = len(A[0]) - 1
    r = np.concatenate([np.repeat(np.arange(n), np.diff(A[0])), [0, n - 1]])
    c = np.concatenate([A[1], [n - 1, 0]])
    v = np.concatenate([A[2], [0.0, 0.0]])
    S = sp.csr_matrix((v, (r, c)), shape=(n, n))
    S.sort_indices()
    assert S.nnz == len(A[1]) + 2
    return (S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64))


# entries above which an operator's FP64 stream (12 B each) is non-temporal
NARROW_NNZ = (192 << 20) // 12


def test_narrow_columns_bitexact(ctx):
    """A square operator whose FP64 stream exceeds 192 MB and whose column
    offsets fit 16 bits is streamed through its column-minus-row copy: SpMV,
    residual and Jacobi sweeps give the bits of the 32-bit-column kernels
    (the same operator with two far explicit zeros), and stay within the row
    bound of the oracle."""
    A = ops.poisson3d(140)
    n = len(A[0]) - 1
    assert len(A[1]) > NARROW_NNZ
    Aw = _with_far_zeros(A)
    dA, dW = bc.Csr(ctx, A), bc.Csr(ctx, Aw)
    # the stencil pattern is kept too (tests/test_gpu_columns.py); this test
    # is about the 16-bit copy
    assert dA.column_bytes == 0 and dW.column_bytes == 4
    dA.set_columns(2)
    assert dA.column_bytes == 2
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    dx, df = dev(ctx, x), dev(ctx, f)
    y1, y2 = bc.DeviceArray(ctx, n), bc.DeviceArray(ctx, n)
    bc.spmv(ctx, dA, dx, y1)
    bc.spmv(ctx, dW, dx, y2)
    assert np.array_equal(bits(y1.numpy()), bits(y2.numpy()))
    assert np.all(np.abs(y1.numpy() - orc.spmv(A, x)) <= RTOL_ROW * absmv(A, x) + 1e-300)
    bc.residual(ctx, dA, dx, df, y1)
    bc.residual(ctx, dW, dx, df, y2)
    assert np.array_equal(bits(y1.numpy()), bits(y2.numpy()))
    d = np.abs(ops.to_scipy(A)).sum(axis=1).A1
    dd = dev(ctx, d)
    u1, u2 = dev(ctx, x), dev(ctx, x)
    s1, s2 = bc.DeviceArray(ctx, n), bc.DeviceArray(ctx, n)
    bc.relax(ctx, dA, dd, df, u1, s1, 3)
    bc.relax(ctx, dW, dd, df, u2, s2, 3)
    assert np.array_equal(bits(u1.numpy()), bits(u2.numpy()))


@pytest.mark.parametrize("off,narrow", [(32767, True), (-32768, True), (32768, False),
                                        (-32769, False)])
def test_narrow_columns_offset_limits(ctx, off, narrow):
    """The 16-bit copy is kept exactly when every column - row offset lies
    in [-32768, 32767]: a banded operator above the non-temporal size with
    one extra diagonal at the given offset."""
    import scipy.sparse as sp
    n = 1 << 21
    band = sp.diags([-np.ones(n - 1), 4.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    far = sp.diags([-0.5 * np.ones(n - abs(off))], [off])
    wide = sp.diags([-0.25 * np.ones(n - 2000)] * 2, [2000, -2000])
    S = band + far + wide + 2 * sp.eye(n)
    for k in (3, 5, 7):
        S = S + sp.diags([-0.01 * np.ones(n - k)] * 2, [k, -k])
    S = S.tocsr()
    S.sort_indices()
    A = (S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64))
    assert S.nnz > NARROW_NNZ
    dA = bc.Csr(ctx, A)
    assert dA.column_bytes == (2 if narrow else 4)
    x = np.random.default_rng(12).uniform(-1, 1, n)
    y = bc.DeviceArray(ctx, n)
    bc.spmv(ctx, dA, dev(ctx, x), y)
    assert np.all(np.abs(y.numpy() - orc.spmv(A, x)) <= RTOL_ROW * absmv(A, x) + 1e-300)


@pytest.mark.parametrize("name", ["poisson3d_12", "aniso2d_48", "lognormal3d_14"])
def test_weights_bitexact(ctx, name):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    rng = np.random.default_rng(2)
    for w in (np.ones(n), rng.uniform(0.1, 2.0, n), rng.uniform(-1, 1, n)):
        W = bc.DeviceArray(ctx, len(A[1]))
        bc.weights(ctx, bc.Csr(ctx, A), dev(ctx, w), W)
        assert np.array_equal(bits(W.numpy()), bits(orc.weights(A, w)))


@pytest.mark.parametrize("name", ["poisson3d_12", "poisson2d_33", "lognormal3d_14", "aniso2d_48"])
def test_matching_aggregation_galerkin_bitexact(ctx, name):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    rng = np.random.default_rng(4)
    for w in (np.ones(n), rng.uniform(0.2, 1.5, n)):
        W = orc.weights(A, w)
        Mr_o, Ms_o, nu_o = orc.matching(A, W)
        dA = bc.Csr(ctx, A)
        Mr, Ms = bc.DeviceArray(ctx, n, np.int32), bc.DeviceArray(ctx, n, np.int32)
        nu = bc.matching(ctx, dA, dev(ctx, W), Mr, Ms)
        assert np.array_equal(Mr.numpy(), Mr_o)
        assert np.array_equal(Ms.numpy(), Ms_o)
        assert nu == nu_o
        agg_o, p_o, nc_o = orc.aggregate(Ms_o, w)
        agg, p = bc.DeviceArray(ctx, n, np.int32), bc.DeviceArray(ctx, n)
        nc = bc.aggregate(ctx, n, dev(ctx, Ms_o), dev(ctx, w), agg, p)
        assert nc == nc_o
        assert np.array_equal(agg.numpy(), agg_o)
        assert np.array_equal(bits(p.numpy()), bits(p_o))
        Ac_o = orc.galerkin(A, agg_o, p_o, nc_o)
        Ac = bc.galerkin(ctx, dA, dev(ctx, agg_o), dev(ctx, p_o), nc_o).download()
        assert np.array_equal(Ac[0], Ac_o[0])
        assert np.array_equal(Ac[1], Ac_o[1])
        assert np.array_equal(bits(Ac[2]), bits(Ac_o[2]))


def _compare_hierarchy(B, O, k):
    L = B.levels(k)
    assert L == O.levels(k)
    for l in range(L):
        Ag = B.level_A(k, l)
        Ao = O.level_A(k, l)
        for a, b in zip(Ag[:2], Ao[:2]):
            assert np.array_equal(a, b), (k, l)
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
            assert np.array_equal(ag, ao), (k, l)
            assert np.array_equal(bits(pg), bits(po)), (k, l)


def test_aggregate_rejects_an_invalid_matching(ctx):
    """bcmg_aggregate takes a caller's matching: a partner out of range, a
    row matched to itself or a partner not matched back is BCMG_EINVAL,
    raised before any aggregate is written."""
    n = 8
    w = dev(ctx, np.ones(n))
    good = np.array([1, 0, 3, 2, -1, 6, 5, -1], dtype=np.int32)
    bad = {"out_of_range": [1, 0, 3, 2, -1, 6, 5, 8], "self": [1, 0, 2, -1, -1, 6, 5, -1],
           "not_mutual": [1, 0, 3, 2, 3, 6, 5, -1]}
    agg = bc.DeviceArray(ctx, n, np.int32)
    p = bc.DeviceArray(ctx, n)
    assert bc.aggregate(ctx, n, dev(ctx, good), w, agg, p) == 5
    assert list(agg.numpy()) == [0, 0, 1, 1, 2, 3, 3, 4]
    for name, M in bad.items():
        sentinel = np.full(n, -7, dtype=np.int32)
        agg.upload(sentinel)
        with pytest.raises(bc.BcmgError) as e:
            bc.aggregate(ctx, n, dev(ctx, np.array(M, dtype=np.int32)), w, agg, p)
        assert e.value.code == bc.EINVAL, name
        assert np.array_equal(agg.numpy(), sentinel), name


@pytest.mark.parametrize("name,setting", [
    ("poisson3d_17", "V_L1_2sweeps"),
    ("aniso2d_48", "W_L1_3sweeps"),
    ("lognormal3d_14", "V_L1_2sweeps"),
    ("poisson2d_33", "H_jacobi_1sweep"),
])
def test_hierarchy_and_fcg(ctx, name, setting):
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm_o = orc.params(**SETTINGS[setting])
    prm = bc.make_params(**SETTINGS[setting])
    O = orc.Boot(A, prm_o)
    assert O.rc == 0
    dA = bc.Csr(ctx, A)
    B = bc.Boot(ctx, dA, prm)
    assert B.info()[0] == O.info()[0]
    _compare_hierarchy(B, O, 0)
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    db, dx = dev(ctx, b), bc.DeviceArray(ctx, n)
    it, resid = B.fcg(db, dx)
    x = dx.numpy()
    check_fcg_vs_oracle(x, it, xo, res_o.iterations, A, b, prm.rtol, name + " " + setting)


@pytest.mark.parametrize("name,setting", [
    ("poisson3d_12", "boot3_mult"),
    ("lognormal3d_10", "boot3_sym"),
    ("poisson3d_12", "boot2_add_mixed"),
])
def test_bootstrap_composites(ctx, name, setting):
    """Hierarchies 0 and 1 are built from w = 1 and must be bit-exact; the
    bootstrapped smooth vector (15 composite cycles) carries solve-phase
    rounding, so later hierarchies are compared through iterations."""
    A = OPERATORS[name]()
    n = len(A[0]) - 1
    prm_o = orc.params(**SETTINGS[setting])
    prm = bc.make_params(**SETTINGS[setting])
    O = orc.Boot(A, prm_o)
    B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
    nh_o, er_o = O.info()
    nh, er = B.info()
    assert nh == nh_o
    print("%s %s: estimated ratio %.17g (oracle %.17g)" % (name, setting, er, er_o))
    assert abs(er - er_o) <= 1e-12 * abs(er_o), (er, er_o)
    for k in range(min(nh, 2)):
        _compare_hierarchy(B, O, k)
    b = ops.uniform_pm1(1, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dx = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(dev(ctx, b), dx)
    check_fcg_vs_oracle(dx.numpy(), it, xo, res_o.iterations, A, b, prm.rtol,
                        name + " " + setting)


def test_setup_phase_timing_and_the_shared_second_hierarchy(ctx):
    """With setup timing on, every phase of a three-hierarchy bootstrap is
    timed and the phases fit in the setup's wall time; hierarchy 1 (built
    from w = 1 like hierarchy 0) is not rebuilt: its summary equals hierarchy
    0's and only its inner iterations take time. Without timing the phases
    read 0."""
    # the oracle's bootstrap builds all three hierarchies on this operator
    # (estimated ratio 0.67 > 0.5 after the second)
    A = OPERATORS["aniso2d_48"]()
    dA = bc.Csr(ctx, A)
    prm = bc.make_params(**SETTINGS["boot3_mult"])
    assert orc.Boot(A, orc.params(**SETTINGS["boot3_mult"])).info()[0] == 3
    bc.time_setup(ctx, True)
    try:
        B = bc.Boot(ctx, dA, prm)
    finally:
        bc.time_setup(ctx, False)
    nh, ratio = B.info()
    assert nh == 3
    t = B.setup_timing(-1)
    print("setup phases (ms):", {k: round(v, 3) for k, v in t.items()})
    named = ("weights_ms", "matching_ms", "aggregation_ms", "galerkin_ms", "smoother_ms",
             "inner_ms")
    assert all(t[k] > 0 for k in named), t
    assert sum(t[k] for k in named) <= t["hierarchy_ms"] * 1.01 + 0.05
    assert t["hierarchy_ms"] <= t["total_ms"] * 1.01 + 0.05
    t0, t1, t2 = (B.setup_timing(k) for k in range(3))
    assert t0["inner_ms"] == 0 and t1["inner_ms"] > 0 and t2["inner_ms"] > 0
    assert all(t1[k] == 0 for k in named[:-1]), t1
    assert all(t2[k] > 0 for k in named), t2
    h0, h1 = B.hier_info(0), B.hier_info(1)
    for key in ("levels", "op_cmplx", "op_wcmplx", "avg_cratio"):
        assert h0[key] == h1[key], key
    assert h0["ratio_after"] == 1.0 and B.hier_info(2)["ratio_after"] == ratio
    lv = [B.level_size(0, l) for l in range(h0["levels"])]
    assert abs(h0["op_cmplx"] - sum(z for _, z in lv) / lv[0][1]) <= 1e-15 * h0["op_cmplx"]
    B.free()
    B = bc.Boot(ctx, dA, prm)
    assert all(v == 0 for k, v in B.setup_timing(-1).items() if k != "total_ms")
    B.free()


def test_precond_apply(ctx):
    A = OPERATORS["lognormal3d_10"]()
    n = len(A[0]) - 1
    prm = SETTINGS["boot3_mult"]
    O = orc.Boot(A, orc.params(**prm))
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**dict(prm, max_hrc=1)))
    O1 = orc.Boot(A, orc.params(**dict(prm, max_hrc=1)))
    r = ops.uniform_pm1(2, n)
    x0 = ops.uniform_pm1(3, n)
    dx = dev(ctx, x0)
    B.precond(dev(ctx, r), dx)
    xo = O1.precond(r, x0)
    assert np.linalg.norm(dx.numpy() - xo) <= 1e-12 * np.linalg.norm(xo)
    del O


def test_solve_host_api_and_errors(ctx):
    A = ops.poisson3d(10)
    n = len(A[0]) - 1
    prm = bc.make_params(rtol=1e-6)
    x, info = bc.solve(A, prm, np.ones(n))
    rc, xo, res = orc.solve(A, np.ones(n), orc.params(rtol=1e-6))
    assert rc == 0
    check_fcg_vs_oracle(x, info["iterations"], xo, res.iterations, A, np.ones(n), 1e-6,
                        "solve() poisson 10^3")
    with pytest.raises(bc.BcmgError) as e:
        bc.solve(A, prm, np.zeros(n))
    assert e.value.code == bc.EZERORHS


def test_graph_replay_matches_eager_launches(ctx):
    """FCG replayed from captured graphs gives the same bits as host launches,
    over several solves (graph reuse) and a change of output vector."""
    A = OPERATORS["lognormal3d_14"]()
    n = len(A[0]) - 1
    B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
    b = dev(ctx, ops.uniform_pm1(9, n))
    bc.use_graphs(ctx, False)
    x_e = bc.DeviceArray(ctx, n)
    it_e, r_e = B.fcg(b, x_e)
    bc.use_graphs(ctx, True)
    for rep in range(3):
        x_g = bc.DeviceArray(ctx, n)
        it_g, r_g = B.fcg(b, x_g)
        assert it_g == it_e and r_g == r_e
        assert np.array_equal(bits(x_g.numpy()), bits(x_e.numpy()))


@pytest.mark.parametrize("setting", ["boot3_mult", "boot3_sym", "boot2_add_mixed"])
def test_bootstrap_same_hierarchies_with_graphs_on_and_off(ctx, setting):
    """The bootstrap builds the same hierarchies bit for bit whether the
    context captures graphs or not (its composite applications launch
    directly either way), for the three composite kinds."""
    A = OPERATORS["lognormal3d_14"]()
    dA = bc.Csr(ctx, A)
    prm = bc.make_params(**SETTINGS[setting])
    bc.use_graphs(ctx, False)
    Be = bc.Boot(ctx, dA, prm)
    bc.use_graphs(ctx, True)
    Bg = bc.Boot(ctx, dA, prm)
    assert Bg.info() == Be.info() and Be.info()[0] >= 2
    for k in range(Be.info()[0]):
        assert Bg.levels(k) == Be.levels(k)
        for l in range(Be.levels(k)):
            g, e = Bg.level_A(k, l), Be.level_A(k, l)
            assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1]), (k, l)
            assert np.array_equal(bits(g[2]), bits(e[2])), (k, l)


def permuted_poisson(N, seed):
    """3D Poisson with rows and columns permuted at random: no band, entries
    anywhere in a row."""
    import scipy.sparse as sp
    S = ops.to_scipy(ops.poisson3d(N)).tocsr()
    p = np.random.default_rng(seed).permutation(S.shape[0])
    T = S[p][:, p].tocsr()
    T.sort_indices()
    return (T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data.astype(np.float64))


def test_randomly_ordered_operator(ctx):
    """A randomly permuted operator (columns up to n apart, gathers without
    locality): row kernels, hierarchy bits and FCG against the oracle."""
    A = permuted_poisson(34, 5)
    n = len(A[0]) - 1
    rows = np.repeat(np.arange(n), np.diff(A[0]))
    assert np.abs(A[1] - rows).max() > 32767
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    dA = bc.Csr(ctx, A)
    dx, df, dy = dev(ctx, x), dev(ctx, f), bc.DeviceArray(ctx, n)
    bound = RTOL_ROW * (absmv(A, x) + np.abs(f)) + 1e-300
    bc.spmv(ctx, dA, dx, dy)
    assert np.all(np.abs(dy.numpy() - orc.spmv(A, x)) <= bound)
    bc.residual(ctx, dA, dx, df, dy)
    assert np.all(np.abs(dy.numpy() - (f - orc.spmv(A, x))) <= bound)
    prm = SETTINGS["V_L1_2sweeps"]
    O = orc.Boot(A, orc.params(**prm))
    B = bc.Boot(ctx, dA, bc.make_params(**prm))
    _compare_hierarchy(B, O, 0)
    b = ops.uniform_pm1(0, n)
    rc, xo, res_o, _ = O.fcg(b)
    assert rc == 0
    dxs = bc.DeviceArray(ctx, n)
    it, _ = B.fcg(dev(ctx, b), dxs)
    check_fcg_vs_oracle(dxs.numpy(), it, xo, res_o.iterations, A, b, prm["rtol"], "permuted")


def test_context_on_a_caller_stream(ctx):
    """A context on a stream the caller created gives the same bits as one on
    its own stream, and leaves the stream alive when it closes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1)) == 0
    try:
        A = OPERATORS["lognormal3d_10"]()
        n = len(A[0]) - 1
        b = ops.uniform_pm1(8, n)
        prm = bc.make_params(**SETTINGS["V_L1_2sweeps"])
        B0 = bc.Boot(ctx, bc.Csr(ctx, A), prm)
        x0 = bc.DeviceArray(ctx, n)
        it0, r0 = B0.fcg(dev(ctx, b), x0)
        c2 = bc.Context(0, stream=s.value)
        B = bc.Boot(c2, bc.Csr(c2, A), prm)
        x = bc.DeviceArray(c2, n)
        it, r = B.fcg(bc.DeviceArray.from_numpy(c2, b), x)
        assert it == it0 and r == r0
        assert np.array_equal(bits(x.numpy()), bits(x0.numpy()))
        c2.close()
        assert hip.hipStreamSynchronize(s) == 0
        with pytest.raises(bc.BcmgError) as e:
            bc.Context(0, stream=0)
    finally:
        hip.hipStreamDestroy(s)
