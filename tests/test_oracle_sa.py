"""Pins of the oracle's smoothed aggregation (aggrtype 1, CPU only).

The reference codes smoothed aggregation (src/AMG/matchingAggregation.cu:
353-390) but cannot run it: CSRm::infinityNorm asserts (src/matrix/CSR.cu:
402-404). The restatement takes the norm the reference left commented out
there (:405-416: the largest sum of |a_ij| over a row) and is pinned by:
- its sparse product against scipy (values) and a product of absolute values
  (the symbolic pattern, which cancellation cannot thin), plus cases where
  only the stated summation order gives the asserted bits;
- every level of a built hierarchy against a scipy restatement of
  P = (I - omega D^-1 A) P_tent, R = P^T, A_c = (R A) P with
  omega = 4 / (3 ||D^-1 A||_inf), and the closed form omega = 2/3 of the
  Poisson stencils (||D^-1 A||_inf = 2 in every interior row);
- the aggregates and level sizes, which are the unsmoothed coarsening's;
- one cycle against a dense restatement, and FCG against a direct solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from bootcmatchg_amd import operators as ops
from oracle import oracle as orc

from common import bits


def csr_of(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))


def to_sp(t, m):
    row, col, val = t
    return sp.csr_matrix((val, col, row), shape=(len(row) - 1, m))


def pattern(M):
    """Structure of a CSR matrix as a set of (row, col) pairs."""
    M = sp.csr_matrix(M)
    r = np.repeat(np.arange(M.shape[0]), np.diff(M.indptr))
    return set(zip(r.tolist(), M.indices.tolist()))


def rand_csr(rng, n, m, density):
    M = sp.random(n, m, density=density, random_state=rng, format="csr")
    M.data = rng.uniform(-1, 1, M.nnz)
    return M


# ------------------------------------------------------------- product

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spgemm_vs_scipy(seed):
    rng = np.random.default_rng(seed)
    L = rand_csr(rng, 57, 40, 0.12)
    R = rand_csr(rng, 40, 33, 0.15)
    crow, ccol, cval = orc.spgemm(csr_of(L), csr_of(R), 33)
    C = to_sp((crow, ccol, cval), 33)
    ref = (L @ R).toarray()
    assert np.allclose(C.toarray(), ref, rtol=1e-13, atol=1e-15)
    # symbolic pattern: the product of absolute values cannot cancel
    assert pattern(C) == pattern(abs(L) @ abs(R))
    for i in range(C.shape[0]):
        assert np.all(np.diff(ccol[crow[i]:crow[i + 1]]) > 0)


def test_spgemm_empty_rows_and_rectangular():
    L = sp.csr_matrix(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 0.0]]))
    R = sp.csr_matrix(np.array([[0.0, 2.0, 0.0, 0.0, 3.0], [0.0, 0.0, 0.0, 0.0, 0.0]]))
    crow, ccol, cval = orc.spgemm(csr_of(L), csr_of(R), 5)
    assert crow.tolist() == [0, 0, 2, 2]
    assert ccol.tolist() == [1, 4] and cval.tolist() == [2.0, 3.0]


def test_spgemm_keeps_cancelled_entries():
    """C(0,0) = 1 * 1 + 1 * (-1) = 0 stays in the pattern."""
    L = csr_of(np.array([[1.0, 1.0]]))
    R = csr_of(np.array([[1.0], [-1.0]]))
    crow, ccol, cval = orc.spgemm(L, R, 1)
    assert crow.tolist() == [0, 1] and ccol.tolist() == [0] and cval[0] == 0.0


def test_spgemm_summation_order():
    """(0 + 1e16) + 1 = 1e16 in double precision, so the left-to-right order
    over the left row's entries gives 0 where another order gives 1."""
    L = csr_of(np.array([[1.0, 1.0, 1.0]]))
    R = csr_of(np.array([[1e16], [1.0], [-1e16]]))
    _, _, cval = orc.spgemm(L, R, 1)
    assert cval[0] == 0.0
    L2 = csr_of(np.array([[1.0, 1.0, 1.0]]))
    R2 = csr_of(np.array([[1.0], [1e16], [-1e16]]))
    _, _, cval2 = orc.spgemm(L2, R2, 1)
    assert cval2[0] == 0.0                  # (0 + 1) + 1e16 = 1e16, - 1e16
    # within one right row: entries in storage order (duplicate columns)
    Lr = (np.array([0, 1], np.int32), np.array([0], np.int32), np.array([1.0]))
    Rr = (np.array([0, 3], np.int32), np.array([0, 0, 0], np.int32),
          np.array([1e16, 1.0, -1e16]))
    _, ccol, cval = orc.spgemm(Lr, Rr, 1)
    assert ccol.tolist() == [0] and cval[0] == 0.0
    Rr2 = (Rr[0], Rr[1], np.array([1e16, -1e16, 1.0]))
    _, _, cval = orc.spgemm(Lr, Rr2, 1)
    assert cval[0] == 1.0


# ------------------------------------------------------------ hierarchy

SA = dict(aggrtype=1, prerelax_sweeps=2, postrelax_sweeps=2, relaxnumber_coarse=16)


def check_levels(B, A, relax_type):
    nl = B.levels(0)
    assert nl >= 3, nl
    As = [to_sp(B.level_A(0, l), len(B.level_A(0, l)[0]) - 1) for l in range(nl)]
    assert np.array_equal(As[0].toarray(), ops.to_scipy(A).toarray())
    for l in range(nl - 1):
        Al = As[l]
        n, nc = Al.shape[0], As[l + 1].shape[0]
        agg, p = B.level_P(0, l)
        Pt = sp.csr_matrix((p, (np.arange(n), agg)), shape=(n, nc))
        d = Al.diagonal()
        Dl, _ = B.level_DM(0, l)
        assert np.array_equal(bits(Dl), bits(d))
        S = sp.diags(1.0 / d) @ Al
        omega = 4.0 / (3.0 * abs(S).sum(axis=1).max())
        row, col, val, smoothed, om = B.level_Pcsr(0, l)
        assert smoothed
        assert abs(om - omega) <= 1e-14 * omega
        P = to_sp((row, col, val), nc)
        Sm = sp.identity(n) - omega * S
        Pref = Sm @ Pt
        assert pattern(P) == pattern(abs(Sm) @ abs(Pt) + abs(Pt))
        assert abs(P - Pref).max() <= 1e-14 * abs(Pref).max()
        Acref = (Pref.T @ Al) @ Pref
        assert pattern(As[l + 1]) == pattern(abs(Pref).T @ abs(Al) @ abs(Pref))
        err = abs(As[l + 1] - Acref).max()
        assert err <= 1e-12 * abs(Acref).max(), (l, err)
        # symmetric operator, symmetric product
        assert abs(As[l + 1] - As[l + 1].T).max() <= 1e-12 * abs(Acref).max()
    return As


@pytest.mark.parametrize("name,sweeps", [("poisson2d", 2), ("poisson3d", 1), ("lognormal", 1),
                                         ("aniso", 2), ("lognormal_jacobi", 1)])
def test_smoothed_levels_vs_scipy(name, sweeps):
    A = {"poisson2d": lambda: ops.poisson2d(64), "poisson3d": lambda: ops.poisson3d(14),
         "lognormal": lambda: ops.lognormal3d(12, 4.0, seed=5),
         "lognormal_jacobi": lambda: ops.lognormal3d(12, 4.0, seed=5),
         "aniso": lambda: ops.aniso2d(64, 0.01)}[name]()
    relax = 0 if name.endswith("jacobi") else 4
    B = orc.Boot(A, orc.params(**dict(SA, aggrsweeps=sweeps, relax_type=relax,
                                      coarse_solver=relax)))
    check_levels(B, A, relax)


@pytest.mark.parametrize("A", [ops.poisson2d(30), ops.poisson3d(10)])
def test_omega_closed_form_poisson(A):
    """Interior rows of the 5- and 7-point Laplacians: |a_ii| = sum of the
    off-diagonal |a_ij|, so ||D^-1 A||_inf = 2 and omega = 4 / 6."""
    B = orc.Boot(A, orc.params(**SA))
    assert B.level_Pcsr(0, 0)[4] == 4.0 / 6.0


def test_aggregates_and_sizes_are_the_unsmoothed_ones():
    A = ops.lognormal3d(12, 4.0, seed=5)
    Bs = orc.Boot(A, orc.params(**SA))
    Bu = orc.Boot(A, orc.params(**dict(SA, aggrtype=0)))
    assert Bs.levels(0) == Bu.levels(0)
    for l in range(Bs.levels(0)):
        assert Bs.level_A(0, l)[0].shape == Bu.level_A(0, l)[0].shape
    agg_s, p_s = Bs.level_P(0, 0)
    agg_u, p_u = Bu.level_P(0, 0)
    assert np.array_equal(agg_s, agg_u) and np.array_equal(bits(p_s), bits(p_u))
    # the unsmoothed composite reports the tentative prolongator as CSR
    row, col, val, smoothed, _ = Bu.level_Pcsr(0, 0)
    assert not smoothed and np.array_equal(col, agg_u) and np.array_equal(row, np.arange(len(row)))


def dense_sa_cycle(B, r, pre=2, post=2, coarse=16):
    nl = B.levels(0)
    As = [ops.to_scipy(B.level_A(0, l)).toarray() for l in range(nl)]
    Ps = []
    for l in range(nl - 1):
        row, col, val, _, _ = B.level_Pcsr(0, l)
        Ps.append(to_sp((row, col, val), As[l + 1].shape[0]).toarray())
    Ds = [B.level_DM(0, l)[1] for l in range(nl)]

    def relax(l, f, x, k):
        for _ in range(k):
            x = x + (f - As[l] @ x) / Ds[l]
        return x

    def cyc(l, f, x):
        if l == nl - 1:
            return relax(l, f, x, coarse)
        x = relax(l, f, x, pre)
        fc = Ps[l].T @ (f - As[l] @ x)
        xc = cyc(l + 1, fc, np.zeros(len(fc)))
        x = x + Ps[l] @ xc
        return relax(l, f, x, post)
    return cyc(0, r, np.zeros(len(r)))


def test_smoothed_cycle_matches_dense_restatement():
    A = ops.lognormal3d(12, 4.0, seed=9)
    n = len(A[0]) - 1
    B = orc.Boot(A, orc.params(**dict(SA, aggrsweeps=1)))
    assert B.levels(0) >= 3
    r = ops.uniform_pm1(2, n)
    z = B.precond(r)
    zd = dense_sa_cycle(B, r)
    assert np.linalg.norm(z - zd) <= 1e-12 * np.linalg.norm(zd)
    rng = np.random.default_rng(1)
    u, v = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert abs(u @ B.precond(v) - v @ B.precond(u)) <= 1e-12 * abs(u @ B.precond(v))


@pytest.mark.parametrize("name", ["poisson", "lognormal", "aniso"])
def test_smoothed_fcg_vs_direct(name):
    A = {"poisson": lambda: ops.poisson3d(14), "lognormal": lambda: ops.lognormal3d(12, 4.0, 1),
         "aniso": lambda: ops.aniso2d(48, 0.01)}[name]()
    S = ops.to_scipy(A).tocsc()
    b = ops.uniform_pm1(4, S.shape[0])
    rc, x, res = orc.solve(A, b, orc.params(**dict(SA, rtol=1e-8)))
    assert rc == 0
    xd = spla.spsolve(S, b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    assert np.linalg.norm(b - S @ x) <= 1.01e-8 * np.linalg.norm(b)
    rcu, _, resu = orc.solve(A, b, orc.params(**dict(SA, aggrtype=0, rtol=1e-8)))
    assert rcu == 0
    print(name, "SA", res.iterations, "unsmoothed", resu.iterations)
    assert res.iterations <= resu.iterations


def test_smoothed_bootstrap_runs_composites():
    A = ops.lognormal3d(10, 4.0, seed=3)
    b = ops.uniform_pm1(1, len(A[0]) - 1)
    prm = orc.params(**dict(SA, max_hrc=3, conv_ratio=0.3, solver_type=1, rtol=1e-8))
    rc, x, res = orc.solve(A, b, prm)
    assert rc == 0 and res.n_hrc >= 2
    assert np.linalg.norm(b - ops.to_scipy(A) @ x) <= 1.01e-8 * np.linalg.norm(b)


def test_aggr_type_1_in_a_settings_file(tmp_path):
    """The reference's parser takes aggr_type 1 (src/AMG/AMG.cu:480-526)."""
    f = tmp_path / "s.txt"
    f.write_text("\n".join(["0", "1", "0.8", "2", "1", "39", "0", "4", "4", "16", "4", "4",
                            "2000", "1e-8"]) + "\n")
    rc, p = orc.params_from_file(str(f))
    assert rc == 0 and p.aggrtype == 1
