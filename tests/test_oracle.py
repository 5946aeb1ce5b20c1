"""Pins of the serial oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path and
cannot be built here (CUDA only), so the oracle is pinned by what the
mathematics fixes: closed forms, textbook equivalences (the suitor fixed
point is the greedy matching), brute force on tiny inputs, scipy products and
solvers, and a dense restatement of one multigrid cycle. Each check is chosen
so that a plausible slip (a transposed operand, a dropped term, a wrong tie
rule or summation order) fails it.
"""
import itertools
import math

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


def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


# ------------------------------------------------------------------ log

def ulps(a, b):
    return abs(int(np.float64(a).view(np.int64)) - int(np.float64(b).view(np.int64)))


def test_log_within_one_ulp_of_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0.5, 2.0, 20000),
                         np.exp(rng.uniform(-700, 700, 20000)),
                         1.0 + rng.uniform(-1e-6, 1e-6, 5000),
                         [5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]])
    worst = max(ulps(orc.log(x), math.log(x)) for x in xs)
    assert worst <= 1
    assert orc.log(1.0) == 0.0
    assert orc.log(0.0) == -math.inf and math.isnan(orc.log(-1.0)) and orc.log(math.inf) == math.inf
    assert orc.log(2.0) == math.log(2.0)


# ------------------------------------------------------------ operators

def test_splitmix64_known_answers():
    # published first outputs of SplitMix64 from state 0
    z = ops.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = ops.uniform_pm1(5, 100000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    g = ops.normal(3, 200000, sigma=2.0)
    assert abs(g.std() - 2.0) < 0.02 and abs(g.mean()) < 0.02


@pytest.mark.parametrize("N", [1, 2, 5, 8])
def test_poisson3d_closed_form(N):
    row, col, val = ops.poisson3d(N)
    n = N ** 3
    assert len(col) == 7 * n - 6 * N * N
    L = lap1d(N)
    I = sp.identity(N)
    K = sp.kron(sp.kron(I, I), L) + sp.kron(sp.kron(I, L), I) + sp.kron(sp.kron(L, I), I)
    S = ops.to_scipy((row, col, val))
    assert abs(S - K).max() == 0
    assert (S.diagonal() == 6).all()


@pytest.mark.parametrize("N,eps", [(4, 0.01), (9, 1.0), (16, 0.5)])
def test_aniso2d_closed_form(N, eps):
    row, col, val = ops.aniso2d(N, eps)
    n = N * N
    assert len(col) == 5 * n - 4 * N
    L = lap1d(N)
    I = sp.identity(N)
    K = sp.kron(I, L) + eps * sp.kron(L, I)
    assert abs(ops.to_scipy((row, col, val)) - K).max() <= 1e-15


def test_lognormal3d_properties():
    A = ops.lognormal3d(6, 4.0, seed=2)
    S = ops.to_scipy(A)
    assert len(A[1]) == 7 * 216 - 6 * 36
    assert abs(S - S.T).max() == 0                       # bitwise symmetric
    assert (S.diagonal() > 0).all() and (S.data[S.indices != np.repeat(np.arange(216), np.diff(A[0]))] < 0).all()
    rowsum = np.asarray(S.sum(axis=1)).ravel()
    assert (rowsum >= -1e-9 * S.diagonal()).all()        # weakly diagonally dominant
    assert np.linalg.eigvalsh(S.toarray()).min() > 0     # SPD


# --------------------------------------------------------- SpMV, relax

def test_spmv_and_relax_vs_scipy():
    rng = np.random.default_rng(1)
    A = ops.lognormal3d(7, 4.0, seed=1)
    S = ops.to_scipy(A)
    x = rng.uniform(-1, 1, S.shape[0])
    assert np.allclose(orc.spmv(A, x), S @ x, rtol=1e-14, atol=1e-14 * abs(S).max())
    d = S.diagonal()
    f = rng.uniform(-1, 1, S.shape[0])
    u = x.copy()
    for _ in range(3):
        u = u + (f - S @ u) / d
    assert np.allclose(orc.relax(A, d, f, x, 3), u, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------- weights

def test_weights_closed_form_constant_coefficients():
    """w = 1 and a constant-coefficient operator: AH_ij = 1 - 2 a_ij/(a_ii+a_jj)
    is the same for every edge, so m = AH and every W is log(1/0.999)."""
    for A in (ops.poisson3d(5), ops.aniso2d(6, 1.0)):
        n = len(A[0]) - 1
        W = orc.weights(A, np.ones(n))
        diag = np.repeat(np.arange(n), np.diff(A[0])) == A[1]
        assert (W[diag] == 0).all()
        off = W[~diag]
        assert (bits(off) == bits(off[0])).all()
        assert abs(off[0] - math.log(1 / 0.999)) <= 2e-15


def test_weights_anisotropic_two_values():
    """5-point with couplings 1 and eps: AH = 1 + 1/(2+2eps) (x) and
    1 + eps/(2+2eps) (y); m is the y value."""
    eps = 0.01
    A = ops.aniso2d(7, eps)
    n = len(A[0]) - 1
    W = orc.weights(A, np.ones(n))
    ahx = 1 + 2 * 1 / (2 * (2 + 2 * eps))
    ahy = 1 + 2 * eps / (2 * (2 + 2 * eps))
    r = np.repeat(np.arange(n), np.diff(A[0]))
    c = A[1]
    xw = W[np.abs(c - r) == 1]
    yw = W[np.abs(c - r) == 7]
    assert np.allclose(xw, math.log(ahx / (0.999 * ahy)), rtol=1e-13)
    assert np.allclose(yw, math.log(1 / 0.999), rtol=1e-13)


def test_weights_invariances():
    """AH is invariant under w -> 2^k w and A -> 2^k A (exact scalings), and
    W is symmetric when A is."""
    rng = np.random.default_rng(3)
    A = ops.lognormal3d(5, 4.0, seed=5)
    n = len(A[0]) - 1
    w = rng.uniform(-1, 1, n)
    W = orc.weights(A, w)
    assert np.array_equal(bits(W), bits(orc.weights(A, 8.0 * w)))
    assert np.array_equal(bits(W), bits(orc.weights((A[0], A[1], A[2] * 0.25), w)))
    S = sp.csr_matrix((W, A[1], A[0]), shape=(n, n))
    assert abs(S - S.T).max() == 0


# -------------------------------------------------------------- matching

def greedy_matching(n, edges):
    """Greedy over edges sorted by (W, max id, min id) descending; W <= 0 skipped."""
    M = -np.ones(n, dtype=np.int64)
    for w, u, v in sorted(edges, key=lambda e: (e[0], max(e[1], e[2]), min(e[1], e[2])),
                          reverse=True):
        if w > 0 and M[u] < 0 and M[v] < 0:
            M[u], M[v] = v, u
    return M


def random_graph(rng, n, p, values):
    rows = {i: {i: 0.0} for i in range(n)}
    edges = []
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < p:
                w = values(rng)
                rows[u][v] = w
                rows[v][u] = w
                edges.append((w, u, v))
    row = [0]
    col, W = [], []
    for i in range(n):
        for j in sorted(rows[i]):
            col.append(j)
            W.append(rows[i][j])
        row.append(len(col))
    A = (np.array(row, np.int32), np.array(col, np.int32), np.ones(len(col)))
    return A, np.array(W), edges


@pytest.mark.parametrize("values", ["ties", "distinct", "signed"])
def test_suitor_equals_greedy_and_is_order_independent(values):
    rng = np.random.default_rng(11)
    gen = {"ties": lambda r: float(r.integers(1, 4)),
           "distinct": lambda r: float(r.uniform(0.1, 1.0)),
           "signed": lambda r: float(r.integers(-2, 3))}[values]
    for trial in range(60):
        n = int(rng.integers(2, 40))
        A, W, edges = random_graph(rng, n, float(rng.uniform(0.05, 0.5)), gen)
        Mr, Ms, nu = orc.matching(A, W)
        G = greedy_matching(n, edges)
        assert nu == 0
        assert np.array_equal(Mr, G) and np.array_equal(Ms, G)
        for _ in range(3):
            Mr2, _, _ = orc.matching(A, W, order=rng.permutation(n))
            assert np.array_equal(Mr2, Mr)


def test_greedy_is_the_unique_locally_dominant_matching_bruteforce():
    """Exhaustive over all matchings of tiny graphs: the greedy matching is the
    only one in which every edge is dominated by a matched edge of higher key
    at one of its ends (so the tie rule is what pins the result)."""
    rng = np.random.default_rng(12)
    key = lambda w, u, v: (w, max(u, v), min(u, v))
    for trial in range(40):
        n = int(rng.integers(2, 8))
        A, W, edges = random_graph(rng, n, 0.6, lambda r: float(r.integers(1, 3)))
        pos = [e for e in edges if e[0] > 0]
        G = greedy_matching(n, edges)
        found = []
        for k in range(len(pos) + 1):
            for sub in itertools.combinations(pos, k):
                ends = [x for e in sub for x in e[1:]]
                if len(ends) != len(set(ends)):
                    continue
                mk = {}
                for e in sub:
                    mk[e[1]] = key(*e)
                    mk[e[2]] = key(*e)
                ok = all(any(x in mk and mk[x] >= key(*e) for x in e[1:]) for e in pos)
                if ok:
                    M = -np.ones(n, dtype=np.int64)
                    for w, u, v in sub:
                        M[u], M[v] = v, u
                    found.append(M)
        assert len(found) == 1 and np.array_equal(found[0], G)
        assert np.array_equal(orc.matching(A, W)[1], G)


def test_unsym_fix_on_asymmetric_weights():
    """With asymmetric W the suitor array need not be symmetric; the fix keeps
    mutual pairs and pairs v -> u (u < v) in ascending v when both are free."""
    rng = np.random.default_rng(13)
    seen = 0
    for trial in range(300):
        n = int(rng.integers(3, 15))
        A, W, _ = random_graph(rng, n, 0.5, lambda r: 1.0)
        W = W * rng.uniform(0.5, 1.5, len(W))          # asymmetric
        diag = np.repeat(np.arange(n), np.diff(A[0])) == A[1]
        W[diag] = 0
        Mr, Ms, nu = orc.matching(A, W)
        for perm in range(2):
            assert np.array_equal(orc.matching(A, W, order=rng.permutation(n))[0], Mr)
        exp = -np.ones(n, dtype=np.int64)
        for v in range(n):
            u = Mr[v]
            if u >= 0 and Mr[u] == v:
                exp[v] = u
        for v in range(n):
            u = Mr[v]
            if u >= 0 and exp[v] == -1 and u < v and exp[u] == -1:
                exp[u], exp[v] = v, u
        assert np.array_equal(Ms, exp)
        assert all(Ms[Ms[v]] == v for v in range(n) if Ms[v] >= 0)
        seen += nu > 0
    assert seen > 0


# ----------------------------------------------------------- aggregation

def test_aggregation_invariants():
    rng = np.random.default_rng(14)
    for trial in range(30):
        n = int(rng.integers(2, 60))
        A, W, edges = random_graph(rng, n, 0.3, lambda r: float(r.uniform(0.1, 1)))
        w = rng.uniform(-1, 1, n)
        w[rng.random(n) < 0.1] = 0.0                   # bad singles / bad pairs
        _, Ms, _ = orc.matching(A, W)
        agg, p, nc = orc.aggregate(Ms, w)
        pairs = [(v, Ms[v]) for v in range(n)
                 if Ms[v] > v and math.hypot(w[v], w[Ms[v]]) > 2.220446049250313e-16]
        assert nc == n - len(pairs)
        first = [np.flatnonzero(agg == I).min() for I in range(nc)]
        assert first == sorted(first)                    # canonical numbering
        for v, u in pairs:
            assert agg[v] == agg[u]
            nrm = math.sqrt(w[v] ** 2 + w[u] ** 2)
            assert p[v] == w[v] / nrm and p[u] == w[u] / nrm
        for I in range(nc):
            mem = np.flatnonzero(agg == I)
            if len(mem) == 1:
                v = mem[0]
                assert p[v] == (np.sign(w[v]) if abs(w[v]) > 2.220446049250313e-16 else 0.0)
            else:
                assert abs(np.sum(p[mem] ** 2) - 1) < 1e-15   # R P = I on the aggregate


# -------------------------------------------------------------- Galerkin

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_galerkin_vs_scipy(seed):
    rng = np.random.default_rng(seed)
    A = ops.lognormal3d(6, 4.0, seed=seed)
    n = len(A[0]) - 1
    w = rng.uniform(0.5, 1.5, n)
    _, Ms, _ = orc.matching(A, orc.weights(A, w))
    agg, p, nc = orc.aggregate(Ms, w)
    rc, cc, vc = orc.galerkin(A, agg, p, nc)
    P = sp.csr_matrix((p, agg, np.arange(n + 1)), shape=(n, nc))
    S = ops.to_scipy(A)
    C = (P.T @ (S @ P)).tocsr()
    C.sort_indices()
    # pattern: every (I, J) with a structural path, zeros included
    pat = ((abs(P).T @ (abs(S) @ abs(P))) > 0).tocsr()
    pat.sort_indices()
    assert np.array_equal(rc, pat.indptr) and np.array_equal(cc, pat.indices)
    D = sp.csr_matrix((vc, cc, rc), shape=(nc, nc))
    assert abs(D - C).max() <= 1e-13 * abs(C).max()


def test_galerkin_fixed_order():
    """The coarse value is 0 + p_i1 (0 + a t1 + ...) + p_i2 (...), in that order:
    a case where another association rounds differently."""
    # 2 fine rows in one aggregate, 3 in row 0 mapping to the same coarse column
    M = sp.csr_matrix(np.array([[1.0, 1e-16, 1.0], [1e-16, 1.0, 0.0], [1.0, 0.0, 3.0]]))
    A = csr_of(M)
    agg = np.array([0, 0, 1], np.int32)
    p = np.array([1.0, 1.0, 1.0])
    rc, cc, vc = orc.galerkin(A, agg, p, 2)
    ap0 = (0.0 + 1.0 * 1.0) + 1e-16 * 1.0                  # row 0, J = 0
    ap1 = (0.0 + 1e-16 * 1.0) + 1.0 * 1.0                  # row 1, J = 0
    exp00 = (0.0 + 1.0 * ap0) + 1.0 * ap1
    assert vc[0] == exp00


# --------------------------------------------------------- cycles, FCG

def dense_cycle(B, r, x0):
    """One V-cycle of one hierarchy written with dense matrices."""
    nl = B.levels(0)
    As = [ops.to_scipy(B.level_A(0, l)).toarray() for l in range(nl)]
    Ps = []
    for l in range(nl - 1):
        agg, p = B.level_P(0, l)
        P = np.zeros((len(agg), As[l + 1].shape[0]))
        P[np.arange(len(agg)), agg] = p
        Ps.append(P)
    Ds = [B.level_DM(0, l)[1] for l in range(nl)]             # L1 divisors

    def relax(l, f, x, k):
        for _ in range(k):
            x = x + (f - As[l] @ x) / Ds[l]
        return x

    def cyc(l, f, x):
        if l == nl - 1:
            return relax(l, f, x, 16)
        x = relax(l, f, x, 2)
        fc = Ps[l].T @ (f - As[l] @ x)
        xc = cyc(l + 1, fc, np.zeros(len(fc)))
        x = x + Ps[l] @ xc
        return relax(l, f, x, 2)
    return cyc(0, r, x0)


def test_cycle_matches_dense_restatement_and_is_symmetric():
    A = ops.lognormal3d(6, 4.0, seed=9)
    n = len(A[0]) - 1
    prm = orc.params(prerelax_sweeps=2, postrelax_sweeps=2, relaxnumber_coarse=16)
    B = orc.Boot(A, prm)
    assert B.levels(0) >= 2
    rng = np.random.default_rng(2)
    r, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert np.allclose(B.precond(r, x0), dense_cycle(B, r, x0), rtol=1e-12, atol=1e-12)
    u, v = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert abs(u @ B.precond(v) - v @ B.precond(u)) <= 1e-12 * abs(u @ B.precond(v))


@pytest.mark.parametrize("name", ["poisson", "lognormal", "aniso"])
def test_fcg_vs_scipy_cg_and_direct(name):
    A = {"poisson": lambda: ops.poisson3d(12), "lognormal": lambda: ops.lognormal3d(10, 4.0, 1),
         "aniso": lambda: ops.aniso2d(40, 0.01)}[name]()
    S = ops.to_scipy(A).tocsc()
    n = S.shape[0]
    b = ops.uniform_pm1(4, n)
    prm = orc.params(rtol=1e-8)
    B = orc.Boot(A, prm)
    rc, x, res, hist = B.fcg(b, history=200)
    assert rc == 0
    xd = spla.spsolve(S, b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    assert np.linalg.norm(b - S @ x) <= 1.01e-8 * np.linalg.norm(b)
    # the same preconditioner in scipy's CG (V-cycle with equal pre/post
    # Jacobi sweeps is symmetric): iteration counts within 2
    M = spla.LinearOperator((n, n), matvec=lambda v: B.precond(v))
    its = []
    spla.cg(S, b, rtol=1e-8, atol=0.0, M=M, maxiter=500, callback=lambda xk: its.append(1))
    assert abs(len(its) - (res.iterations - 1)) <= 2, (len(its), res.iterations)
    # the recorded residual history decreases overall
    assert hist[-1] < 1e-8 * res.delta0 and len(hist) == res.iterations - 1


def test_bootstrap_second_hierarchy_repeats_the_first():
    """The reference builds hierarchy 1 from w = 1 again (no inner iterations
    run before it: src/BOOT/bootstrap.cu:68-75), so it equals hierarchy 0."""
    A = ops.poisson3d(8)
    B = orc.Boot(A, orc.params(max_hrc=2, conv_ratio=0.5))
    assert B.info()[0] == 2
    for l in range(B.levels(0)):
        a0, a1 = B.level_A(0, l), B.level_A(1, l)
        assert all(np.array_equal(x, y) for x, y in zip(a0, a1))
    assert np.array_equal(B.w(1), np.ones(len(A[0]) - 1))


def test_zero_rhs():
    A = ops.poisson3d(4)
    n = len(A[0]) - 1
    rc, _, _ = orc.solve(A, np.zeros(n), orc.params())
    assert rc == -3


def dense_kcycle(B, r, pre=2, post=2, coarse=16):
    """The K-cycle (cycle_type 3) written with dense matrices: the coarse
    correction of a level whose next level is not the coarsest is the
    Galerkin projection of A_m x = f_m onto span{d0, d1} - what two flexible
    CG steps compute - with d0 = B f, w = f - alpha A d0 (alpha = f.d0 /
    d0.A d0), d1 = B w, B the cycle from the next level; only d0 when
    ||w|| <= 0.25 ||f|| (src/SOLVER/FCG.cu:620-723)."""
    nl = B.levels(0)
    As = [ops.to_scipy(B.level_A(0, l)).toarray() for l in range(nl)]
    Ps = []
    for l in range(nl - 1):
        agg, p = B.level_P(0, l)
        P = np.zeros((len(agg), As[l + 1].shape[0]))
        P[np.arange(len(agg)), agg] = p
        Ps.append(P)
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
        if l + 1 == nl - 1:
            xc = cyc(l + 1, fc, np.zeros(len(fc)))
        else:
            m = l + 1
            d0 = cyc(m, fc, np.zeros(len(fc)))
            alpha = (fc @ d0) / (d0 @ As[m] @ d0)
            w = fc - alpha * (As[m] @ d0)
            if np.linalg.norm(w) <= 0.25 * np.linalg.norm(fc):
                xc = alpha * d0
            else:
                d1 = cyc(m, w, np.zeros(len(fc)))
                V = np.stack([d0, d1], axis=1)
                xc = V @ np.linalg.solve(V.T @ As[m] @ V, V.T @ fc)
        x = x + Ps[l] @ xc
        return relax(l, f, x, post)
    return cyc(0, r, np.zeros(len(r)))


KCYCLE = dict(prerelax_sweeps=2, postrelax_sweeps=2, relaxnumber_coarse=16, cycle_type=3)


def test_kcycle_on_two_levels_is_the_v_cycle():
    """With two levels the next level is the coarsest: no inner iteration."""
    A = ops.poisson3d(10)
    Bk = orc.Boot(A, orc.params(**KCYCLE))
    Bv = orc.Boot(A, orc.params(**dict(KCYCLE, cycle_type=0)))
    assert Bk.levels(0) == 2
    r = ops.uniform_pm1(3, len(A[0]) - 1)
    assert np.array_equal(bits(Bk.precond(r)), bits(Bv.precond(r)))


@pytest.mark.parametrize("name", ["poisson", "lognormal"])
def test_kcycle_matches_galerkin_projection(name):
    A = {"poisson": lambda: ops.poisson3d(14),
         "lognormal": lambda: ops.lognormal3d(12, 4.0, seed=5)}[name]()
    n = len(A[0]) - 1
    B = orc.Boot(A, orc.params(**dict(KCYCLE, aggrsweeps=1)))
    assert B.levels(0) >= 4      # two levels with an inner iteration
    r = ops.uniform_pm1(6, n)
    z = B.precond(r)
    zd = dense_kcycle(B, r)
    assert np.linalg.norm(z - zd) <= 1e-10 * np.linalg.norm(zd)


def test_kcycle_fcg_converges_faster_than_v():
    A = ops.lognormal3d(14, 4.0, seed=7)
    n = len(A[0]) - 1
    b = ops.uniform_pm1(0, n)
    res = {}
    for ct in (0, 3):
        B = orc.Boot(A, orc.params(**dict(KCYCLE, cycle_type=ct, aggrsweeps=1, rtol=1e-8)))
        rc, x, r, _ = B.fcg(b)
        assert rc == 0
        assert np.linalg.norm(b - ops.to_scipy(A) @ x) <= 1.01e-8 * np.linalg.norm(b)
        res[ct] = r.iterations
    assert res[3] <= res[0], res


def test_kcycle_is_rejected_in_a_settings_file(tmp_path):
    """The reference's parser accepts cycle types 0..2 only (AMG.cu:506-508)."""
    f = tmp_path / "s.txt"
    f.write_text("\n".join(["0", "1", "0.8", "2", "0", "39", "3", "4", "4", "16", "4", "4",
                            "2000", "1e-8"]) + "\n")
    assert orc.params_from_file(str(f))[0] == -2          # ORC_ESETTINGS


# ------------------------------------------- H/W cycles, composites, bootstrap
#
# An independent restatement (scipy.sparse, not the oracle's code) of the
# recursion of src/SOLVER/GAMG_cycle.cu:25-140 with the grid-sweep counts of
# src/AMG/AMG.cu:371-394, of the composites of src/SOLVER/FCG.cu:112-187 and of
# the bootstrap inner iterations of src/BOOT/bootstrap.cu:5-44. Only the level
# operators and prolongators are taken from the oracle (they are pinned above
# against scipy's Galerkin product); smoother divisors are recomputed here.
# Each restatement takes a named slip, and every test checks that the slipped
# variant disagrees with the oracle, so a plausible misreading shared by the
# oracle and the GPU library would fail it.

class SparseHierarchy:
    """Level operators A_l, prolongators P_l, diagonals and L1 row sums."""

    def __init__(self, B, k):
        nl = B.levels(k)
        self.A = [ops.to_scipy(B.level_A(k, l)).tocsr() for l in range(nl)]
        self.P = []
        for l in range(nl - 1):
            agg, p = B.level_P(k, l)
            nf, nc = self.A[l].shape[0], self.A[l + 1].shape[0]
            self.P.append(sp.csr_matrix((p, (np.arange(nf), agg)), shape=(nf, nc)))
        self.D = [A.diagonal() for A in self.A]
        self.M = [np.asarray(abs(A).sum(axis=1)).ravel() for A in self.A]

    @property
    def levels(self):
        return len(self.A)


def visits(cycle_type, l, L, slip=None):
    """How often level l + 1 (0-based) is visited from level l of L levels.
    V: once. W: twice, but the level above the coarsest visits it once (the
    break of GAMG_cycle.cu:107-108). H: twice from even 0-based levels
    (AMG.cu:381-386), with the same break."""
    g = 1
    if cycle_type == 2:
        g = 2
    elif cycle_type == 1:
        g = 2 if l % 2 == (1 if slip == "h_odd_levels" else 0) else 1
    if l + 1 == L - 1 and slip != "no_coarsest_break":
        g = 1
    return g


def sparse_cycle(H, f, x, s, l=0, slip=None):
    """One cycle from level l: pre-smoothing, restricted residual, recursion,
    prolongation, post-smoothing; the coarsest level gets relaxnumber_coarse
    sweeps of the coarse smoother. s: cycle_type, pre, post, coarse,
    relax_type, coarse_solver."""
    L = H.levels

    def relax(l, f, x, k, t):
        d = H.M[l] if t == 4 else H.D[l]
        for _ in range(k):
            x = x + (f - H.A[l] @ x) / d
        return x

    if l == L - 1:
        return relax(l, f, x, s["coarse"], s["coarse_solver"])
    x = relax(l, f, x, s["pre"], s["relax_type"])
    fc = H.P[l].T @ (f - H.A[l] @ x)
    xc = np.zeros(len(fc))
    for _ in range(visits(s["cycle_type"], l, L, slip)):
        xc = sparse_cycle(H, fc, xc, s, l + 1, slip)
    x = x + H.P[l] @ xc
    return relax(l, f, x, s["post"], s["relax_type"])


def sparse_composite(Hs, r, x, s, solver_type, slip=None):
    """Composite of hierarchies: 0 multiplicative, 1 symmetrized (forward,
    then backward), 2 additive (mean of the cycles, each from the input x)."""
    run = lambda k, x: sparse_cycle(Hs[k], r, x, s)
    order = list(range(len(Hs)))
    if solver_type == 0:
        for k in (order[::-1] if slip == "reversed" else order):
            x = run(k, x)
        return x
    if solver_type == 1:
        for k in order + (order if slip == "forward_twice" else order[::-1]):
            x = run(k, x)
        return x
    acc = np.zeros(len(r))
    xin = x
    for k in order:
        xk = run(k, xin if slip != "chained" else x)
        x = xk
        acc = acc + xk
    return acc / (1 if slip == "sum" else len(Hs))


def cycle_settings(prm):
    return dict(cycle_type=prm.cycle_type, pre=prm.prerelax_sweeps, post=prm.postrelax_sweeps,
                coarse=prm.relaxnumber_coarse, relax_type=prm.relax_type,
                coarse_solver=prm.coarse_solver)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


DEEP = lambda: ops.lognormal3d(16, 4.0, seed=11)        # 5 levels with one sweep


FOUR = lambda: ops.lognormal3d(12, 4.0, seed=3)         # 4 levels with one sweep


@pytest.mark.parametrize("cycle_type,relax_type,coarse_solver,slips,op", [
    (1, 4, 4, ["h_odd_levels"], DEEP),
    (1, 0, 4, ["h_odd_levels", "no_coarsest_break"], FOUR),
    (1, 4, 0, ["h_odd_levels", "no_coarsest_break"], FOUR),
    (2, 4, 0, ["no_coarsest_break"], DEEP),
    (2, 0, 0, ["no_coarsest_break"], FOUR),
    (0, 4, 4, [], DEEP),
])
def test_hw_cycles_match_restatement(cycle_type, relax_type, coarse_solver, slips, op):
    A = op()
    n = len(A[0]) - 1
    prm = orc.params(aggrsweeps=1, cycle_type=cycle_type, relax_type=relax_type,
                     coarse_solver=coarse_solver, prerelax_sweeps=2, postrelax_sweeps=1,
                     relaxnumber_coarse=6)
    B = orc.Boot(A, prm)
    assert B.rc == 0 and B.levels(0) >= 4
    H = SparseHierarchy(B, 0)
    L = H.levels
    for slip in slips:            # every named slip changes the visit schedule here
        assert any(visits(cycle_type, l, L) != visits(cycle_type, l, L, slip)
                   for l in range(L - 1)), slip
    s = cycle_settings(prm)
    rng = np.random.default_rng(cycle_type * 10 + relax_type)
    r, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for xin in (np.zeros(n), x0):
        z = B.precond(r, xin)
        assert rel(z, sparse_cycle(H, r, xin, s)) <= 1e-12
        for slip in slips:
            assert rel(z, sparse_cycle(H, r, xin, s, slip=slip)) > 1e-6, slip


def test_w_cycle_visit_counts_on_five_levels():
    """The restatement's schedule against the textbook counts: with 5 levels a
    W-cycle visits level l (0-based) 2^l times, except the coarsest, visited
    as often as the level above it: 1, 2, 4, 8, 8; the H-cycle doubles from
    levels 0 and 2 only: 1, 2, 2, 4, 4."""
    L = 5
    counts = [1]
    for l in range(L - 1):
        counts.append(counts[-1] * visits(2, l, L))
    assert counts == [1, 2, 4, 8, 8]
    counts = [1]
    for l in range(L - 1):
        counts.append(counts[-1] * visits(1, l, L))
    assert counts == [1, 2, 2, 4, 4]                   # H-cycle: even levels double


@pytest.mark.parametrize("solver_type,slips", [
    (0, ["reversed"]),
    (1, ["forward_twice"]),
    (2, ["chained", "sum"]),
])
@pytest.mark.parametrize("cycle_type", [0, 2])
def test_three_hierarchy_composites_match_restatement(solver_type, slips, cycle_type):
    A = DEEP()
    n = len(A[0]) - 1
    prm = orc.params(solver_type=solver_type, max_hrc=3, conv_ratio=1e-6, aggrsweeps=1,
                     cycle_type=cycle_type, prerelax_sweeps=2, postrelax_sweeps=2,
                     relaxnumber_coarse=8)
    B = orc.Boot(A, prm)
    assert B.rc == 0 and B.info()[0] == 3
    Hs = [SparseHierarchy(B, k) for k in range(3)]
    assert all(H.levels >= 4 for H in Hs)
    # hierarchy 2 is built from the bootstrapped w, so it differs from 0
    assert not np.array_equal(Hs[2].P[0].data, Hs[0].P[0].data)
    s = cycle_settings(prm)
    rng = np.random.default_rng(solver_type)
    r, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for xin in (np.zeros(n), x0):
        z = B.precond(r, xin)
        assert rel(z, sparse_composite(Hs, r, xin, s, solver_type)) <= 1e-12
        for slip in slips:
            assert rel(z, sparse_composite(Hs, r, xin, s, solver_type, slip)) > 1e-6, slip


def inner_iterations(Hs, A, s, solver_type, slip=None):
    """x = 1; 15 times x <- B(0; x); ratio ||x_15||_A / ||x_14||_A; w = x / ||x||_A."""
    n = A.shape[0]
    norm = (lambda v: np.linalg.norm(v)) if slip == "two_norm" else (lambda v: np.sqrt(v @ (A @ v)))
    x = np.ones(n)
    old = norm(x)
    ratio = 1.0
    for _ in range(14 if slip == "fourteen" else 15):
        x = sparse_composite(Hs, np.zeros(n), x, s, solver_type)
        new = norm(x)
        ratio, old = new / old, new
    scale = np.linalg.norm(x) if slip == "two_norm_scale" else np.sqrt(x @ (A @ x))
    return x / scale, ratio


@pytest.mark.parametrize("solver_type", [0, 1, 2])
def test_bootstrap_inner_iterations_match_restatement(solver_type):
    """w(2) (the vector hierarchy 2 is built from) comes from the composite of
    hierarchies 0 and 1; the final estimated ratio from that of all three."""
    A = DEEP()
    S = ops.to_scipy(A)
    prm = orc.params(solver_type=solver_type, max_hrc=3, conv_ratio=1e-6, aggrsweeps=1,
                     prerelax_sweeps=2, postrelax_sweeps=2, relaxnumber_coarse=8)
    B = orc.Boot(A, prm)
    assert B.rc == 0 and B.info()[0] == 3
    Hs = [SparseHierarchy(B, k) for k in range(3)]
    s = cycle_settings(prm)
    n = S.shape[0]
    assert np.array_equal(B.w(0), np.ones(n)) and np.array_equal(B.w(1), np.ones(n))
    w2, _ = inner_iterations(Hs[:2], S, s, solver_type)
    assert rel(B.w(2), w2) <= 1e-13
    assert abs(np.sqrt(w2 @ (S @ w2)) - 1.0) <= 1e-12
    _, ratio = inner_iterations(Hs, S, s, solver_type)
    est = B.info()[1]
    assert abs(est - ratio) <= 1e-13 * ratio
    # a slip must miss by ten times the tolerance the restatement meets (one
    # application fewer moves the power iteration's ratio by 7e-12 to 4e-5)
    for slip in ("two_norm", "fourteen"):
        _, rs = inner_iterations(Hs, S, s, solver_type, slip)
        assert abs(est - rs) > 1e-12 * ratio, slip
    for slip in ("two_norm_scale", "fourteen"):
        ws, _ = inner_iterations(Hs[:2], S, s, solver_type, slip)
        assert rel(B.w(2), ws) > 1e-12, slip
