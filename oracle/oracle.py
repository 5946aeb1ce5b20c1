"""ctypes binding of the serial CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    """Same field order as the reference's struct params (src/AMG/AMG.cu:20-37)."""
    _fields_ = [
        ("solver_type", ctypes.c_int),
        ("max_hrc", ctypes.c_int),
        ("conv_ratio", ctypes.c_double),
        ("matchtype", ctypes.c_int),
        ("aggrsweeps", ctypes.c_int),
        ("aggrtype", ctypes.c_int),
        ("max_levels", ctypes.c_int),
        ("cycle_type", ctypes.c_int),
        ("coarse_solver", ctypes.c_int),
        ("relax_type", ctypes.c_int),
        ("prerelax_sweeps", ctypes.c_int),
        ("postrelax_sweeps", ctypes.c_int),
        ("itnlim", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("relaxnumber_coarse", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("delta0", ctypes.c_double),
        ("n_hrc", ctypes.c_int),
        ("estimated_ratio", ctypes.c_double),
        ("setup_seconds", ctypes.c_double),
        ("solve_seconds", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_log.restype = ctypes.c_double
        L.orc_log.argtypes = [ctypes.c_double]
        L.orc_setup.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
                                ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]
        L.orc_boot_free.argtypes = [ctypes.c_void_p]
        L.orc_boot_free.restype = None
        for name in ("orc_boot_info", "orc_hier_levels", "orc_level_size", "orc_level_A",
                     "orc_level_P", "orc_level_DM", "orc_boot_w", "orc_precond_apply",
                     "orc_fcg", "orc_level_Pinfo", "orc_level_Pcsr", "orc_spgemm"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def params(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, aggrtype=0,
           max_levels=39, cycle_type=0, coarse_solver=4, relax_type=4,
           relaxnumber_coarse=16, prerelax_sweeps=4, postrelax_sweeps=4,
           itnlim=2000, rtol=1e-6):
    return Params(solver_type, max_hrc, conv_ratio, 3, aggrsweeps, aggrtype, max_levels,
                  cycle_type, coarse_solver, relax_type, prerelax_sweeps,
                  postrelax_sweeps, itnlim, rtol, relaxnumber_coarse)


def params_from_file(path):
    p = Params()
    rc = lib().orc_params_from_file(path.encode(), ctypes.byref(p))
    return rc, p


def log(x):
    return lib().orc_log(float(x))


def _csr(A):
    row, col, val = A
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return row, col, val


def weights(A, w):
    row, col, val = _csr(A)
    n = len(row) - 1
    w = np.ascontiguousarray(w, dtype=np.float64)
    W = np.zeros(len(col))
    rc = lib().orc_weights(n, _p(row, _i32p), _p(col, _i32p), _p(val, _f64p), _p(w, _f64p),
                           _p(W, _f64p))
    assert rc == 0, rc
    return W


def matching(A, W, order=None):
    row, col, _ = _csr(A)
    n = len(row) - 1
    W = np.ascontiguousarray(W, dtype=np.float64)
    Mr = np.empty(n, dtype=np.int32)
    Ms = np.empty(n, dtype=np.int32)
    nu = ctypes.c_int(0)
    op = None
    if order is not None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        op = _p(order, _i32p)
    rc = lib().orc_matching(n, _p(row, _i32p), _p(col, _i32p), _p(W, _f64p), op,
                            _p(Mr, _i32p), _p(Ms, _i32p), ctypes.byref(nu))
    assert rc == 0, rc
    return Mr, Ms, nu.value


def aggregate(Ms, w):
    n = len(Ms)
    Ms = np.ascontiguousarray(Ms, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    agg = np.empty(n, dtype=np.int32)
    p = np.empty(n)
    nc = ctypes.c_int(0)
    rc = lib().orc_aggregate(n, _p(Ms, _i32p), _p(w, _f64p), _p(agg, _i32p), _p(p, _f64p),
                             ctypes.byref(nc))
    assert rc == 0, rc
    return agg, p, nc.value


def galerkin(A, agg, p, nc):
    row, col, val = _csr(A)
    n = len(row) - 1
    agg = np.ascontiguousarray(agg, dtype=np.int32)
    p = np.ascontiguousarray(p, dtype=np.float64)
    rc_ = np.empty(nc + 1, dtype=np.int32)
    nnz = ctypes.c_int(0)
    L = lib()
    rc = L.orc_galerkin(n, _p(row, _i32p), _p(col, _i32p), _p(val, _f64p), _p(agg, _i32p),
                        _p(p, _f64p), nc, _p(rc_, _i32p), None, None, ctypes.byref(nnz))
    assert rc == 0, rc
    cc = np.empty(nnz.value, dtype=np.int32)
    vc = np.empty(nnz.value)
    rc = L.orc_galerkin(n, _p(row, _i32p), _p(col, _i32p), _p(val, _f64p), _p(agg, _i32p),
                        _p(p, _f64p), nc, _p(rc_, _i32p), _p(cc, _i32p), _p(vc, _f64p),
                        ctypes.byref(nnz))
    assert rc == 0, rc
    return rc_, cc, vc


def spgemm(Lm, Rm, m_right):
    """C = L R for CSR triples (L is n x len(R row)-1, R has m_right columns),
    in the fixed order of smoothed aggregation's products."""
    lrow, lcol, lval = [np.ascontiguousarray(a) for a in Lm]
    rrow, rcol, rval = [np.ascontiguousarray(a) for a in Rm]
    lrow = lrow.astype(np.int32)
    lcol = lcol.astype(np.int32)
    rrow = rrow.astype(np.int32)
    rcol = rcol.astype(np.int32)
    lval = lval.astype(np.float64)
    rval = rval.astype(np.float64)
    nl = len(lrow) - 1
    nr = len(rrow) - 1
    L = lib()
    crow = np.empty(nl + 1, dtype=np.int32)
    nnz = ctypes.c_int(0)
    args = (nl, nr, _p(lrow, _i32p), _p(lcol, _i32p), _p(lval, _f64p), nr, m_right,
            _p(rrow, _i32p), _p(rcol, _i32p), _p(rval, _f64p), _p(crow, _i32p))
    rc = L.orc_spgemm(*args, None, None, ctypes.byref(nnz))
    if rc:
        raise ValueError("orc_spgemm: %d" % rc)
    ccol = np.empty(nnz.value, dtype=np.int32)
    cval = np.empty(nnz.value)
    assert L.orc_spgemm(*args, _p(ccol, _i32p), _p(cval, _f64p), ctypes.byref(nnz)) == 0
    return crow, ccol, cval


def spmv(A, x):
    row, col, val = _csr(A)
    n = len(row) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(n)
    rc = lib().orc_spmv(n, _p(row, _i32p), _p(col, _i32p), _p(val, _f64p), _p(x, _f64p),
                        _p(y, _f64p))
    assert rc == 0
    return y


def relax(A, d, f, u, sweeps):
    row, col, val = _csr(A)
    n = len(row) - 1
    d = np.ascontiguousarray(d, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.array(u, dtype=np.float64)
    rc = lib().orc_relax(n, _p(row, _i32p), _p(col, _i32p), _p(val, _f64p), _p(d, _f64p),
                         _p(f, _f64p), _p(u, _f64p), sweeps)
    assert rc == 0
    return u


class Boot:
    """An oracle setup (bootstrap composite of AMG hierarchies)."""

    def __init__(self, A, prm):
        row, col, val = _csr(A)
        self._keep = (row, col, val)
        self.n = len(row) - 1
        h = ctypes.c_void_p()
        self.rc = lib().orc_setup(self.n, len(col), _p(row, _i32p), _p(col, _i32p),
                                  _p(val, _f64p), ctypes.byref(prm), ctypes.byref(h))
        self.h = h if self.rc == 0 else None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_boot_free(self.h)
            self.h = None

    def info(self):
        nh = ctypes.c_int(0)
        er = ctypes.c_double(0)
        lib().orc_boot_info(self.h, ctypes.byref(nh), ctypes.byref(er))
        return nh.value, er.value

    def levels(self, k=0):
        return lib().orc_hier_levels(self.h, k)

    def level_A(self, k, l):
        n = ctypes.c_int(0)
        nnz = ctypes.c_int(0)
        assert lib().orc_level_size(self.h, k, l, ctypes.byref(n), ctypes.byref(nnz)) == 0
        row = np.empty(n.value + 1, dtype=np.int32)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value)
        assert lib().orc_level_A(self.h, k, l, _p(row, _i32p), _p(col, _i32p),
                                 _p(val, _f64p)) == 0
        return row, col, val

    def level_P(self, k, l):
        n = ctypes.c_int(0)
        assert lib().orc_level_size(self.h, k, l, ctypes.byref(n), None) == 0
        agg = np.empty(n.value, dtype=np.int32)
        p = np.empty(n.value)
        assert lib().orc_level_P(self.h, k, l, _p(agg, _i32p), _p(p, _f64p)) == 0
        return agg, p

    def level_Pcsr(self, k, l):
        """(row, col, val, smoothed, omega) of level l's prolongator."""
        nnz = ctypes.c_int(0)
        sm = ctypes.c_int(0)
        om = ctypes.c_double(0)
        assert lib().orc_level_Pinfo(self.h, k, l, ctypes.byref(nnz), ctypes.byref(sm),
                                     ctypes.byref(om)) == 0
        n = ctypes.c_int(0)
        assert lib().orc_level_size(self.h, k, l, ctypes.byref(n), None) == 0
        row = np.empty(n.value + 1, dtype=np.int32)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value)
        assert lib().orc_level_Pcsr(self.h, k, l, _p(row, _i32p), _p(col, _i32p),
                                    _p(val, _f64p)) == 0
        return row, col, val, bool(sm.value), om.value

    def level_DM(self, k, l):
        n = ctypes.c_int(0)
        assert lib().orc_level_size(self.h, k, l, ctypes.byref(n), None) == 0
        D = np.empty(n.value)
        M = np.empty(n.value)
        hd = ctypes.c_int(0)
        hm = ctypes.c_int(0)
        assert lib().orc_level_DM(self.h, k, l, _p(D, _f64p), _p(M, _f64p), ctypes.byref(hd),
                                  ctypes.byref(hm)) == 0
        return (D if hd.value else None), (M if hm.value else None)

    def w(self, k):
        w = np.empty(self.n)
        assert lib().orc_boot_w(self.h, k, _p(w, _f64p)) == 0
        return w

    def precond(self, r, x=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        x = np.zeros(self.n) if x is None else np.array(x, dtype=np.float64)
        assert lib().orc_precond_apply(self.h, _p(r, _f64p), _p(x, _f64p)) == 0
        return x

    def fcg(self, b, history=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        res = Result()
        hist = np.zeros(max(history, 1))
        rc = lib().orc_fcg(self.h, _p(b, _f64p), _p(x, _f64p), ctypes.byref(res),
                           _p(hist, _f64p), history)
        return rc, x, res, hist[:min(history, max(res.iterations - 1, 0))]


def solve(A, b, prm):
    row, col, val = _csr(A)
    n = len(row) - 1
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    res = Result()
    rc = lib().orc_solve(n, len(col), _p(row, _i32p), _p(col, _i32p), _p(val, _f64p),
                         _p(b, _f64p), ctypes.byref(prm), _p(x, _f64p), ctypes.byref(res))
    return rc, x, res
