"""bootcmatchg_amd - MI355X (gfx950) AMG-preconditioned flexible CG.

Python mirror of BootCMatchG's solve path over the C-ABI in include/bcmg.h
(libbcmg.so, hand-written HIP kernels). The names follow the reference's
facade (src/BootCMatchG.h): a settings record read by `params_from_file`
(AMG::Params::initFromFile, src/AMG/AMG.cu:480-526) and `solve(A, p, rhs)`
(src/SOLVER/solver.cu:20-129). Errors raise BcmgError carrying the C status
code instead of terminating the process as the reference does.

There is no CPU fallback: if libbcmg.so is missing or no gfx950 device is
present, the calls that need the GPU raise.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbcmg.so")

OK = 0
EINVAL = -1
ESETTINGS = -2
EZERORHS = -3
EUNSUP = -4
EBREAKDOWN = -5
EHIP = -6
ENOMEM = -7
ENODEV = -8

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class BcmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bcmg error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    """bcmg_params: field order of the reference's struct params
    (src/AMG/AMG.cu:20-37)."""
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

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def make_params(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, aggrtype=0,
                max_levels=39, cycle_type=0, coarse_solver=4, relax_type=4,
                relaxnumber_coarse=16, prerelax_sweeps=4, postrelax_sweeps=4,
                itnlim=2000, rtol=1e-6):
    """Settings in the order of the 14-line AMGsettings file."""
    return Params(solver_type, max_hrc, conv_ratio, 3, aggrsweeps, aggrtype, max_levels,
                  cycle_type, coarse_solver, relax_type, prerelax_sweeps, postrelax_sweeps,
                  itnlim, rtol, relaxnumber_coarse)


_lib = None


def lib():
    """Load libbcmg.so (built by __graft_entry__.build or `make -C csrc`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BcmgError(ENODEV, "libbcmg.so is not built (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.bcmg_last_error.restype = ctypes.c_char_p
        L.bcmg_hier_levels.restype = ctypes.c_int
        for name in ("bcmg_ctx_create", "bcmg_setup", "bcmg_csr_upload", "bcmg_galerkin",
                     "bcmg_dmalloc"):
            getattr(L, name).restype = ctypes.c_int
        L.bcmg_dmalloc.argtypes = [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.bcmg_dfree.argtypes = [_vp, _vp]
        L.bcmg_memcpy_h2d.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.bcmg_memcpy_d2h.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.bcmg_memcpy_d2d.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        _lib = L
    return _lib


PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2


class SetupTimes(ctypes.Structure):
    """bcmg_setup_times: the setup's phases in ms."""
    _fields_ = [(f, ctypes.c_double) for f in (
        "weights_ms", "matching_ms", "aggregation_ms", "galerkin_ms", "smoother_ms",
        "inner_ms", "hierarchy_ms", "total_ms")]


def _check(rc):
    if rc != OK:
        raise BcmgError(rc, lib().bcmg_last_error().decode(errors="replace"))
    return rc


def params_from_file(path):
    """AMG::Params::initFromFile (src/AMG/AMG.cu:480-526)."""
    p = Params()
    _check(lib().bcmg_params_from_file(os.fsencode(path), ctypes.byref(p)))
    return p


def _as(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """One GPU and one HIP stream (replaces Handles, src/utility/handles.cu).
    stream: a hipStream_t (an int address, e.g. torch's
    `torch.cuda.Stream().cuda_stream`) the caller created on `device`; the
    context then enqueues all its work there and leaves the stream alive."""

    def __init__(self, device=0, stream=None):
        self.h = _vp()
        self._kids = []
        if stream is None:
            _check(lib().bcmg_ctx_create(device, ctypes.byref(self.h)))
        else:
            _check(lib().bcmg_ctx_create_on_stream(device, _vp(int(stream)),
                                                   ctypes.byref(self.h)))

    def _adopt(self, obj):
        """Track an object that lives in this context: close() frees the
        live ones first, newest first (a server before its composites, a
        composite before its matrix), so nothing outlives its context."""
        self._kids.append(weakref.ref(obj))
        if len(self._kids) > 4096:
            self._kids = [k for k in self._kids if k() is not None]

    def close(self):
        if getattr(self, "h", None):
            for ref in reversed(self._kids):
                obj = ref()
                if obj is not None:
                    obj._release()
            self._kids = []
            lib().bcmg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module's globals (lib) may already be gone
        # and the process is ending: a failed release is not reported there;
        # explicit free()/close() calls still raise
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().bcmg_sync(self.h))


class DeviceArray:
    """A device buffer of a numpy dtype owned by a Context."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.ptr = _vp()
        _check(lib().bcmg_dmalloc(ctx.h, max(self.n, 1) * self.dtype.itemsize,
                                  ctypes.byref(self.ptr)))
        ctx._adopt(self)

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = _as(a, self.dtype)
        assert a.size == self.n
        _check(lib().bcmg_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib().bcmg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None) and self.ptr.value and self.ctx.h:
            lib().bcmg_dfree(self.ctx.h, self.ptr)
        self.ptr = _vp()

    _release = free

    def __del__(self):
        # at interpreter exit the module's globals (lib) may already be gone
        # and the process is ending: a failed release is not reported there;
        # explicit free()/close() calls still raise
        try:
            self.free()
        except Exception:
            pass


class Csr:
    """Device CSR (CSRm::copyToDevice, src/matrix/CSR.cu:194-217)."""

    def __init__(self, ctx, A=None, handle=None):
        self.ctx = ctx
        self.h = _vp()
        if handle is not None:
            self.h = handle
        else:
            row, col, val = A
            row, col, val = _as(row, np.int32), _as(col, np.int32), _as(val, np.float64)
            n = len(row) - 1
            _check(lib().bcmg_csr_upload(ctx.h, n, len(col), row.ctypes.data_as(_i32p),
                                         col.ctypes.data_as(_i32p), val.ctypes.data_as(_f64p),
                                         ctypes.byref(self.h)))
        ctx._adopt(self)
        n = ctypes.c_int()
        nnz = ctypes.c_int()
        _check(lib().bcmg_csr_info(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @property
    def column_bytes(self):
        """Bytes per column index the streamed row kernels read (0 with the
        stencil pattern, 2 with the 16-bit column-minus-row copy, else 4)."""
        b = ctypes.c_int()
        _check(lib().bcmg_csr_column_bytes(self.h, ctypes.byref(b)))
        return b.value

    def set_columns(self, nbytes):
        """Restrict the row kernels to a column form: 0 the stencil pattern,
        2 the 16-bit offsets, 4 the int32 columns (same results)."""
        _check(lib().bcmg_csr_set_columns(self.h, nbytes))

    def download(self):
        row = np.empty(self.n + 1, np.int32)
        col = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz)
        _check(lib().bcmg_csr_download(self.h, row.ctypes.data_as(_i32p),
                                       col.ctypes.data_as(_i32p), val.ctypes.data_as(_f64p)))
        return row, col, val

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().bcmg_csr_free(self.h)
        self.h = None

    _release = free

    def __del__(self):
        # at interpreter exit the module's globals (lib) may already be gone
        # and the process is ending: a failed release is not reported there;
        # explicit free()/close() calls still raise
        try:
            self.free()
        except Exception:
            pass


class Boot:
    """A built composite of AMG hierarchies (Bootstrap::bootstrap,
    src/BOOT/bootstrap.cu:46-90)."""

    def __init__(self, ctx, A, prm):
        self.ctx = ctx
        self.A = A
        self.h = _vp()
        _check(lib().bcmg_setup(ctx.h, A.h, ctypes.byref(prm), ctypes.byref(self.h)))
        ctx._adopt(self)

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().bcmg_boot_free(self.h)
        self.h = None

    _release = free

    def __del__(self):
        # at interpreter exit the module's globals (lib) may already be gone
        # and the process is ending: a failed release is not reported there;
        # explicit free()/close() calls still raise
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        nh = ctypes.c_int()
        er = ctypes.c_double()
        _check(lib().bcmg_boot_info(self.h, ctypes.byref(nh), ctypes.byref(er)))
        return nh.value, er.value

    def hier_info(self, k=0):
        """Hierarchy k: levels, operator complexity, its W-cycle form, the
        average coarsening ratio and the estimated ratio once k was added
        (AMG::Hierarchy::printInfo, src/AMG/AMG.cu:241-251)."""
        lv = ctypes.c_int()
        d = [ctypes.c_double() for _ in range(4)]
        _check(lib().bcmg_hier_info(self.h, k, ctypes.byref(lv), *[ctypes.byref(x) for x in d]))
        return dict(levels=lv.value, op_cmplx=d[0].value, op_wcmplx=d[1].value,
                    avg_cratio=d[2].value, ratio_after=d[3].value)

    def setup_timing(self, k=-1):
        """Event-timed setup phases in ms (recorded while time_setup(ctx, True)
        was in force when the composite was built); k = -1 sums the
        composite, k >= 0 gives hierarchy k."""
        t = SetupTimes()
        _check(lib().bcmg_setup_timing(self.h, k, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in SetupTimes._fields_}

    def levels(self, k=0):
        r = lib().bcmg_hier_levels(self.h, k)
        if r < 0:
            _check(r)
        return r

    def level_size(self, k, l):
        n = ctypes.c_int()
        nnz = ctypes.c_int()
        _check(lib().bcmg_level_size(self.h, k, l, ctypes.byref(n), ctypes.byref(nnz)))
        return n.value, nnz.value

    def level_A(self, k, l):
        n, nnz = self.level_size(k, l)
        row = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz)
        _check(lib().bcmg_level_A(self.h, k, l, row.ctypes.data_as(_i32p),
                                  col.ctypes.data_as(_i32p), val.ctypes.data_as(_f64p)))
        return row, col, val

    def level_P(self, k, l):
        n, _ = self.level_size(k, l)
        agg = np.empty(n, np.int32)
        p = np.empty(n)
        _check(lib().bcmg_level_P(self.h, k, l, agg.ctypes.data_as(_i32p),
                                  p.ctypes.data_as(_f64p)))
        return agg, p

    def level_Pcsr(self, k, l):
        """(row, col, val, smoothed, omega) of level l's prolongator."""
        n, _ = self.level_size(k, l)
        nnz = ctypes.c_int()
        sm = ctypes.c_int()
        om = ctypes.c_double()
        _check(lib().bcmg_level_Pinfo(self.h, k, l, ctypes.byref(nnz), ctypes.byref(sm),
                                      ctypes.byref(om)))
        row = np.empty(n + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value)
        _check(lib().bcmg_level_Pcsr(self.h, k, l, row.ctypes.data_as(_i32p),
                                     col.ctypes.data_as(_i32p), val.ctypes.data_as(_f64p)))
        return row, col, val, bool(sm.value), om.value

    def level_DM(self, k, l):
        n, _ = self.level_size(k, l)
        D = np.empty(n)
        M = np.empty(n)
        hd = ctypes.c_int()
        hm = ctypes.c_int()
        _check(lib().bcmg_level_DM(self.h, k, l, D.ctypes.data_as(_f64p),
                                   M.ctypes.data_as(_f64p), ctypes.byref(hd), ctypes.byref(hm)))
        return (D if hd.value else None), (M if hm.value else None)

    def fcg(self, b_dev, x_dev):
        """FCG v3 from x0 = 0 on device arrays; returns (iterations, residual)."""
        it = ctypes.c_int()
        res = ctypes.c_double()
        _check(lib().bcmg_fcg(self.ctx.h, self.h, b_dev.ptr, x_dev.ptr, ctypes.byref(it),
                              ctypes.byref(res)))
        return it.value, res.value

    def fcg_multi(self, bs, xs):
        """FCG on several right-hand sides at once (device arrays); returns
        (iterations, residuals, status) arrays, column j equal to fcg(bs[j])."""
        k = len(bs)
        if len(xs) != k:
            raise ValueError("need as many outputs as right-hand sides")
        bp = (_vp * k)(*[b.ptr.value for b in bs])
        xp = (_vp * k)(*[x.ptr.value for x in xs])
        it = np.zeros(k, np.int32)
        res = np.zeros(k)
        st = np.zeros(k, np.int32)
        _check(lib().bcmg_fcg_multi(self.ctx.h, self.h, k, bp, xp, it.ctypes.data_as(_i32p),
                                    res.ctypes.data_as(_f64p), st.ctypes.data_as(_i32p)))
        return it, res, st

    def precond(self, r_dev, x_dev):
        _check(lib().bcmg_precond_apply(self.ctx.h, self.h, r_dev.ptr, x_dev.ptr))

    def set_precision(self, precision):
        """Precision of the preconditioner: PREC_FP64 (the drop-in), PREC_FP32
        or PREC_BF16 (bfloat16 operator values on levels >= 1); FCG stays
        FP64 with a true-residual check in the reduced precisions."""
        _check(lib().bcmg_set_precision(self.h, int(precision)))

    def restarts(self):
        """True-residual restarts made by the last fcg()."""
        r = ctypes.c_int()
        _check(lib().bcmg_fcg_restarts(self.h, ctypes.byref(r)))
        return r.value


class Server:
    """Serving loop with continuous batching over cached composites
    (bcmg_server_*): requests are admitted into column slots of their
    operator between iterations; each result equals Boot.fcg on its b."""

    def __init__(self, ctx, hbm_budget_bytes=64 << 30, slots_per_operator=8):
        self.ctx = ctx
        self.h = _vp()
        self._io = {}
        self._boots = []
        _check(lib().bcmg_server_create(ctx.h, ctypes.c_size_t(int(hbm_budget_bytes)),
                                        int(slots_per_operator), ctypes.byref(self.h)))
        ctx._adopt(self)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().bcmg_server_destroy(self.h)
        self.h = None

    _release = close

    def __del__(self):
        # at interpreter exit the module's globals (lib) may already be gone
        # and the process is ending: a failed release is not reported there;
        # explicit free()/close() calls still raise
        try:
            self.close()
        except Exception:
            pass

    def add(self, boot):
        op = ctypes.c_int()
        _check(lib().bcmg_server_add_operator(self.h, boot.h, ctypes.byref(op)))
        self._boots.append(boot)
        return op.value

    def submit(self, op, b_dev, x_dev):
        rid = ctypes.c_longlong()
        _check(lib().bcmg_server_submit(self.h, int(op), b_dev.ptr, x_dev.ptr,
                                        ctypes.byref(rid)))
        self._io[rid.value] = (b_dev, x_dev)
        return rid.value

    def step(self):
        c = ctypes.c_int()
        _check(lib().bcmg_server_step(self.h, ctypes.byref(c)))
        return c.value

    def result(self, rid):
        """None while pending, else dict(iterations, residual, status, latency_ms)."""
        done, it, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        res, lat = ctypes.c_double(), ctypes.c_double()
        _check(lib().bcmg_server_result(self.h, ctypes.c_longlong(rid), ctypes.byref(done),
                                        ctypes.byref(it), ctypes.byref(res), ctypes.byref(st),
                                        ctypes.byref(lat)))
        if not done.value:
            return None
        self._io.pop(rid, None)
        return dict(iterations=it.value, residual=res.value, status=st.value,
                    latency_ms=lat.value)

    def completed(self, cap=4096):
        """ids of requests finished since the last call"""
        ids = np.zeros(cap, np.int64)
        cnt = ctypes.c_int()
        _check(lib().bcmg_server_completed(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                           cap, ctypes.byref(cnt)))
        return ids[:cnt.value].tolist()

    def info(self):
        q, r, nl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        by = ctypes.c_size_t()
        _check(lib().bcmg_server_info(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(by),
                                      ctypes.byref(nl)))
        return dict(queued=q.value, running=r.value, bytes=by.value, lanes=nl.value)


def solve(A, p, rhs):
    """solve() of the reference (src/SOLVER/solver.cu:20-129) on host arrays:
    setup + FCG from x0 = 0. Returns (x, info) with info = dict(iterations,
    residual, setup_ms, solve_ms)."""
    row, col, val = A
    row, col, val = _as(row, np.int32), _as(col, np.int32), _as(val, np.float64)
    n = len(row) - 1
    rhs = _as(rhs, np.float64)
    x = np.empty(n)
    it = ctypes.c_int()
    res = ctypes.c_double()
    tsu = ctypes.c_double()
    tso = ctypes.c_double()
    _check(lib().bcmg_solve(n, len(col), row.ctypes.data_as(_i32p), col.ctypes.data_as(_i32p),
                            val.ctypes.data_as(_f64p), ctypes.byref(p),
                            rhs.ctypes.data_as(_f64p), x.ctypes.data_as(_f64p),
                            ctypes.byref(it), ctypes.byref(res), ctypes.byref(tsu),
                            ctypes.byref(tso)))
    return x, dict(iterations=it.value, residual=res.value, setup_ms=tsu.value,
                   solve_ms=tso.value)


# --------------------------------------------------------- building blocks

def spmv(ctx, A, x_dev, y_dev):
    _check(lib().bcmg_spmv(ctx.h, A.h, x_dev.ptr, y_dev.ptr))


def residual(ctx, A, x_dev, f_dev, y_dev):
    _check(lib().bcmg_residual(ctx.h, A.h, x_dev.ptr, f_dev.ptr, y_dev.ptr))


def relax(ctx, A, d_dev, f_dev, u_dev, u2_dev, sweeps):
    _check(lib().bcmg_relax(ctx.h, A.h, d_dev.ptr, f_dev.ptr, u_dev.ptr, u2_dev.ptr, sweeps))


# rows of the static device array that holds bcmg_relax_noalloc's second iterate
RELAX_NOALLOC_MAX_ROWS = 4194304


def relax_noalloc(ctx, A, d_dev, f_dev, u_dev, sweeps):
    """relax() without a workspace and without device allocation; operators
    of at most RELAX_NOALLOC_MAX_ROWS rows."""
    _check(lib().bcmg_relax_noalloc(ctx.h, A.h, d_dev.ptr, f_dev.ptr, u_dev.ptr, sweeps))


def weights(ctx, A, w_dev, W_dev):
    _check(lib().bcmg_weights(ctx.h, A.h, w_dev.ptr, W_dev.ptr))


def matching(ctx, A, W_dev, Mraw_dev, Msym_dev):
    nu = ctypes.c_int()
    _check(lib().bcmg_matching(ctx.h, A.h, W_dev.ptr, Mraw_dev.ptr, Msym_dev.ptr,
                               ctypes.byref(nu)))
    return nu.value


def aggregate(ctx, n, Msym_dev, w_dev, agg_dev, p_dev):
    nc = ctypes.c_int()
    _check(lib().bcmg_aggregate(ctx.h, n, Msym_dev.ptr, w_dev.ptr, agg_dev.ptr, p_dev.ptr,
                                ctypes.byref(nc)))
    return nc.value


def use_graphs(ctx, enable):
    """Replay FCG iterations from captured HIP graphs (default) or launch
    every kernel from the host."""
    _check(lib().bcmg_use_graphs(ctx.h, 1 if enable else 0))


def time_smoother(ctx, enable):
    """Bracket every level-0 Jacobi sweep (from a non-zero iterate) with
    HIP events until disabled; read with smoother_timing()."""
    _check(lib().bcmg_time_smoother(ctx.h, 1 if enable else 0))


def time_setup(ctx, enable):
    """Time the phases of the next setups with HIP events (Boot.setup_timing)."""
    _check(lib().bcmg_time_setup(ctx.h, 1 if enable else 0))


def smoother_timing(ctx):
    tot = ctypes.c_double()
    cnt = ctypes.c_int()
    _check(lib().bcmg_smoother_timing(ctx.h, ctypes.byref(tot), ctypes.byref(cnt)))
    return tot.value, cnt.value


def time_spmv(ctx, A, x_dev, y_dev, reps):
    ms = ctypes.c_double()
    _check(lib().bcmg_time_spmv(ctx.h, A.h, x_dev.ptr, y_dev.ptr, reps, ctypes.byref(ms)))
    return ms.value


def spgemm(ctx, Lm, Rm, m_right, slots_per_pass=0):
    """C = L R for host CSR triples (uploaded here); returns C as a Csr
    handle (rectangular: n = L's rows, columns 0..m_right-1)."""
    lrow, lcol, lval = _as(Lm[0], np.int32), _as(Lm[1], np.int32), _as(Lm[2], np.float64)
    rrow, rcol, rval = _as(Rm[0], np.int32), _as(Rm[1], np.int32), _as(Rm[2], np.float64)
    nl, ml = len(lrow) - 1, len(rrow) - 1
    dev = [DeviceArray.from_numpy(ctx, a) for a in (lrow, lcol, lval, rrow, rcol, rval)]
    h = _vp()
    try:
        _check(lib().bcmg_spgemm(ctx.h, nl, ml, m_right, *[d.ptr for d in dev],
                                 ctypes.c_longlong(slots_per_pass), ctypes.byref(h)))
    finally:
        for d in dev:
            d.free()
    return Csr(ctx, handle=h)


def galerkin(ctx, A, agg_dev, p_dev, nc):
    h = _vp()
    _check(lib().bcmg_galerkin(ctx.h, A.h, agg_dev.ptr, p_dev.ptr, nc, ctypes.byref(h)))
    return Csr(ctx, handle=h)
