"""Open-loop load generator for the serving loop (bcmg_server_*).

The serving workload: a pool of 8 cached operators, requests arriving as a
Poisson process at a given rate, each naming an operator (uniform) and a
right-hand side uniform[-1, 1] from its own SplitMix64 stream (request k of a
trace uses stream seed0 + k). Right-hand sides are made on the host before
a replay starts; latency runs from a request's scheduled arrival to the
moment its solution is back in host memory, so it includes queueing, the
upload of b and the download of x, and an overloaded server shows it in p99.

The same arrival trace is replayed through two servers:
- `served`: bcmg_server (continuous batching into column slots, one stream
  per operator, captured iterations), in the composites' precision;
- `plain`: one solve at a time with Boot.fcg in arrival order - the
  reference's own way of solving (one solve() per right-hand side).
After a replay every result's true residual ||b - A x|| / ||b|| is computed
on the host (outside the timing) for the requests selected by `check_every`.
"""
import time

import numpy as np

from . import operators as ops

# (name, generator): the pool's 8 operators
POOL = [
    ("poisson2d_256", lambda: ops.poisson2d(256)),
    ("aniso2d_1024_eps0.01", lambda: ops.aniso2d(1024, 0.01)),
    ("poisson2d_1024", lambda: ops.poisson2d(1024)),
    ("poisson3d_32", lambda: ops.poisson3d(32)),
    ("poisson3d_64", lambda: ops.poisson3d(64)),
    ("poisson3d_96", lambda: ops.poisson3d(96)),
    ("poisson3d_128", lambda: ops.poisson3d(128)),
    ("lognormal3d_64_s2_4", lambda: ops.lognormal3d(64, 4.0, seed=0)),
]


class Pool:
    """The operators, their cached composites, scipy copies for residual
    checks and free lists of device vectors per operator."""

    def __init__(self, bc, ctx, settings, precision, names=None):
        self.bc, self.ctx = bc, ctx
        self.ops = []
        for name, gen in POOL:
            if names is not None and name not in names:
                continue
            A = gen()
            n = len(A[0]) - 1
            dA = bc.Csr(ctx, A)
            B = bc.Boot(ctx, dA, bc.make_params(**settings))
            B.set_precision(precision)
            self.ops.append(dict(name=name, n=n, nnz=len(A[1]), A=dA, B=B, S=ops.to_scipy(A),
                                 free=[], bfree=[]))

    def set_precision(self, precision):
        for o in self.ops:
            o["B"].set_precision(precision)

    def out(self, op):
        o = self.ops[op]
        return o["free"].pop() if o["free"] else self.bc.DeviceArray(self.ctx, o["n"])

    def give_back(self, op, x):
        self.ops[op]["free"].append(x)

    def rhs_buffer(self, op):
        o = self.ops[op]
        return o["bfree"].pop() if o["bfree"] else self.bc.DeviceArray(self.ctx, o["n"])

    def give_back_rhs(self, op, b):
        self.ops[op]["bfree"].append(b)

    def rhs(self, opsel, seed0):
        """Host right-hand sides of a trace: request k uses stream seed0 + k."""
        return [ops.uniform_pm1(seed0 + k, self.ops[int(op)]["n"]) for k, op in enumerate(opsel)]


def trace(rate, n_ops, max_requests, max_seconds, seed=0):
    """Poisson arrivals: (times in s, operator ids)."""
    rng = np.random.default_rng(seed)
    horizon = min(max_seconds, max_requests / rate)
    t = np.cumsum(rng.exponential(1.0 / rate, int(rate * horizon * 2) + 16))
    t = t[t < horizon]
    return t, rng.integers(0, n_ops, len(t))


def _stats(lat_ms, t_first, t_last, n, residuals):
    lat = np.asarray(lat_ms)
    span = max(t_last - t_first, 1e-9)
    return {"requests": int(n), "solves_per_s": n / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "residuals_checked": len(residuals),
            "max_true_rel_residual": float(max(residuals)) if residuals else None}


def _residuals(pool, opsel, bh, xh, check_every):
    out = []
    for k in range(0, len(bh), check_every):
        S = pool.ops[int(opsel[k])]["S"]
        out.append(float(np.linalg.norm(bh[k] - S @ xh[k]) / np.linalg.norm(bh[k])))
    return out


def run_served(pool, server, arrivals, opsel, bh, check_every=1):
    """Replay a trace through the serving loop (b uploaded at submission, x
    downloaded at completion); returns latency statistics."""
    N = len(arrivals)
    rid_info = {}
    lat = [None] * N
    xh = [None] * N
    i = done = 0
    t0 = time.perf_counter()
    t_last = 0.0
    while done < N:
        now = time.perf_counter() - t0
        while i < N and arrivals[i] <= now:
            op = int(opsel[i])
            x = pool.out(op)
            b = pool.rhs_buffer(op)
            b.upload(bh[i])
            rid = server.submit(op, b, x)
            rid_info[rid] = (i, op, x, b)
            i += 1
        if rid_info:
            server.step()
            for rid in server.completed():
                k, op, x, b = rid_info.pop(rid)
                r = server.result(rid)
                assert r is not None and r["status"] == 0, r
                xh[k] = x.numpy()
                t = time.perf_counter() - t0
                lat[k] = (t - arrivals[k]) * 1e3
                pool.give_back(op, x)
                pool.give_back_rhs(op, b)
                done += 1
                t_last = t
        elif i < N:
            time.sleep(max(0.0, min(1e-4, arrivals[i] - now)))
    return _stats(lat, arrivals[0], t_last, N, _residuals(pool, opsel, bh, xh, check_every))


def run_plain(pool, arrivals, opsel, bh, check_every=1):
    """Replay a trace one solve at a time (upload b, Boot.fcg, download x, in
    arrival order)."""
    N = len(arrivals)
    lat = [None] * N
    xh = [None] * N
    t0 = time.perf_counter()
    t_last = 0.0
    for k in range(N):
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        op = int(opsel[k])
        o = pool.ops[op]
        x = pool.out(op)
        b = pool.rhs_buffer(op)
        b.upload(bh[k])
        o["B"].fcg(b, x)
        xh[k] = x.numpy()
        t = time.perf_counter() - t0
        lat[k] = (t - arrivals[k]) * 1e3
        pool.give_back(op, x)
        pool.give_back_rhs(op, b)
        t_last = t
    return _stats(lat, arrivals[0], t_last, N, _residuals(pool, opsel, bh, xh, check_every))
