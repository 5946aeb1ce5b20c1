"""Open-loop load generator for the serving loop (bcmg_server_*).

The serving workload: a pool of 8 cached operators, requests arriving as a
Poisson process at a given rate, each naming an operator (uniform) and a
right-hand side uniform[-1, 1] from a seeded SplitMix64 stream. Latency is
measured from a request's scheduled arrival to its completion (queueing
included), so an overloaded server shows it in p99.

The same arrival trace is replayed through two servers:
- `served`: bcmg_server (continuous batching into column slots, one stream
  per operator, captured iterations), in the composites' precision;
- `plain`: one solve at a time with Boot.fcg in arrival order - the
  reference's own way of solving (one solve() per right-hand side).
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
RHS_PER_OP = 4


class Pool:
    """The operators, their cached composites and device right-hand sides."""

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
            rhs = [bc.DeviceArray.from_numpy(ctx, ops.uniform_pm1(s, n))
                   for s in range(RHS_PER_OP)]
            self.ops.append(dict(name=name, n=n, nnz=len(A[1]), A=dA, B=B, rhs=rhs, free=[]))

    def set_precision(self, precision):
        for o in self.ops:
            o["B"].set_precision(precision)

    def out(self, op):
        o = self.ops[op]
        return o["free"].pop() if o["free"] else self.bc.DeviceArray(self.ctx, o["n"])

    def give_back(self, op, x):
        self.ops[op]["free"].append(x)


def trace(rate, n_ops, max_requests, max_seconds, seed=0):
    """Poisson arrivals: (times in s, operator ids, rhs ids)."""
    rng = np.random.default_rng(seed)
    horizon = min(max_seconds, max_requests / rate)
    t = np.cumsum(rng.exponential(1.0 / rate, int(rate * horizon * 2) + 16))
    t = t[t < horizon]
    return t, rng.integers(0, n_ops, len(t)), rng.integers(0, RHS_PER_OP, len(t))


def _stats(lat_ms, t_first, t_last, n):
    lat = np.asarray(lat_ms)
    span = max(t_last - t_first, 1e-9)
    return {"requests": int(n), "solves_per_s": n / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


def run_served(pool, server, arrivals, opsel, rhsel):
    """Replay a trace through the serving loop; returns latency statistics."""
    N = len(arrivals)
    rid_info = {}
    lat = [None] * N
    i = done = 0
    t0 = time.perf_counter()
    t_last = 0.0
    while done < N:
        now = time.perf_counter() - t0
        while i < N and arrivals[i] <= now:
            op = int(opsel[i])
            x = pool.out(op)
            rid = server.submit(op, pool.ops[op]["rhs"][int(rhsel[i])], x)
            rid_info[rid] = (i, op, x)
            i += 1
        if rid_info:
            server.step()
            for rid in server.completed():
                t = time.perf_counter() - t0
                k, op, x = rid_info.pop(rid)
                r = server.result(rid)
                assert r is not None and r["status"] == 0, r
                lat[k] = (t - arrivals[k]) * 1e3
                pool.give_back(op, x)
                done += 1
                t_last = t
        elif i < N:
            time.sleep(max(0.0, min(1e-4, arrivals[i] - now)))
    return _stats(lat, arrivals[0], t_last, N)


def run_plain(pool, arrivals, opsel, rhsel):
    """Replay a trace one solve at a time (Boot.fcg in arrival order)."""
    N = len(arrivals)
    lat = [None] * N
    t0 = time.perf_counter()
    t_last = 0.0
    for k in range(N):
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        op = int(opsel[k])
        o = pool.ops[op]
        x = pool.out(op)
        o["B"].fcg(o["rhs"][int(rhsel[k])], x)
        t = time.perf_counter() - t0
        lat[k] = (t - arrivals[k]) * 1e3
        pool.give_back(op, x)
        t_last = t
    return _stats(lat, arrivals[0], t_last, N)
