"""The serving loop (bcmg_server_*): continuous batching over cached
composites. A request's result (x bits, iterations, residual) must be the
same served alone and served in a mix - other operators, other requests
arriving while it runs, slots and device memory running short - and equal
to the single solve of its right-hand side."""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops

from common import OPERATORS, SETTINGS, bits

pytestmark = pytest.mark.gpu

CASES = [("poisson3d_17", "V_L1_2sweeps", bc.PREC_FP64),
         ("aniso2d_48", "W_L1_3sweeps", bc.PREC_FP32),
         ("lognormal3d_14", "boot3_mult", bc.PREC_BF16),
         ("poisson2d_33", "H_jacobi_1sweep", bc.PREC_FP32)]


@pytest.fixture(scope="module")
def ctx():
    c = bc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def boots(ctx):
    out = []
    for name, setting, prec in CASES:
        A = OPERATORS[name]()
        B = bc.Boot(ctx, bc.Csr(ctx, A), bc.make_params(**SETTINGS[setting]))
        B.set_precision(prec)
        out.append((B, len(A[0]) - 1))
    return out


def requests(boots, per_op=5):
    reqs = []
    for op, (_, n) in enumerate(boots):
        for q in range(per_op):
            b = ops.uniform_pm1(1000 * op + q, n) * (10.0 ** (q % 3 - 1))
            if op == 1 and q == 3:
                b = np.zeros(n)
            reqs.append((op, b))
    return reqs


def serve(ctx, boots, reqs, arrivals, slots=8, budget=64 << 30):
    """Run the loop; request i is submitted at step arrivals[i]."""
    S = bc.Server(ctx, budget, slots)
    ids = [S.add(B) for B, _ in boots]
    assert ids == list(range(len(boots)))
    xs = [bc.DeviceArray(ctx, boots[op][1]) for op, _ in reqs]
    rid = [None] * len(reqs)
    out = [None] * len(reqs)
    step = 0
    while any(o is None for o in out):
        for i, (op, b) in enumerate(reqs):
            if rid[i] is None and arrivals[i] <= step:
                rid[i] = S.submit(op, bc.DeviceArray.from_numpy(ctx, b), xs[i])
        S.step()
        for i in range(len(reqs)):
            if rid[i] is not None and out[i] is None:
                r = S.result(rid[i])
                if r is not None:
                    out[i] = (r, xs[i].numpy())
        step += 1
        assert step < 20000
    info = S.info()
    S.close()
    return out, info


def test_alone_equals_mix_equals_single(ctx, boots):
    reqs = requests(boots)
    alone = [serve(ctx, boots, [r], [0])[0][0] for r in reqs]
    rng = np.random.default_rng(0)
    arrivals = rng.integers(0, 40, len(reqs))      # many join while others run
    mixed, _ = serve(ctx, boots, reqs, arrivals, slots=3)
    for i, ((op, b), (ra, xa), (rm, xm)) in enumerate(zip(reqs, alone, mixed)):
        assert (ra["iterations"], ra["residual"], ra["status"]) == \
               (rm["iterations"], rm["residual"], rm["status"]), i
        assert np.array_equal(bits(xa), bits(xm)), i
        B, n = boots[op]
        if not np.any(b):
            assert rm["status"] == bc.EZERORHS and not np.any(xm)
            continue
        x1 = bc.DeviceArray(ctx, n)
        it1, res1 = B.fcg(bc.DeviceArray.from_numpy(ctx, b), x1)
        assert (it1, res1) == (rm["iterations"], rm["residual"]), i
        assert np.array_equal(bits(x1.numpy()), bits(xm)), i
        assert rm["latency_ms"] > 0


def test_budget_closes_idle_lanes(ctx, boots):
    """A device-memory budget one byte short of every operator's one-slot
    lane at once: lanes open as operators get work and idle ones close to
    make room; the results do not change."""
    reqs = requests(boots, per_op=2)
    full, _ = serve(ctx, boots, reqs, [0] * len(reqs), slots=2)
    lane_bytes = []
    for op in range(len(boots)):
        _, info = serve(ctx, boots, [reqs[2 * op]], [0], slots=1)
        lane_bytes.append(info["bytes"])
    budget = sum(lane_bytes) - 1
    tight, info = serve(ctx, boots, reqs, [0] * len(reqs), slots=1, budget=budget)
    assert info["bytes"] <= budget
    assert info["lanes"] < len(boots)
    for (rf, xf), (rt, xt) in zip(full, tight):
        assert rf["iterations"] == rt["iterations"] and rf["residual"] == rt["residual"]
        assert np.array_equal(bits(xf), bits(xt))


def _hip():
    import ctypes
    return ctypes.CDLL("libamdhip64.so")


def device_bytes_in_use():
    import ctypes
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert _hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return total.value - free.value


# four operators of about 90 000 rows each, all of different sizes (so that
# no two lanes share the pool's size classes)
SIMILAR = [lambda: ops.poisson3d(44), lambda: ops.poisson3d(46), lambda: ops.poisson2d(300),
           lambda: ops.aniso2d(310, 0.01)]


def test_closed_lanes_give_their_memory_back(ctx):
    """A budget that holds one full lane at a time: requests of another
    operator close idle lanes, and a closed lane's device memory leaves the
    process, so the memory in use grows by at most the budget (not by the
    sum of every lane ever opened)."""
    mats = [g() for g in SIMILAR]
    prm = bc.make_params(**SETTINGS["V_L1_2sweeps"])
    # lane sizes, measured in the module's context
    sizes = []
    for A in mats:
        B = bc.Boot(ctx, bc.Csr(ctx, A), prm)
        n = len(A[0]) - 1
        _, info = serve(ctx, [(B, n)], [(0, ops.uniform_pm1(1, n))], [0], slots=8)
        sizes.append(info["bytes"])
        B.free()
    assert max(sizes) < 1.1 * min(sizes), sizes
    budget = int(1.3 * max(sizes))
    c2 = bc.Context(0)
    boots = []
    for A in mats:
        B = bc.Boot(c2, bc.Csr(c2, A), prm)
        boots.append((B, len(A[0]) - 1))
    c2.sync()
    used0 = device_bytes_in_use()
    reqs = [(op % 4, ops.uniform_pm1(50 + op, boots[op % 4][1])) for op in range(12)]
    arrivals = [40 * i for i in range(12)]      # one operator after the other
    out, info = serve(c2, boots, reqs, arrivals, slots=8, budget=budget)
    used1 = device_bytes_in_use()
    print("lane bytes %s, budget %d, device memory grown by %d" % (sizes, budget, used1 - used0))
    # lanes narrower than 8 slots may share the budget; what they hold, and
    # what the device gave the process, stays within it
    assert info["bytes"] <= budget
    assert used1 - used0 <= budget + (32 << 20), (used1 - used0, budget)
    for (op, b), (r, x) in zip(reqs, out):
        assert r["status"] == bc.OK
        B, n = boots[op]
        x1 = bc.DeviceArray(c2, n)
        it1, res1 = B.fcg(bc.DeviceArray.from_numpy(c2, b), x1)
        assert (it1, res1) == (r["iterations"], r["residual"])
        assert np.array_equal(bits(x1.numpy()), bits(x))
    c2.close()


def test_b_written_on_the_callers_stream_right_before_submit():
    """A context on the caller's stream: b is copied into place on that
    stream (behind a queue of large copies) just before submit and step; the
    served result equals the single solve of b bit for bit."""
    import ctypes
    hip = _hip()
    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1)) == 0
    try:
        c2 = bc.Context(0, stream=s.value)
        A = OPERATORS["poisson3d_17"]()
        n = len(A[0]) - 1
        B = bc.Boot(c2, bc.Csr(c2, A), bc.make_params(**SETTINGS["V_L1_2sweeps"]))
        b = ops.uniform_pm1(77, n)
        src = bc.DeviceArray.from_numpy(c2, b)
        dst = bc.DeviceArray.from_numpy(c2, np.zeros(n))
        big = 64 << 20
        bsrc = bc.DeviceArray(c2, big)
        bdst = bc.DeviceArray(c2, big)
        c2.sync()
        S = bc.Server(c2, 64 << 30, 4)
        S.add(B)
        x = bc.DeviceArray(c2, n)
        for _ in range(4):                    # keep the stream busy
            assert hip.hipMemcpyAsync(bdst.ptr, bsrc.ptr, ctypes.c_size_t(8 * big),
                                      ctypes.c_int(3), s) == 0
        assert hip.hipMemcpyAsync(dst.ptr, src.ptr, ctypes.c_size_t(8 * n), ctypes.c_int(3),
                                  s) == 0
        rid = S.submit(0, dst, x)
        r = None
        for _ in range(20000):
            S.step()
            r = S.result(rid)
            if r is not None:
                break
        assert r is not None and r["status"] == bc.OK
        x1 = bc.DeviceArray(c2, n)
        it1, res1 = B.fcg(src, x1)
        assert (it1, res1) == (r["iterations"], r["residual"])
        assert np.array_equal(bits(x1.numpy()), bits(x.numpy()))
        S.close()
        c2.close()
    finally:
        hip.hipStreamDestroy(s)


def test_budget_below_any_lane_fails_the_requests(ctx, boots):
    """No lane of an operator fits a tiny budget even with every other lane
    closed: its requests end with BCMG_ENOMEM instead of staying queued."""
    S = bc.Server(ctx, 1024, 4)
    S.add(boots[0][0])
    n = boots[0][1]
    x = bc.DeviceArray(ctx, n)
    rids = [S.submit(0, bc.DeviceArray.from_numpy(ctx, ops.uniform_pm1(q, n)), x)
            for q in range(3)]
    done = S.step()
    assert done == 3
    for rid in rids:
        r = S.result(rid)
        assert r is not None and r["status"] == bc.ENOMEM
    assert S.info()["queued"] == 0 and S.info()["lanes"] == 0
    S.close()


POOL_RTOL = 1e-8
POOL_SETTINGS = dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, aggrtype=0,
                     max_levels=39, cycle_type=0, coarse_solver=4, relax_type=4,
                     relaxnumber_coarse=16, prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000,
                     rtol=POOL_RTOL)
# members of the pool the oracle solves in seconds
POOL_ORACLE = ("poisson3d_32", "poisson3d_64", "poisson2d_256")


@pytest.mark.parametrize("prec", [bc.PREC_FP64, bc.PREC_FP32])
def test_serving_pool_of_eight_operators(ctx, prec):
    """The serving workload's pool (2D 5-point 256^2 and 1024^2 with eps 1
    and 0.01, 3D 7-point 32^3 to 128^3, log-normal 64^3) at the drop-in
    settings with rtol 1e-8: 40 requests with distinct seeded right-hand
    sides, submitted a few at a time while earlier ones run. Every result
    meets rtol on its true residual and equals the single solve of its
    right-hand side bit for bit; on the 32^3, 64^3 and 256^2 members the
    FP64 results also agree with the oracle."""
    from bootcmatchg_amd import serving as sv
    from oracle import oracle as orc
    from common import check_fcg_vs_oracle
    pool = sv.Pool(bc, ctx, POOL_SETTINGS, prec)
    names = [o["name"] for o in pool.ops]
    assert len(names) == 8
    rng = np.random.default_rng(4)
    opsel = rng.integers(0, 8, 40)
    opsel[:8] = np.arange(8)                       # every operator served
    reqs = [(int(op), ops.uniform_pm1(5000 + i, pool.ops[int(op)]["n"]))
            for i, op in enumerate(opsel)]
    arrivals = np.sort(rng.integers(0, 60, len(reqs)))
    boots = [(o["B"], o["n"]) for o in pool.ops]
    print("pool", names, "precision", prec, flush=True)
    out, _ = serve(ctx, boots, reqs, arrivals, slots=16)
    oracle_done = set()
    mats = {}
    for i, ((op, b), (r, x)) in enumerate(zip(reqs, out)):
        o = pool.ops[op]
        assert r["status"] == 0, (i, o["name"], r)
        if op not in mats:
            A = o["A"].download()
            mats[op] = (A, ops.to_scipy(A))
        A, S = mats[op]
        tr = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
        assert tr <= POOL_RTOL * (1 + 1e-3), (i, o["name"], tr)
        x1 = bc.DeviceArray(ctx, o["n"])
        it1, res1 = o["B"].fcg(bc.DeviceArray.from_numpy(ctx, b), x1)
        assert (it1, res1) == (r["iterations"], r["residual"]), (i, o["name"])
        assert np.array_equal(bits(x1.numpy()), bits(x)), (i, o["name"])
        if prec == bc.PREC_FP64 and o["name"] in POOL_ORACLE and o["name"] not in oracle_done:
            oracle_done.add(o["name"])
            rc, xo, ro = orc.solve(A, b, orc.params(**POOL_SETTINGS))
            assert rc == 0
            check_fcg_vs_oracle(x, r["iterations"], xo, ro.iterations, A, b, POOL_RTOL,
                                "served %s" % o["name"])
    if prec == bc.PREC_FP64:
        assert oracle_done == set(POOL_ORACLE)
    for o in pool.ops:
        o["B"].free()
