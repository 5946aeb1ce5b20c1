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
