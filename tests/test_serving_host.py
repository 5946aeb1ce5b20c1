"""Host-side pieces of the serving form (no GPU): the open-loop arrival
trace and the latency statistics of bootcmatchg_amd/serving.py."""
import numpy as np

from bootcmatchg_amd import serving as sv


def test_trace_is_a_seeded_poisson_process():
    t, ops_ = sv.trace(1000.0, 8, 5000, 100.0, seed=3)
    assert len(t) == len(ops_)
    assert np.all(np.diff(t) > 0) and t[0] > 0 and t[-1] < 5.0   # horizon 5000 / 1000 s
    gaps = np.diff(np.concatenate([[0.0], t]))
    # exponential gaps with mean 1/rate: mean within 5 %, coefficient of variation ~1
    assert abs(gaps.mean() * 1000.0 - 1.0) < 0.05
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.08
    assert set(np.unique(ops_)) == set(range(8))
    t2, o2 = sv.trace(1000.0, 8, 5000, 100.0, seed=3)
    assert np.array_equal(t, t2) and np.array_equal(ops_, o2)


def test_trace_horizon_caps_requests_and_time():
    t, _ = sv.trace(10.0, 8, 400, 3.0, seed=1)     # 3 s at 10/s
    assert t[-1] < 3.0 and 10 <= len(t) <= 60
    t, _ = sv.trace(4000.0, 8, 400, 3.0, seed=1)   # 400 requests at 4000/s
    assert t[-1] < 0.1 and 300 <= len(t) <= 500


def test_stats():
    st = sv._stats([1.0, 2.0, 3.0, 100.0], 0.0, 2.0, 4, [1e-9, 3e-9])
    assert st["requests"] == 4 and st["solves_per_s"] == 2.0
    assert st["residuals_checked"] == 2 and st["max_true_rel_residual"] == 3e-9
    assert st["p50_ms"] == 2.5 and 97.0 < st["p99_ms"] <= 100.0


def test_pool_lists_eight_operators():
    assert len(sv.POOL) == 8 and len({n for n, _ in sv.POOL}) == 8
