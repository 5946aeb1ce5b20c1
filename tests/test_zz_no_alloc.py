"""bcmg_relax_noalloc (no device allocation, static scratch) against
bcmg_relax on the same inputs: bit-identical results on two calls in a row,
the second on a smaller operator, and a clear error beyond the scratch size.
"""
import numpy as np
import pytest

import bootcmatchg_amd as bc
from bootcmatchg_amd import operators as ops

from common import bits

pytestmark = pytest.mark.gpu


def l1_divisors(A):
    return np.abs(ops.to_scipy(A)).sum(axis=1).A1


def relax_both(ctx, A, sweeps, seed):
    n = len(A[0]) - 1
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    dA = bc.Csr(ctx, A)
    dd = bc.DeviceArray.from_numpy(ctx, l1_divisors(A))
    df = bc.DeviceArray.from_numpy(ctx, f)
    ua = bc.DeviceArray.from_numpy(ctx, u0)
    u2 = bc.DeviceArray(ctx, n)
    ub = bc.DeviceArray.from_numpy(ctx, u0)
    print("relax vs relax_noalloc: n=%d nnz=%d sweeps=%d columns %d B"
          % (n, dA.nnz, sweeps, dA.column_bytes), flush=True)
    bc.relax(ctx, dA, dd, df, ua, u2, sweeps)
    bc.relax_noalloc(ctx, dA, dd, df, ub, sweeps)
    a, b = ua.numpy(), ub.numpy()
    for arr in (u2, ua, ub, df, dd):
        arr.free()
    dA.free()
    return a, b, u0


def test_two_calls_in_a_row_second_smaller():
    ctx = bc.Context(0)
    try:
        # 160^3 Poisson: 4,096,000 rows (within the scratch), streamed through
        # the 16-bit column copy; an odd sweep count ends in the scratch and
        # is copied back
        a, b, u0 = relax_both(ctx, ops.poisson3d(160), 3, 0)
        assert not np.array_equal(a, u0)
        assert np.array_equal(bits(a), bits(b))
        # smaller, even sweep count: the scratch still holds the first call's
        # iterate beyond these rows
        a, b, u0 = relax_both(ctx, ops.poisson3d(17), 4, 1)
        assert not np.array_equal(a, u0)
        assert np.array_equal(bits(a), bits(b))
    finally:
        ctx.close()


def test_operator_beyond_the_scratch_is_refused():
    ctx = bc.Context(0)
    try:
        n = bc.RELAX_NOALLOC_MAX_ROWS + 1
        A = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
        dA = bc.Csr(ctx, A)
        v = bc.DeviceArray.from_numpy(ctx, np.ones(n))
        with pytest.raises(bc.BcmgError) as e:
            bc.relax_noalloc(ctx, dA, v, v, v, 1)
        assert e.value.code == bc.EINVAL
        assert "static scratch" in str(e.value)
        v.free()
        dA.free()
    finally:
        ctx.close()
