"""Shared helpers of the test suite (operators, settings presets)."""
import numpy as np

from bootcmatchg_amd import operators as ops


def bits(a):
    """View float64 data as int64 so that equality is bit equality."""
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


# small operators the oracle handles in well under a second
OPERATORS = {
    "poisson3d_12": lambda: ops.poisson3d(12),
    "poisson3d_17": lambda: ops.poisson3d(17),
    "aniso2d_48": lambda: ops.aniso2d(48, 0.01),
    "poisson2d_33": lambda: ops.poisson2d(33),
    "lognormal3d_10": lambda: ops.lognormal3d(10, 4.0, seed=3),
    "lognormal3d_14": lambda: ops.lognormal3d(14, 4.0, seed=7),
}

# settings presets named by what they exercise (14 fields, file order)
SETTINGS = {
    "V_L1_2sweeps": dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, cycle_type=0,
                         coarse_solver=4, relax_type=4, relaxnumber_coarse=16,
                         prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000, rtol=1e-8),
    "W_L1_3sweeps": dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=3, cycle_type=2,
                         coarse_solver=4, relax_type=4, relaxnumber_coarse=16,
                         prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000, rtol=1e-8),
    "H_jacobi_1sweep": dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=1, cycle_type=1,
                            coarse_solver=0, relax_type=0, relaxnumber_coarse=20,
                            prerelax_sweeps=1, postrelax_sweeps=1, itnlim=2000, rtol=1e-8),
    "boot3_mult": dict(solver_type=0, max_hrc=3, conv_ratio=0.5, aggrsweeps=2, cycle_type=0,
                       coarse_solver=4, relax_type=4, relaxnumber_coarse=16,
                       prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000, rtol=1e-8),
    "boot3_sym": dict(solver_type=1, max_hrc=3, conv_ratio=0.3, aggrsweeps=2, cycle_type=0,
                      coarse_solver=4, relax_type=4, relaxnumber_coarse=16,
                      prerelax_sweeps=2, postrelax_sweeps=2, itnlim=2000, rtol=1e-8),
    "boot2_add_mixed": dict(solver_type=2, max_hrc=2, conv_ratio=0.5, aggrsweeps=2, cycle_type=1,
                            coarse_solver=0, relax_type=4, relaxnumber_coarse=10,
                            prerelax_sweeps=3, postrelax_sweeps=2, itnlim=2000, rtol=1e-8),
    "K_L1_1sweep_agg": dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=1, cycle_type=3,
                            coarse_solver=4, relax_type=4, relaxnumber_coarse=16,
                            prerelax_sweeps=2, postrelax_sweeps=2, itnlim=2000, rtol=1e-8),
}


# FCG on the GPU against the oracle's FCG (both FP64). The kernels sum rows in
# another order than the oracle (fused multiply-adds, lane-strided partial
# sums), so the two runs' iterates differ by rounding only: with equal
# iteration counts the solutions agree to SOL_RTOL (observed: 1e-16 to 1e-15).
# The counts may differ by up to 2 (the reference's own bar, enforced against
# the oracle's count); then both runs stopped at rtol, and the GPU solution's
# true residual is checked against rtol instead.
SOL_RTOL = 1e-12


def check_fcg_vs_oracle(x, it, xo, it_o, A, b, rtol, label=""):
    rel = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    print("%s: iterations %d (oracle %d), relative difference to the oracle %.2e"
          % (label, it, it_o, rel), flush=True)
    assert abs(it - it_o) <= 2, (label, it, it_o)
    if it == it_o:
        assert rel <= SOL_RTOL, (label, rel)
    else:
        S = ops.to_scipy(A)
        tr = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
        assert tr <= rtol * (1 + 1e-6), (label, tr)
    return rel
