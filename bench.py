"""Benchmark of the AMG-preconditioned FCG drop-in path on one MI355X.

Workload (one step = one complete solve(): bootstrap setup of the AMG
hierarchy + FCG v3 to rtol 1e-8, as the reference's solve() does,
src/SOLVER/solver.cu:20-129): 3D 7-point Poisson on a 160^3 grid
(4,096,000 rows, 28,518,400 nnz, int32 indices, FP64 values), b uniform[-1,1]
from SplitMix64 stream 0, settings 0,1,0.8,2,0,39,0,4,4,16,4,4,2000,1e-8.
The operator and b are resident in HBM before the timed region.

Printed JSON line: value = solves/s of the whole FP64 drop-in step; extras
give the setup/solve split, the rate with a cached hierarchy, the dominant
kernel's roofline (level-0 Jacobi/L1 sweep, timed with HIP events inside the
timed region), the serial CPU oracle on the same workload, and the precision
form: the same step with the hierarchy converted to FP32 (or bfloat16 values
on levels >= 1) inside FP64 FCG with a true-residual check, on this workload
and on the bootstrap composite workload (cell-coefficient diffusion on 128^3,
k = exp(g), g ~ N(0, 4), max_hrc 3, rho 0.5); the multi-right-hand-side
form (32 columns through every launch, anisotropic 2048^2, W-cycle); and the
serving loop (8 cached operators, Poisson arrivals, continuous batching)
against one FP64 solve at a time on the same arrival trace; the small
case the reference runs on a CPU (32^3, b = 1, rtol 1e-6) on the GPU and on
the serial oracle; and smoothed
aggregation (aggr_type 1) against the unsmoothed drop-in on 32^3, 64^3,
100^3 and 160^3 Poisson.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("solves/s & p50/p99 latency, mixed vs FP64 AMG-FCG at rtol 1e-8; SpMV GB/s")
HBM_PEAK_GBS = 8000.0
GRID = 160
SETTINGS = dict(solver_type=0, max_hrc=1, conv_ratio=0.8, aggrsweeps=2, aggrtype=0,
                max_levels=39, cycle_type=0, coarse_solver=4, relax_type=4,
                relaxnumber_coarse=16, prerelax_sweeps=4, postrelax_sweeps=4, itnlim=2000,
                rtol=1e-8)


# Measured HBM bytes per level-0 sweep launch at 160^3 (rocprofv3 FETCH_SIZE
# x 1.9987 + WRITE_SIZE x 1.000, factors from a plain copy kernel of known
# size; summary: profiles/jacobi_sweep_poisson160_traffic.md). Re-measured
# whenever the sweep kernel changes; reported only for the 160^3 workload
# streamed in the stencil-pattern column form (the library's choice for it).
SWEEP_TRAFFIC_BYTES_160 = 372.3e6


C2_GRID = 128
C2_SETTINGS = dict(SETTINGS, max_hrc=3, conv_ratio=0.5)
PREC_NAMES = {0: "fp64", 1: "fp32", 2: "bf16"}


def run_form(bc, ctx, dA, db, dx, prm, prec, steps, warmup):
    """`steps` timed full solves (setup, precision conversion, FCG, free)
    after `warmup` untimed ones, then `steps` FCG solves on one cached
    hierarchy. Returns the figures and the last solution."""
    lat, setup, solve = [], [], []
    it = restarts = 0
    for k in range(warmup + steps):
        ctx.sync()
        t0 = time.perf_counter()
        B = bc.Boot(ctx, dA, prm)
        if prec:
            B.set_precision(prec)
        t1 = time.perf_counter()
        it, _ = B.fcg(db, dx)
        t2 = time.perf_counter()
        restarts = B.restarts()
        B.free()
        if k >= warmup:
            lat.append((t2 - t0) * 1e3)
            setup.append((t1 - t0) * 1e3)
            solve.append((t2 - t1) * 1e3)
    B = bc.Boot(ctx, dA, prm)
    if prec:
        B.set_precision(prec)
    B.fcg(db, dx)
    ctx.sync()
    tc0 = time.perf_counter()
    for _ in range(steps):
        B.fcg(db, dx)
    ctx.sync()
    tc1 = time.perf_counter()
    B.free()
    lat = np.array(lat)
    return {"solves_per_s": 1e3 / float(np.mean(lat)),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "setup_ms": float(np.mean(setup)), "solve_ms": float(np.mean(solve)),
            "iterations": it, "restarts": restarts,
            "cached_hierarchy_solves_per_s": steps / (tc1 - tc0)}, dx.numpy()


def precision_forms(bc, ctx, A, b, settings, steps, warmup, x_ref=None, ref_name=None):
    """FP64, FP32 and BF16 forms on one workload; errors against x_ref (or
    the FP64 GPU solution when x_ref is None)."""
    from bootcmatchg_amd import operators as ops
    n = len(A[0]) - 1
    prm = bc.make_params(**settings)
    dA = bc.Csr(ctx, A)
    db = bc.DeviceArray.from_numpy(ctx, b)
    dx = bc.DeviceArray(ctx, n)
    S = ops.to_scipy(A)
    nb = np.linalg.norm(b)
    out = {}
    xs = {}
    for prec in (0, 1, 2):
        r, x = run_form(bc, ctx, dA, db, dx, prm, prec, steps, warmup)
        r["true_rel_residual"] = float(np.linalg.norm(b - S @ x) / nb)
        xs[prec] = x
        out[PREC_NAMES[prec]] = r
    ref = x_ref if x_ref is not None else xs[0]
    for prec in (0, 1, 2):
        out[PREC_NAMES[prec]]["rel_err_vs_" + (ref_name or "fp64_gpu")] = float(
            np.linalg.norm(xs[prec] - ref) / np.linalg.norm(ref))
    for k in ("fp32", "bf16"):
        out[k]["speedup_vs_fp64"] = out[k]["solves_per_s"] / out["fp64"]["solves_per_s"]
        out[k]["cached_speedup_vs_fp64"] = (out[k]["cached_hierarchy_solves_per_s"]
                                            / out["fp64"]["cached_hierarchy_solves_per_s"])
    del dA, db, dx
    return out


C3_GRID = 2048
C3_RHS = 32
C3_SETTINGS = dict(SETTINGS, aggrsweeps=3, cycle_type=2)


def multi_rhs(bc, ctx, steps):
    """32 right-hand sides on one hierarchy (anisotropic 2D, W-cycle): FP32
    hierarchy inside FP64 FCG, all columns in every launch, against the
    same right-hand sides solved one at a time in FP64 and in FP32."""
    from bootcmatchg_amd import operators as ops
    A = ops.aniso2d(C3_GRID, 0.01)
    n = len(A[0]) - 1
    S = ops.to_scipy(A)
    dA = bc.Csr(ctx, A)
    ctx.sync()
    t0 = time.perf_counter()
    B = bc.Boot(ctx, dA, bc.make_params(**C3_SETTINGS))
    ctx.sync()
    setup_ms = (time.perf_counter() - t0) * 1e3
    bh = [ops.uniform_pm1(j, n) for j in range(C3_RHS)]
    bs = [bc.DeviceArray.from_numpy(ctx, b) for b in bh]
    xs = [bc.DeviceArray(ctx, n) for _ in range(C3_RHS)]
    out = {"workload": "aniso2d 5-pt %d^2 eps 0.01, %d rhs uniform[-1,1] seeds 0..%d, "
                       "settings 0,1,0.8,3,0,39,2,4,4,16,4,4,2000,1e-8"
                       % (C3_GRID, C3_RHS, C3_RHS - 1),
           "rows": n, "nnz": len(A[1]), "setup_ms": setup_ms}
    single = {}
    for prec in (0, 1):
        B.set_precision(prec)
        B.fcg(bs[0], xs[0])
        ctx.sync()
        t0 = time.perf_counter()
        its = []
        for j in range(4):
            its.append(B.fcg(bs[j], xs[j])[0])
        ctx.sync()
        single[prec] = (time.perf_counter() - t0) / 4
        out["single_%s_ms_per_rhs" % PREC_NAMES[prec]] = single[prec] * 1e3
    B.set_precision(1)
    B.fcg_multi(bs, xs)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        it, res, st = B.fcg_multi(bs, xs)
    ctx.sync()
    per = (time.perf_counter() - t0) / steps
    rel = [float(np.linalg.norm(bh[j] - S @ xs[j].numpy()) / np.linalg.norm(bh[j]))
           for j in range(C3_RHS)]
    levels = [B.level_size(0, l) for l in range(B.levels(0))]
    mb = multi_solve_bytes(levels, [int(v) for v in it], C3_SETTINGS["prerelax_sweeps"],
                           C3_SETTINGS["postrelax_sweeps"], C3_SETTINGS["relaxnumber_coarse"],
                           cycle_visits(len(levels), C3_SETTINGS["cycle_type"]))
    out["solve_roofline"] = {
        "bytes_per_batch": mb, "bytes_per_rhs": mb / C3_RHS,
        "bound_rhs_solves_per_s": HBM_PEAK_GBS * 1e9 / (mb / C3_RHS),
        "achieved_rhs_solves_per_s": C3_RHS / per,
        "frac": (mb / per) / (HBM_PEAK_GBS * 1e9),
        "note": "matrix data counted once per launch for all columns over the longest "
                "column's iterations, vectors per column over its own iterations; FP32 "
                "hierarchy and cycle vectors, FP64 FCG vectors, 32-bit indices"}
    out.update({"multi_fp32_ms_per_batch": per * 1e3,
                "multi_fp32_ms_per_rhs": per * 1e3 / C3_RHS,
                "rhs_solves_per_s": C3_RHS / per,
                "speedup_vs_single_fp64": single[0] * C3_RHS / per,
                "speedup_vs_single_fp32": single[1] * C3_RHS / per,
                "iterations_min_max": [int(it.min()), int(it.max())],
                "max_true_rel_residual": max(rel)})
    B.free()
    return out


SERVE_RATES = (10, 50, 200, 500, 1000, 2000)


def serving(bc, ctx, quick=False):
    """The serving workload: 8 cached operators, Poisson arrivals, uniform
    operator choice, a distinct seeded right-hand side per request; the
    served loop (continuous batching) in FP32 and FP64 against one FP64 solve
    at a time on the same traces. Latency includes the upload of b and the
    download of x. A saturation trace (400 requests offered at 4000/s) gives
    the throughput; the rate sweep gives p50/p99 from light load to overload
    (at most 200 requests or 4 s of arrivals per rate). quick: only the
    served FP32 loop on a 200-request trace (for profiling)."""
    from bootcmatchg_amd import serving as sv
    pool = sv.Pool(bc, ctx, SETTINGS, 0)
    out = {"operators": [{"name": o["name"], "rows": o["n"], "nnz": o["nnz"]} for o in pool.ops],
           "settings": "0,1,0.8,2,0,39,0,4,4,16,4,4,2000,1e-8",
           "slots_per_operator": 16,
           "latency": "scheduled arrival to solution in host memory (queueing, upload of b, "
                      "solve, download of x); b of request k is uniform[-1,1] stream k"}
    if quick:
        pool.set_precision(1)
        S = bc.Server(ctx, 96 << 30, 16)
        for o in pool.ops:
            S.add(o["B"])
        arr, opsel = sv.trace(4000, len(pool.ops), 200, 3.0, seed=1)
        out["served_fp32"] = sv.run_served(pool, S, arr, opsel, pool.rhs(opsel, 0), 8)
        S.close()
        return out
    arr, opsel = sv.trace(4000, len(pool.ops), 400, 3.0, seed=1)
    bh = pool.rhs(opsel, 0)
    sweep = {}
    traces = {}
    for rate in SERVE_RATES:
        a, o_ = sv.trace(rate, len(pool.ops), 200, 4.0, seed=rate)
        traces[rate] = (a, o_, pool.rhs(o_, 100000 * rate))
        sweep[str(rate)] = {}
    out["saturation_trace"] = "%d requests offered at 4000/s" % len(arr)
    out["plain_fp64_one_at_a_time"] = sv.run_plain(pool, arr, opsel, bh)
    for rate in SERVE_RATES:
        a, o_, b_ = traces[rate]
        sweep[str(rate)]["plain_fp64"] = sv.run_plain(pool, a, o_, b_, 4)
    for prec in (1, 0):
        pool.set_precision(prec)
        S = bc.Server(ctx, 96 << 30, 16)
        for o in pool.ops:
            S.add(o["B"])
        a2, o2 = sv.trace(4000, len(pool.ops), 64, 1.0, seed=9)
        sv.run_served(pool, S, a2, o2, pool.rhs(o2, 7000000), 8)  # lanes opened, graphs captured
        out["served_%s" % PREC_NAMES[prec]] = sv.run_served(pool, S, arr, opsel, bh)
        for rate in SERVE_RATES:
            a, o_, b_ = traces[rate]
            sweep[str(rate)]["served_%s" % PREC_NAMES[prec]] = sv.run_served(pool, S, a, o_, b_, 4)
        out["device_bytes_%s" % PREC_NAMES[prec]] = S.info()["bytes"]
        S.close()
    out["rate_sweep"] = sweep
    plain = out["plain_fp64_one_at_a_time"]["solves_per_s"]
    out["served_fp32_vs_plain_fp64"] = out["served_fp32"]["solves_per_s"] / plain
    out["served_fp64_vs_plain_fp64"] = out["served_fp64"]["solves_per_s"] / plain
    out["served_fp32_vs_served_fp64"] = (out["served_fp32"]["solves_per_s"]
                                         / out["served_fp64"]["solves_per_s"])
    for o in pool.ops:
        o["B"].free()
    return out


def small_case(bc, ctx, steps):
    """The reference's CPU-runnable case: 3D Poisson 32^3, b = 1, x0 = 0,
    the drop-in settings with rtol 1e-6; full solves on the GPU against the
    serial oracle on one core (median of 5 after one warm-up)."""
    from bootcmatchg_amd import operators as ops
    from oracle import oracle as orc
    A = ops.poisson3d(32)
    n = len(A[0]) - 1
    b = np.ones(n)
    settings = dict(SETTINGS, rtol=1e-6)
    prm = bc.make_params(**settings)
    dA = bc.Csr(ctx, A)
    db = bc.DeviceArray.from_numpy(ctx, b)
    dx = bc.DeviceArray(ctx, n)
    lat = []
    for k in range(steps + 1):
        ctx.sync()
        t0 = time.perf_counter()
        B = bc.Boot(ctx, dA, prm)
        it, _ = B.fcg(db, dx)
        t1 = time.perf_counter()
        B.free()
        if k:
            lat.append(t1 - t0)
    x = dx.numpy()
    cpu = []
    for k in range(6):
        t0 = time.perf_counter()
        rc, xo, res = orc.solve(A, b, orc.params(**settings))
        t1 = time.perf_counter()
        if k:
            cpu.append(t1 - t0)
    return {"workload": "poisson3d_7pt_32^3 b=1 rtol 1e-6, settings "
                        "0,1,0.8,2,0,39,0,4,4,16,4,4,2000,1e-6",
            "gpu_solves_per_s": 1.0 / float(np.median(lat)), "gpu_iterations": it,
            "cpu_oracle_solves_per_s": 1.0 / float(np.median(cpu)), "cpu_cores": 1,
            "cpu_iterations": res.iterations,
            "solution_rel_diff": float(np.linalg.norm(x - xo) / np.linalg.norm(xo))}


SA_GRIDS = (32, 64, 100, 160)


def smoothed(bc, ctx, steps, grids=SA_GRIDS):
    """Smoothed aggregation (aggr_type 1, omega = 4 / (3 ||D^-1 A||_inf))
    against the unsmoothed drop-in: 3D 7-point Poisson at each grid, the
    drop-in settings with rtol 1e-8, b uniform[-1,1] seed 0; full solves
    (setup + FCG) timed per step."""
    from bootcmatchg_amd import operators as ops
    out = {"settings": "0,1,0.8,2,{0|1},39,0,4,4,16,4,4,2000,1e-8"}
    for N in grids:
        A = ops.poisson3d(N)
        n = len(A[0]) - 1
        b = ops.uniform_pm1(0, n)
        S = ops.to_scipy(A)
        dA = bc.Csr(ctx, A)
        db = bc.DeviceArray.from_numpy(ctx, b)
        dx = bc.DeviceArray(ctx, n)
        res = {}
        xs = {}
        for agg in (0, 1):
            prm = bc.make_params(**dict(SETTINGS, aggrtype=agg))
            setup, solve = [], []
            for k in range(steps + 1):
                ctx.sync()
                t0 = time.perf_counter()
                B = bc.Boot(ctx, dA, prm)
                ctx.sync()
                t1 = time.perf_counter()
                it, _ = B.fcg(db, dx)
                t2 = time.perf_counter()
                lv = [B.level_size(0, l) for l in range(B.levels(0))]
                B.free()
                if k > 0:
                    setup.append((t1 - t0) * 1e3)
                    solve.append((t2 - t1) * 1e3)
            x = dx.numpy()
            xs[agg] = x
            res["smoothed" if agg else "unsmoothed"] = {
                "setup_ms": float(np.mean(setup)), "solve_ms": float(np.mean(solve)),
                "solves_per_s": 1e3 / (float(np.mean(setup)) + float(np.mean(solve))),
                "iterations": it, "levels": len(lv),
                "operator_complexity": sum(z[1] for z in lv) / lv[0][1],
                "true_rel_residual": float(np.linalg.norm(b - S @ x) / np.linalg.norm(b))}
        res["solution_rel_diff"] = float(np.linalg.norm(xs[1] - xs[0]) / np.linalg.norm(xs[0]))
        out["poisson3d_%d^3" % N] = res
        del dA, db, dx
    return out


def sweep_bytes(n, nnz, cb=4):
    """Algorithmic HBM bytes of one Jacobi/L1 sweep u' = u + (f - A u) / d:
    val + col (8 + cb B/nnz, cb = 4 for 32-bit columns, 2 for the 16-bit
    column-minus-row copy), row pointer, u (gathered, counted once), f, d, u'."""
    return (8 + cb) * nnz + 4 * (n + 1) + 4 * 8 * n


def spmv_bytes(n, nnz, cb=4):
    """y = A x: val + col, row pointer, x (counted once), y."""
    return (8 + cb) * nnz + 4 * (n + 1) + 2 * 8 * n


def streamed_bytes(n, nnz, cb, vectors):
    """Bytes a level-0 row kernel streams in the column form it reads: 8-byte
    values, then cb-byte columns and the row pointers, or (cb = 0, the
    stencil pattern) one mask byte per row and two row pointers per 256-row
    block; plus `vectors` 8-byte vectors of n."""
    if cb == 0:
        return 8 * nnz + n + 8 * ((n + 255) // 256) + vectors * 8 * n
    return (8 + cb) * nnz + 4 * (n + 1) + vectors * 8 * n


# Measured HBM bytes per level-0 SpMV launch at 160^3 (same PMC runs and
# factors as the sweep: profiles/jacobi_sweep_poisson160_traffic.md).
SPMV_TRAFFIC_BYTES_160 = 305.9e6
MALL_BYTES = 256 << 20


def solve_bytes(levels, iterations, pre, post, coarse):
    """Algorithmic HBM bytes of one FCG solve with a V-cycle, counted the
    reference algorithm's way: per level l < L-1, (pre + post) sweeps, the
    residual, the restriction, the zeroing of the coarse iterate and the
    prolongation; `coarse` sweeps on the last level; per FCG iteration the
    fine SpMV, the triple product (32n), two merged axpys (80n), the norm (8n)
    and one zeroing (8n). levels = [(n_l, nnz_l)]."""
    cyc = 0
    L = len(levels)
    for l, (n, nnz) in enumerate(levels):
        if l == L - 1:
            cyc += coarse * sweep_bytes(n, nnz)
            continue
        nc = levels[l + 1][0]
        cyc += (pre + post) * sweep_bytes(n, nnz)
        cyc += spmv_bytes(n, nnz) + 8 * n            # residual f - A u
        cyc += 12 * n + 4 * (nc + 1) + 8 * n + 8 * nc  # restriction R r (one entry per fine row)
        cyc += 8 * nc                                 # coarse iterate zeroed
        cyc += 28 * n + 8 * nc                        # prolongation u += P u_c
    n0, nnz0 = levels[0]
    fcg = spmv_bytes(n0, nnz0) + (32 + 80 + 8 + 8) * n0
    return iterations * (cyc + fcg)


def cycle_visits(L, cycle_type, max_levels=39):
    """Visits of each level in one cycle (src/SOLVER/GAMG_cycle.cu:103-111
    with the grid sweeps of src/AMG/AMG.cu:371-394): level j + 1 is entered
    num_grid_sweeps[j] times per visit of level j, once when it is the
    coarsest."""
    ml = max_levels - 1
    g = [1] * ml
    if cycle_type == 1:
        g = [2 if i % 2 == 0 else 1 for i in range(ml)]
    elif cycle_type == 2:
        g = [2 if i < ml - 1 else 1 for i in range(ml)]
    v = [1]
    for j in range(L - 1):
        v.append(v[j] * (1 if j + 1 == L - 1 else g[j]))
    return v


def multi_solve_bytes(levels, its, pre, post, coarse, visits, sv=4):
    """Algorithmic HBM bytes of a batched solve of len(its) right-hand sides
    on one hierarchy whose operators, divisors and prolongators hold sv-byte
    values (4: FP32) with 32-bit indices, cycle vectors sv bytes per column
    entry, FCG vectors FP64: every launch reads the matrix data once for all
    columns (counted for max(its) iterations) and each column's vectors
    (counted for its own iterations). Per level visit as solve_bytes; per
    FCG iteration the fine SpMV, the triple product, two merged axpys, the
    norm, one zeroing and the conversions of r into and z out of the cycle's
    precision."""
    L = len(levels)
    cm = cv = 0
    for j, (n, nnz) in enumerate(levels):
        v = visits[j]
        mat = nnz * (sv + 4) + 4 * (n + 1)
        if j == L - 1:
            cm += v * coarse * (mat + sv * n)
            cv += v * coarse * 3 * sv * n
            continue
        nc = levels[j + 1][0]
        cm += v * ((pre + post) * (mat + sv * n) + mat + 2 * n * (sv + 4) + 4 * (nc + 1))
        cv += v * ((pre + post) * 3 * sv * n + 3 * sv * n + sv * (n + nc) + sv * nc
                   + sv * (nc + 2 * n))
    n0, nnz0 = levels[0]
    fm = 12 * nnz0 + 4 * (n0 + 1)
    fv = 16 * n0 + (32 + 80 + 8 + 8) * n0 + 2 * (8 + sv) * n0
    return max(its) * (cm + fm) + sum(its) * (cv + fv)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32/BF16 precision-form measurements")
    ap.add_argument("--no-multi", action="store_true",
                    help="skip the 32-right-hand-side measurement")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving-loop measurement")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every FCG kernel from the host instead of replaying "
                         "captured HIP graphs (used for kernel-trace profiling)")
    ap.add_argument("--no-smoothed", action="store_true",
                    help="skip the smoothed-aggregation measurement")
    ap.add_argument("--only", choices=("multi", "serving", "smoothed", "small"), default=None,
                    help="run one extra section alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path has no multi-GPU form here "
              "(each GPU would serve its own operators and requests)", file=sys.stderr)
        sys.exit(2)

    import bootcmatchg_amd as bc
    from bootcmatchg_amd import operators as ops

    if args.only:
        ctx = bc.Context(0)
        if args.no_graphs:
            bc.use_graphs(ctx, False)
        if args.only == "multi":
            res = multi_rhs(bc, ctx, max(1, args.steps // 2))
        elif args.only == "smoothed":
            res = smoothed(bc, ctx, max(1, args.steps // 2))
        elif args.only == "small":
            res = small_case(bc, ctx, args.steps)
        else:
            res = serving(bc, ctx, quick=args.steps <= 1)
        print(json.dumps({args.only: res}))
        ctx.close()
        return

    N = args.grid
    A = ops.poisson3d(N)
    n = len(A[0]) - 1
    nnz = len(A[1])
    b = ops.uniform_pm1(0, n)
    prm = bc.make_params(**SETTINGS)

    ctx = bc.Context(0)
    if args.no_graphs:
        bc.use_graphs(ctx, False)
    dA = bc.Csr(ctx, A)
    db = bc.DeviceArray.from_numpy(ctx, b)
    dx = bc.DeviceArray(ctx, n)

    def step(record=None):
        t0 = time.perf_counter()
        B = bc.Boot(ctx, dA, prm)
        t1 = time.perf_counter()
        it, res = B.fcg(db, dx)
        t2 = time.perf_counter()
        levels = [B.level_size(0, l) for l in range(B.levels(0))]
        B.free()
        if record is not None:
            record.append(dict(setup_ms=(t1 - t0) * 1e3, solve_ms=(t2 - t1) * 1e3,
                               total_ms=(t2 - t0) * 1e3, iterations=it, residual=res,
                               levels=levels))
        return it

    for _ in range(args.warmup):
        step()
    ctx.sync()
    rec = []
    bc.time_smoother(ctx, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(rec)
    ctx.sync()
    t1 = time.perf_counter()
    sm_ms, sm_launches = bc.smoother_timing(ctx)
    bc.time_smoother(ctx, False)
    elapsed = t1 - t0
    lat = np.array([r["total_ms"] for r in rec])

    # setup phases, event-timed on the context stream (separate setups: the
    # events stay out of the timed loop)
    bc.time_setup(ctx, True)
    phases = []
    for _ in range(args.steps):
        Bp = bc.Boot(ctx, dA, prm)
        phases.append(Bp.setup_timing(-1))
        Bp.free()
    bc.time_setup(ctx, False)
    setup_phases = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    named = ("weights_ms", "matching_ms", "aggregation_ms", "galerkin_ms", "smoother_ms",
             "inner_ms")
    setup_phases["other_ms"] = setup_phases["total_ms"] - sum(setup_phases[k] for k in named)
    setup_phases["note"] = ("means over %d setups with HIP events around each phase; "
                            "total_ms is the host wall time of bcmg_setup, other_ms the "
                            "rest (A-norm, sizes read back, tile analysis, buffers)"
                            % args.steps)

    # cached hierarchy: FCG only (the serving form's unit of work)
    B = bc.Boot(ctx, dA, prm)
    B.fcg(db, dx)
    ctx.sync()
    tc0 = time.perf_counter()
    for _ in range(args.steps):
        it_c, _ = B.fcg(db, dx)
    ctx.sync()
    tc1 = time.perf_counter()
    x = dx.numpy()
    B.free()
    S = ops.to_scipy(A)
    true_rel = float(np.linalg.norm(b - S @ x) / np.linalg.norm(b))

    # fine-level SpMV bandwidth (algorithmic bytes / event-timed launch)
    dy = bc.DeviceArray(ctx, n)
    spmv_ms = bc.time_spmv(ctx, dA, db, dy, 50)
    cb = dA.column_bytes
    # algorithmic bytes count 32-bit columns (the operator's own format); the
    # library may stream a 16-bit copy (cb = 2) and then moves fewer bytes
    spmv_gbs = spmv_bytes(n, nnz) / (spmv_ms * 1e-3) / 1e9

    # whole FCG solve against the HBM roofline (cached hierarchy: no setup)
    sb = solve_bytes([tuple(z) for z in rec[-1]["levels"]], rec[-1]["iterations"],
                     SETTINGS["prerelax_sweeps"], SETTINGS["postrelax_sweeps"],
                     SETTINGS["relaxnumber_coarse"])
    cached_rate = args.steps / (tc1 - tc0)
    solve_rl = {"bytes_per_solve": sb, "bound_solves_per_s": HBM_PEAK_GBS * 1e9 / sb,
                "achieved_solves_per_s": cached_rate,
                "frac": cached_rate * sb / (HBM_PEAK_GBS * 1e9),
                "note": "FCG with a cached hierarchy; bytes counted the reference "
                        "algorithm's way, 32-bit columns on every level (the library "
                        "streams level 0 in its stencil-pattern form, one mask byte per "
                        "row, so it moves fewer bytes than this count); setup excluded, "
                        "time only"}

    sweep_avg_ms = sm_ms / max(sm_launches, 1)
    sweep_b = sweep_bytes(n, nnz)
    achieved = sweep_b / (sweep_avg_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        tcpu0 = time.perf_counter()
        rc, xo_cpu, ores = orc.solve(A, b, orc.params(**SETTINGS))
        tcpu1 = time.perf_counter()
        cpu = {"value": 1.0 / (tcpu1 - tcpu0), "unit": "solves/s", "cores": 1,
               "host_cores": os.cpu_count(), "kind": "port",
               "sample": "one full solve (setup %.2f s + FCG %.2f s, %d iterations) of the "
                         "same 160^3 workload by the serial C++ oracle, 1 thread"
                         % (ores.setup_seconds, ores.solve_seconds, ores.iterations),
               "iterations": ores.iterations,
               "solution_rel_diff_vs_gpu": float(np.linalg.norm(x - xo_cpu)
                                                 / np.linalg.norm(xo_cpu))}

    precision = None
    if not args.no_precision:
        xo = None if cpu is None else xo_cpu
        precision = {"c1_poisson160": precision_forms(
            bc, ctx, A, b, SETTINGS, args.steps, 1, x_ref=xo,
            ref_name="fp64_oracle" if xo is not None else None)}
        from bootcmatchg_amd import operators as ops2
        A2 = ops2.lognormal3d(C2_GRID, 4.0, seed=0)
        b2 = ops2.uniform_pm1(0, len(A2[0]) - 1)
        xo2 = None
        if cpu is not None:
            rc2, xo2, ores2 = orc.solve(A2, b2, orc.params(**C2_SETTINGS))
            assert rc2 == 0
        precision["c2_lognormal128_boot3"] = precision_forms(
            bc, ctx, A2, b2, C2_SETTINGS, max(2, args.steps // 2), 1, x_ref=xo2,
            ref_name="fp64_oracle" if xo2 is not None else None)
        if xo2 is not None:
            precision["c2_lognormal128_boot3"]["oracle_iterations"] = ores2.iterations
        precision["c2_lognormal128_boot3"]["workload"] = (
            "cell-coefficient diffusion 128^3, k=exp(g), g~N(0,4) seed 0, harmonic faces; "
            "b uniform[-1,1] seed 0; settings 0,3,0.5,2,0,39,0,4,4,16,4,4,2000,1e-8")

    multi = None if args.no_multi else multi_rhs(bc, ctx, max(1, args.steps // 2))
    served = None if args.no_serving else serving(bc, ctx)
    sa = None if args.no_smoothed else smoothed(bc, ctx, max(1, args.steps // 2))
    small = None if args.no_cpu_baseline else small_case(bc, ctx, args.steps)

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "poisson3d_7pt_%d^3 fp64 drop-in solve (setup+FCG) rtol 1e-8" % N,
                   "rows": n, "nnz": nnz, "rhs": "uniform[-1,1] splitmix64 seed 0",
                   "settings": "0,1,0.8,2,0,39,0,4,4,16,4,4,2000,1e-8"},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "setup_ms": float(np.mean([r["setup_ms"] for r in rec])),
        "setup_phases_ms": setup_phases,
        "solve_ms": float(np.mean([r["solve_ms"] for r in rec])),
        "iterations": rec[-1]["iterations"],
        "levels": rec[-1]["levels"],
        "true_rel_residual": true_rel,
        "cached_hierarchy_solves_per_s": args.steps / (tc1 - tc0),
        "spmv_gbs": spmv_gbs,
        "spmv_frac_of_hbm": spmv_gbs / HBM_PEAK_GBS,
        "spmv": {"kernel": "level-0 y = A x (k_rows_pattern<OP_SPMV> in the stencil-pattern "
                           "form, else k_rows_staged<OP_SPMV>)",
                 "bytes_per_launch": spmv_bytes(n, nnz), "avg_launch_ms": spmv_ms,
                 "achieved_gbs": spmv_gbs, "column_bytes": cb,
                 "streamed_bytes_per_launch": streamed_bytes(n, nnz, cb, 2),
                 "traffic_bytes": SPMV_TRAFFIC_BYTES_160 if (N == GRID and cb == 0) else None,
                 "working_set_bytes": spmv_bytes(n, nnz, cb), "mall_bytes": MALL_BYTES},
        "solve_roofline": solve_rl,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": SWEEP_TRAFFIC_BYTES_160 if (N == GRID and cb == 0) else None,
                     "kernel": "level-0 Jacobi/L1 sweep (k_rows_pattern<OP_JACOBI, non-temporal> "
                               "in the stencil-pattern form, else k_rows_staged<OP_JACOBI>)",
                     "bytes_per_launch": sweep_b, "avg_launch_ms": sweep_avg_ms,
                     "launches": sm_launches, "column_bytes": cb,
                     "streamed_bytes_per_launch": streamed_bytes(n, nnz, cb, 4),
                     "note": "achieved = algorithmic bytes (8-byte values, 32-bit columns, "
                             "row pointers, u gathered once, f, d, u') / event-timed launch; "
                             "the library streams the column form column_bytes names "
                             "(0: the stencil pattern, one mask byte per row), "
                             "streamed_bytes_per_launch, and traffic is the measured HBM "
                             "bytes"},
        "cpu_baseline": cpu,
        "precision_form": precision,
        "multi_rhs": multi,
        "serving": served,
        "smoothed_aggregation": sa,
        "small_case": small,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
