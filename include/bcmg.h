/*
 * bcmg.h - C-ABI of the gfx950 (MI355X) AMG-preconditioned flexible CG.
 *
 * This is the drop-in boundary for BootCMatchG's setup + solve path. Each
 * entry point names the reference interface it replaces (file:line in the
 * BootCMatchG source tree). Conventions:
 *   - every function returns BCMG_OK (0) or a negative BCMG_E* code; the
 *     reference calls exit()/assert instead (see the notes per function);
 *     bcmg_last_error() gives the message of the calling thread's last error;
 *   - matrices are CSR, 0-based, int32 indices, FP64 values, columns ascending
 *     and an explicit diagonal entry in every row (src/matrix/CSR.h:13-28,
 *     src/utility/setting.h:20-25);
 *   - "device pointer" means memory of the context's GPU (bcmg_dmalloc or any
 *     HIP allocation on that device); all work is ordered on the context's
 *     stream and entry points return after it is complete unless noted;
 *   - a context is used by one host thread at a time.
 */
#ifndef BCMG_H
#define BCMG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BCMG_OK 0
#define BCMG_EINVAL -1       /* bad argument or malformed matrix */
#define BCMG_ESETTINGS -2    /* settings file missing, short or out of range */
#define BCMG_EZERORHS -3     /* ||b - A x0|| <= DBL_EPSILON ||b||: the reference exits (FCG.cu:459-462) */
#define BCMG_EUNSUP -4       /* a setting or size this library does not run */
#define BCMG_EBREAKDOWN -5   /* non-finite residual (the reference asserts: FCG.cu:567) */
#define BCMG_EHIP -6         /* HIP runtime error */
#define BCMG_ENOMEM -7
#define BCMG_ENODEV -8       /* no usable gfx950 device */

/* The 14 settings of the AMGsettings file plus matchtype, in the field order
   of the reference's struct params (src/AMG/AMG.cu:20-37). */
typedef struct {
    int solver_type;        /* 0 multiplicative, 1 symmetrized, 2 additive */
    int max_hrc;
    double conv_ratio;      /* rho */
    int matchtype;          /* always 3 (AMG.cu:497) */
    int aggrsweeps;
    int aggrtype;           /* 0 unsmoothed, 1 smoothed (omega = 4 / (3 ||D^-1 A||_inf)) */
    int max_levels;
    int cycle_type;         /* 0 V, 1 H, 2 W */
    int coarse_solver;      /* 0 Jacobi, 4 L1 */
    int relax_type;         /* 0 Jacobi, 4 L1 */
    int prerelax_sweeps;
    int postrelax_sweeps;
    int itnlim;
    double rtol;
    int relaxnumber_coarse;
} bcmg_params;

typedef struct bcmg_ctx_s *bcmg_ctx;     /* one GPU, one stream */
typedef struct bcmg_csr_s *bcmg_csr;     /* a device-resident CSR matrix */
typedef struct bcmg_boot_s *bcmg_boot;   /* a built composite (1..max_hrc hierarchies) */

const char *bcmg_last_error(void);

/* Parse a 14-line AMGsettings file. Replaces AMG::Params::initFromFile
   (src/AMG/AMG.cu:480-526, parser src/AMG/utility.cu:7-25); the reference
   exits on a missing file or a value outside the allowed sets (AMG.cu:8-18). */
int bcmg_params_from_file(const char *path, bcmg_params *out);

/* Context on GPU `device`. Replaces Handles::init / Handles::free
   (src/utility/handles.cu:12-48): one HIP stream, no library handles. */
int bcmg_ctx_create(int device, bcmg_ctx *out);
/* A context whose work is ordered on a stream the caller created (a
   hipStream_t, passed as a pointer; not the null stream) on `device`: every
   call on it enqueues there, so the library composes with the caller's own
   work and dependencies; bcmg_ctx_destroy leaves the stream alive. The
   reference runs on the default stream (src/utility/handles.cu:12-48). */
int bcmg_ctx_create_on_stream(int device, void *stream, bcmg_ctx *out);
int bcmg_ctx_destroy(bcmg_ctx ctx);
int bcmg_sync(bcmg_ctx ctx);

/* Device memory helpers (Vector::init, copyToDevice, copyToHost, free: src/matrix/vector.cu). */
int bcmg_dmalloc(bcmg_ctx ctx, size_t bytes, void **ptr);
int bcmg_dfree(bcmg_ctx ctx, void *ptr);
int bcmg_memcpy_h2d(bcmg_ctx ctx, void *dst_dev, const void *src_host, size_t bytes);
int bcmg_memcpy_d2h(bcmg_ctx ctx, void *dst_host, const void *src_dev, size_t bytes);
int bcmg_memcpy_d2d(bcmg_ctx ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* Copy a host CSR to the device. Replaces CSRm::copyToDevice
   (src/matrix/CSR.cu:194-217). The input is validated (sorted columns,
   diagonal present); the reference leaves such inputs undefined. */
int bcmg_csr_upload(bcmg_ctx ctx, int n, int nnz, const int *row, const int *col,
                    const double *val, bcmg_csr *out);
int bcmg_csr_download(bcmg_csr A, int *row, int *col, double *val);
int bcmg_csr_info(bcmg_csr A, int *n, int *nnz);
/* Bytes per column index that the FP64 row kernels read for A. A square
   operator can get narrower copies of its columns at upload: 16-bit offsets
   (column minus row) when they all fit and it has more than 16.7 M entries
   (an FP64 stream above 192 MB), and a stencil pattern when it has more
   than 11.2 M entries (above 128 MB) and its offsets take at most 8
   distinct values (one byte per row: which offsets of a sorted table the
   row has). The row kernels read the narrowest one that is kept: 0 the
   pattern (no per-entry column bytes, one byte per row, and only two row
   pointers per 256 rows), 2 the 16-bit offsets (FP32/BF16 value copies use
   them when their 32-bit-column stream also exceeds 192 MB), else 4.
   Results are the same bits in every form. No counterpart in the reference
   (which always streams 32-bit columns); it lets a caller price a kernel's
   memory traffic. */
int bcmg_csr_column_bytes(bcmg_csr A, int *bytes);
/* Restricts the row kernels on A to a column form: bytes 0 (the pattern),
   2 (the 16-bit offsets) or 4 (the int32 columns); BCMG_EUNSUP when A has no
   such copy. Composites set up on A before the call keep the launches they
   captured. No counterpart in the reference: for A/B measurements and tests
   of the column forms. */
int bcmg_csr_set_columns(bcmg_csr A, int bytes);
/* Replaces CSRm::free (src/matrix/CSR.cu:160-181). */
int bcmg_csr_free(bcmg_csr A);

/* AMG setup: the bootstrap composite of up to max_hrc hierarchies built by
   compatible weighted matching. Replaces Bootstrap::bootstrap
   (src/BOOT/bootstrap.cu:46-90) with AMG::BootBuildData::initByParams,
   AMG::ApplyData::initByParams/setGridSweeps (src/AMG/AMG.cu:359-448).
   A must stay alive while the result is used. */
int bcmg_setup(bcmg_ctx ctx, bcmg_csr A, const bcmg_params *p, bcmg_boot *out);
/* Replaces AMG::Boot::free (src/AMG/AMG.cu:465-469). */
int bcmg_boot_free(bcmg_boot B);

/* Precision of the preconditioner (an extension: the reference is FP64
   only). BCMG_PREC_FP64 is the drop-in. BCMG_PREC_FP32 stores every level's
   operator, smoother divisors and prolongator in FP32 and runs the cycles in
   FP32; BCMG_PREC_BF16 also stores the operator values of levels >= 1 as
   bfloat16 (FP32 arithmetic). In both, FCG's SpMV, inner products, scalar
   recurrence and updates stay FP64, and when the recursive residual reaches
   rtol the FP64 true residual b - A x is checked; FCG restarts from it if it
   is still above rtol ||b||. The hierarchy itself is built in FP64 (same
   aggregates as the drop-in) and converted. */
#define BCMG_PREC_FP64 0
#define BCMG_PREC_FP32 1
#define BCMG_PREC_BF16 2
int bcmg_set_precision(bcmg_boot B, int precision);
/* true-residual restarts made by the last bcmg_fcg on B (0 in FP64, where
   the stopping test is the reference's recursive one); in a reduced
   precision the resid bcmg_fcg returns is the checked true ||b - A x||.
   No counterpart in the reference. */
int bcmg_fcg_restarts(bcmg_boot B, int *restarts);

/* FCG v3 from x0 = 0 with the composite preconditioner; b and x are device
   pointers of length n. Replaces flexibileConjugateGradients_v3
   (src/SOLVER/FCG.cu:433-585). iters follows the reference's count
   (loops + 1, FCG.cu:569); resid is the last recursive ||r||. */
int bcmg_fcg(bcmg_ctx ctx, bcmg_boot B, const double *b_dev, double *x_dev, int *iters,
             double *resid);

/* k right-hand sides on one composite at once (an extension: the
   reference solves one). b_dev[j], x_dev[j] are device pointers of length
   n; per column j: iters[j], resid[j] as bcmg_fcg reports them and
   status[j] (BCMG_OK, BCMG_EZERORHS: x = 0, BCMG_EBREAKDOWN). Every launch
   serves all columns (up to 32 per block; more are done 32 at a time), and
   column j's results equal bcmg_fcg on b_dev[j] alone bit for bit. */
int bcmg_fcg_multi(bcmg_ctx ctx, bcmg_boot B, int k, const double *const *b_dev,
                   double *const *x_dev, int *iters, double *resid, int *status);

/* ---- serving loop (an extension: continuous batching of solves) ---------
   A server serves solves on composites built beforehand (cached on the
   device). Requests name an operator and give device pointers b (read when
   the request is admitted) and x (written when it completes). Each call of
   bcmg_server_step waits for the iteration the previous step queued,
   retires finished requests, admits queued requests into free column slots
   (requests arriving while others run join at the next iteration), and
   queues one iteration on every operator with active slots, each operator
   on a stream of its own. An operator's slots (up to slots_per_operator,
   at most 32) take device memory only while it has work, within
   hbm_budget_bytes; idle operators give theirs back when another needs it.
   A request's result equals bcmg_fcg on its b alone, bit for bit. The
   composites must stay alive and keep their precision while served; destroy
   the server before them and before its context. */
typedef struct bcmg_server_s *bcmg_server;
int bcmg_server_create(bcmg_ctx ctx, size_t hbm_budget_bytes, int slots_per_operator,
                       bcmg_server *out);
int bcmg_server_destroy(bcmg_server S);
int bcmg_server_add_operator(bcmg_server S, bcmg_boot B, int *op_id);
int bcmg_server_submit(bcmg_server S, int op_id, const double *b_dev, double *x_dev,
                       long long *req_id);
/* one round of the loop; *completed = requests finished in it */
int bcmg_server_step(bcmg_server S, int *completed);
/* done = 0 while pending; a finished result is returned once (then
   forgotten); latency_ms runs from submit to completion */
int bcmg_server_result(bcmg_server S, long long req_id, int *done, int *iters, double *resid,
                       int *status, double *latency_ms);
int bcmg_server_info(bcmg_server S, int *queued, int *running, size_t *bytes_in_use,
                     int *lanes);
/* ids of requests finished since the last call (up to cap, oldest first) */
int bcmg_server_completed(bcmg_server S, long long *ids, int cap, int *count);

/* setup + FCG + free on host arrays: the whole of solve()
   (src/SOLVER/solver.cu:20-129), with x returned in host memory.
   setup_ms / solve_ms (may be NULL) are the two phases' wall times. */
int bcmg_solve(int n, int nnz, const int *row, const int *col, const double *val,
               const bcmg_params *p, const double *rhs, double *x, int *iters, double *resid,
               double *setup_ms, double *solve_ms);

/* ---- inspection of a built composite (host outputs) ---------------------- */
int bcmg_boot_info(bcmg_boot B, int *n_hrc, double *estimated_ratio);
/* number of levels of hierarchy k (AMG hierarchy->num_levels, AMG.cu:51) */
int bcmg_hier_levels(bcmg_boot B, int k);
int bcmg_level_size(bcmg_boot B, int k, int l, int *n, int *nnz);
/* hierarchy k's summary (AMG::Hierarchy::finalize_cmplx / finalize_wcmplx
   and the average coarsening ratio, src/AMG/AMG.cu:218-237,
   src/AMG/matchingAggregation.cu:330-333,406; printed by
   AMG::Hierarchy::printInfo, src/AMG/AMG.cu:241-251): levels, operator
   complexity sum_l nnz_l / nnz_0, its W-cycle form sum_l 2^l nnz_l / nnz_0,
   the mean of n_{l-1} / n_l over the coarsenings, and the bootstrap's
   estimated ratio once hierarchy k was added (src/BOOT/bootstrap.cu:76).
   Any output may be NULL. */
int bcmg_hier_info(bcmg_boot B, int k, int *levels, double *op_cmplx, double *op_wcmplx,
                   double *avg_cratio, double *ratio_after);
int bcmg_level_A(bcmg_boot B, int k, int l, int *row, int *col, double *val);
/* prolongator from level l to l+1: coarse index and value per fine row */
int bcmg_level_P(bcmg_boot B, int k, int l, int *agg, double *p);
/* prolongator from level l to l+1 as CSR (n_l + 1 row pointers, nnz
   entries): the smoothed one (I - omega D^-1 A) P_tent when aggrtype is 1
   (src/AMG/matchingAggregation.cu:353-390), else the tentative one (one
   entry per row). omega = 4 / (3 ||D^-1 A||_inf); 0 when not smoothed. */
int bcmg_level_Pinfo(bcmg_boot B, int k, int l, int *nnz, int *smoothed, double *omega);
int bcmg_level_Pcsr(bcmg_boot B, int k, int l, int *row, int *col, double *val);
/* smoother divisors: D = diag(A), M = L1 row sums; has* tell which exist */
int bcmg_level_DM(bcmg_boot B, int k, int l, double *D, double *M, int *hasD, int *hasM);

/* ---- building blocks (device pointers), for parity tests ---------------- */
/* y = A x. Replaces CSRm::CSRVector_product_adaptive_miniwarp with alpha 1,
   beta 0 (src/matrix/CSR.cu:926-960). */
int bcmg_spmv(bcmg_ctx ctx, bcmg_csr A, const double *x, double *y);
/* y = f - A x (the OP 2 form of the same kernel, GAMG_cycle.cu:59-67). */
int bcmg_residual(bcmg_ctx ctx, bcmg_csr A, const double *x, const double *f, double *y);
/* `sweeps` Jacobi sweeps u <- u + (f - A u) / d, result in u; u2 is scratch.
   Replaces jacobi_adaptive_miniwarp (src/AMG/relaxation.cu:211-285). */
int bcmg_relax(bcmg_ctx ctx, bcmg_csr A, const double *d, const double *f, double *u,
               double *u2, int sweeps);
/* bcmg_relax without the u2 argument and without any device allocation:
   the second iterate lives in a static device array of
   BCMG_RELAX_NOALLOC_MAX_ROWS doubles (32 MiB) that each GPU the library
   runs on holds until the process ends. No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same kernels and bits as bcmg_relax. An operator with more rows returns
   BCMG_EINVAL. All calls on one GPU share that GPU's array: calls that
   overlap on one GPU (two threads or two contexts) race. */
#define BCMG_RELAX_NOALLOC_MAX_ROWS 4194304
int bcmg_relax_noalloc(bcmg_ctx ctx, bcmg_csr A, const double *d, const double *f, double *u,
                       int sweeps);
/* matching weights W aligned with A's nnz (diagonal slots 0). Replaces
   Matching::makeAH + toMaximumProductMatrix (src/AMG/matching.cu:293-472). */
int bcmg_weights(bcmg_ctx ctx, bcmg_csr A, const double *w, double *W);
/* suitor matching and symmetrisation given W. Replaces the body of
   Matching::suitor (src/AMG/matching.cu:489-496, suitor.cu:427-454,
   unsymMatching.cu:58-71). n_unsym = number of non-mutual suitor pairs. */
int bcmg_matching(bcmg_ctx ctx, bcmg_csr A, const double *W, int *M_raw, int *M_sym,
                  int *n_unsym);
/* pairwise aggregation of a symmetric matching. Replaces
   matchingPairAggregation's aggregation (src/AMG/matchingAggregation.cu:181-208);
   aggregates are numbered by their smallest fine index. */
int bcmg_aggregate(bcmg_ctx ctx, int n, const int *M_sym, const double *w, int *agg, double *p,
                   int *n_coarse);
/* A_c = R (A P) for a pairwise P (every aggregate has 1 or 2 fine rows).
   Replaces the two CSRm::CSRCSR_product calls of matchingAggregation
   (src/AMG/matchingAggregation.cu:241-244). */
int bcmg_galerkin(bcmg_ctx ctx, bcmg_csr A, const int *agg, const double *p, int n_coarse,
                  bcmg_csr *Ac);
/* C = L R (nl x ml times ml x mr; device CSR arrays), each entry the sum of
   its products in the order of L's row, then R's row (storage order), the
   pattern symbolic with columns ascending. Replaces CSRm::CSRCSR_product
   (src/matrix/CSR.cu:1069-1079) in smoothed aggregation's three products
   (src/AMG/matchingAggregation.cu:375,383,387). C may be rectangular: read it
   with bcmg_csr_info / bcmg_csr_download. Rows are processed in batches whose
   temporary slots (per row: min(its products, mr) plus a packing weight of
   products + 1 for rows of at most 1024 products) stay within slots_per_pass
   (0: 2^28); a slot takes 12 bytes of device memory. */
int bcmg_spgemm(bcmg_ctx ctx, int nl, int ml, int mr, const int *lrow, const int *lcol,
                const double *lval, const int *rrow, const int *rcol, const double *rval,
                long long slots_per_pass, bcmg_csr *C);
/* x <- B(r; x): one application of the composite preconditioner with x as
   the initial guess. Replaces preconditionApply (src/SOLVER/FCG.cu:112-187). */
int bcmg_precond_apply(bcmg_ctx ctx, bcmg_boot B, const double *r, double *x);

/* Replay the FCG iteration as a captured HIP graph (default on). Results are
   identical either way; off runs every launch from the host. */
int bcmg_use_graphs(bcmg_ctx ctx, int enable);

/* ---- measurement -------------------------------------------------------- */
/* When enabled, each run of consecutive fine-level (level 0) smoother sweeps
   launched by bcmg_fcg is bracketed by one pair of HIP events on the context
   stream (nothing else runs between them); bcmg_smoother_timing reads (and
   resets) the summed time and the number of sweeps it covers. */
int bcmg_time_smoother(bcmg_ctx ctx, int enable);
int bcmg_smoother_timing(bcmg_ctx ctx, double *total_ms, int *launches);
/* Phases of bcmg_setup timed with HIP events on the context stream while
   bcmg_time_setup(ctx, 1) is in force (0 otherwise). The reference times
   its setup as a whole (bootstrap_time, src/SOLVER/solver.cu:76-87) and the
   products of the coarsening (TOTAL_MUL_TIME, src/AMG/matchingAggregation.cu:
   229-248,262-267,421); this splits it by the kernels' roles. */
typedef struct {
    double weights_ms;      /* matching weights AH and log (Matching::makeAH, toMaximumProductMatrix) */
    double matching_ms;     /* suitor matching and symmetrisation (approx_match_gpu_suitor, unsymFix) */
    double aggregation_ms;  /* aggregates, P, R = P^T, P composition, w restriction (and smoothing of P) */
    double galerkin_ms;     /* coarse operators R (A P): the reference's TOTAL_MUL_TIME */
    double smoother_ms;     /* relaxPrepare: diagonal and L1 row sums */
    double inner_ms;        /* bootstrap inner iterations (Bootstrap::innerIterations) */
    double hierarchy_ms;    /* each hierarchy's whole build including its inner iterations */
    double total_ms;        /* k = -1: host wall time of bcmg_setup; k >= 0: hierarchy_ms */
} bcmg_setup_times;
int bcmg_time_setup(bcmg_ctx ctx, int enable);
/* k = -1: sums over the composite; k >= 0: hierarchy k (a hierarchy that
   shares the first one's storage shows only its inner iterations) */
int bcmg_setup_timing(bcmg_boot B, int k, bcmg_setup_times *out);

/* `reps` back-to-back launches of y = A x on the context stream between two
   HIP events; avg_ms is the mean time per launch. */
int bcmg_time_spmv(bcmg_ctx ctx, bcmg_csr A, const double *x, double *y, int reps,
                   double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* BCMG_H */
