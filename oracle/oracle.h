/*
 * oracle.h - C interface of the serial CPU oracle for BootCMatchG's
 * AMG-preconditioned flexible CG (setup + solve).
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is what the HIP library's results are
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it. The product (bootcmatchg_amd, include/) never
 * links or calls it.
 *
 * All matrices are CSR, 0-based, int32 indices, FP64 values, columns sorted
 * ascending, with an explicit diagonal entry in every row (the reference
 * requires this: src/AMG/matching.cu:216-218, src/matrix/CSR.cu:320-354).
 * Functions return 0 on success and a negative ORC_E* code on failure.
 */
#ifndef BCMG_ORACLE_H
#define BCMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EINVAL -1     /* bad argument / malformed matrix */
#define ORC_ESETTINGS -2  /* settings file missing, short or out of range */
#define ORC_EZERORHS -3   /* ||b - A x0|| <= DBL_EPSILON ||b|| (reference exits) */
#define ORC_EUNSUP -4     /* a feature the restatement does not cover */
#define ORC_EBREAKDOWN -5 /* non-finite residual */

/* Field order of the reference's struct params (src/AMG/AMG.cu:20-37). */
typedef struct {
    int solver_type;
    int max_hrc;
    double conv_ratio;
    int matchtype;
    int aggrsweeps;
    int aggrtype;
    int max_levels;
    int cycle_type;
    int coarse_solver;
    int relax_type;
    int prerelax_sweeps;
    int postrelax_sweeps;
    int itnlim;
    double rtol;
    int relaxnumber_coarse;
} orc_params;

typedef struct orc_boot_s orc_boot;

typedef struct {
    int iterations;        /* reference: *num_iter = iter + 1 (FCG.cu:569) */
    double residual;       /* last recursive ||r|| (FCG.cu:558) */
    double delta0;         /* ||b - A x0|| */
    int n_hrc;
    double estimated_ratio;
    double setup_seconds;
    double solve_seconds;
} orc_result;

/* settings (src/AMG/AMG.cu:480-526, src/AMG/utility.cu:7-25) */
int orc_params_from_file(const char *path, orc_params *out);

/* portable log shared with the device code (include/bcmg_log.h) */
double orc_log(double x);

/* building blocks on one level ------------------------------------------ */
/* W aligned with A's nnz; diagonal slots get 0 (never matched).
   matching.cu:141-162, 200-238, 386-396, 401-431 */
int orc_weights(int n, const int *row, const int *col, const double *val,
                const double *w, double *W);
/* suitor matching (suitor.cu:118-392) then symmetrisation (unsymMatching.cu);
   M_raw = suitor array, M_sym = symmetric matching. order = NULL processes
   vertices 0..n-1, else in the given order (the result must not depend on it). */
int orc_matching(int n, const int *row, const int *col, const double *W,
                 const int *order, int *M_raw, int *M_sym, int *n_unsym);
/* pairwise aggregation with canonical numbering (matchingAggregation.cu:92-137) */
int orc_aggregate(int n, const int *M_sym, const double *w, int *agg,
                  double *p, int *n_coarse);
/* Galerkin product R (A P) for P with one entry per row (fixed order).
   First call with col_c = NULL to get nnz_c and row_c; second call fills. */
int orc_galerkin(int n, const int *row, const int *col, const double *val,
                 const int *agg, const double *p, int n_c,
                 int *row_c, int *col_c, double *val_c, int *nnz_c);

/* C = L R (nl x ml times nr x mr, ml == nr) in the fixed order of smoothed
   aggregation's products. First call with ccol = NULL for crow and nnz_c. */
int orc_spgemm(int nl, int ml, const int *lrow, const int *lcol, const double *lval,
               int nr, int mr, const int *rrow, const int *rcol, const double *rval,
               int *crow, int *ccol, double *cval, int *nnz_c);

/* kernels of the solve phase ------------------------------------------- */
int orc_spmv(int n, const int *row, const int *col, const double *val,
             const double *x, double *y);
int orc_relax(int n, const int *row, const int *col, const double *val,
              const double *d, const double *f, double *u, int sweeps);

/* whole path ------------------------------------------------------------ */
int orc_setup(int n, int nnz, const int *row, const int *col, const double *val,
              const orc_params *p, orc_boot **out);
void orc_boot_free(orc_boot *b);
int orc_boot_info(const orc_boot *b, int *n_hrc, double *estimated_ratio);
int orc_hier_levels(const orc_boot *b, int k);
/* sizes of level l of hierarchy k */
int orc_level_size(const orc_boot *b, int k, int l, int *n, int *nnz);
int orc_level_A(const orc_boot *b, int k, int l, int *row, int *col, double *val);
/* prolongator of level l -> l+1 (composed over the aggregation sweeps) */
int orc_level_P(const orc_boot *b, int k, int l, int *agg, double *p);
/* prolongator of level l as CSR: the smoothed one when aggrtype is 1, else
   the tentative one (one entry per row); omega = 4 / (3 ||D^-1 A||_inf) */
int orc_level_Pinfo(const orc_boot *b, int k, int l, int *nnz, int *smoothed, double *omega);
int orc_level_Pcsr(const orc_boot *b, int k, int l, int *row, int *col, double *val);
/* smoother divisors: D = diag, M = L1 row sums (NULL if not prepared) */
int orc_level_DM(const orc_boot *b, int k, int l, double *D, double *M, int *hasD, int *hasM);
/* smooth vector w used to build hierarchy k */
int orc_boot_w(const orc_boot *b, int k, double *w);
/* x <- B(rhs; x) : one application of the composite preconditioner */
int orc_precond_apply(const orc_boot *b, const double *rhs, double *x);
/* FCG v3 with x0 given in x (normally zero); x is overwritten */
int orc_fcg(const orc_boot *b, const double *rhs, double *x, orc_result *res,
            double *resid_history, int history_len);
/* setup + FCG from x0 = 0 (reference solve(), src/SOLVER/solver.cu:20-129) */
int orc_solve(int n, int nnz, const int *row, const int *col, const double *val,
              const double *rhs, const orc_params *p, double *x, orc_result *res);

#ifdef __cplusplus
}
#endif

#endif
