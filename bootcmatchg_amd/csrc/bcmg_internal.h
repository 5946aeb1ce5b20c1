/*
 * bcmg_internal.h - internal types of the gfx950 AMG-FCG library.
 *
 * Host code (setup orchestration, cycles, FCG) and the kernel launch wrappers
 * share these declarations. Kernels live in kernels_*.hip; each launch wrapper
 * takes the stream explicitly. Everything here stays behind the C-ABI of
 * include/bcmg.h: no exceptions cross it, no torch types appear in it.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <unordered_map>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/bcmg.h"

namespace bcmg {

/* ------------------------------------------------------------ errors */

struct Error : public std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define BCMG_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (void)hipGetLastError(); /* do not leave it for the next check */       \
            throw ::bcmg::Error(BCMG_EHIP, std::string(#call) + ": " +              \
                                               hipGetErrorString(e_) + " @" +       \
                                               __FILE__ + ":" + std::to_string(__LINE__)); \
        }                                                                           \
    } while (0)

#define BCMG_CHECK_LAUNCH() BCMG_HIP(hipGetLastError())

/* ------------------------------------------------------- device memory */

/* Caching allocator of one context. Blocks are kept per size class (four
   classes per power of two) and handed out again in stream order: every
   kernel of a context runs on its one stream, so a block released while a
   kernel still reads it is only reused by work queued after that kernel.
   hipMalloc/hipFree of large buffers cost milliseconds and hipFree
   synchronises the device; the setup allocates and drops hundreds of
   temporaries, so it goes through the pool. */
struct Pool {
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> cls_;
    size_t cached = 0, live = 0;
    /* while set, released blocks are freed at once instead of cached (the
       serving loop gives a closed lane's memory back to the device) */
    bool free_released = false;
    static size_t size_class(size_t bytes);
    void *alloc(size_t bytes);
    void release(void *p);
    void trim();
    ~Pool();
};

/* Allocations made on this thread while a PoolScope is active come from its
   pool (the C-ABI opens one per call with the context's pool). */
Pool *current_pool();
struct PoolScope {
    Pool *prev;
    explicit PoolScope(Pool *p);
    ~PoolScope();
};

void *dev_alloc(Pool *pool, size_t bytes);
void dev_free(Pool *pool, void *p);

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    Pool *pool = nullptr;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), pool(o.pool) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; pool = o.pool;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        pool = current_pool();
        p = static_cast<T *>(dev_alloc(pool, count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) { if (count > n || !p) alloc(count); }
    void release() {
        if (p) dev_free(pool, p);
        p = nullptr;
        n = 0;
    }
    T *get() const { return p; }
};

/* -------------------------------------------------------------- matrix */

/* which stored values a row kernel reads */
enum MatType { MT_F64 = 0, MT_F32 = 1, MT_BF16 = 2 };

/* most distinct column offsets of a stencil-pattern operator */
constexpr int PAT_MAX = 8;

/* CSR, 0-based, int32 indices, FP64 values; columns ascending with a diagonal
   entry in every row (the layout of src/matrix/CSR.h:13-28). */
struct DevCsr {
    int n = 0, m = 0, nnz = 0;
    DBuf<int> row, col;
    DBuf<double> val;
    DBuf<float> val32;           /* reduced-precision copies of val (optional) */
    DBuf<unsigned short> val16;  /* bfloat16 bit patterns */
    /* fill val32 (bf16 = false) or val16 (bf16 = true) from val */
    void make_lowp(hipStream_t s, bool bf16);
    /* narrow copy of the columns, col - row in 16 bits, kept when every entry
       of a square operator whose FP64 stream (12 B per entry) is read
       non-temporally fits (banded and stencil orderings); the row kernels
       read it instead of col whenever the int32-column stream of the value
       type they use would be non-temporal */
    DBuf<short> col16;
    bool narrow = false;
    /* stencil pattern of the same operators when their column offsets take
       at most PAT_MAX distinct values: poff holds them ascending and bit k of
       pmask[r] is set when row r has column r + poff[k]; a row's entries are
       its set bits in ascending k, which is the CSR order, so the values stay
       where they are and the row kernels need neither the columns nor the
       row pointers beyond one per 256-row block (pk = 0: no pattern) */
    DBuf<unsigned char> pmask;
    int pk = 0;
    int poff[PAT_MAX] = {};
    bool cols_analysed = false; /* narrow copy and pattern considered */
    /* the narrowest column form the row kernels may read: 2 the pattern,
       1 the 16-bit offsets, 0 the int32 columns (bcmg_csr_set_columns) */
    int col_mode = 2;
    bool use_pattern() const { return pk > 0 && col_mode >= 2; }
    bool use_narrow() const { return narrow && col_mode >= 1; }
    int group = 8;          /* lanes per row of the row-group kernels */
    int tile = 0;           /* max nnz of a 256-row block; 0 = not analysed, -1 pending */
    int maxrow = 0;         /* longest row (valid with tile) */
    DBuf<int> tile_dev;     /* device copy while pending: {tile, maxrow} */
};

/* Lanes per row for an average row length (a power of two in [1, 64]). */
int lanes_for(int n, long long nnz);
struct Ctx;
/* choose the row kernel of A (LDS row blocks when a block's nnz fits) */
void analyze_rows(Ctx &c, DevCsr &A);
/* queue the tile computation of A (A.tile = -1 until resolve_tiles) */
void launch_block_nnz(Ctx &c, DevCsr &A);
void resolve_tiles(Ctx &c, const std::vector<DevCsr *> &mats);
/* build the narrow column copy of every square operator in mats above the
   non-temporal size whose column offsets all fit 16 bits, and its stencil
   pattern when those offsets take at most PAT_MAX values (one round trip
   for the lot) */
void make_narrow(Ctx &c, const std::vector<DevCsr *> &mats);

/* ------------------------------------------------------------- context */

struct Ctx {
    int device = 0;
    Pool pool;
    hipStream_t stream = nullptr;
    bool own_stream = true;     /* false: the caller's stream, not destroyed here */
    /* pinned host scalars used for the few device->host reads */
    double *h_scalars = nullptr;
    int *h_ints = nullptr;
    /* reduction scratch */
    DBuf<double> red;           /* partial sums (per block) + results */
    DBuf<int> iscratch;
    /* timing of the dominant kernel (optional) */
    bool time_smoother = false;
    /* event timing of the setup phases (optional) */
    bool time_setup = false;
    bool use_graphs = true;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_launches;  /* launches between each pair of events */
    hipEvent_t next_event();
    /* grow the reduction scratch for an n-row operator (outside any graph
       capture) */
    void ensure_reductions(int n);
    ~Ctx();
};

/* ---------------------------------------------------------- hierarchy */

struct Level {
    DevCsr *Aext = nullptr;     /* level 0: the caller's matrix */
    DevCsr Aown;                /* coarse levels: owned */
    const DevCsr &mat() const { return Aext ? *Aext : Aown; }
    DBuf<double> D, M;          /* diag(A), sum_j |a_ij| (may be empty) */
    bool hasD = false, hasM = false;
    DBuf<float> D32, M32, p32;  /* FP32 copies for the reduced-precision cycle */
    /* prolongator to the next level: one entry per fine row */
    DBuf<int> agg;
    DBuf<double> p;
    /* R = P^T as CSR (n_next rows) */
    DevCsr R;
    /* smoothed aggregation: the prolongator (I - omega D^-1 A) P_tent,
       stored NEGATED so that the residual row operation f - (-P) x gives
       the prolongation f + P x exactly (negation commutes with rounding);
       agg/p keep the tentative prolongator */
    bool smoothed = false;
    double omega = 0.0;
    DevCsr Pneg;
};

struct Hierarchy {
    std::vector<Level> L;
    double op_cmplx = 0, op_wcmplx = 0, avg_cratio = 0;
    int num_levels() const { return (int)L.size(); }
};

/* phases of the setup, timed with events on the context stream when
   Ctx::time_setup is set (the field order of bcmg_setup_times) */
enum SetupPhase {
    PH_WEIGHTS = 0,     /* matching weights AH and their logarithms */
    PH_MATCHING,        /* suitor matching and its symmetrisation */
    PH_AGGREGATION,     /* aggregates, P, R = P^T, P composition, w restriction */
    PH_GALERKIN,        /* coarse operators R (A P) */
    PH_SMOOTHER,        /* smoother divisors (diagonal, L1 row sums) */
    PH_INNER,           /* bootstrap inner iterations */
    PH_HIER,            /* the whole build of one hierarchy with its inner iterations */
    PH_COUNT
};

/* work vectors of the cycles in one compute type, per level, sized to the
   largest hierarchy */
template <typename VT>
struct CycleBufs {
    std::vector<DBuf<VT>> xa, xb, f;
    /* K-cycle inner iteration per level: w, d0, d1, A d0, A d1 */
    std::vector<DBuf<VT>> kw, kd0, kd1, kv, kv1;
    DBuf<VT> res;               /* residual buffer, size n0 */
    DBuf<VT> xsave, xadd;       /* additive composite */
};

struct MultiWork;
void destroy_multi(MultiWork *m);
void drop_multi_graph(MultiWork *m);

struct Boot {
    bcmg_params prm{};
    int coarse_solver = 4;
    std::vector<int> num_grid_sweeps;
    /* the composite's hierarchies; slot 1 shares slot 0's storage when it is
       built from the same smooth vector (w = 1, src/BOOT/bootstrap.cu:68-75) */
    std::vector<std::shared_ptr<Hierarchy>> H;
    double estimated_ratio = 1.0;
    /* estimated ratio after each hierarchy was added (printed per hierarchy
       by the reference, src/BOOT/bootstrap.cu:76) */
    std::vector<double> ratio_after;
    /* per hierarchy: event-timed phases (ms), zero unless timed */
    std::vector<std::vector<double>> t_phase;
    DevCsr *A0 = nullptr;       /* fine operator (owned by the caller's handle) */
    int prec = BCMG_PREC_FP64;  /* storage/compute precision of the cycles */
    int restarts = 0;           /* true-residual restarts of the last FCG */
    int fcg_calls = 0;          /* FCG solves started on this composite */
    CycleBufs<double> b64;
    CycleBufs<float> b32;
    DBuf<float> r32, z32;       /* level-0 input/output of an FP32 cycle */
    /* FCG vectors */
    DBuf<double> v, w, d, q, r;
    DBuf<double> scal;          /* FCG scalars on device */
    DBuf<double> kscal;         /* K-cycle scalars, KS_STRIDE per level */
    /* setup timing (ms) */
    double t_setup_ms = 0;
    /* captured FCG iterations (even, odd) and the output vector they use */
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const double *graph_x = nullptr;
    const double *graph_red = nullptr;
    /* multi-column work space (bcmg_fcg_multi and the serving loop) */
    MultiWork *multi = nullptr;
    void drop_graphs();
    void ensure_work();
    Boot() = default;
    Boot(const Boot &) = delete;
    Boot &operator=(const Boot &) = delete;
    ~Boot();
};

/* stored values a level's kernels read: FP64 for the drop-in; FP32, or
   bfloat16 on levels >= 1, for the reduced-precision preconditioner */
inline MatType level_mt(const Boot &B, int l)
{
    if (B.prec == BCMG_PREC_FP64) return MT_F64;
    if (B.prec == BCMG_PREC_BF16 && l > 0) return MT_BF16;
    return MT_F32;
}
/* per-precision views of a level's smoother divisor and prolongator values */
template <typename VT>
inline const VT *divisor(const Level &lv, int type)
{
    if constexpr (std::is_same<VT, double>::value)
        return type == 4 ? lv.M.get() : lv.D.get();
    else
        return type == 4 ? lv.M32.get() : lv.D32.get();
}
template <typename VT>
inline const VT *pvals(const Level &lv)
{
    if constexpr (std::is_same<VT, double>::value)
        return lv.p.get();
    else
        return lv.p32.get();
}

/* ----------------------------------------------------- launch wrappers */
/* solve-phase kernels (kernels_solve.hip). VT is the vector/compute type
   (double or float); MatType picks which stored values of a DevCsr a row
   kernel reads. */
/* OP_JACOBI_Z: the first two Jacobi sweeps from a zero iterate in one pass
   (the first sweep's iterate f / d is formed at each gather and for the row
   itself; x is not read) */
enum RowOp { OP_SPMV = 0, OP_RESID = 1, OP_JACOBI = 2, OP_JACOBI_Z = 3 };
template <typename VT>
void launch_rows(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x, const VT *f,
                 const VT *d, VT *y);
void launch_rows(hipStream_t s, const DevCsr &A, RowOp op, const double *x, const double *f,
                 const double *d, double *y);
template <typename VT>
void launch_jacobi_zero(hipStream_t s, int n, const VT *f, const VT *d, VT *y);
template <typename VT>
void launch_prolong(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc, VT *x);
/* y = x + P x_c with y another buffer than x */
template <typename VT>
void launch_prolong_to(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc,
                       const VT *x, VT *y);
template <typename VT>
void launch_fill(hipStream_t s, int n, VT v, VT *x);
template <typename VT>
void launch_copy(hipStream_t s, int n, const VT *a, VT *b);
/* y = alpha x + y */
template <typename VT>
void launch_axpy(hipStream_t s, int n, VT alpha, const VT *x, VT *y);
template <typename VT>
void launch_scale(hipStream_t s, int n, VT alpha, VT *x);
void launch_d2f(hipStream_t s, size_t n, const double *a, float *b);
void launch_f2d(hipStream_t s, size_t n, const float *a, double *b);
/* deterministic dot products: up to 3 dots of (a_k, b) ; results to out[k] (device) */
void launch_dots(Ctx &c, int n, int k, const double *a0, const double *a1, const double *a2,
                 const double *b, double *out);
/* FCG v3 scalar recurrences on device; s = [alpha, beta, gamma, delta, theta, theta2, l2, ...] */
void launch_fcg_scalars(hipStream_t s, double *scal);
/* z = -theta zo + z; x += theta2 z; t = -theta to + t; r += -theta2 t; scal[6] = ||r|| */
void launch_fcg_update(Ctx &c, int n, double *scal, const double *zo, double *z, double *x,
                       const double *to, double *t, double *r);
/* first FCG step: x += theta v; r -= theta w; ||r|| (theta from scal[4]) */
void launch_fcg_first(Ctx &c, int n, double *scal, const double *v, double *x,
                      const double *w, double *r);
/* r = b - A x (FP64, y is scratch), scal[6] = ||r|| */
void launch_true_residual(Ctx &c, const DevCsr &A, const double *b, const double *x, double *y,
                          double *r, double *scal);
/* t = A z with the FCG products of z (k = 2: r.z, t.z; k = 3: also to.z) and
   the scalar recurrence, all on the device */
void launch_spmv_dots(Ctx &c, const DevCsr &A, int k, const double *z, double *t,
                      const double *r, const double *to, double *scal);
/* K-cycle inner iteration pieces: products of VT vectors in double into
   out[0..k); alpha and the breakdown flag; w -= alpha v; the coefficients;
   x = c0 d0 + c1 d1 (device scalars ks, layout in kernels_solve.hip) */
constexpr int KS_STRIDE = 16;
constexpr int KS_FLAG_GLOBAL = 15; /* level 0's flag slot: any level's breakdown */
template <typename VT>
void launch_kdots(Ctx &c, int n, int k, const VT *a0, const VT *a1, const VT *a2, const VT *b,
                  double *out);
void launch_kcycle_alpha(hipStream_t s, double *ks, double *gflag);
void launch_kcycle_coef(hipStream_t s, double *ks, double rtol);
template <typename VT>
void launch_kcycle_axpy(hipStream_t s, int n, const double *ks, const VT *v, VT *w);
template <typename VT>
void launch_kcycle_combine(hipStream_t s, int n, const double *ks, const VT *d0, const VT *d1,
                           VT *x);

/* doubles of reduction scratch needed for an operator of n rows */
size_t red_size_for(int n);
/* all sweeps of a small level in one workgroup (when relax_block_ok(A));
   with ncols > 1, workgroup j does column j of f/xin/xout (element i of
   column j at j * cs + i * rs), skipped when act[j] == 0 */
template <typename VT>
bool relax_block_ok(const DevCsr &A);
/* whether the first two sweeps from zero run as one OP_JACOBI_Z pass on A */
template <typename VT>
bool fuse_zero_sweeps(const DevCsr &A, MatType mt);
/* work folded into a single-column one-workgroup relaxation (same
   arithmetic as its separate launches): with agg set, the prolongation
   x = x_in + P x_c (x_in = 0 when zero) before the sweeps; with fc set, the
   residual restricted to the next level after them, fc = R (f - A x) over
   the nc rows of R (CSR rrp, rci, rv) */
template <typename VT>
struct BlockFuse {
    const int *agg = nullptr;
    const VT *p = nullptr, *xc = nullptr;
    int nc = 0;
    const int *rrp = nullptr, *rci = nullptr;
    const VT *rv = nullptr;
    VT *fc = nullptr;
};
template <typename VT>
void launch_relax_block(hipStream_t s, const DevCsr &A, MatType mt, const VT *d, const VT *f,
                        const VT *xin, VT *xout, int sweeps, bool zero, int ncols = 1,
                        long cs = 0, int rs = 1, const int *act = nullptr,
                        const BlockFuse<VT> &fu = BlockFuse<VT>());
/* whether the row kernels sum A's rows in storage order from 0 (the staged
   kernels; otherwise the row-group kernels), the order fused kernels repeat */
template <typename VT>
bool rows_sequential(const DevCsr &A);
/* y = one Jacobi/L1 sweep from x + P x_c (the prolongation formed where it
   is read; the bits of a prolongation and a sweep launch); false when A
   takes the row-group kernels (then nothing is launched) */
template <typename VT>
bool launch_jacobi_prolong(hipStream_t s, const DevCsr &A, MatType mt, const int *agg,
                           const VT *p, const VT *xc, const VT *x, const VT *f, const VT *d,
                           VT *y);
/* the residual of a small cache-resident level restricted to the next one
   in one launch, y = R (f - A x), the bits of the residual and restriction
   launches; false when A or R takes the row-group kernels (then nothing is
   launched) */
template <typename VT>
bool launch_resid_restrict(hipStream_t s, const DevCsr &A, MatType mt, const DevCsr &R,
                           const VT *x, const VT *f, VT *y);

/* ---------------------------------------- multi-column kernels (kernels_multi.hip)
   Column blocks: column j of a level's block at base + j * ld. Each kernel
   applies per column the arithmetic of its single-column counterpart (so a
   column's bits do not depend on the other columns); act (device, one int
   per column, nullptr = all) skips columns. */
constexpr int MULTI_MAX_COLS = 32;
template <typename VT>
void launch_rows_multi(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x,
                       long ldx, const VT *f, const VT *d, VT *y, long ldy, int nc,
                       const int *act);
template <typename VT>
void launch_jacobi_zero_multi(hipStream_t s, int n, const VT *f, const VT *d, VT *y, long ld,
                              int nc, const int *act);
template <typename VT>
void launch_prolong_multi(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc,
                          long ldc, VT *x, long ld, int nc, const int *act);
template <typename VT>
void launch_fill_multi(hipStream_t s, int n, VT v, VT *x, long ld, int nc, const int *act);
template <typename S, typename D>
void launch_copy_multi(hipStream_t s, int n, const S *a, long lda, D *b, long ldb, int nc,
                       const int *act);
template <typename VT>
void launch_axpy_multi(hipStream_t s, int n, VT alpha, const VT *x, VT *y, long ld, int nc,
                       const int *act);
template <typename VT>
void launch_scale_multi(hipStream_t s, int n, VT alpha, VT *x, long ld, int nc, const int *act);

/* interleaved blocks (element (i, j) at i * 2^lkp + j): the cycle's layout */
template <typename VT>
void launch_rows_il(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x,
                    const VT *f, const VT *d, VT *y, int lkp, int nc, const int *act);
/* the last Jacobi/L1 sweep of a cycle with its result widened into the
   column-major FP64 block z (column j at z + j ldz); false when the operator
   takes the group kernel or the block is narrower than one vector (the
   caller then sweeps and converts) */
template <typename VT>
bool launch_jacobi_il_c64(hipStream_t s, const DevCsr &A, MatType mt, const VT *x, const VT *f,
                          const VT *d, double *z, long ldz, int lkp, int nc, const int *act);
template <typename VT>
void launch_jacobi_zero_il(hipStream_t s, int n, const VT *f, const VT *d, VT *y, int lkp,
                           int nc, const int *act);
template <typename VT>
void launch_prolong_il(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc, VT *x,
                       int lkp, int nc, const int *act);
template <typename VT>
void launch_fill_il(hipStream_t s, int n, VT v, VT *x, int lkp, int nc, const int *act);
template <typename VT>
void launch_copy_il(hipStream_t s, int n, const VT *a, VT *b, int lkp, int nc, const int *act);
template <typename VT>
void launch_axpy_il(hipStream_t s, int n, VT alpha, const VT *x, VT *y, int lkp, int nc,
                    const int *act);
template <typename VT>
void launch_scale_il(hipStream_t s, int n, VT alpha, VT *x, int lkp, int nc, const int *act);
/* column-major (leading dimension ld) <-> interleaved, converting S -> D */
template <typename S, typename D>
void launch_to_il(hipStream_t s, int n, const S *a, long lda, D *b, int lkp, int nc,
                  const int *act);
template <typename S, typename D>
void launch_from_il(hipStream_t s, int n, const S *a, D *b, long ldb, int lkp, int nc,
                    const int *act);

/* per-column FCG state on the device (FP64 blocks of leading dimension ld):
   direction pairs (V[s], W[s]), the newest in pair par[j]; Z the
   preconditioner output; R residual, X iterate, B right-hand side, Y
   scratch; scal 8 doubles per column (the single-column layout, [7] = the
   initial ||r||^2); act/it/tol the stopping test of each column. */
struct MultiFcg {
    double *V[2] = {nullptr, nullptr}, *W[2] = {nullptr, nullptr};
    double *Z = nullptr, *R = nullptr, *X = nullptr, *B = nullptr, *Y = nullptr;
    long ld = 0;
    int nc = 0;
    double *scal = nullptr;
    int *act = nullptr, *par = nullptr, *it = nullptr;
    double *tol = nullptr;
    int maxit = 0;
    double *kflag = nullptr;   /* K-cycle breakdown flag of column j at [j KS_STRIDE] */
};
/* work space of the multi-column solve on one composite: kmax column slots
   of every cycle and FCG vector, per-slot FCG state, pinned staging */
struct MultiWork {
    int kmax = 0;
    int prec = -1;             /* precision the float buffers were sized for */
    std::vector<long> ld;      /* per level (largest hierarchy) */
    CycleBufs<double> b64;
    CycleBufs<float> b32;
    DBuf<float> r32, z32;      /* level-0 input/output of an FP32 cycle */
    DBuf<double> r64, z64;     /* the same for an FP64 cycle */
    std::vector<long> nl;      /* rows per level (largest hierarchy) */
    int lkp = 0;               /* cycle blocks interleave 2^lkp columns */
    bool prec64 = false;       /* FP64 cycle buffers allocated */
    DBuf<double> vec;          /* V0 V1 W0 W1 Z R X B Y */
    DBuf<double> scal, tol, part;
    DBuf<int> flags;           /* act, par, it, mask: kmax each */
    DBuf<double> kscal;        /* K-cycle scalars: level l, column j at (l kmax + j) KS_STRIDE */
    MultiFcg F;
    int *mask = nullptr;       /* device mask of admission / check steps */
    /* pinned staging */
    double *h_scal = nullptr, *h_tol = nullptr;
    int *h_flags = nullptr;
    /* captured iterations, one per span of slots in use (2^0 .. 2^5) */
    hipGraphExec_t gexec[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int gprec = -1;
    void drop_graph()
    {
        for (auto &g : gexec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    ~MultiWork()
    {
        drop_graph();
        if (h_scal) (void)hipHostFree(h_scal);
        if (h_tol) (void)hipHostFree(h_tol);
        if (h_flags) (void)hipHostFree(h_flags);
    }
};

/* the K-cycle inner iteration on interleaved blocks: per column the
   arithmetic of launch_kdots / launch_kcycle_* (ks: column j at
   ks + j KS_STRIDE) */
template <typename VT>
void launch_kdots_il(hipStream_t s, int n, int k, const VT *a0, const VT *a1, const VT *a2,
                     const VT *b, int lkp, int nc, const int *act, double *part, double *ks,
                     int at);
void launch_kcycle_alpha_il(hipStream_t s, double *ks, double *gflag, int nc, const int *act);
void launch_kcycle_coef_il(hipStream_t s, double *ks, double rtol, int nc, const int *act);
template <typename VT>
void launch_kcycle_axpy_il(hipStream_t s, int n, const double *ks, const VT *v, VT *w, int lkp,
                           int nc, const int *act);
template <typename VT>
void launch_kcycle_combine_il(hipStream_t s, int n, const double *ks, const VT *d0, const VT *d1,
                              VT *x, int lkp, int nc, const int *act);

MultiWork &ensure_multi_into(Ctx &c, Boot &B, MultiWork &M, int kmax);
MultiWork &multi_work(Ctx &c, Boot &B, MultiWork *&holder, int kmax);
/* device bytes a MultiWork of kmax slots takes on composite B */
size_t multi_work_bytes(const Boot &B, int kmax);

enum SlotState { SLOT_FREE = 0, SLOT_RUN = 1, SLOT_DONE = 2 };
struct SlotInfo {
    int state = SLOT_FREE;
    int status = BCMG_OK;
    bool zero_rhs = false;
    int iters = 0, restarts = 0;
    double l2 = 0.0, delta0 = 0.0;
};

/* Per-slot FCG over a MultiWork: admit right-hand sides into free slots,
   launch one iteration of all active slots (queued), wait, collect the slots
   that stopped (true-residual restarts in a reduced precision), retire done
   slots. A slot's results are those of fcg_solve on its right-hand side. */
struct SlotEngine {
    Ctx &c;
    Boot &B;
    MultiWork &M;
    std::vector<SlotInfo> slots;
    bool launched = false;
    SlotEngine(Ctx &c, Boot &B, MultiWork &M);
    bool any_active() const;
    int free_slots() const;
    void admit(const std::vector<int> &js, const std::vector<const double *> &bs);
    void launch();
    void wait();
    int collect();
    void retire(int j, double *x);
};

/* the serving loop (server.hip) */
struct Server;
Server *server_create(Ctx &c, size_t budget, int slots);
void server_destroy(Server *S);
int server_add(Server &S, Boot &B);
long long server_submit(Server &S, int op, const double *b, double *x);
int server_step(Server &S);
bool server_result(Server &S, long long id, int *done, int *iters, double *resid, int *status,
                   double *latency_ms);
void server_info(const Server &S, int *queued, int *running, size_t *bytes, int *lanes);
int server_completed(Server &S, long long *ids, int cap);

/* doubles of partials scratch the multi-column FCG kernels need */
size_t multi_part_size(int n, int nc);
/* scal[8 j + 7] = ||R_j||^2 for masked columns */
void launch_norm2_multi(Ctx &c, int n, MultiFcg &M, const int *mask, double *part);
/* t = A Z with the products and recurrence (k = 2: first step into pair 0;
   k = 3: iteration into pair par ^ 1) */
void launch_spmv_dots_multi(Ctx &c, const DevCsr &A, int k, MultiFcg &M, const int *mask,
                            double *part);
/* iteration update of the active columns, then their stopping test */
void launch_fcg_update_multi(Ctx &c, int n, MultiFcg &M, double *part);
/* first step of the masked columns (direction Z moves to pair 0) */
void launch_fcg_first_multi(Ctx &c, int n, MultiFcg &M, const int *mask, double *part);
/* R = B - A X and scal[8 j + 6] = ||R_j|| for masked columns (Y scratch) */
void launch_true_residual_multi(Ctx &c, const DevCsr &A, MultiFcg &M, const int *mask,
                                double *part);

/* setup-phase kernels (kernels_setup.hip) */
void launch_diag(hipStream_t s, const DevCsr &A, double *D);
void launch_abs_rowsum(hipStream_t s, const DevCsr &A, double *M);
/* W aligned with A's nnz (diag slots 0) */
void match_weights(Ctx &c, const DevCsr &A, const double *w, double *W);
/* suitor + symmetrisation: Msym (and M raw); returns number of non-mutual v */
int match_suitor(Ctx &c, const DevCsr &A, const double *W, int *M, int *Msym, bool want_count);
/* canonical pairwise aggregation; returns n_c. members[2*I], [2*I+1] (or -1).
   validate: check that Msym is a symmetric matching first (BCMG_EINVAL if
   not) - for matchings that do not come from match_suitor */
int aggregate_pairs(Ctx &c, int n, const int *Msym, const double *w, int *agg, double *p,
                    int *members, bool validate = false);
/* A_c = R (A P) in the fixed order, pairwise P given by (agg, p, members) */
void galerkin_pairwise(Ctx &c, const DevCsr &A, const int *agg, const double *p,
                       const int *members, int nc, DevCsr &Ac);
/* w_c(I) = 0 + sum_{members} p_i w_i */
void restrict_w(hipStream_t s, int nc, const int *members, const double *p, const double *w,
                double *wc);
/* composed P: aggP[i] = agg[aggP[i]], pP[i] = 0 + pP[i] * p[aggP_old[i]] */
void compose_P(hipStream_t s, int n0, int *aggP, double *pP, const int *agg, const double *p);
/* R = P^T (rows ascending in fine index) */
void transpose_P(Ctx &c, int n, int nc, const int *agg, const double *p, DevCsr &R);
/* members of a given pairwise aggregation (throws if not pairwise) */
void members_from_agg(Ctx &c, int n, const int *agg, int nc, int *members);
/* exclusive scan of int32 counts into out[0..n], out[n] = total; returns total */
long long exclusive_scan(Ctx &c, const int *in, int *out, int n, bool read_total = true);

/* smoothed aggregation (kernels_spgemm.hip) */
/* a CSR operand by pointers: n rows, m columns */
struct CsrRef {
    int n = 0, m = 0, nnz = 0;
    const int *rp = nullptr, *ci = nullptr;
    const double *v = nullptr;
};
inline CsrRef ref_of(const DevCsr &A)
{
    return CsrRef{A.n, A.m, A.nnz, A.row.get(), A.col.get(), A.val.get()};
}
/* C = L R, symbolic pattern, columns ascending, each entry 0 + its products
   in the order of L's row then R's row (storage order); rows are taken in
   batches of at most per_pass output slots (min(products, columns of C) per
   row) plus packing weights (0: the library's default);
   right_unique: every row of R has distinct columns */
void spgemm(Ctx &c, const CsrRef &L, const CsrRef &R, DevCsr &C, long long per_pass = 0,
            bool right_unique = true);
/* T = A^T, each row in ascending column order */
void transpose_csr(Ctx &c, const CsrRef &A, DevCsr &T);
/* S = I - omega D^-1 A in A's pattern, omega = 4 / (3 ||D^-1 A||_inf) */
double sa_smoother(Ctx &c, const DevCsr &A, const double *D, DBuf<double> &S);
void launch_iota(hipStream_t s, int n, int *a);
void launch_negate(hipStream_t s, long long n, double *v);

/* ------------------------------------------------------------ host algo */
/* switch a built composite to BCMG_PREC_* (converts the hierarchy copies) */
void set_precision(Ctx &c, Boot &B, int prec);
void build_boot(Ctx &c, DevCsr &A, const bcmg_params &p, Boot &B);
void precond_apply(Ctx &c, Boot &B, const double *r, double *x, bool x_is_zero);
int fcg_solve(Ctx &c, Boot &B, const double *b, double *x, int *iters, double *resid,
              double *delta0_out);
/* k right-hand sides at once (device pointers b[j], x[j]); per column the
   iterations, final residual and status (BCMG_OK, BCMG_EZERORHS,
   BCMG_EBREAKDOWN) of fcg_solve on that column alone, bit for bit */
void fcg_multi(Ctx &c, Boot &B, int k, const double *const *b, double *const *x, int *iters,
               double *resid, int *status);
double a_norm(Ctx &c, const DevCsr &A, const double *x, double *tmp);

}  // namespace bcmg
