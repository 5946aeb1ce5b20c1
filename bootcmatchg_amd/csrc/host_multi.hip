/*
 * host_multi.hip - the multi-column solve: several right-hand sides on one
 * composite, each with its own FCG state, sharing every launch.
 *
 * Column slots: a MultiWork holds kmax slots; every vector of the cycles and
 * of FCG is a column block (slot j at base + j * ld). A slot is admitted
 * (b copied in, ||b||, first FCG step), iterates while its flag in `act` is
 * set (the stopping test runs on the device), and is retired (x copied
 * out). The kernels apply per column the arithmetic of the single-column
 * path, so a slot's iterations, residual and x are those of fcg_solve on
 * that right-hand side alone, whatever the other slots hold; only the
 * traffic of the matrices and the launch overhead are shared.
 *
 * One iteration of all active slots is one HIP graph replay (the slot
 * choices live on the device: act, par), followed by one read of the flags.
 */
#include "bcmg_internal.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace bcmg {

void destroy_multi(MultiWork *m) { delete m; }
void drop_multi_graph(MultiWork *m)
{
    if (m) m->drop_graph();
}

namespace {

inline long round_ld(size_t n) { return (long)((n + 63) / 64 * 64); }

/* cycle blocks: level l holds n_l rows of 2^lkp interleaved columns */
template <typename VT>
void ensure_mbufs(const Boot &B, const MultiWork &M, CycleBufs<VT> &W)
{
    const size_t L = M.nl.size();
    if (W.xa.size() < L) {
        W.xa.resize(L);
        W.xb.resize(L);
        W.f.resize(L);
    }
    for (size_t l = 0; l < L; l++) {
        const size_t sz = (size_t)M.nl[l] << M.lkp;
        W.xa[l].ensure(sz);
        W.xb[l].ensure(sz);
        W.f[l].ensure(sz);
    }
    const size_t s0 = (size_t)M.nl[0] << M.lkp;
    W.res.ensure(s0);
    if (B.prm.solver_type == 2) {
        W.xadd.ensure(s0);
        W.xsave.ensure(s0);
    }
    if (B.prm.cycle_type == 3) {
        for (auto *v : {&W.kw, &W.kd0, &W.kd1, &W.kv, &W.kv1}) {
            if (v->size() < L) v->resize(L);
            for (size_t l = 0; l < L; l++) (*v)[l].ensure((size_t)M.nl[l] << M.lkp);
        }
    }
}

inline int log2_cols(int k)
{
    int l = 0;
    while ((1 << l) < k) l++;
    return l;
}

} // namespace

/* work space for kmax slots (kept and grown across calls) */
MultiWork &ensure_multi_into(Ctx &c, Boot &B, MultiWork &M, int kmax)
{
    if (kmax < 1 || kmax > MULTI_MAX_COLS)
        throw Error(BCMG_EINVAL, "multi-column solve: 1..32 columns per block");
    const int n = B.A0->n;
    const bool grow = kmax > M.kmax;
    if (grow) {
        M.drop_graph();
        M.kmax = kmax;
        size_t maxL = 0;
        for (auto &h : B.H) maxL = std::max(maxL, h->L.size());
        M.ld.assign(maxL, 0);
        M.nl.assign(maxL, 0);
        for (size_t l = 0; l < maxL; l++) {
            size_t nl = 0;
            for (auto &h : B.H)
                if (l < h->L.size()) nl = std::max(nl, (size_t)h->L[l].mat().n);
            M.ld[l] = round_ld(nl);
            M.nl[l] = (long)nl;
        }
        M.lkp = log2_cols(kmax);
        const size_t blk = (size_t)kmax * M.ld[0];
        M.vec.alloc(9 * blk);
        M.scal.alloc((size_t)kmax * 8);
        M.tol.alloc(kmax);
        M.part.alloc(multi_part_size(n, kmax));
        /* act, par, it, mask; the pad lets kernels read MULTI_MAX_COLS flags
           from act or mask unconditionally */
        M.flags.alloc(4 * (size_t)kmax + MULTI_MAX_COLS);
        if (M.h_scal) (void)hipHostFree(M.h_scal);
        if (M.h_tol) (void)hipHostFree(M.h_tol);
        if (M.h_flags) (void)hipHostFree(M.h_flags);
        M.h_scal = M.h_tol = nullptr;
        M.h_flags = nullptr;
        BCMG_HIP(hipHostMalloc((void **)&M.h_scal, (size_t)kmax * 8 * sizeof(double)));
        BCMG_HIP(hipHostMalloc((void **)&M.h_tol, (size_t)kmax * sizeof(double)));
        BCMG_HIP(hipHostMalloc((void **)&M.h_flags, 4 * (size_t)kmax * sizeof(int)));
        double *v = M.vec.get();
        MultiFcg &F = M.F;
        F.V[0] = v;
        F.V[1] = v + blk;
        F.W[0] = v + 2 * blk;
        F.W[1] = v + 3 * blk;
        F.Z = v + 4 * blk;
        F.R = v + 5 * blk;
        F.X = v + 6 * blk;
        F.B = v + 7 * blk;
        F.Y = v + 8 * blk;
        F.ld = M.ld[0];
        F.scal = M.scal.get();
        F.tol = M.tol.get();
        F.act = M.flags.get();
        F.par = F.act + kmax;
        F.it = F.act + 2 * kmax;
        M.mask = F.act + 3 * kmax;
        /* no slot active, no stale flags */
        BCMG_HIP(hipMemsetAsync(M.flags.get(), 0, (4 * (size_t)kmax + MULTI_MAX_COLS) * sizeof(int),
                                c.stream));
        std::memset(M.h_flags, 0, 4 * (size_t)kmax * sizeof(int));
        M.prec = -1;
        M.prec64 = false;
    }
    if (B.prec == BCMG_PREC_FP64 && !M.prec64) {
        ensure_mbufs<double>(B, M, M.b64);
        M.r64.ensure((size_t)M.nl[0] << M.lkp);
        M.z64.ensure((size_t)M.nl[0] << M.lkp);
        M.prec64 = true;
    }
    if (B.prm.cycle_type == 3) {
        /* K-cycle scalars per level and column; level 0's flag slot of a
           column is its breakdown flag (cleared at admission) */
        const size_t need = M.nl.size() * (size_t)M.kmax * KS_STRIDE;
        if (M.kscal.n < need || !M.kscal.get()) {
            M.kscal.alloc(need);
            BCMG_HIP(hipMemsetAsync(M.kscal.get(), 0, need * sizeof(double), c.stream));
        }
        M.F.kflag = M.kscal.get() + KS_FLAG_GLOBAL;
    } else {
        M.F.kflag = nullptr;
    }
    M.F.nc = M.kmax;
    M.F.maxit = B.prm.itnlim;
    if (B.prec != BCMG_PREC_FP64 && M.prec != B.prec) {
        ensure_mbufs<float>(B, M, M.b32);
        M.r32.ensure((size_t)M.nl[0] << M.lkp);
        M.z32.ensure((size_t)M.nl[0] << M.lkp);
        M.prec = B.prec;
    }
    return M;
}

namespace {

template <typename VT>
CycleBufs<VT> &mbufs(MultiWork &M);
template <>
CycleBufs<double> &mbufs<double>(MultiWork &M) { return M.b64; }
template <>
CycleBufs<float> &mbufs<float>(MultiWork &M) { return M.b32; }

template <typename VT>
struct MState {
    VT *x = nullptr, *x2 = nullptr;
    bool zero = true;
};

template <typename VT>
struct MRun {
    Ctx &c;
    Boot &B;
    MultiWork &M;
    const Hierarchy &H;
    std::vector<MState<VT>> &st;
    const VT *f0;
    VT *final_out;
    int nc;
    int lkp;    /* interleave stride 2^lkp of this application */
    const int *act;
    /* when set, the sweep that writes final_out writes its result widened
       into this column-major FP64 block instead (and sets *fused) */
    double *z64 = nullptr;
    long ldz = 0;
    bool *fused = nullptr;
};

/* relax_level of host_amg.hip on an interleaved block */
template <typename VT>
void relax_level_m(MRun<VT> &R, const Level &lv, MatType mt, int type, int sweeps, const VT *f,
                   MState<VT> &s, VT *final_out)
{
    Ctx &c = R.c;
    const int lkp = R.lkp;
    const DevCsr &A = lv.mat();
    const VT *d = divisor<VT>(lv, type);
    if (sweeps == 0) return;
    if (relax_block_ok<VT>(A)) {
        VT *out = (final_out && s.x != final_out) ? final_out : s.x;
        launch_relax_block<VT>(c.stream, A, mt, d, f, s.x, out, sweeps, s.zero, R.nc, 1,
                               1 << lkp, R.act);
        if (out != s.x) {
            s.x2 = s.x;
            s.x = out;
        }
        s.zero = false;
        return;
    }
    for (int k = 0; k < sweeps; k++) {
        VT *out = s.x2;
        const bool last = final_out && k == sweeps - 1;
        if (last && s.x != final_out) out = final_out;
        if (s.zero) {
            launch_jacobi_zero_il<VT>(c.stream, A.n, f, d, out, lkp, R.nc, R.act);
        } else if (last && R.z64 &&
                   launch_jacobi_il_c64<VT>(c.stream, A, mt, s.x, f, d, R.z64, R.ldz, lkp, R.nc,
                                            R.act)) {
            /* the result is in z64 only; final_out is not written */
            out = final_out;
            *R.fused = true;
        } else {
            launch_rows_il<VT>(c.stream, A, mt, OP_JACOBI, s.x, f, d, out, lkp, R.nc, R.act);
        }
        s.x2 = s.x;
        s.x = out;
        s.zero = false;
    }
}

/* inner-iteration tolerance of the K-cycle (src/SOLVER/GAMG_cycle.cu:5) */
constexpr double KCYCLE_RTOL_M = 0.25;

template <typename VT>
void cycle_m(MRun<VT> &R, int l, const VT *fin = nullptr);

/* kcycle_inner of host_amg.hip on interleaved blocks: per column, at most
   two flexible CG steps on A_m x = f_m preconditioned by the cycle from level
   m + 1 (src/SOLVER/FCG.cu:620-723), every decision on the device */
template <typename VT>
void kcycle_inner_m(MRun<VT> &R, int m)
{
    Ctx &c = R.c;
    Boot &B = R.B;
    MultiWork &M = R.M;
    CycleBufs<VT> &W = mbufs<VT>(M);
    const DevCsr &A = R.H.L[m].mat();
    const MatType mt = level_mt(B, m);
    const int n = A.n, lkp = R.lkp, nc = R.nc;
    const int *act = R.act;
    double *ks = M.kscal.get() + (size_t)m * M.kmax * KS_STRIDE;
    double *gflag = M.kscal.get() + KS_FLAG_GLOBAL;
    double *part = M.part.get();
    VT *w = W.kw[m].get(), *d0 = W.kd0[m].get(), *d1 = W.kd1[m].get();
    VT *v = W.kv[m].get(), *v1 = W.kv1[m].get();
    MState<VT> &s = R.st[m];
    launch_copy_il<VT>(c.stream, n, W.f[m].get(), w, lkp, nc, act);
    launch_kdots_il<VT>(c.stream, n, 1, w, nullptr, nullptr, w, lkp, nc, act, part, ks, 0);
    auto apply = [&](VT *out) {
        s.zero = true;
        cycle_m<VT>(R, m + 1, w);
        if (s.zero)
            launch_fill_il<VT>(c.stream, n, VT(0), out, lkp, nc, act);
        else
            launch_copy_il<VT>(c.stream, n, s.x, out, lkp, nc, act);
    };
    apply(d0);
    launch_kdots_il<VT>(c.stream, n, 1, w, nullptr, nullptr, d0, lkp, nc, act, part, ks, 1);
    launch_rows_il<VT>(c.stream, A, mt, OP_SPMV, d0, nullptr, nullptr, v, lkp, nc, act);
    launch_kdots_il<VT>(c.stream, n, 1, d0, nullptr, nullptr, v, lkp, nc, act, part, ks, 2);
    launch_kcycle_alpha_il(c.stream, ks, gflag, nc, act);
    launch_kcycle_axpy_il<VT>(c.stream, n, ks, v, w, lkp, nc, act);
    launch_kdots_il<VT>(c.stream, n, 1, w, nullptr, nullptr, w, lkp, nc, act, part, ks, 4);
    apply(d1);
    launch_rows_il<VT>(c.stream, A, mt, OP_SPMV, d1, nullptr, nullptr, v1, lkp, nc, act);
    launch_kdots_il<VT>(c.stream, n, 3, v, v1, w, d1, lkp, nc, act, part, ks, 5);
    launch_kcycle_coef_il(c.stream, ks, KCYCLE_RTOL_M, nc, act);
    launch_kcycle_combine_il<VT>(c.stream, n, ks, d0, d1, s.x, lkp, nc, act);
    s.zero = false;
}

/* cycle of host_amg.hip on interleaved blocks (src/SOLVER/GAMG_cycle.cu:25-140);
   fin, when given, is the level's right-hand side (an inner K-cycle
   application) */
template <typename VT>
void cycle_m(MRun<VT> &R, int l, const VT *fin)
{
    Ctx &c = R.c;
    Boot &B = R.B;
    CycleBufs<VT> &W = mbufs<VT>(R.M);
    const int lkp = R.lkp;
    const Hierarchy &H = R.H;
    const int L = H.num_levels();
    const Level &lv = H.L[l - 1];
    const MatType mt = level_mt(B, l - 1);
    MState<VT> &s = R.st[l - 1];
    const VT *f = fin ? fin : ((l == 1) ? R.f0 : W.f[l - 1].get());
    const bool top = (l == 1);
    if (l == L) {
        relax_level_m<VT>(R, lv, mt, B.coarse_solver, B.prm.relaxnumber_coarse, f, s,
                          top ? R.final_out : nullptr);
        return;
    }
    relax_level_m<VT>(R, lv, mt, B.prm.relax_type, B.prm.prerelax_sweeps, f, s, nullptr);
    const DevCsr &A = lv.mat();
    const VT *rsrc = f;
    if (!s.zero) {
        launch_rows_il<VT>(c.stream, A, mt, OP_RESID, s.x, f, nullptr, W.res.get(), lkp, R.nc,
                           R.act);
        rsrc = W.res.get();
    }
    launch_rows_il<VT>(c.stream, lv.R, std::is_same<VT, double>::value ? MT_F64 : MT_F32,
                       OP_SPMV, rsrc, nullptr, nullptr, W.f[l].get(), lkp, R.nc, R.act);
    MState<VT> &sc = R.st[l];
    sc.zero = true;
    if (B.prm.cycle_type == 3 && l != L - 1) {
        kcycle_inner_m<VT>(R, l);
    } else {
        for (int i = 1; i <= B.num_grid_sweeps[l - 1]; i++) {
            cycle_m<VT>(R, l + 1);
            if (l == L - 1) break;
        }
    }
    if (!sc.zero) {
        if (s.zero) {
            launch_fill_il<VT>(c.stream, A.n, VT(0), s.x, lkp, R.nc, R.act);
            s.zero = false;
        }
        if (lv.smoothed) {
            /* x = x - (-P) x_c: the stored negated prolongator */
            launch_rows_il<VT>(c.stream, lv.Pneg,
                               std::is_same<VT, double>::value ? MT_F64 : MT_F32, OP_RESID,
                               sc.x, s.x, nullptr, s.x2, lkp, R.nc, R.act);
            std::swap(s.x, s.x2);
        } else {
            launch_prolong_il<VT>(c.stream, A.n, lv.agg.get(), pvals<VT>(lv), sc.x, s.x, lkp,
                                  R.nc, R.act);
        }
    }
    relax_level_m<VT>(R, lv, mt, B.prm.relax_type, B.prm.postrelax_sweeps, f, s,
                      top ? R.final_out : nullptr);
}

template <typename VT>
void run_hierarchy_m(MRun<VT> &R, int k)
{
    CycleBufs<VT> &W = mbufs<VT>(R.M);
    const Hierarchy &H = *R.B.H[k];
    for (int l = 1; l < H.num_levels(); l++) {
        R.st[l].x = W.xa[l].get();
        R.st[l].x2 = W.xb[l].get();
        R.st[l].zero = true;
    }
    MRun<VT> Rk{R.c, R.B, R.M, H, R.st, R.f0, R.final_out, R.nc, R.lkp, R.act, R.z64, R.ldz,
                R.fused};
    cycle_m<VT>(Rk, 1);
}

/* precond_apply_t of host_amg.hip from x = 0 on interleaved blocks. With
   z64, a multiplicative or symmetrized composite may deliver its result
   widened into that column-major FP64 block instead of x; *fused tells. */
template <typename VT>
void precond_m(Ctx &c, Boot &B, MultiWork &M, const VT *r, VT *x, const int *act,
               double *z64 = nullptr, long ldz = 0, bool *fused = nullptr)
{
    if (fused) *fused = false;
    CycleBufs<VT> &W = mbufs<VT>(M);
    const int n = B.A0->n;
    const int nc = M.F.nc; /* columns in use */
    /* the cycle's blocks are rebuilt from the FCG blocks at every
       application, so their stride follows the columns in use: a lightly
       used block stays dense */
    const int lkp = log2_cols(nc);
    const int nh = (int)B.H.size();
    std::vector<MState<VT>> st(M.nl.size());
    if (B.prm.solver_type == 0 || B.prm.solver_type == 1) {
        st[0].x = x;
        st[0].x2 = W.xa[0].get();
        st[0].zero = true;
        std::vector<int> order;
        for (int k = 0; k < nh; k++) order.push_back(k);
        if (B.prm.solver_type == 1)
            for (int k = nh - 1; k >= 0; k--) order.push_back(k);
        for (size_t t = 0; t < order.size(); t++) {
            const bool last = (t + 1 == order.size());
            MRun<VT> R{c,   B,   M,   *B.H[order[t]], st,   r, last ? x : nullptr,
                       nc,  lkp, act, last ? z64 : nullptr, ldz, fused};
            run_hierarchy_m<VT>(R, order[t]);
        }
        if (fused && *fused) return;
        if (st[0].zero)
            launch_fill_il<VT>(c.stream, n, VT(0), x, lkp, nc, act);
        else if (st[0].x != x)
            launch_copy_il<VT>(c.stream, n, st[0].x, x, lkp, nc, act);
    } else {
        launch_fill_il<VT>(c.stream, n, VT(0), W.xadd.get(), lkp, nc, act);
        for (int k = 0; k < nh; k++) {
            st[0].x = W.xb[0].get();
            st[0].x2 = W.xa[0].get();
            st[0].zero = true;
            MRun<VT> R{c, B, M, *B.H[k], st, r, nullptr, nc, lkp, act};
            run_hierarchy_m<VT>(R, k);
            if (st[0].zero) launch_fill_il<VT>(c.stream, n, VT(0), st[0].x, lkp, nc, act);
            launch_axpy_il<VT>(c.stream, n, VT(1), st[0].x, W.xadd.get(), lkp, nc, act);
        }
        launch_scale_il<VT>(c.stream, n, VT(1) / (VT)nh, W.xadd.get(), lkp, nc, act);
        launch_copy_il<VT>(c.stream, n, W.xadd.get(), x, lkp, nc, act);
    }
}

/* Z = B(R) for the masked slots: R is turned into the cycle's interleaved
   layout (rounded to FP32 for a reduced-precision hierarchy), the cycles run,
   and the result comes back column-major in FP64: written there by the
   cycle's last sweep where that sweep can, converted after otherwise */
void precond_block(Ctx &c, Boot &B, MultiWork &M, const int *act)
{
    const int n = B.A0->n;
    const long ld = M.F.ld;
    const int nc = M.F.nc, lkp = log2_cols(nc);
    bool fused = false;
    if (B.prec == BCMG_PREC_FP64) {
        launch_to_il<double, double>(c.stream, n, M.F.R, ld, M.r64.get(), lkp, nc, act);
        precond_m<double>(c, B, M, M.r64.get(), M.z64.get(), act, M.F.Z, ld, &fused);
        if (!fused)
            launch_from_il<double, double>(c.stream, n, M.z64.get(), M.F.Z, ld, lkp, nc, act);
        return;
    }
    launch_to_il<double, float>(c.stream, n, M.F.R, ld, M.r32.get(), lkp, nc, act);
    precond_m<float>(c, B, M, M.r32.get(), M.z32.get(), act, M.F.Z, ld, &fused);
    if (!fused) launch_from_il<float, double>(c.stream, n, M.z32.get(), M.F.Z, ld, lkp, nc, act);
}

void upload_flags(Ctx &c, MultiWork &M)
{
    BCMG_HIP(hipMemcpyAsync(M.flags.get(), M.h_flags, 4 * (size_t)M.kmax * sizeof(int),
                            hipMemcpyHostToDevice, c.stream));
}

void read_back(Ctx &c, MultiWork &M, bool flags, bool scal)
{
    if (flags)
        BCMG_HIP(hipMemcpyAsync(M.h_flags, M.flags.get(), 3 * (size_t)M.kmax * sizeof(int),
                                hipMemcpyDeviceToHost, c.stream));
    if (scal)
        BCMG_HIP(hipMemcpyAsync(M.h_scal, M.scal.get(), (size_t)M.kmax * 8 * sizeof(double),
                                hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
}

/* first FCG step of the slots whose host mask entry is set (the flags are
   uploaded first; act/par/it of those slots come from h_flags) */
void first_step(Ctx &c, Boot &B, MultiWork &M)
{
    const DevCsr &A = *B.A0;
    upload_flags(c, M);
    precond_block(c, B, M, M.mask);
    launch_spmv_dots_multi(c, A, 2, M.F, M.mask, M.part.get());
    launch_fcg_first_multi(c, A.n, M.F, M.mask, M.part.get());
}

/* one FCG iteration of every active slot (src/SOLVER/FCG.cu:497-565) */
void enqueue_iteration_m(Ctx &c, Boot &B, MultiWork &M)
{
    const DevCsr &A = *B.A0;
    precond_block(c, B, M, M.F.act);
    launch_spmv_dots_multi(c, A, 3, M.F, M.F.act, M.part.get());
    launch_fcg_update_multi(c, A.n, M.F, M.part.get());
    BCMG_HIP(hipMemcpyAsync(M.h_flags, M.flags.get(), 3 * (size_t)M.kmax * sizeof(int),
                            hipMemcpyDeviceToHost, c.stream));
}

} // namespace

/* ------------------------------------------------------------ slot engine */

SlotEngine::SlotEngine(Ctx &c_, Boot &B_, MultiWork &M_) : c(c_), B(B_), M(M_)
{
    slots.assign(M.kmax, SlotInfo());
}

bool SlotEngine::any_active() const
{
    for (int j = 0; j < M.kmax; j++)
        if (slots[j].state == SLOT_RUN && M.h_flags[j]) return true;
    return false;
}

int SlotEngine::free_slots() const
{
    int f = 0;
    for (auto &s : slots) f += s.state == SLOT_FREE;
    return f;
}

/* Admission of right-hand sides into free slots (fcg_solve's start,
   src/SOLVER/FCG.cu:445-486): B = b, R = b, X = 0, delta0 = ||b||, the
   stopping tolerance rtol delta0, then the first step; a slot with b = 0 is
   done at once with BCMG_EZERORHS (x = 0). Every admitted slot then runs at
   least one iteration (FCG.cu:488-497). Synchronises once (for ||b||). */
void SlotEngine::admit(const std::vector<int> &js, const std::vector<const double *> &bs)
{
    if (js.empty()) return;
    MultiFcg &F = M.F;
    const int n = B.A0->n, km = M.kmax;
    const long ld = F.ld;
    int *h_act = M.h_flags, *h_par = M.h_flags + km, *h_it = M.h_flags + 2 * km,
        *h_mask = M.h_flags + 3 * km;
    F.nc = km; /* admission launches span every slot (masked) */
    for (int j = 0; j < km; j++) h_mask[j] = 0;
    for (size_t q = 0; q < js.size(); q++) {
        const int j = js[q];
        if (j < 0 || j >= km || slots[j].state != SLOT_FREE)
            throw Error(BCMG_EINVAL, "admission into a slot that is not free");
        h_mask[j] = 1;
        h_act[j] = 0;
        h_par[j] = 0;
        h_it[j] = 0;
        launch_copy<double>(c.stream, n, bs[q], F.B + (size_t)j * ld);
        if (F.kflag)
            BCMG_HIP(hipMemsetAsync(F.kflag + (size_t)j * KS_STRIDE, 0, sizeof(double),
                                    c.stream));
    }
    upload_flags(c, M);
    launch_copy_multi<double, double>(c.stream, n, F.B, ld, F.R, ld, km, M.mask);
    launch_fill_multi<double>(c.stream, n, 0.0, F.X, ld, km, M.mask);
    launch_norm2_multi(c, n, F, M.mask, M.part.get());
    read_back(c, M, false, true);
    bool any = false;
    for (int j : js) {
        SlotInfo &S = slots[j];
        S = SlotInfo();
        S.delta0 = std::sqrt(M.h_scal[(size_t)j * 8 + 7]);
        if (S.delta0 <= DBL_EPSILON * S.delta0) {
            /* the reference exits on a zero initial residual (FCG.cu:459-462) */
            S.state = SLOT_DONE;
            S.status = BCMG_EZERORHS;
            S.zero_rhs = true;
            h_mask[j] = 0;
            continue;
        }
        S.state = SLOT_RUN;
        M.h_tol[j] = B.prm.rtol * S.delta0;
        h_act[j] = 1;
        any = true;
    }
    if (!any) return;
    BCMG_HIP(hipMemcpyAsync(M.tol.get(), M.h_tol, (size_t)km * sizeof(double),
                            hipMemcpyHostToDevice, c.stream));
    first_step(c, B, M);
    /* h_flags stays untouched until the next synchronisation */
}

/* One iteration of every active slot, queued (graph replay); wait() ends it.
   The launches span the slots 0 .. 2^s - 1 that hold the highest active
   slot (admission fills the lowest free slots), one captured graph per
   span: a lane with few requests does not pay for its idle slots. */
void SlotEngine::launch()
{
    int hi = -1;
    for (int j = 0; j < M.kmax; j++)
        if (slots[j].state == SLOT_RUN && M.h_flags[j]) hi = j;
    if (hi < 0) return;
    int cls = 0;
    while ((1 << cls) < hi + 1) cls++;
    const int span = std::min(1 << cls, M.kmax);
    M.F.nc = span;
    if (M.gprec != B.prec) M.drop_graph();
    if (c.use_graphs) {
        hipGraphExec_t &ge = M.gexec[cls];
        if (!ge) {
            hipGraph_t g = nullptr;
            BCMG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_iteration_m(c, B, M);
            } catch (...) {
                (void)hipStreamEndCapture(c.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            BCMG_HIP(hipStreamEndCapture(c.stream, &g));
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            BCMG_HIP(e);
            M.gprec = B.prec;
        }
        BCMG_HIP(hipGraphLaunch(ge, c.stream));
    } else {
        enqueue_iteration_m(c, B, M);
    }
    launched = true;
}

void SlotEngine::wait()
{
    BCMG_HIP(hipStreamSynchronize(c.stream));
    launched = false;
}

/* Slots whose stopping test ended their iterations: FP64 (the reference's
   recursive test) makes them done; in a reduced precision the true residual
   decides between done and a restart from it (fcg_solve's rule, per slot).
   Returns the number of slots that became done. */
int SlotEngine::collect()
{
    const int km = M.kmax;
    int *h_act = M.h_flags, *h_par = M.h_flags + km, *h_it = M.h_flags + 2 * km,
        *h_mask = M.h_flags + 3 * km;
    const bool lowp = B.prec != BCMG_PREC_FP64;
    const int max_iter = B.prm.itnlim;
    int done = 0;
    M.F.nc = km;
    for (;;) {
        std::vector<int> stopped;
        for (int j = 0; j < km; j++)
            if (slots[j].state == SLOT_RUN && !h_act[j]) stopped.push_back(j);
        if (stopped.empty()) break;
        read_back(c, M, false, true);
        for (int j = 0; j < km; j++) h_mask[j] = 0;
        bool check = false;
        for (int j : stopped) {
            SlotInfo &S = slots[j];
            S.l2 = M.h_scal[(size_t)j * 8 + 6];
            S.iters = h_it[j] + 1;
            if (!lowp || !std::isfinite(S.l2)) {
                S.state = SLOT_DONE;
                S.status = std::isfinite(S.l2) ? BCMG_OK : BCMG_EBREAKDOWN;
                done++;
                continue;
            }
            h_mask[j] = 1;
            check = true;
        }
        if (!check) break;
        upload_flags(c, M);
        launch_true_residual_multi(c, *B.A0, M.F, M.mask, M.part.get());
        read_back(c, M, false, true);
        bool restart = false;
        for (int j : stopped) {
            if (!h_mask[j]) continue;
            SlotInfo &S = slots[j];
            S.l2 = M.h_scal[(size_t)j * 8 + 6];
            if (S.l2 <= M.h_tol[j] || h_it[j] >= max_iter || !std::isfinite(S.l2)) {
                S.state = SLOT_DONE;
                S.status = std::isfinite(S.l2) ? BCMG_OK : BCMG_EBREAKDOWN;
                done++;
                h_mask[j] = 0;
                continue;
            }
            h_par[j] = 0;
            restart = true;
        }
        if (!restart) break;
        /* restart from the true residual: the first step again */
        first_step(c, B, M);
        read_back(c, M, false, true);
        for (int j : stopped) {
            if (!h_mask[j]) continue;
            SlotInfo &S = slots[j];
            S.restarts++;
            B.restarts++;
            h_it[j]++;
            S.l2 = M.h_scal[(size_t)j * 8 + 6];
            S.iters = h_it[j] + 1;
            h_act[j] = (S.l2 > M.h_tol[j] && h_it[j] < max_iter) ? 1 : 0;
            h_mask[j] = 0;
        }
        upload_flags(c, M);
        /* slots whose restart already meets the test go round again */
    }
    return done;
}

/* copy a done slot's x out (queued on the slot engine's stream) and free it */
void SlotEngine::retire(int j, double *x)
{
    SlotInfo &S = slots[j];
    if (S.state != SLOT_DONE) throw Error(BCMG_EINVAL, "retiring a slot that is not done");
    const int n = B.A0->n;
    if (x) {
        if (S.zero_rhs)
            launch_fill<double>(c.stream, n, 0.0, x);
        else
            launch_copy<double>(c.stream, n, M.F.X + (size_t)j * M.F.ld, x);
    }
    S.state = SLOT_FREE;
}

MultiWork &multi_work(Ctx &c, Boot &B, MultiWork *&holder, int kmax)
{
    if (!holder) holder = new MultiWork();
    return ensure_multi_into(c, B, *holder, kmax);
}

size_t multi_work_bytes(const Boot &B, int kmax)
{
    size_t maxL = 0;
    for (auto &h : B.H) maxL = std::max(maxL, h->L.size());
    const size_t kp = (size_t)1 << log2_cols(kmax);
    /* per level: x, x2, f (and, for the K-cycle, its five inner vectors) */
    const size_t per_level = B.prm.cycle_type == 3 ? 8 : 3;
    size_t cyc = 0, n0 = 0;
    for (size_t l = 0; l < maxL; l++) {
        size_t nl = 0;
        for (auto &h : B.H)
            if (l < h->L.size()) nl = std::max(nl, (size_t)h->L[l].mat().n);
        if (l == 0) n0 = nl;
        cyc += per_level * nl * kp;
    }
    cyc += n0 * kp * (3 + (B.prm.solver_type == 2 ? 2 : 0)); /* res, r, z (+ additive) */
    const size_t ksc = B.prm.cycle_type == 3 ? maxL * (size_t)kmax * KS_STRIDE * 8 : 0;
    const size_t vt = B.prec == BCMG_PREC_FP64 ? 8 : 4;
    const size_t fcg = 9 * (size_t)round_ld(n0) * kmax * 8;
    return cyc * vt + fcg + ksc + multi_part_size(B.A0->n, kmax) * 8 + (size_t)kmax * 80;
}

/* k right-hand sides through one slot engine: admit all, iterate until
   every slot is done, copy the results out */
namespace {

void fcg_block(Ctx &c, Boot &B, int k, const double *const *b, double *const *x, int *iters,
               double *resid, int *status)
{
    MultiWork &M = multi_work(c, B, B.multi, k);
    SlotEngine E(c, B, M);
    std::vector<int> js(k);
    std::vector<const double *> bs(k);
    for (int j = 0; j < k; j++) {
        js[j] = j;
        bs[j] = b[j];
    }
    E.admit(js, bs);
    for (;;) {
        E.collect();
        if (!E.any_active()) {
            bool pending = false;
            for (int j = 0; j < k; j++) pending |= E.slots[j].state == SLOT_RUN;
            if (!pending) break;
        }
        E.launch();
        E.wait();
    }
    for (int j = 0; j < k; j++) {
        const SlotInfo &S = E.slots[j];
        iters[j] = S.zero_rhs ? 0 : S.iters;
        resid[j] = S.zero_rhs ? 0.0 : S.l2;
        status[j] = S.status;
        E.retire(j, x[j]);
    }
    BCMG_HIP(hipStreamSynchronize(c.stream));
}

} // namespace

void fcg_multi(Ctx &c, Boot &B, int k, const double *const *b, double *const *x, int *iters,
               double *resid, int *status)
{
    B.restarts = 0;
    for (int j0 = 0; j0 < k; j0 += MULTI_MAX_COLS) {
        const int kk = std::min(MULTI_MAX_COLS, k - j0);
        fcg_block(c, B, kk, b + j0, x + j0, iters + j0, resid + j0, status + j0);
    }
}

} // namespace bcmg
