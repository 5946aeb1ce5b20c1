/*
 * oracle.cpp - serial CPU restatement of BootCMatchG's AMG-FCG path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). It restates, in plain serial C++,
 * what the reference computes on the GPU, with two deliberate determinisations
 * that the HIP library follows bit for bit:
 *   - aggregates are numbered in increasing order of their smallest fine index
 *     (the reference numbers them in atomicAdd arrival order,
 *      src/AMG/matchingAggregation.cu:113,128,133);
 *   - the Galerkin product sums in a fixed order: AP(i,J) over j ascending in
 *     row i, then A_c(I,J) over the members i of I ascending (the reference's
 *     hash SpGEMM adds with shared-memory atomics in arrival order,
 *     src/EXTERNAL/nsparse-master/cuda-c/src/kernel/kernel_spgemm_hash_d.cu:686-692);
 *   - the general sparse product of smoothed aggregation sums every entry
 *     C(i,j) in the order of the products: the left row's entries in storage
 *     order, each over the right row's entries in storage order.
 * The matching weight uses the portable bcmg_log (include/bcmg_log.h) instead
 * of CUDA's log so that both sides produce the same weights.
 * Build with -ffp-contract=off: the setup arithmetic must not be fused.
 */
#include "oracle.h"
#include "../include/bcmg_log.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct Csr {
    int n = 0, m = 0;
    std::vector<int> row, col;
    std::vector<double> val;
    int nnz() const { return (int)col.size(); }
};

struct Level {
    Csr A;
    std::vector<double> D, M;      /* diag(A), sum_j |a_ij| */
    std::vector<int> agg;          /* P: fine row -> coarse column */
    std::vector<double> p;         /* P values (one per fine row) */
    Csr R;                         /* R = P^T, columns ascending */
    /* smoothed aggregation: P = (I - omega D^-1 A) P_tent replaces agg/p in
       the cycle (agg/p keep the tentative prolongator) */
    bool smoothed = false;
    Csr P;
    double omega = 0.0;
};

struct Hierarchy {
    std::vector<Level> L;
    std::vector<double> w;         /* smooth vector the hierarchy was built from */
    double op_cmplx = 0, op_wcmplx = 0, avg_cratio = 0;
};

struct ApplyData {                 /* src/AMG/AMG.cu:77-88, 359-394 */
    int cycle_type = 0, relax_type = 0, relaxnumber_coarse = 1;
    int prerelax = 1, postrelax = 1;
    std::vector<int> num_grid_sweeps;
};

} // namespace

struct orc_boot_s {
    orc_params p;
    ApplyData ad;
    int coarse_solver = 4;
    std::vector<Hierarchy> H;
    double estimated_ratio = 1.0;
    /* K-cycle breakdown seen during an application (the reference exits:
       src/SOLVER/FCG.cu:625-628, 640-643) */
    mutable int kerr = 0;
};

namespace {

/* ---------------------------------------------------------------- basics */

bool valid_csr(int n, int nnz, const int *row, const int *col)
{
    if (n <= 0 || nnz < 0 || !row || (nnz > 0 && !col)) return false;
    if (row[0] != 0 || row[n] != nnz) return false;
    for (int i = 0; i < n; i++) {
        if (row[i + 1] < row[i]) return false;
        bool diag = false;
        for (int j = row[i]; j < row[i + 1]; j++) {
            if (col[j] < 0 || col[j] >= n) return false;
            if (j > row[i] && col[j] <= col[j - 1]) return false;
            if (col[j] == i) diag = true;
        }
        if (!diag) return false;
    }
    return true;
}

Csr make_csr(int n, int nnz, const int *row, const int *col, const double *val)
{
    Csr A;
    A.n = A.m = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + nnz);
    A.val.assign(val, val + nnz);
    return A;
}

/* y = A x, row sums in CSR order (src/matrix/CSR.cu:884-922 computes the same
   sum in a lane-strided order). */
void spmv(const Csr &A, const double *x, double *y)
{
    for (int i = 0; i < A.n; i++) {
        double t = 0.0;
        for (int j = A.row[i]; j < A.row[i + 1]; j++) t += A.val[j] * x[A.col[j]];
        y[i] = t;
    }
}

double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

double nrm2(int n, const double *a) { return std::sqrt(dot(n, a, a)); }

/* sqrt(x^T A x) (src/matrix/CSR.cu:360-368) */
double a_norm(const Csr &A, const std::vector<double> &x)
{
    std::vector<double> t(A.n);
    spmv(A, x.data(), t.data());
    return std::sqrt(dot(A.n, t.data(), x.data()));
}

/* src/matrix/CSR.cu:320-354 */
std::vector<double> diag(const Csr &A)
{
    std::vector<double> D(A.n, 0.0);
    for (int i = 0; i < A.n; i++)
        for (int j = A.row[i]; j < A.row[i + 1]; j++)
            if (A.col[j] == i) D[i] = A.val[j];
    return D;
}

/* src/matrix/CSR.cu:638-652: sum of |a_ij| over the row, CSR order */
std::vector<double> abs_row_sum(const Csr &A)
{
    std::vector<double> S(A.n);
    for (int i = 0; i < A.n; i++) {
        double s = 0.0;
        for (int j = A.row[i]; j < A.row[i + 1]; j++) s += std::fabs(A.val[j]);
        S[i] = s;
    }
    return S;
}

/* ------------------------------------------------------------- matching */

/* C_i = a_ii w_i^2; AH_ij = 1 - 2 a_ij w_i w_j / (C_i + C_j) (or DBL_EPSILON);
   m = min |AH|; W = log(|AH| / (0.999 m)).
   src/AMG/matching.cu:141-162 (_makeC), 200-238 (_makeAH_warp),
   401-431 (AbsMin reduction), 386-396 (_make_w). */
void weights(const Csr &A, const double *w, double *W)
{
    const int n = A.n;
    std::vector<double> C(n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = A.row[i]; j < A.row[i + 1]; j++)
            if (A.col[j] == i) { C[i] = A.val[j] * (w[i] * w[i]); break; }
    double mn = DBL_MAX;
    for (int i = 0; i < n; i++) {
        for (int j = A.row[i]; j < A.row[i + 1]; j++) {
            const int c = A.col[j];
            if (c == i) { W[j] = 0.0; continue; }
            const int lo = c < i ? c : i, hi = c < i ? i : c;
            const double norm = C[lo] + C[hi];
            double ah;
            if (norm > DBL_EPSILON) {
                const double wt = w[lo] * w[hi];
                ah = 1. - ((2. * A.val[j] * wt) / norm);
            } else {
                ah = DBL_EPSILON;
            }
            W[j] = ah;
            const double a = std::fabs(ah);
            mn = a < mn ? a : mn;
        }
    }
    const double den = 0.999 * mn;
    for (int i = 0; i < n; i++)
        for (int j = A.row[i]; j < A.row[i + 1]; j++)
            if (A.col[j] != i) W[j] = bcmg_log(std::fabs(W[j]) / den);
}

/* Suitor matching, processed serially (src/AMG/suitor.cu:118-392).
   Proposer u prefers the neighbour y with the largest (W_uy, y); vertex y
   keeps the proposal with the largest (W_uy, u). Edges with W <= 0 (and NaN)
   are never proposed. A displaced proposer searches again. The final suitor
   array is the proposer-optimal stable outcome and does not depend on the
   order in which vertices are processed. */
void suitor(const Csr &A, const double *W, const int *order, int *s)
{
    const int n = A.n;
    std::vector<double> ws(n, 0.0);
    for (int i = 0; i < n; i++) s[i] = -1;
    for (int t = 0; t < n; t++) {
        int cur = order ? order[t] : t;
        while (cur >= 0) {
            int best = -1;
            double bw = -1.0;
            for (int j = A.row[cur]; j < A.row[cur + 1]; j++) {
                const double wv = W[j];
                if (!(wv > 0.0)) continue;
                const int y = A.col[j];
                if (y == cur) continue;
                if (wv < bw || (wv == bw && y < best)) continue;
                if (wv < ws[y] || (wv == ws[y] && cur < s[y])) continue;
                bw = wv;
                best = y;
            }
            if (best < 0) break;
            const int prev = s[best];
            s[best] = cur;
            ws[best] = bw;
            cur = prev;
        }
    }
}

/* src/AMG/unsymMatching.cu: keep mutual pairs (_setSymPair), then pair a
   non-mutual v -> u with u < v if both are still free (_setUnsymPair). The
   reference resolves conflicts between such v by atomicCAS arrival order; here
   they are resolved in ascending v, which is one of the outcomes the reference
   can produce. */
int unsym_fix(int n, const int *M, int *Ms)
{
    for (int v = 0; v < n; v++) Ms[v] = -1;
    for (int v = 0; v < n; v++) {
        const int u = M[v];
        if (u >= 0 && M[u] == v) { Ms[v] = u; Ms[u] = v; }
    }
    int nonmutual = 0;
    for (int v = 0; v < n; v++) {
        const int u = M[v];
        if (u < 0 || Ms[v] != -1 || u > v) continue;
        nonmutual++;
        if (Ms[u] == -1) { Ms[u] = v; Ms[v] = u; }
    }
    return nonmutual;
}

/* src/AMG/matchingAggregation.cu:92-137 with canonical numbering. */
int aggregate(int n, const int *Ms, const double *w, int *agg, double *p)
{
    int nc = 0;
    for (int v = 0; v < n; v++) {
        const int u = Ms[v];
        if (u != -1) {
            const double wv = w[v], wu = w[u];
            const double nrm = std::sqrt(wv * wv + wu * wu);
            if (nrm > DBL_EPSILON) {
                if (v < u) {
                    agg[v] = agg[u] = nc++;
                    p[v] = wv / nrm;
                    p[u] = wu / nrm;
                }
                continue;
            }
        }
        agg[v] = nc++;
        p[v] = std::fabs(w[v]) > DBL_EPSILON ? w[v] / std::fabs(w[v]) : 0.0;
    }
    return nc;
}

/* members of each coarse row, ascending fine index (R = P^T, CSR.cu:272-285) */
Csr transpose_P(int n, int nc, const int *agg, const double *p)
{
    Csr R;
    R.n = nc;
    R.m = n;
    R.row.assign(nc + 1, 0);
    for (int i = 0; i < n; i++) R.row[agg[i] + 1]++;
    for (int I = 0; I < nc; I++) R.row[I + 1] += R.row[I];
    R.col.resize(n);
    R.val.resize(n);
    std::vector<int> fill(R.row.begin(), R.row.end() - 1);
    for (int i = 0; i < n; i++) {
        const int pos = fill[agg[i]]++;
        R.col[pos] = i;
        R.val[pos] = p[i];
    }
    return R;
}

/* A_c = R (A P), P with exactly one entry per row, fixed summation order:
   AP(i,J) = 0 + sum_{j in row i, agg(j) = J, ascending j} a_ij p_j
   A_c(I,J) = 0 + sum_{i in I ascending, J in AP(i,:)} p_i AP(i,J).
   The pattern keeps structural zeros and columns are ascending
   (src/AMG/matchingAggregation.cu:241-244 and the hash SpGEMM it calls). */
Csr galerkin(const Csr &A, const int *agg, const double *p, int nc)
{
    Csr R = transpose_P(A.n, nc, agg, p);
    Csr Ac;
    Ac.n = Ac.m = nc;
    Ac.row.assign(nc + 1, 0);
    std::map<int, double> acc, ap;
    for (int I = 0; I < nc; I++) {
        acc.clear();
        for (int k = R.row[I]; k < R.row[I + 1]; k++) {
            const int i = R.col[k];
            ap.clear();
            for (int j = A.row[i]; j < A.row[i + 1]; j++) {
                const int J = agg[A.col[j]];
                const double t = A.val[j] * p[A.col[j]];
                auto it = ap.find(J);
                if (it == ap.end()) ap.emplace(J, 0.0 + t);
                else it->second = it->second + t;
            }
            for (const auto &e : ap) {
                const double t = p[i] * e.second;
                auto it = acc.find(e.first);
                if (it == acc.end()) acc.emplace(e.first, 0.0 + t);
                else it->second = it->second + t;
            }
        }
        for (const auto &e : acc) {
            Ac.col.push_back(e.first);
            Ac.val.push_back(e.second);
        }
        Ac.row[I + 1] = (int)Ac.col.size();
    }
    return Ac;
}

/* ---------------------------------------------- smoothed aggregation */

/* C = L Rm (CSRm::CSRCSR_product, src/matrix/CSR.cu:1039-1079, a hash
   SpGEMM). Pattern: the symbolic union of the right rows named by each left
   row (no cancellation drops an entry), columns ascending. Values:
   C(i,j) = 0 + sum of L(i,e) Rm(e,j) over the entries e of L's row i in
   storage order, and within Rm's row e in storage order. */
Csr spgemm(const Csr &L, const Csr &Rm)
{
    Csr C;
    C.n = L.n;
    C.m = Rm.m;
    C.row.assign(L.n + 1, 0);
    std::vector<int> mark(Rm.m, -1), cols;
    std::vector<double> acc(Rm.m, 0.0);
    for (int i = 0; i < L.n; i++) {
        cols.clear();
        for (int e = L.row[i]; e < L.row[i + 1]; e++) {
            const int k = L.col[e];
            for (int f = Rm.row[k]; f < Rm.row[k + 1]; f++) {
                const int j = Rm.col[f];
                const double t = L.val[e] * Rm.val[f];
                if (mark[j] != i) {
                    mark[j] = i;
                    acc[j] = 0.0;
                    cols.push_back(j);
                }
                acc[j] = acc[j] + t;
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) {
            C.col.push_back(j);
            C.val.push_back(acc[j]);
        }
        C.row[i + 1] = (int)C.col.size();
    }
    return C;
}

/* A^T with the entries of each row in ascending column order
   (CSRm::T, src/matrix/CSR.cu:272-285) */
Csr transpose(const Csr &A)
{
    Csr T;
    T.n = A.m;
    T.m = A.n;
    T.row.assign(A.m + 1, 0);
    for (int e = 0; e < A.nnz(); e++) T.row[A.col[e] + 1]++;
    for (int j = 0; j < A.m; j++) T.row[j + 1] += T.row[j];
    T.col.resize(A.nnz());
    T.val.resize(A.nnz());
    std::vector<int> fill(T.row.begin(), T.row.end() - 1);
    for (int i = 0; i < A.n; i++)
        for (int e = A.row[i]; e < A.row[i + 1]; e++) {
            const int pos = fill[A.col[e]]++;
            T.col[pos] = i;
            T.val[pos] = A.val[e];
        }
    return T;
}

/* ||A||_inf = max_i (0 + sum_j |a_ij| in storage order): the body the
   reference left commented out in CSRm::infinityNorm (src/matrix/CSR.cu:
   405-416, _row_sum :386-400); the live function asserts (:402-404). */
double infinity_norm(const Csr &A)
{
    double m = 0.0;
    for (int i = 0; i < A.n; i++) {
        double s = 0.0;
        for (int e = A.row[i]; e < A.row[i + 1]; e++) s += std::fabs(A.val[e]);
        if (s > m) m = s;
    }
    return m;
}

/* P_tent as CSR: one entry per fine row (agg(i), p_i) */
Csr tentative(int n, int nc, const std::vector<int> &agg, const std::vector<double> &p)
{
    Csr P;
    P.n = n;
    P.m = nc;
    P.row.resize(n + 1);
    for (int i = 0; i <= n; i++) P.row[i] = i;
    P.col = agg;
    P.val = p;
    return P;
}

void relax_prepare(Level &lv, int relax_type);

/* Smoothed prolongators for the levels built by the unsmoothed coarsening
   (src/AMG/matchingAggregation.cu:353-390): on each level j,
     S = D^-1 A_j (each row divided by its diagonal, CSR.cu:520-545),
     omega = 4 / (3 ||S||_inf),
     S = -omega S, plus 1 on the diagonal (_applyOmega :64-80),
     P_j = S P_tent, R_j = P_j^T, A_{j+1} = (R_j A_j) P_j,
   and the smoother data of level j + 1 is prepared from the new A_{j+1}
   before level j + 1 is smoothed. The aggregates (and so every level size)
   are the unsmoothed coarsening's. */
void smooth_prolongators(const orc_params &p, Hierarchy &H)
{
    const int levels = (int)H.L.size();
    for (int j = 0; j + 1 < levels; j++) {
        Level &lv = H.L[j];
        Csr S = lv.A;
        for (int i = 0; i < S.n; i++)
            for (int e = S.row[i]; e < S.row[i + 1]; e++) S.val[e] /= lv.D[i];
        const double omega = 4.0 / (3.0 * infinity_norm(S));
        for (int i = 0; i < S.n; i++)
            for (int e = S.row[i]; e < S.row[i + 1]; e++) {
                S.val[e] = -omega * S.val[e];
                if (S.col[e] == i) S.val[e] += 1.;
            }
        const int nc = H.L[j + 1].A.n;
        lv.P = spgemm(S, tentative(S.n, nc, lv.agg, lv.p));
        lv.smoothed = true;
        lv.omega = omega;
        lv.R = transpose(lv.P);
        const Csr T = spgemm(lv.R, lv.A);
        Level &next = H.L[j + 1];
        next.A = spgemm(T, lv.P);
        next.D.clear();
        next.M.clear();
        relax_prepare(next, p.relax_type);
    }
}

/* ---------------------------------------------------------------- setup */

/* src/AMG/matchingAggregation.cu:18-60 */
void relax_prepare(Level &lv, int relax_type)
{
    if (relax_type == 0) {
        lv.D = diag(lv.A);
    } else if (relax_type == 4) {
        lv.D = diag(lv.A);
        lv.M = abs_row_sum(lv.A);
    }
}

struct BuildState {
    int maxcoarsesize = 0;
    int ftcoarse = 1;
};

/* src/AMG/matchingAggregation.cu:211-295: aggrsweeps pairwise sweeps, each
   A_{i+1} = R_i (A_i P_i), w_{i+1} = R_i w_i, P = P P_i. */
Csr matching_aggregation(const orc_params &p, BuildState &bs, const Csr &A0,
                         std::vector<double> &w, std::vector<int> &aggP,
                         std::vector<double> &pP)
{
    Csr Ai = A0;
    std::vector<double> wi = w;
    const int n0 = A0.n;
    aggP.resize(n0);
    pP.resize(n0);
    for (int i = 0; i < n0; i++) { aggP[i] = i; pP[i] = 1.0; }
    bool first = true;
    for (int s = 0; s < p.aggrsweeps; s++) {
        const int n = Ai.n;
        std::vector<double> W(Ai.nnz());
        weights(Ai, wi.data(), W.data());
        std::vector<int> M(n), Ms(n), agg(n);
        std::vector<double> pv(n);
        suitor(Ai, W.data(), nullptr, M.data());
        unsym_fix(n, M.data(), Ms.data());
        const int nc = aggregate(n, Ms.data(), wi.data(), agg.data(), pv.data());
        Csr Ac = galerkin(Ai, agg.data(), pv.data(), nc);
        /* w_c = R w (0 + sum over members ascending) */
        Csr R = transpose_P(n, nc, agg.data(), pv.data());
        std::vector<double> wc(nc);
        for (int I = 0; I < nc; I++) {
            double t = 0.0;
            for (int k = R.row[I]; k < R.row[I + 1]; k++) t = t + R.val[k] * wi[R.col[k]];
            wc[I] = t;
        }
        /* P = P P_i: one entry per row, value 0 + p_prev * p_i */
        if (first) {
            aggP = agg;
            pP = pv;
            first = false;
        } else {
            for (int i = 0; i < n0; i++) {
                const int I = aggP[i];
                pP[i] = 0.0 + pP[i] * pv[I];
                aggP[i] = agg[I];
            }
        }
        const double coarse_ratio = (double)n / (double)nc;
        Ai = std::move(Ac);
        wi = std::move(wc);
        if (coarse_ratio <= 1.2) bs.ftcoarse = 100;
        if ((double)nc <= (double)bs.ftcoarse * bs.maxcoarsesize) break;
    }
    w = std::move(wi);
    return Ai;
}

/* src/AMG/matchingAggregation.cu:298-423 */
int adaptive_coarsening(const orc_params &p, BuildState &bs, const Csr &A,
                        const std::vector<double> &w_in, Hierarchy &H)
{
    std::vector<double> w = w_in;
    H.w = w_in;
    H.L.clear();
    H.L.emplace_back();
    H.L[0].A = A;
    const double normw = a_norm(A, w);
    relax_prepare(H.L[0], p.relax_type);
    bs.ftcoarse = 1;
    double avcoarseratio = 0.0;
    int level = 0;
    if (normw > DBL_EPSILON) {
        for (level = 1; level < p.max_levels;) {
            Level next;
            Level &prev = H.L[level - 1];
            next.A = matching_aggregation(p, bs, prev.A, w, prev.agg, prev.p);
            prev.R = transpose_P(prev.A.n, next.A.n, prev.agg.data(), prev.p.data());
            relax_prepare(next, p.relax_type);
            const double size_coarse = next.A.n;
            avcoarseratio += prev.A.n / size_coarse;
            H.L.push_back(std::move(next));
            level++;
            if (size_coarse <= (double)bs.ftcoarse * bs.maxcoarsesize) break;
        }
    } else {
        /* the reference reaches finalize_level(H, 0) and asserts */
        return ORC_EINVAL;
    }
    if (p.aggrtype == 1 && level > 1) smooth_prolongators(p, H);
    if (p.coarse_solver != p.relax_type) relax_prepare(H.L[level - 1], p.coarse_solver);
    double c = 0, cw = 0;
    for (int i = 0; i < level; i++) {
        c += H.L[i].A.nnz();
        cw += std::pow(2.0, i) * H.L[i].A.nnz();
    }
    H.op_cmplx = c / H.L[0].A.nnz();
    H.op_wcmplx = cw / H.L[0].A.nnz();
    H.avg_cratio = avcoarseratio / (level - 1);
    return ORC_OK;
}

/* ---------------------------------------------------------------- apply */

/* u <- u + (f - A u) / d, k sweeps, Jacobi (src/AMG/relaxation.cu:176-209) */
void relax(const Csr &A, const std::vector<double> &d, const double *f,
           std::vector<double> &u, int k)
{
    std::vector<double> t(A.n);
    for (int s = 0; s < k; s++) {
        for (int i = 0; i < A.n; i++) {
            double T = 0.0;
            for (int j = A.row[i]; j < A.row[i + 1]; j++) T += A.val[j] * u[A.col[j]];
            t[i] = ((-T + f[i]) / d[i]) + u[i];
        }
        u.swap(t);
    }
}

const std::vector<double> &smoother_divisor(const Level &lv, int type)
{
    return type == 4 ? lv.M : lv.D;
}

/* inner-iteration tolerance of the K-cycle (src/SOLVER/GAMG_cycle.cu:5) */
constexpr double KCYCLE_RTOL = 0.25;

void cycle(const orc_boot &B, const Hierarchy &H, std::vector<std::vector<double>> &rhs,
           std::vector<std::vector<double>> &x, int l);

/* one application of the cycle from level lvl (1-based) to g with a zero
   start, in buffers of its own (preconditionApplyK, src/SOLVER/FCG.cu:590-615) */
std::vector<double> apply_from_level(const orc_boot &B, const Hierarchy &H,
                                     const std::vector<double> &g, int lvl)
{
    std::vector<std::vector<double>> rhs(H.L.size()), x(H.L.size());
    rhs[lvl - 1] = g;
    x[lvl - 1].assign(g.size(), 0.0);
    cycle(B, H, rhs, x, lvl);
    return x[lvl - 1];
}

/* K-cycle coarse correction on level m (0-based): at most two flexible CG
   steps on A_m x = f_m preconditioned by the cycle from level m + 1
   (inneritkcycle, src/SOLVER/FCG.cu:620-723); pow(t, 2) is t * t */
void kcycle_inner(const orc_boot &B, const Hierarchy &H, const std::vector<double> &f,
                  std::vector<double> &x, int m)
{
    const Csr &A = H.L[m].A;
    const int n = A.n;
    std::vector<double> w = f;
    const double delta0 = nrm2(n, w.data());
    const std::vector<double> d0 = apply_from_level(B, H, w, m + 1);
    const double delta_old = dot(n, w.data(), d0.data());
    if (!(delta_old > 0.0)) B.kerr = 1;
    std::vector<double> v(n);
    spmv(A, d0.data(), v.data());
    const double tau = dot(n, d0.data(), v.data());
    if (!(tau > 0.0)) B.kerr = 2;
    double alpha = delta_old / tau;
    for (int i = 0; i < n; i++) w[i] = -alpha * v[i] + w[i];
    const double l2 = nrm2(n, w.data());
    x.assign(n, 0.0);
    if (l2 <= KCYCLE_RTOL * delta0) {
        for (int i = 0; i < n; i++) x[i] = alpha * d0[i] + x[i];
        return;
    }
    const std::vector<double> d1 = apply_from_level(B, H, w, m + 1);
    std::vector<double> v1(n);
    spmv(A, d1.data(), v1.data());
    const double tau1 = dot(n, d1.data(), v.data());
    const double tau2 = dot(n, d1.data(), v1.data());
    const double tau3 = dot(n, d1.data(), w.data());
    const double tau4 = tau2 - (tau1 * tau1) / tau;
    alpha = alpha - (tau1 * tau3) / (tau * tau4);
    for (int i = 0; i < n; i++) x[i] = alpha * d0[i] + x[i];
    const double alpha2 = tau3 / tau4;
    for (int i = 0; i < n; i++) x[i] = alpha2 * d1[i] + x[i];
}

/* recursive V/H/W/K cycle (src/SOLVER/GAMG_cycle.cu:25-140); l is 1-based */
void cycle(const orc_boot &B, const Hierarchy &H, std::vector<std::vector<double>> &rhs,
           std::vector<std::vector<double>> &x, int l)
{
    const int L = (int)H.L.size();
    const Level &lv = H.L[l - 1];
    if (l == L) {
        relax(lv.A, smoother_divisor(lv, B.coarse_solver), rhs[l - 1].data(), x[l - 1],
              B.ad.relaxnumber_coarse);
        return;
    }
    relax(lv.A, smoother_divisor(lv, B.ad.relax_type), rhs[l - 1].data(), x[l - 1], B.ad.prerelax);
    const int n = lv.A.n;
    std::vector<double> res(n);
    spmv(lv.A, x[l - 1].data(), res.data());
    for (int i = 0; i < n; i++) res[i] = rhs[l - 1][i] - res[i];
    /* f_{l+1} = R res (0 + sum over members ascending) */
    const int nc = lv.R.n;
    rhs[l].assign(nc, 0.0);
    for (int I = 0; I < nc; I++) {
        double t = 0.0;
        for (int k = lv.R.row[I]; k < lv.R.row[I + 1]; k++) t += lv.R.val[k] * res[lv.R.col[k]];
        rhs[l][I] = t;
    }
    x[l].assign(nc, 0.0);
    if (B.ad.cycle_type == 3 && l != L - 1) {
        kcycle_inner(B, H, rhs[l], x[l], l);
    } else {
        for (int i = 1; i <= B.ad.num_grid_sweeps[l - 1]; i++) {
            cycle(B, H, rhs, x, l + 1);
            if (l == L - 1) break;
        }
    }
    if (lv.smoothed) {
        /* x_l = 1 P x_{l+1} + 1 x_l (src/SOLVER/GAMG_cycle.cu:115-117) */
        const Csr &P = lv.P;
        for (int i = 0; i < n; i++) {
            double t = 0.0;
            for (int e = P.row[i]; e < P.row[i + 1]; e++) t += P.val[e] * x[l][P.col[e]];
            x[l - 1][i] = t + x[l - 1][i];
        }
    } else {
        /* x_l += P x_{l+1} (src/matrix/CSR.cu:965-976) */
        for (int i = 0; i < n; i++) x[l - 1][i] += lv.p[i] * x[l][lv.agg[i]];
    }
    relax(lv.A, smoother_divisor(lv, B.ad.relax_type), rhs[l - 1].data(), x[l - 1], B.ad.postrelax);
}

/* composite preconditioner (src/SOLVER/FCG.cu:112-187) */
void precond_apply(const orc_boot &B, const double *r, std::vector<double> &x)
{
    const int n = (int)x.size();
    const int nh = (int)B.H.size();
    auto run = [&](int k, const std::vector<double> &xin) {
        const Hierarchy &H = B.H[k];
        std::vector<std::vector<double>> rhs(H.L.size()), xs(H.L.size());
        rhs[0].assign(r, r + n);
        xs[0] = xin;
        cycle(B, H, rhs, xs, 1);
        return xs[0];
    };
    if (B.p.solver_type == 0) {
        for (int k = 0; k < nh; k++) x = run(k, x);
    } else if (B.p.solver_type == 1) {
        for (int k = 0; k < nh; k++) x = run(k, x);
        for (int k = nh - 1; k >= 0; k--) x = run(k, x);
    } else {
        std::vector<double> xadd(n, 0.0);
        for (int k = 0; k < nh; k++) {
            std::vector<double> xk = run(k, x);
            for (int i = 0; i < n; i++) xadd[i] = 1.0 * xk[i] + xadd[i];
        }
        const double alpha = 1.0 / (double)nh;
        for (int i = 0; i < n; i++) xadd[i] *= alpha;
        x = xadd;
    }
}

/* src/AMG/AMG.cu:371-394 */
std::vector<int> grid_sweeps(int cycle_type, int max_levels)
{
    int ml = max_levels - 1;
    std::vector<int> g(ml > 0 ? ml : 1, 1);
    if (cycle_type == 1) {
        for (int i = 0; i < ml; i++)
            if (i % 2 == 0) g[i] = 2;
    } else if (cycle_type == 2) {
        for (int i = 0; i < ml - 1; i++) g[i] = 2;
    }
    return g;
}

/* src/BOOT/bootstrap.cu:5-44 */
void inner_iterations(orc_boot &B, const Csr &A, std::vector<double> &w)
{
    std::vector<double> x(A.n, 1.0), rhs(A.n, 0.0);
    double normold = a_norm(A, x), normnew = normold, ratio = 1.0;
    for (int i = 1; i <= 15; i++) {
        precond_apply(B, rhs.data(), x);
        normnew = a_norm(A, x);
        ratio = normnew / normold;
        normold = normnew;
    }
    const double alpha = 1. / normnew;
    for (int i = 0; i < A.n; i++) x[i] *= alpha;
    w = x;
    B.estimated_ratio = ratio;
}

int check_params(const orc_params &p)
{
    if (p.solver_type < 0 || p.solver_type > 2) return ORC_ESETTINGS;
    if (p.aggrtype != 0 && p.aggrtype != 1) return ORC_ESETTINGS;
    /* cycle_type 3 (K-cycle) is rejected by the reference's settings parser
       (src/AMG/AMG.cu:506-508) but coded in its cycle; a parameter block may
       ask for it */
    if (p.cycle_type < 0 || p.cycle_type > 3) return ORC_ESETTINGS;
    if (p.coarse_solver != 0 && p.coarse_solver != 4) return ORC_ESETTINGS;
    if (p.relax_type != 0 && p.relax_type != 4) return ORC_ESETTINGS;
    if (p.max_hrc < 1 || p.aggrsweeps < 1 || p.max_levels < 1) return ORC_EINVAL;
    if (p.prerelax_sweeps < 0 || p.postrelax_sweeps < 0 || p.relaxnumber_coarse < 0)
        return ORC_EINVAL;
    return ORC_OK;
}

/* src/BOOT/bootstrap.cu:46-90 with the defaults of src/AMG/AMG.cu:426-448 */
int bootstrap(const Csr &A, const orc_params &p, orc_boot *B)
{
    B->p = p;
    B->coarse_solver = p.coarse_solver;
    B->ad.cycle_type = p.cycle_type;
    B->ad.relax_type = p.relax_type;
    B->ad.relaxnumber_coarse = p.relaxnumber_coarse;
    B->ad.prerelax = p.prerelax_sweeps;
    B->ad.postrelax = p.postrelax_sweeps;
    B->ad.num_grid_sweeps = grid_sweeps(p.cycle_type, p.max_levels);
    BuildState bs;
    bs.maxcoarsesize = (int)(40 * std::pow((double)A.n, (double)1. / 3.));
    std::vector<double> w(A.n, 1.0);
    B->estimated_ratio = 1.0;
    B->H.clear();
    int num_hrc = 0;
    while (B->estimated_ratio > p.conv_ratio && num_hrc < p.max_hrc) {
        Hierarchy H;
        int rc = adaptive_coarsening(p, bs, A, w, H);
        if (rc) return rc;
        B->H.push_back(std::move(H));
        num_hrc++;
        if (num_hrc > 1) inner_iterations(*B, A, w);
    }
    return ORC_OK;
}

/* FCG v3 (src/SOLVER/FCG.cu:433-585) */
int fcg(const orc_boot &B, const double *rhs, std::vector<double> &x, orc_result *res,
        double *hist, int hist_len)
{
    const Csr &A = B.H[0].L[0].A;
    const int n = A.n;
    const int max_iter = B.p.itnlim;
    const double rtol = B.p.rtol;
    B.kerr = 0;
    std::vector<double> v(n, 0.0), w(n), r(rhs, rhs + n), d, q;
    spmv(A, x.data(), w.data());
    for (int i = 0; i < n; i++) r[i] = -1.0 * w[i] + r[i];
    const double delta0 = nrm2(n, r.data());
    const double rhs_norm = nrm2(n, rhs);
    if (res) res->delta0 = delta0;
    if (delta0 <= DBL_EPSILON * rhs_norm) {
        if (res) res->iterations = 0;
        return ORC_EZERORHS;
    }
    precond_apply(B, r.data(), v);
    spmv(A, v.data(), w.data());
    double alpha = dot(n, r.data(), v.data());
    double beta = dot(n, w.data(), v.data());
    double delta = beta;
    double theta = alpha / delta;
    double gamma;
    for (int i = 0; i < n; i++) x[i] = theta * v[i] + x[i];
    for (int i = 0; i < n; i++) r[i] = -theta * w[i] + r[i];
    double l2 = nrm2(n, r.data());
    int iter = 0;
    d = v;
    q = w;
    do {
        const int idx = iter % 2;
        std::vector<double> &z = idx == 0 ? v : d;      /* new direction */
        std::vector<double> &t = idx == 0 ? w : q;      /* A z */
        std::vector<double> &zo = idx == 0 ? d : v;     /* previous direction */
        std::vector<double> &to = idx == 0 ? q : w;     /* A zo */
        std::fill(z.begin(), z.end(), 0.0);
        precond_apply(B, r.data(), z);
        spmv(A, z.data(), t.data());
        alpha = dot(n, r.data(), z.data());
        beta = dot(n, t.data(), z.data());
        gamma = dot(n, to.data(), z.data());
        theta = gamma / delta;
        delta = beta - (gamma * gamma) / delta;
        const double theta2 = alpha / delta;
        for (int i = 0; i < n; i++) {
            const double zi = -theta * zo[i] + z[i];
            x[i] = theta2 * zi + x[i];
            z[i] = zi;
        }
        for (int i = 0; i < n; i++) {
            const double ti = -theta * to[i] + t[i];
            r[i] = -theta2 * ti + r[i];
            t[i] = ti;
        }
        l2 = nrm2(n, r.data());
        if (hist && iter < hist_len) hist[iter] = l2;
        iter++;
    } while (l2 > rtol * delta0 && iter < max_iter);
    if (res) {
        res->iterations = iter + 1;
        res->residual = l2;
    }
    if (!std::isfinite(l2) || B.kerr) return ORC_EBREAKDOWN;
    return ORC_OK;
}

} // namespace

/* ================================================================ C ABI */

extern "C" {

int orc_params_from_file(const char *path, orc_params *out)
{
    if (!path || !out) return ORC_EINVAL;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return ORC_ESETTINGS;
    char buf[1025];
    bool ok = true;
    auto next_token = [&](void) -> const char * {
        if (!std::fgets(buf, 1024, fp)) { ok = false; return nullptr; }
        char *save = nullptr;
        const char *tok = strtok_r(buf, "%", &save);
        if (!tok) ok = false;
        return tok;
    };
    auto get_int = [&](void) -> int {
        int v = 0;
        const char *t = next_token();
        if (!t || std::sscanf(t, "%d", &v) != 1) ok = false;
        return v;
    };
    auto get_dbl = [&](void) -> double {
        double v = 0;
        const char *t = next_token();
        if (!t || std::sscanf(t, "%lf", &v) != 1) ok = false;
        return v;
    };
    orc_params p;
    std::memset(&p, 0, sizeof(p));
    p.solver_type = get_int();
    p.max_hrc = get_int();
    p.conv_ratio = get_dbl();
    p.matchtype = 3;
    p.aggrsweeps = get_int();
    p.aggrtype = get_int();
    p.max_levels = get_int();
    p.cycle_type = get_int();
    p.coarse_solver = get_int();
    p.relax_type = get_int();
    p.relaxnumber_coarse = get_int();
    p.prerelax_sweeps = get_int();
    p.postrelax_sweeps = get_int();
    p.itnlim = get_int();
    p.rtol = get_dbl();
    std::fclose(fp);
    if (!ok) return ORC_ESETTINGS;
    if (p.solver_type < 0 || p.solver_type > 2 || (p.aggrtype != 0 && p.aggrtype != 1) ||
        p.cycle_type < 0 || p.cycle_type > 2 || (p.coarse_solver != 0 && p.coarse_solver != 4) ||
        (p.relax_type != 0 && p.relax_type != 4))
        return ORC_ESETTINGS;
    *out = p;
    return ORC_OK;
}

double orc_log(double x) { return bcmg_log(x); }

int orc_weights(int n, const int *row, const int *col, const double *val, const double *w,
                double *W)
{
    if (!valid_csr(n, row ? row[n] : -1, row, col) || !val || !w || !W) return ORC_EINVAL;
    Csr A = make_csr(n, row[n], row, col, val);
    weights(A, w, W);
    return ORC_OK;
}

int orc_matching(int n, const int *row, const int *col, const double *W, const int *order,
                 int *M_raw, int *M_sym, int *n_unsym)
{
    if (!valid_csr(n, row ? row[n] : -1, row, col) || !W || !M_raw || !M_sym) return ORC_EINVAL;
    Csr A;
    A.n = A.m = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + row[n]);
    suitor(A, W, order, M_raw);
    const int nu = unsym_fix(n, M_raw, M_sym);
    if (n_unsym) *n_unsym = nu;
    return ORC_OK;
}

int orc_aggregate(int n, const int *M_sym, const double *w, int *agg, double *p, int *n_coarse)
{
    if (n <= 0 || !M_sym || !w || !agg || !p || !n_coarse) return ORC_EINVAL;
    *n_coarse = aggregate(n, M_sym, w, agg, p);
    return ORC_OK;
}

int orc_galerkin(int n, const int *row, const int *col, const double *val, const int *agg,
                 const double *p, int n_c, int *row_c, int *col_c, double *val_c, int *nnz_c)
{
    if (!valid_csr(n, row ? row[n] : -1, row, col) || !val || !agg || !p || n_c <= 0 || !row_c ||
        !nnz_c)
        return ORC_EINVAL;
    Csr A = make_csr(n, row[n], row, col, val);
    Csr Ac = galerkin(A, agg, p, n_c);
    std::copy(Ac.row.begin(), Ac.row.end(), row_c);
    *nnz_c = Ac.nnz();
    if (col_c) std::copy(Ac.col.begin(), Ac.col.end(), col_c);
    if (val_c) std::copy(Ac.val.begin(), Ac.val.end(), val_c);
    return ORC_OK;
}

int orc_spmv(int n, const int *row, const int *col, const double *val, const double *x, double *y)
{
    if (n <= 0 || !row || !col || !val || !x || !y) return ORC_EINVAL;
    Csr A = make_csr(n, row[n], row, col, val);
    spmv(A, x, y);
    return ORC_OK;
}

int orc_relax(int n, const int *row, const int *col, const double *val, const double *d,
              const double *f, double *u, int sweeps)
{
    if (n <= 0 || !row || !col || !val || !d || !f || !u || sweeps < 0) return ORC_EINVAL;
    Csr A = make_csr(n, row[n], row, col, val);
    std::vector<double> dd(d, d + n), uu(u, u + n);
    relax(A, dd, f, uu, sweeps);
    std::copy(uu.begin(), uu.end(), u);
    return ORC_OK;
}

int orc_setup(int n, int nnz, const int *row, const int *col, const double *val,
              const orc_params *p, orc_boot **out)
{
    if (!p || !out || !val || !valid_csr(n, nnz, row, col)) return ORC_EINVAL;
    int rc = check_params(*p);
    if (rc) return rc;
    Csr A = make_csr(n, nnz, row, col, val);
    orc_boot *B = new orc_boot();
    rc = bootstrap(A, *p, B);
    if (rc) {
        delete B;
        return rc;
    }
    *out = B;
    return ORC_OK;
}

void orc_boot_free(orc_boot *b) { delete b; }

int orc_boot_info(const orc_boot *b, int *n_hrc, double *estimated_ratio)
{
    if (!b) return ORC_EINVAL;
    if (n_hrc) *n_hrc = (int)b->H.size();
    if (estimated_ratio) *estimated_ratio = b->estimated_ratio;
    return ORC_OK;
}

int orc_hier_levels(const orc_boot *b, int k)
{
    if (!b || k < 0 || k >= (int)b->H.size()) return ORC_EINVAL;
    return (int)b->H[k].L.size();
}

static const Level *get_level(const orc_boot *b, int k, int l)
{
    if (!b || k < 0 || k >= (int)b->H.size()) return nullptr;
    if (l < 0 || l >= (int)b->H[k].L.size()) return nullptr;
    return &b->H[k].L[l];
}

int orc_level_size(const orc_boot *b, int k, int l, int *n, int *nnz)
{
    const Level *lv = get_level(b, k, l);
    if (!lv) return ORC_EINVAL;
    if (n) *n = lv->A.n;
    if (nnz) *nnz = lv->A.nnz();
    return ORC_OK;
}

int orc_level_A(const orc_boot *b, int k, int l, int *row, int *col, double *val)
{
    const Level *lv = get_level(b, k, l);
    if (!lv) return ORC_EINVAL;
    if (row) std::copy(lv->A.row.begin(), lv->A.row.end(), row);
    if (col) std::copy(lv->A.col.begin(), lv->A.col.end(), col);
    if (val) std::copy(lv->A.val.begin(), lv->A.val.end(), val);
    return ORC_OK;
}

int orc_level_P(const orc_boot *b, int k, int l, int *agg, double *p)
{
    const Level *lv = get_level(b, k, l);
    if (!lv || lv->agg.empty()) return ORC_EINVAL;
    if (agg) std::copy(lv->agg.begin(), lv->agg.end(), agg);
    if (p) std::copy(lv->p.begin(), lv->p.end(), p);
    return ORC_OK;
}

int orc_level_Pinfo(const orc_boot *b, int k, int l, int *nnz, int *smoothed, double *omega)
{
    const Level *lv = get_level(b, k, l);
    if (!lv || lv->agg.empty()) return ORC_EINVAL;
    if (nnz) *nnz = lv->smoothed ? lv->P.nnz() : (int)lv->agg.size();
    if (smoothed) *smoothed = lv->smoothed ? 1 : 0;
    if (omega) *omega = lv->omega;
    return ORC_OK;
}

int orc_level_Pcsr(const orc_boot *b, int k, int l, int *row, int *col, double *val)
{
    const Level *lv = get_level(b, k, l);
    if (!lv || lv->agg.empty()) return ORC_EINVAL;
    const Csr P = lv->smoothed ? lv->P
                               : tentative((int)lv->agg.size(), lv->R.n, lv->agg, lv->p);
    if (row) std::copy(P.row.begin(), P.row.end(), row);
    if (col) std::copy(P.col.begin(), P.col.end(), col);
    if (val) std::copy(P.val.begin(), P.val.end(), val);
    return ORC_OK;
}

int orc_spgemm(int nl, int ml, const int *lrow, const int *lcol, const double *lval, int nr,
               int mr, const int *rrow, const int *rcol, const double *rval, int *crow,
               int *ccol, double *cval, int *nnz_c)
{
    if (nl < 0 || ml != nr || mr < 0 || !lrow || !rrow || !crow || !nnz_c) return ORC_EINVAL;
    Csr L, R;
    L.n = nl;
    L.m = ml;
    L.row.assign(lrow, lrow + nl + 1);
    L.col.assign(lcol, lcol + lrow[nl]);
    L.val.assign(lval, lval + lrow[nl]);
    R.n = nr;
    R.m = mr;
    R.row.assign(rrow, rrow + nr + 1);
    R.col.assign(rcol, rcol + rrow[nr]);
    R.val.assign(rval, rval + rrow[nr]);
    for (int c : L.col)
        if (c < 0 || c >= nr) return ORC_EINVAL;
    for (int c : R.col)
        if (c < 0 || c >= mr) return ORC_EINVAL;
    const Csr C = spgemm(L, R);
    *nnz_c = C.nnz();
    std::copy(C.row.begin(), C.row.end(), crow);
    if (ccol) std::copy(C.col.begin(), C.col.end(), ccol);
    if (cval) std::copy(C.val.begin(), C.val.end(), cval);
    return ORC_OK;
}

int orc_level_DM(const orc_boot *b, int k, int l, double *D, double *M, int *hasD, int *hasM)
{
    const Level *lv = get_level(b, k, l);
    if (!lv) return ORC_EINVAL;
    if (hasD) *hasD = !lv->D.empty();
    if (hasM) *hasM = !lv->M.empty();
    if (D && !lv->D.empty()) std::copy(lv->D.begin(), lv->D.end(), D);
    if (M && !lv->M.empty()) std::copy(lv->M.begin(), lv->M.end(), M);
    return ORC_OK;
}

int orc_boot_w(const orc_boot *b, int k, double *w)
{
    if (!b || k < 0 || k >= (int)b->H.size() || !w) return ORC_EINVAL;
    std::copy(b->H[k].w.begin(), b->H[k].w.end(), w);
    return ORC_OK;
}

int orc_precond_apply(const orc_boot *b, const double *rhs, double *x)
{
    if (!b || !rhs || !x) return ORC_EINVAL;
    const int n = b->H[0].L[0].A.n;
    std::vector<double> xx(x, x + n);
    precond_apply(*b, rhs, xx);
    std::copy(xx.begin(), xx.end(), x);
    return ORC_OK;
}

int orc_fcg(const orc_boot *b, const double *rhs, double *x, orc_result *res, double *hist,
            int hist_len)
{
    if (!b || !rhs || !x) return ORC_EINVAL;
    const int n = b->H[0].L[0].A.n;
    std::vector<double> xx(x, x + n);
    int rc = fcg(*b, rhs, xx, res, hist, hist_len);
    std::copy(xx.begin(), xx.end(), x);
    if (res) {
        res->n_hrc = (int)b->H.size();
        res->estimated_ratio = b->estimated_ratio;
    }
    return rc;
}

int orc_solve(int n, int nnz, const int *row, const int *col, const double *val,
              const double *rhs, const orc_params *p, double *x, orc_result *res)
{
    if (!rhs || !x) return ORC_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    orc_boot *B = nullptr;
    int rc = orc_setup(n, nnz, row, col, val, p, &B);
    if (rc) return rc;
    auto t1 = std::chrono::steady_clock::now();
    std::vector<double> xx(n, 0.0);
    orc_result r;
    std::memset(&r, 0, sizeof(r));
    rc = fcg(*B, rhs, xx, &r, nullptr, 0);
    auto t2 = std::chrono::steady_clock::now();
    std::copy(xx.begin(), xx.end(), x);
    r.n_hrc = (int)B->H.size();
    r.estimated_ratio = B->estimated_ratio;
    r.setup_seconds = std::chrono::duration<double>(t1 - t0).count();
    r.solve_seconds = std::chrono::duration<double>(t2 - t1).count();
    if (res) *res = r;
    delete B;
    return rc;
}

} // extern "C"
