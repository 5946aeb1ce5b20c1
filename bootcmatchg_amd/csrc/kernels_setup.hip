/*
 * kernels_setup.hip - AMG setup kernels for gfx950: matching weights, suitor
 * matching, pairwise aggregation, Galerkin product, transposes and scans.
 *
 * The setup is integer work plus a little FP64 arithmetic whose results must
 * be reproducible bit for bit against the serial restatement of the method:
 *   - this file is compiled with -ffp-contract=off, so a*b + c is always a
 *     rounded product followed by a rounded sum, as in the restatement;
 *   - the matching weight uses bcmg_log (include/bcmg_log.h), the same
 *     function the restatement calls;
 *   - aggregates are numbered by a prefix sum over "leader" flags (the
 *     smallest fine index of each aggregate), not by atomic arrival order;
 *   - the Galerkin product sums every coarse entry in one fixed order (fine
 *     row of the aggregate ascending, then column position ascending), not
 *     with floating-point atomics.
 * The suitor matching is lock-free: a vertex's current suitor and the weight
 * it offered live in one 64-bit word (edge index << 32 | suitor id) updated
 * with compare-and-swap, so no lane ever waits on another lane of its wave.
 */
#include "bcmg_internal.h"
#include "../../include/bcmg_log.h"

#include <cfloat>

namespace bcmg {

namespace {

constexpr int BS = 256;
constexpr int SCAN_CHUNK = BS * 4;

inline int blocks_for(long long work) { return (int)((work + BS - 1) / BS); }

/* ------------------------------------------------------------ scan */

__global__ __launch_bounds__(BS) void k_scan_blocks(int n, const int *__restrict__ in,
                                                    int *__restrict__ bsum)
{
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * 4;
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (base + k < n) s += in[base + k];
    __shared__ int sh[BS];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

/* one block: exclusive scan of nb block sums, total -> boff[nb] */
__global__ __launch_bounds__(BS) void k_scan_top(int nb, int *__restrict__ boff)
{
    __shared__ int sh[BS];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += BS) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? boff[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < BS; o <<= 1) {
            const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) boff[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == BS - 1) carry += sh[BS - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) boff[nb] = carry;
}

__global__ __launch_bounds__(BS) void k_scan_apply(int n, const int *__restrict__ in,
                                                   const int *__restrict__ boff,
                                                   int *__restrict__ out)
{
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    __shared__ int sh[BS];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < BS; o <<= 1) {
        const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = boff[blockIdx.x] + sh[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = boff[gridDim.x];
}

/* --------------------------------------------------- diagonal, row sums */

__global__ __launch_bounds__(BS) void k_diag(int n, const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const double *__restrict__ av,
                                             double *__restrict__ D)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++)
        if (ci[j] == i) d = av[j];
    D[i] = d;
}

/* sum_j |a_ij| in CSR order (src/matrix/CSR.cu:638-652) */
__global__ __launch_bounds__(BS) void k_abs_rowsum(int n, const int *__restrict__ rp,
                                                   const double *__restrict__ av,
                                                   double *__restrict__ M)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) s += fabs(av[j]);
    M[i] = s;
}

/* ------------------------------------------------------ matching weights */

/* C_i = a_ii w_i^2 (src/AMG/matching.cu:141-162) */
__global__ __launch_bounds__(BS) void k_weights_C(int n, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  const double *__restrict__ av,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ C)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double c = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++)
        if (ci[j] == i) {
            c = av[j] * (w[i] * w[i]);
            break;
        }
    C[i] = c;
}

/* The weight kernels run over entries, not rows: block b owns rows
   [256 b, 256 b + 256) and its threads walk the block's entries with
   consecutive lanes on consecutive entries (coalesced); an entry's row is
   found by binary search in the block's row pointers, kept in LDS. */
__device__ __forceinline__ int row_in_block(const int *__restrict__ srp, int nr, int e)
{
    int lo = 0, hi = nr - 1; /* largest q with srp[q] <= e (empty rows skipped) */
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (srp[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int load_block_rows(const int *__restrict__ rp, int n, int *srp,
                                               int &r0)
{
    r0 = blockIdx.x * BS;
    const int nr = min(BS, n - r0);
    for (int q = threadIdx.x; q <= nr; q += BS) srp[q] = rp[r0 + q];
    __syncthreads();
    return nr;
}

/* AH_ij = 1 - 2 a_ij w_i w_j / (C_i + C_j) or DBL_EPSILON
   (src/AMG/matching.cu:200-238) into W (diagonal slots 0); per block the
   minimum of |AH| as a bit pattern (non-negative doubles order like their
   unsigned bit patterns) into bmin[b] */
__global__ __launch_bounds__(BS) void k_weights_AH(int n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const double *__restrict__ av,
                                                   const double *__restrict__ w,
                                                   const double *__restrict__ C,
                                                   double *__restrict__ W,
                                                   unsigned long long *__restrict__ bmin)
{
    __shared__ int srp[BS + 1];
    __shared__ unsigned long long sm[BS / 64];
    int r0;
    const int nr = load_block_rows(rp, n, srp, r0);
    unsigned long long mb = 0x7FEFFFFFFFFFFFFFull; /* DBL_MAX */
    for (int e = srp[0] + threadIdx.x; e < srp[nr]; e += BS) {
        const int i = r0 + row_in_block(srp, nr, e);
        const int c = ci[e];
        if (c == i) {
            W[e] = 0.0;
            continue;
        }
        const int lo = c < i ? c : i, hi = c < i ? i : c;
        const double norm = C[lo] + C[hi];
        double ah;
        if (norm > DBL_EPSILON) {
            const double wt = w[lo] * w[hi];
            ah = 1. - ((2. * av[e] * wt) / norm);
        } else {
            ah = DBL_EPSILON;
        }
        W[e] = ah;
        const unsigned long long bts = __double_as_longlong(fabs(ah));
        mb = bts < mb ? bts : mb;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mb, o, 64);
        mb = t < mb ? t : mb;
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < BS / 64; q++) mb = sm[q] < mb ? sm[q] : mb;
        bmin[blockIdx.x] = mb < sm[0] ? mb : sm[0];
    }
}

/* minimum of the per-block minima (one block; order-independent) */
__global__ __launch_bounds__(BS) void k_min_final(int nb, const unsigned long long *__restrict__ bmin,
                                                  unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long sm[BS / 64];
    unsigned long long mb = 0x7FEFFFFFFFFFFFFFull;
    for (int q = threadIdx.x; q < nb; q += BS) mb = bmin[q] < mb ? bmin[q] : mb;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mb, o, 64);
        mb = t < mb ? t : mb;
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < BS / 64; q++) mb = sm[q] < mb ? sm[q] : mb;
        out[0] = mb;
    }
}

/* W = log(|AH| / (0.999 m)) (src/AMG/matching.cu:386-396); diagonal slots 0 */
__global__ __launch_bounds__(BS) void k_weights_log(int n, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const unsigned long long *__restrict__ minbits,
                                                    double *__restrict__ W)
{
    __shared__ int srp[BS + 1];
    int r0;
    const int nr = load_block_rows(rp, n, srp, r0);
    const double m = __longlong_as_double(*minbits);
    const double den = 0.999 * m;
    for (int e = srp[0] + threadIdx.x; e < srp[nr]; e += BS) {
        const int i = r0 + row_in_block(srp, nr, e);
        if (ci[e] != i) W[e] = bcmg_log(fabs(W[e]) / den);
    }
}

/* ------------------------------------------------------------- suitor */

__device__ __forceinline__ double key_w(unsigned long long sv, const double *__restrict__ W)
{
    const int j = (int)(unsigned)(sv >> 32);
    return j < 0 ? 0.0 : W[j];
}

__device__ __forceinline__ int key_u(unsigned long long sv) { return (int)(unsigned)(sv & 0xFFFFFFFFull); }

/* One thread per initial proposer. A proposer picks the neighbour y with the
   largest (W_uy, y) whose current offer it beats (larger W, or equal W and a
   larger proposer id: src/AMG/suitor.cu:118-214), then installs itself with
   CAS; a displaced suitor is searched for by the same thread
   (suitor.cu:49-114). Every successful CAS strictly raises a vertex's offer,
   so the loop ends; the final state is the proposer-optimal stable outcome,
   which does not depend on the interleaving. A step costs dependent memory
   round trips, and the displacement chains of tied weights are followed
   step by step, so each step saves what it can: the offer read during the
   scan is the CAS's expected value (a stale one only fails the CAS, which
   returns the current offer), and the row bounds of the proposer that the
   CAS would displace are loaded while the CAS is in flight. */
__global__ __launch_bounds__(BS) void k_suitor(int n, const int *__restrict__ rp,
                                               const int *__restrict__ ci,
                                               const double *__restrict__ W,
                                               unsigned long long *S)
{
    constexpr int SB = 8; /* neighbours examined per batch of loads */
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    /* blocks start in index order, so thread t takes vertex n-1-t: the high
       ids (which win every tie) propose first and fewer offers get undone */
    int cur = n - 1 - t;
    int j0 = rp[cur], j1 = rp[cur + 1];
    while (cur >= 0) {
        int best = -1, bj = -1;
        double bw = -1.0;
        unsigned long long bsv = ~0ull;
        for (int jb = j0; jb < j1; jb += SB) {
            /* issue all loads of the batch before any use: one round trip for
               W and the columns, one for the neighbours' offers, one for the
               weights of those offers */
            double wv[SB];
            int y[SB];
            unsigned long long sv[SB];
            double wy[SB];
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const int j = jb + k;
                wv[k] = j < j1 ? W[j] : 0.0;
                y[k] = j < j1 ? ci[j] : cur;
            }
#pragma unroll
            for (int k = 0; k < SB; k++)
                sv[k] = (wv[k] > 0.0 && y[k] != cur)
                            ? __hip_atomic_load(&S[y[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : ~0ull;
#pragma unroll
            for (int k = 0; k < SB; k++) wy[k] = key_w(sv[k], W);
#pragma unroll
            for (int k = 0; k < SB; k++) {
                const double w = wv[k];
                if (!(w > 0.0) || y[k] == cur) continue;
                if (w < bw || (w == bw && y[k] < best)) continue;
                if (w < wy[k] || (w == wy[k] && cur < key_u(sv[k]))) continue;
                bw = w;
                best = y[k];
                bj = jb + k;
                bsv = sv[k];
            }
        }
        if (best < 0) break;
        const unsigned long long mine =
            ((unsigned long long)(unsigned)bj << 32) | (unsigned long long)(unsigned)cur;
        unsigned long long old = bsv;
        int next = -2, nj0 = 0, nj1 = 0;
        while (true) {
            const double wo = key_w(old, W);
            const int uo = key_u(old); /* -1: no offer yet */
            if (bw > wo || (bw == wo && cur > uo)) {
                if (uo >= 0) {
                    nj0 = rp[uo];
                    nj1 = rp[uo + 1];
                }
                const unsigned long long prev = atomicCAS(&S[best], old, mine);
                if (prev == old) {
                    next = uo;
                    break;
                }
                old = prev;
            } else {
                break;
            }
        }
        if (next == -2) continue; /* outbid meanwhile: search again */
        cur = next;               /* displaced suitor, or -1 */
        j0 = nj0;
        j1 = nj1;
    }
}

__global__ __launch_bounds__(BS) void k_fill_u64(int n, unsigned long long v,
                                                 unsigned long long *__restrict__ S)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) S[i] = v;
}

__global__ __launch_bounds__(BS) void k_suitor_extract(int n,
                                                       const unsigned long long *__restrict__ S,
                                                       int *__restrict__ M)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) M[i] = key_u(S[i]);
}

/* mutual pairs (src/AMG/unsymMatching.cu:32-50) + non-mutual candidates
   v -> u with u < v (unsymMatching.cu:3-30) */
__global__ __launch_bounds__(BS) void k_sym_pairs(int n, const int *__restrict__ M,
                                                  int *__restrict__ Ms, int *__restrict__ flag)
{
    const int v = blockIdx.x * BS + threadIdx.x;
    if (v >= n) return;
    const int u = M[v];
    const bool mutual = u >= 0 && M[u] == v;
    Ms[v] = mutual ? u : -1;
    flag[v] = (u >= 0 && !mutual && u < v) ? 1 : 0;
}

__global__ __launch_bounds__(BS) void k_compact(int n, const int *__restrict__ flag,
                                                const int *__restrict__ pos,
                                                int *__restrict__ list)
{
    const int v = blockIdx.x * BS + threadIdx.x;
    if (v < n && flag[v]) list[pos[v]] = v;
}

/* the non-mutual candidates in ascending v, one at a time; the count is
   read on the device (pos[n] of the scan), so no host round trip */
__global__ void k_unsym_serial(const int *__restrict__ cntp, const int *__restrict__ list,
                               const int *__restrict__ M, int *__restrict__ Ms)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int cnt = *cntp;
    for (int t = 0; t < cnt; t++) {
        const int v = list[t];
        const int u = M[v];
        if (Ms[u] == -1 && Ms[v] == -1) {
            Ms[u] = v;
            Ms[v] = u;
        }
    }
}

/* ---------------------------------------------------------- aggregation */

__device__ __forceinline__ bool good_pair(int v, int u, const double *__restrict__ w, double *nrm)
{
    if (u < 0) return false;
    const double wv = w[v], wu = w[u];
    *nrm = sqrt(wv * wv + wu * wu);
    return *nrm > DBL_EPSILON;
}

/* err (optional, for matchings from outside the library): raised when v's
   partner is out of range, v itself, or not matched back to v; such a v
   leads nothing and w[u] is not read */
__global__ __launch_bounds__(BS) void k_agg_leader(int n, const int *__restrict__ Ms,
                                                   const double *__restrict__ w,
                                                   int *__restrict__ flag,
                                                   int *__restrict__ err)
{
    const int v = blockIdx.x * BS + threadIdx.x;
    if (v >= n) return;
    const int u = Ms[v];
    if (err && (u >= n || u == v || (u >= 0 && Ms[u] != v))) {
        atomicOr(err, 1);
        flag[v] = 0;
        return;
    }
    double nrm;
    const bool good = good_pair(v, u, w, &nrm);
    flag[v] = (!good || v < u) ? 1 : 0;
}

/* src/AMG/matchingAggregation.cu:92-137 with ids from the leader prefix sum */
__global__ __launch_bounds__(BS) void k_agg_assign(int n, const int *__restrict__ Ms,
                                                   const double *__restrict__ w,
                                                   const int *__restrict__ flag,
                                                   const int *__restrict__ ids,
                                                   int *__restrict__ agg, double *__restrict__ p,
                                                   int *__restrict__ members)
{
    const int v = blockIdx.x * BS + threadIdx.x;
    if (v >= n || !flag[v]) return;
    const int id = ids[v];
    const int u = Ms[v];
    double nrm;
    if (good_pair(v, u, w, &nrm)) {
        agg[v] = id;
        agg[u] = id;
        p[v] = w[v] / nrm;
        p[u] = w[u] / nrm;
        members[2 * id] = v;
        members[2 * id + 1] = u;
    } else {
        const double wv = w[v];
        agg[v] = id;
        p[v] = fabs(wv) > DBL_EPSILON ? wv / fabs(wv) : 0.0;
        members[2 * id] = v;
        members[2 * id + 1] = -1;
    }
}

/* members of each aggregate from an arbitrary pairwise agg (test entry) */
__global__ __launch_bounds__(BS) void k_members_place(int n, const int *__restrict__ agg, int nc,
                                                      int *__restrict__ cnt,
                                                      int *__restrict__ members,
                                                      int *__restrict__ err)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    const int I = agg[i];
    if (I < 0 || I >= nc) {
        atomicOr(err, 1);
        return;
    }
    const int slot = atomicAdd(&cnt[I], 1);
    if (slot < 2)
        members[2 * I + slot] = i;
    else
        atomicOr(err, 2);
}

__global__ __launch_bounds__(BS) void k_members_fix(int nc, const int *__restrict__ cnt,
                                                    int *__restrict__ members,
                                                    int *__restrict__ err)
{
    const int I = blockIdx.x * BS + threadIdx.x;
    if (I >= nc) return;
    const int c = cnt[I];
    if (c == 0) {
        atomicOr(err, 4);
        members[2 * I] = members[2 * I + 1] = -1;
    } else if (c == 1) {
        members[2 * I + 1] = -1;
    } else if (c == 2) {
        const int a = members[2 * I], b = members[2 * I + 1];
        if (a > b) {
            members[2 * I] = b;
            members[2 * I + 1] = a;
        }
    }
}

/* ---------------------------------------------------- Galerkin product */

__global__ __launch_bounds__(BS) void k_stage_len(int nc, const int *__restrict__ rp,
                                                  const int *__restrict__ members,
                                                  int *__restrict__ len)
{
    const int I = blockIdx.x * BS + threadIdx.x;
    if (I >= nc) return;
    const int m0 = members[2 * I], m1 = members[2 * I + 1];
    int l = rp[m0 + 1] - rp[m0];
    if (m1 >= 0) l += rp[m1 + 1] - rp[m1];
    len[I] = l;
}

/* expand row i of A P into the segment of its aggregate: coarse column
   agg(j) and product a_ij p_j, in column order; member 0 first */
template <int G>
__global__ __launch_bounds__(BS) void k_stage(int n, const int *__restrict__ rp,
                                              const int *__restrict__ ci,
                                              const double *__restrict__ av,
                                              const int *__restrict__ agg,
                                              const double *__restrict__ p,
                                              const int *__restrict__ members,
                                              const int *__restrict__ stoff,
                                              int *__restrict__ Jst, double *__restrict__ tst)
{
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (i >= n) return;
    const int I = agg[i];
    const int m0 = members[2 * I];
    const int base = stoff[I] + (i == m0 ? 0 : rp[m0 + 1] - rp[m0]);
    const int j0 = rp[i], j1 = rp[i + 1];
    for (int j = j0 + lane; j < j1; j += G) {
        const int c = ci[j];
        Jst[base + (j - j0)] = agg[c];
        tst[base + (j - j0)] = av[j] * p[c];
    }
}

/* first-occurrence flags of the coarse columns of each segment, and the
   number of distinct columns (the coarse row length). A segment of at most
   G entries is read once, one entry per lane, and compared through
   shuffles; a longer one by the group's loop over the staged arrays. */
template <int G>
__global__ __launch_bounds__(BS) void k_rap_count(int nc, const int *__restrict__ stoff,
                                                  const int *__restrict__ Jst,
                                                  unsigned char *__restrict__ first,
                                                  int *__restrict__ cnt)
{
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int I = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (I >= nc) return; /* the whole group: its lanes share I */
    const int s0 = stoff[I], s1 = stoff[I + 1];
    int c = 0;
    if (s1 - s0 <= G) {
        const bool valid = lane < s1 - s0;
        const int J = valid ? Jst[s0 + lane] : -1;
        bool f = valid;
        for (int k = 0; k < G; k++) {
            const int Jk = __shfl(J, k, G);
            if (k < lane && Jk == J) f = false;
        }
        if (valid) first[s0 + lane] = f ? 1 : 0;
        c = f ? 1 : 0;
    } else {
        for (int e = s0 + lane; e < s1; e += G) {
            const int J = Jst[e];
            bool f = true;
            for (int e2 = s0; e2 < e; e2++)
                if (Jst[e2] == J) {
                    f = false;
                    break;
                }
            first[e] = f ? 1 : 0;
            c += f ? 1 : 0;
        }
    }
#pragma unroll
    for (int o = G >> 1; o > 0; o >>= 1) c += __shfl_xor(c, o, G);
    if (lane == 0) cnt[I] = c;
}

/* coarse entries in ascending column order, values summed in the fixed order:
   for each distinct column J of the segment, its rank among the distinct
   columns and the two members' sums of their staged products, each in
   staging order. A segment of at most G entries is read once and scanned
   through shuffles (the same additions in the same order); a longer one by
   the group's loop over the staged arrays. */
template <int G>
__global__ __launch_bounds__(BS) void k_rap_fill(int nc, const int *__restrict__ rp,
                                                 const int *__restrict__ members,
                                                 const double *__restrict__ p,
                                                 const int *__restrict__ stoff,
                                                 const int *__restrict__ Jst,
                                                 const double *__restrict__ tst,
                                                 const unsigned char *__restrict__ first,
                                                 const int *__restrict__ rpc,
                                                 int *__restrict__ colc, double *__restrict__ valc)
{
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int I = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (I >= nc) return; /* the whole group: its lanes share I */
    const int s0 = stoff[I], s1 = stoff[I + 1];
    const int m0 = members[2 * I], m1 = members[2 * I + 1];
    const int len0 = rp[m0 + 1] - rp[m0]; /* member 0's entries come first */
    const double p0 = p[m0];
    const double p1 = m1 >= 0 ? p[m1] : 0.0;
    const int out0 = rpc[I];
    if (s1 - s0 <= G) {
        const bool valid = lane < s1 - s0;
        const int J = valid ? Jst[s0 + lane] : -1;
        const double t = valid ? tst[s0 + lane] : 0.0;
        const int f = valid ? first[s0 + lane] : 0;
        int rank = 0;
        double sa = 0.0, sb = 0.0;
        bool ha = false, hb = false;
        for (int k = 0; k < G; k++) {
            const int Jk = __shfl(J, k, G);
            const double tk = __shfl(t, k, G);
            const int fk = __shfl(f, k, G);
            rank += (fk && Jk < J) ? 1 : 0;
            if (Jk == J) {
                if (k < len0) {
                    sa = sa + tk;
                    ha = true;
                } else {
                    sb = sb + tk;
                    hb = true;
                }
            }
        }
        if (valid && f) {
            double acc = 0.0;
            if (ha) acc = acc + p0 * sa;
            if (hb) acc = acc + p1 * sb;
            colc[out0 + rank] = J;
            valc[out0 + rank] = acc;
        }
        return;
    }
    const int split = s0 + len0;
    for (int e = s0 + lane; e < s1; e += G) {
        if (!first[e]) continue;
        const int J = Jst[e];
        /* one pass over the staged entries: the rank of J among the row's
           distinct columns, and the two members' sums in staging order */
        int rank = 0;
        double sa = 0.0, sb = 0.0;
        bool ha = false, hb = false;
        for (int e2 = s0; e2 < s1; e2++) {
            const int J2 = Jst[e2];
            rank += (J2 < J && first[e2]) ? 1 : 0;
            if (J2 == J) {
                if (e2 < split) {
                    sa = sa + tst[e2];
                    ha = true;
                } else {
                    sb = sb + tst[e2];
                    hb = true;
                }
            }
        }
        double acc = 0.0;
        if (ha) acc = acc + p0 * sa;
        if (hb) acc = acc + p1 * sb;
        colc[out0 + rank] = J;
        valc[out0 + rank] = acc;
    }
}

/* w_c(I) = 0 + sum over members ascending of p_i w_i */
__global__ __launch_bounds__(BS) void k_restrict_w(int nc, const int *__restrict__ members,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ wc)
{
    const int I = blockIdx.x * BS + threadIdx.x;
    if (I >= nc) return;
    const int m0 = members[2 * I], m1 = members[2 * I + 1];
    double t = 0.0;
    t = t + p[m0] * w[m0];
    if (m1 >= 0) t = t + p[m1] * w[m1];
    wc[I] = t;
}

__global__ __launch_bounds__(BS) void k_compose(int n0, int *__restrict__ aggP,
                                                double *__restrict__ pP,
                                                const int *__restrict__ agg,
                                                const double *__restrict__ p)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n0) return;
    const int I = aggP[i];
    pP[i] = 0.0 + pP[i] * p[I];
    aggP[i] = agg[I];
}

/* ------------------------------------------------------------ transpose */

__global__ __launch_bounds__(BS) void k_count(int n, const int *__restrict__ agg,
                                              int *__restrict__ cnt)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) atomicAdd(&cnt[agg[i]], 1);
}

__global__ __launch_bounds__(BS) void k_tfill(int n, const int *__restrict__ agg,
                                              const double *__restrict__ p,
                                              int *__restrict__ cursor, int *__restrict__ col,
                                              double *__restrict__ val)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    const int pos = atomicAdd(&cursor[agg[i]], 1);
    col[pos] = i;
    val[pos] = p[i];
}

/* sort each (short) row by column: insertion sort */
__global__ __launch_bounds__(BS) void k_sort_rows(int nr, const int *__restrict__ rp,
                                                  int *__restrict__ col, double *__restrict__ val)
{
    const int I = blockIdx.x * BS + threadIdx.x;
    if (I >= nr) return;
    const int a = rp[I], b = rp[I + 1];
    for (int k = a + 1; k < b; k++) {
        const int c = col[k];
        const double v = val[k];
        int m = k - 1;
        while (m >= a && col[m] > c) {
            col[m + 1] = col[m];
            val[m + 1] = val[m];
            m--;
        }
        col[m + 1] = c;
        val[m + 1] = v;
    }
}

int group_for(double avg)
{
    int G = 4;
    while (G < 64 && G < avg) G <<= 1;
    return G;
}

} // namespace

/* ================================================================ host */

namespace {

/* exclusive scan of n ints with the block offsets in boff (at least
   ceil(n / SCAN_CHUNK) + 1 ints); out[n] = total */
long long exclusive_scan_into(Ctx &c, const int *in, int *out, int n, int *boff, bool read_total)
{
    if (n <= 0) {
        BCMG_HIP(hipMemsetAsync(out, 0, sizeof(int), c.stream));
        return 0;
    }
    const int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    k_scan_blocks<<<nb, BS, 0, c.stream>>>(n, in, boff);
    k_scan_top<<<1, BS, 0, c.stream>>>(nb, boff);
    k_scan_apply<<<nb, BS, 0, c.stream>>>(n, in, boff, out);
    BCMG_CHECK_LAUNCH();
    if (!read_total) return -1;
    BCMG_HIP(hipMemcpyAsync(c.h_ints, out + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    return c.h_ints[0];
}

} // namespace

long long exclusive_scan(Ctx &c, const int *in, int *out, int n, bool read_total)
{
    const int nb = n > 0 ? (n + SCAN_CHUNK - 1) / SCAN_CHUNK : 0;
    c.iscratch.ensure((size_t)nb + 1);
    return exclusive_scan_into(c, in, out, n, c.iscratch.get(), read_total);
}

void launch_diag(hipStream_t s, const DevCsr &A, double *D)
{
    k_diag<<<blocks_for(A.n), BS, 0, s>>>(A.n, A.row.get(), A.col.get(), A.val.get(), D);
    BCMG_CHECK_LAUNCH();
}

void launch_abs_rowsum(hipStream_t s, const DevCsr &A, double *M)
{
    k_abs_rowsum<<<blocks_for(A.n), BS, 0, s>>>(A.n, A.row.get(), A.val.get(), M);
    BCMG_CHECK_LAUNCH();
}

void match_weights(Ctx &c, const DevCsr &A, const double *w, double *W)
{
    const int n = A.n;
    if (n <= 0) return;
    DBuf<double> C(n);
    const int nb = blocks_for(n);
    DBuf<unsigned long long> mb(nb + 1);
    k_weights_C<<<nb, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(), A.val.get(), w, C.get());
    k_weights_AH<<<nb, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(), A.val.get(), w, C.get(), W,
                                          mb.get() + 1);
    k_min_final<<<1, BS, 0, c.stream>>>(nb, mb.get() + 1, mb.get());
    k_weights_log<<<nb, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(), mb.get(), W);
    BCMG_CHECK_LAUNCH();
}

namespace {

/* device memory of one suitor matching beyond its inputs and outputs: the
   offer word of every vertex, the non-mutual flags, their prefix sums, the
   list of non-mutual candidates and the scan's block offsets. Every word is
   written before it is read, so nothing needs clearing. */
struct SuitorWork {
    unsigned long long *S;
    int *flag, *pos, *list, *boff;
};

int match_suitor_with(Ctx &c, const DevCsr &A, const double *W, int *M, int *Msym,
                      const SuitorWork &w, bool want_count)
{
    const int n = A.n;
    k_fill_u64<<<blocks_for(n), BS, 0, c.stream>>>(n, ~0ull, w.S);
    k_suitor<<<blocks_for(n), BS, 0, c.stream>>>(n, A.row.get(), A.col.get(), W, w.S);
    k_suitor_extract<<<blocks_for(n), BS, 0, c.stream>>>(n, w.S, M);
    k_sym_pairs<<<blocks_for(n), BS, 0, c.stream>>>(n, M, Msym, w.flag);
    BCMG_CHECK_LAUNCH();
    exclusive_scan_into(c, w.flag, w.pos, n, w.boff, false);
    k_compact<<<blocks_for(n), BS, 0, c.stream>>>(n, w.flag, w.pos, w.list);
    k_unsym_serial<<<1, 64, 0, c.stream>>>(w.pos + n, w.list, M, Msym);
    BCMG_CHECK_LAUNCH();
    if (!want_count) return -1;
    BCMG_HIP(hipMemcpyAsync(c.h_ints, w.pos + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    return c.h_ints[0];
}

} // namespace

int match_suitor(Ctx &c, const DevCsr &A, const double *W, int *M, int *Msym, bool want_count)
{
    const int n = A.n;
    DBuf<unsigned long long> S(n);
    DBuf<int> flag(n), pos(n + 1), list(n), boff((n + SCAN_CHUNK - 1) / SCAN_CHUNK + 1);
    return match_suitor_with(c, A, W, M, Msym,
                             SuitorWork{S.get(), flag.get(), pos.get(), list.get(), boff.get()},
                             want_count);
}

int aggregate_pairs(Ctx &c, int n, const int *Msym, const double *w, int *agg, double *p,
                    int *members, bool validate)
{
    DBuf<int> flag(n), ids(n + 1), err;
    if (validate) {
        err.alloc(1);
        BCMG_HIP(hipMemsetAsync(err.get(), 0, sizeof(int), c.stream));
    }
    k_agg_leader<<<blocks_for(n), BS, 0, c.stream>>>(n, Msym, w, flag.get(), err.get());
    BCMG_CHECK_LAUNCH();
    if (validate) {
        BCMG_HIP(hipMemcpyAsync(c.h_ints, err.get(), sizeof(int), hipMemcpyDeviceToHost,
                                c.stream));
        BCMG_HIP(hipStreamSynchronize(c.stream));
        if (c.h_ints[0] != 0)
            throw Error(BCMG_EINVAL, "aggregation: M_sym is not a symmetric matching (a "
                                     "partner out of range, the row itself, or not matched back)");
    }
    const long long nc = exclusive_scan(c, flag.get(), ids.get(), n);
    k_agg_assign<<<blocks_for(n), BS, 0, c.stream>>>(n, Msym, w, flag.get(), ids.get(), agg, p,
                                                      members);
    BCMG_CHECK_LAUNCH();
    return (int)nc;
}

void members_from_agg(Ctx &c, int n, const int *agg, int nc, int *members)
{
    DBuf<int> cnt(nc), err(1);
    BCMG_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int) * nc, c.stream));
    BCMG_HIP(hipMemsetAsync(err.get(), 0, sizeof(int), c.stream));
    k_members_place<<<blocks_for(n), BS, 0, c.stream>>>(n, agg, nc, cnt.get(), members, err.get());
    k_members_fix<<<blocks_for(nc), BS, 0, c.stream>>>(nc, cnt.get(), members, err.get());
    BCMG_CHECK_LAUNCH();
    BCMG_HIP(hipMemcpyAsync(c.h_ints, err.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    if (c.h_ints[0] != 0)
        throw Error(BCMG_EINVAL, "aggregation is not pairwise (every coarse index needs 1 or 2 fine rows)");
}

#define BCMG_GROUP_SWITCH(G, KERNEL, GRID, ...)                                         \
    switch (G) {                                                                       \
    case 4: KERNEL<4><<<GRID, BS, 0, c.stream>>>(__VA_ARGS__); break;                  \
    case 8: KERNEL<8><<<GRID, BS, 0, c.stream>>>(__VA_ARGS__); break;                  \
    case 16: KERNEL<16><<<GRID, BS, 0, c.stream>>>(__VA_ARGS__); break;                \
    case 32: KERNEL<32><<<GRID, BS, 0, c.stream>>>(__VA_ARGS__); break;                \
    default: KERNEL<64><<<GRID, BS, 0, c.stream>>>(__VA_ARGS__); break;                \
    }

void galerkin_pairwise(Ctx &c, const DevCsr &A, const int *agg, const double *p,
                       const int *members, int nc, DevCsr &Ac)
{
    const int n = A.n;
    DBuf<int> len(nc), stoff(nc + 1);
    k_stage_len<<<blocks_for(nc), BS, 0, c.stream>>>(nc, A.row.get(), members, len.get());
    BCMG_CHECK_LAUNCH();
    /* the segments of all aggregates together hold every entry of A once */
    exclusive_scan(c, len.get(), stoff.get(), nc, false);
    const long long tot = A.nnz;
    DBuf<int> Jst(tot);
    DBuf<double> tst(tot);
    DBuf<unsigned char> first(tot);
    const int Gs = group_for((double)A.nnz / (double)n);
    BCMG_GROUP_SWITCH(Gs, k_stage, blocks_for((long long)n * Gs), n, A.row.get(), A.col.get(),
                      A.val.get(), agg, p, members, stoff.get(), Jst.get(), tst.get());
    const int Gr = group_for((double)tot / (double)nc);
    DBuf<int> cnt(nc);
    BCMG_GROUP_SWITCH(Gr, k_rap_count, blocks_for((long long)nc * Gr), nc, stoff.get(), Jst.get(),
                      first.get(), cnt.get());
    BCMG_CHECK_LAUNCH();
    Ac.n = Ac.m = nc;
    Ac.row.alloc(nc + 1);
    exclusive_scan(c, cnt.get(), Ac.row.get(), nc, false);
    /* one round trip for the coarse nnz and the row-block tile */
    launch_block_nnz(c, Ac);
    BCMG_HIP(hipMemcpyAsync(c.h_ints, Ac.row.get() + nc, sizeof(int), hipMemcpyDeviceToHost,
                            c.stream));
    BCMG_HIP(hipMemcpyAsync(c.h_ints + 1, Ac.tile_dev.get(), 2 * sizeof(int),
                            hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    const long long nnzc = c.h_ints[0];
    Ac.nnz = (int)nnzc;
    Ac.tile = c.h_ints[1] > 0 ? c.h_ints[1] : 1;
    Ac.maxrow = c.h_ints[2];
    Ac.group = lanes_for(nc, nnzc);
    Ac.col.alloc(nnzc);
    Ac.val.alloc(nnzc);
    BCMG_GROUP_SWITCH(Gr, k_rap_fill, blocks_for((long long)nc * Gr), nc, A.row.get(), members, p,
                      stoff.get(), Jst.get(), tst.get(), first.get(), Ac.row.get(), Ac.col.get(),
                      Ac.val.get());
    BCMG_CHECK_LAUNCH();
}

void restrict_w(hipStream_t s, int nc, const int *members, const double *p, const double *w,
                double *wc)
{
    k_restrict_w<<<blocks_for(nc), BS, 0, s>>>(nc, members, p, w, wc);
    BCMG_CHECK_LAUNCH();
}

void compose_P(hipStream_t s, int n0, int *aggP, double *pP, const int *agg, const double *p)
{
    k_compose<<<blocks_for(n0), BS, 0, s>>>(n0, aggP, pP, agg, p);
    BCMG_CHECK_LAUNCH();
}

void transpose_P(Ctx &c, int n, int nc, const int *agg, const double *p, DevCsr &R)
{
    DBuf<int> cnt(nc), cursor(nc + 1);
    BCMG_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int) * nc, c.stream));
    k_count<<<blocks_for(n), BS, 0, c.stream>>>(n, agg, cnt.get());
    BCMG_CHECK_LAUNCH();
    R.n = nc;
    R.m = n;
    R.nnz = n;
    R.row.alloc(nc + 1);
    R.col.alloc(n);
    R.val.alloc(n);
    exclusive_scan(c, cnt.get(), R.row.get(), nc, false);
    BCMG_HIP(hipMemcpyAsync(cursor.get(), R.row.get(), sizeof(int) * (nc + 1),
                            hipMemcpyDeviceToDevice, c.stream));
    k_tfill<<<blocks_for(n), BS, 0, c.stream>>>(n, agg, p, cursor.get(), R.col.get(), R.val.get());
    k_sort_rows<<<blocks_for(nc), BS, 0, c.stream>>>(nc, R.row.get(), R.col.get(), R.val.get());
    BCMG_CHECK_LAUNCH();
    /* tile: read later together with the other pending tiles */
    launch_block_nnz(c, R);
}

} // namespace bcmg
