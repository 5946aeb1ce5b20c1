/*
 * kernels_spgemm.hip - the sparse products of smoothed aggregation on gfx950
 * (src/AMG/matchingAggregation.cu:353-390): the smoother S = I - omega D^-1 A,
 * the general product C = L R and the general transpose.
 *
 * C = L R is computed row by row in one fixed summation order: every entry
 * C(i,j) is 0 + the products L(i,e) R(e,j), taken over L's row i in storage
 * order and, within it, over R's row e in storage order (the reference's hash
 * SpGEMM adds with shared-memory atomics in arrival order, so it fixes no
 * order). A row's products are expanded into LDS under a 64-bit key
 * (row, column, position), bitonic-sorted, and each run of one (row, column)
 * is summed by one lane in position order:
 *   - rows with at most SG_CAP products are packed into workgroups by a
 *     prefix sum of (products + 1): a workgroup owns the rows whose prefix
 *     starts in its SG_CAP-slot range, and since such a row ends within the
 *     next SG_CAP slots, 2 SG_CAP keys hold all of them;
 *   - a longer row is taken by a workgroup SG_CH products at a time: each
 *     chunk's column runs continue their columns' sums in a dense accumulator
 *     of the workgroup (one FP64 value and one row stamp per column of C), and
 *     the touched columns are put in order at the end of the row; when the
 *     right rows are long (SG_SERIAL_MIN products per left entry or more) and
 *     hold distinct columns, no sort is needed: the left entries are taken
 *     one at a time, each adding its products column-parallel.
 * The pattern is symbolic (a cancelled entry stays), columns ascending. The
 * file is compiled with -ffp-contract=off: products and sums are rounded
 * separately, as in the serial restatement.
 */
#include "bcmg_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace bcmg {

namespace {

constexpr int BS = 256;
constexpr int SG_CAP = 1024;          /* products of a packed row, at most */
constexpr int SG_KEYS = 2 * SG_CAP;   /* keys of a packed workgroup */
constexpr int SG_CH = 2048;           /* products per chunk of a long row */
constexpr int SG_SORT = 2 * SG_CH;    /* touched columns sorted in LDS, at most */
constexpr long long SG_DENSE_BYTES = 2ll << 30; /* dense accumulators, all workgroups */
constexpr int SG_DENSE_BLOCKS = 1024;
constexpr int SG_SERIAL_MIN = 64;     /* products per left entry for the serial mode */
constexpr int SG_LDS_COLS = 4096;     /* columns of C for LDS accumulators */
constexpr int SHORT_ROW = 64;         /* transpose rows ranked by one lane */
constexpr int RANK_CHUNK = 4096;      /* columns per LDS load when ranking */
constexpr unsigned long long SG_NONE = ~0ull;
constexpr int POS_BITS = 11;          /* position field: < 2048 */
constexpr int ROW_SHIFT = 42;         /* row field above 31 column bits */
constexpr long long SG_PER_PASS = 1ll << 28; /* slots + weights per batch of rows */

inline int blocks_for(long long w) { return (int)((w + BS - 1) / BS); }

/* first index t in [0, n) with a[t] >= x (n if none) */
__device__ inline int lower_bound_i(const int *a, int n, int x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* exclusive prefix sum of v over the workgroup; *tot = the sum of all */
__device__ inline int block_scan_excl(int v, int *sh, int *tot)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; w++) {
        const int t = sh[w];
        if (w < wv) pre += t;
        all += t;
    }
    __syncthreads();
    *tot = all;
    return pre + x - v;
}

/* ascending bitonic sort of key[0, N), N a power of two */
template <typename K>
__device__ inline void bitonic_sort(K *key, int N)
{
    for (int k = 2; k <= N; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < (N >> 1); t += BS) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int l = i + j;
                const K a = key[i], b = key[l];
                const bool up = (i & k) == 0;
                if ((a > b) == up) {
                    key[i] = b;
                    key[l] = a;
                }
            }
            __syncthreads();
        }
}

__device__ inline int pow2_at_least(int v)
{
    int N = 2;
    while (N < v) N <<= 1;
    return N;
}

/* per row: products u, output slots sl = min(products, columns of C) (a
   row has at most that many entries), packing weight ws, the long-row
   list; the total of slots plus weights (what a batch of rows is bounded
   by); *big = 1 if a row has 2^30 products or more */
__global__ __launch_bounds__(BS) void k_sg_count(int n, const int *__restrict__ Lrp,
                                                 const int *__restrict__ Lci,
                                                 const int *__restrict__ Rrp, long long m,
                                                 int *__restrict__ u, int *__restrict__ sl,
                                                 int *__restrict__ ws, int *__restrict__ cnt,
                                                 int *__restrict__ large,
                                                 int *__restrict__ nlarge,
                                                 unsigned long long *__restrict__ total,
                                                 int *__restrict__ big)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    long long s = 0, slots = 0;
    if (i < n) {
        for (int e = Lrp[i]; e < Lrp[i + 1]; e++) {
            const int k = Lci[e];
            s += Rrp[k + 1] - Rrp[k];
        }
        if (s >= (1ll << 30)) atomicOr(big, 1);
        u[i] = (int)(s < (1ll << 30) ? s : (1ll << 30));
        slots = s < m ? s : m;
        sl[i] = (int)slots;
        cnt[i] = 0;
        if (s <= SG_CAP) {
            ws[i] = (int)s + 1;
        } else {
            ws[i] = 0;
            large[atomicAdd(nlarge, 1)] = i;
        }
    }
    /* the batch cost of a row: its slots plus its packing weight */
    unsigned long long t = (unsigned long long)slots + (i < n && s <= SG_CAP ? s + 1 : 0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(total, t);
}

/* the packed rows of one workgroup (see the file comment) */
__global__ __launch_bounds__(BS) void k_sg_small(int n, const int *__restrict__ Lrp,
                                                 const int *__restrict__ Lci,
                                                 const double *__restrict__ Lv,
                                                 const int *__restrict__ Rrp,
                                                 const int *__restrict__ Rci,
                                                 const double *__restrict__ Rv,
                                                 const int *__restrict__ ws,
                                                 const int *__restrict__ soff,
                                                 const int *__restrict__ uoff,
                                                 int *__restrict__ tcol, double *__restrict__ tval,
                                                 int *__restrict__ cnt)
{
    __shared__ unsigned long long key[SG_KEYS];
    __shared__ double val[SG_KEYS];
    __shared__ int rowid[SG_CAP], rowfirst[SG_CAP];
    __shared__ int sh[BS / 64];
    __shared__ int rng[3];
    const int lo = blockIdx.x * SG_CAP;
    if (threadIdx.x == 0) {
        const int r0 = lower_bound_i(soff, n, lo);
        const int r1 = lower_bound_i(soff, n, lo + SG_CAP);
        rng[0] = r0;
        rng[1] = r1;
        rng[2] = soff[r1] - lo;        /* slots used, <= 2 SG_CAP */
    }
    for (int t = threadIdx.x; t < SG_KEYS; t += BS) key[t] = SG_NONE;
    __syncthreads();
    const int r0 = rng[0], r1 = rng[1];
    const int N = pow2_at_least(rng[2]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int r = r0 + wv; r < r1; r += BS / 64) {
        if (ws[r] == 0) continue;      /* a long row: a workgroup of its own */
        const int base = soff[r] - lo;
        if (lane == 0) rowid[base] = r;
        int off = base;
        const int e1 = Lrp[r + 1];
        for (int e0 = Lrp[r]; e0 < e1; e0 += 64) {
            const int e = e0 + lane;
            int len = 0, f0 = 0;
            double lv = 0.0;
            if (e < e1) {
                const int k = Lci[e];
                f0 = Rrp[k];
                len = Rrp[k + 1] - f0;
                lv = Lv[e];
            }
            int x = len;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            /* the window's products spread over the lanes: product q belongs
               to the first entry whose inclusive prefix exceeds q */
            const int total = __shfl(x, 63, 64);
            for (int q0 = 0; q0 < total; q0 += 64) {
                const int q = q0 + lane;
                int lo = 0, hi = 63;
#pragma unroll
                for (int st = 0; st < 6; st++) {
                    const int mid = (lo + hi) >> 1;
                    const int v = __shfl(x, mid, 64);
                    if (v <= q)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                const int xe = __shfl(x, lo, 64), le = __shfl(len, lo, 64);
                const int fe = __shfl(f0, lo, 64);
                const double ve = __shfl(lv, lo, 64);
                if (q < total) {
                    const int f = fe + (q - (xe - le));
                    const int pos = off + q;
                    key[pos] = ((unsigned long long)base << ROW_SHIFT) |
                               ((unsigned long long)Rci[f] << POS_BITS) | (unsigned)pos;
                    val[pos] = ve * Rv[f];
                }
            }
            off += total;
        }
    }
    __syncthreads();
    bitonic_sort(key, N);
    constexpr int PT = SG_KEYS / BS;
    const int t0 = threadIdx.x * PT;
    int nh = 0;
#pragma unroll
    for (int q = 0; q < PT; q++) {
        const int t = t0 + q;
        if (t < N) {
            const unsigned long long k = key[t];
            if (k != SG_NONE && (t == 0 || (key[t - 1] >> POS_BITS) != (k >> POS_BITS))) nh++;
        }
    }
    int tot;
    const int uid0 = block_scan_excl(nh, sh, &tot);
    int uid = uid0;
#pragma unroll
    for (int q = 0; q < PT; q++) {
        const int t = t0 + q;
        if (t < N) {
            const unsigned long long k = key[t];
            if (k != SG_NONE && (t == 0 || (key[t - 1] >> POS_BITS) != (k >> POS_BITS))) {
                if (t == 0 || (key[t - 1] >> ROW_SHIFT) != (k >> ROW_SHIFT))
                    rowfirst[k >> ROW_SHIFT] = uid;
                uid++;
            }
        }
    }
    __syncthreads();
    uid = uid0;
    for (int q = 0; q < PT; q++) {
        const int t = t0 + q;
        if (t >= N) break;
        const unsigned long long k = key[t];
        if (k == SG_NONE || (t > 0 && (key[t - 1] >> POS_BITS) == (k >> POS_BITS))) continue;
        const unsigned long long g = k >> POS_BITS;
        double s = 0.0;
        int t2 = t;
        for (; t2 < N && (key[t2] >> POS_BITS) == g; t2++)
            s = s + val[key[t2] & ((1u << POS_BITS) - 1)];
        const int base = (int)(k >> ROW_SHIFT);
        const int r = rowid[base];
        const int idx = uid - rowfirst[base];
        const int o = uoff[r] + idx;
        tcol[o] = (int)(g & 0x7fffffffull);
        tval[o] = s;
        if (t2 >= N || (key[t2] >> ROW_SHIFT) != (unsigned long long)base) cnt[r] = idx + 1;
        uid++;
    }
}

/* Long rows, one workgroup each in turn (a grid-stride loop over the list):
   the row's products SG_CH at a time, sorted by (column, position); each
   column run continues its column's sum in the workgroup's dense
   accumulator (acc, valid where stamp holds the row's id, so nothing is
   cleared between rows) in position order. The touched columns are listed
   as they appear, then put in ascending order: a sort in LDS up to SG_SORT
   of them, else a scan of the stamps over the touched column range. Row ids
   in the stamps are global (row_base + r): batches share the accumulators. */
__global__ __launch_bounds__(BS) void k_sg_dense(int nlarge, const int *__restrict__ large,
                                                 const int *__restrict__ Lrp,
                                                 const int *__restrict__ Lci,
                                                 const double *__restrict__ Lv,
                                                 const int *__restrict__ Rrp,
                                                 const int *__restrict__ Rci,
                                                 const double *__restrict__ Rv,
                                                 const int *__restrict__ u,
                                                 const int *__restrict__ uoff,
                                                 int *__restrict__ tcol, double *__restrict__ tval,
                                                 double *acc_all, int *stamp_all, long long m,
                                                 int row_base, int serial_min, int skip_serial,
                                                 int *__restrict__ cnt)
{
    __shared__ unsigned long long key[SG_CH];   /* also the sort buffer of SG_SORT ints */
    __shared__ double val[SG_CH];
    __shared__ int sh[BS / 64];
    __shared__ int ws0[BS + 1], wf0[BS];        /* a window of left entries */
    __shared__ double wlv[BS];
    __shared__ int ntouch, cmin, cmax;
    double *const acc = acc_all + (long long)blockIdx.x * m;
    int *const stamp = stamp_all + (long long)blockIdx.x * m;
    int *const skey = reinterpret_cast<int *>(key);
    for (int q = blockIdx.x; q < nlarge; q += gridDim.x) {
        const int r = large[q];
        const int id = row_base + r;
        const int U = u[r];
        const int e_end = Lrp[r + 1];
        int ew0 = Lrp[r], pb0 = 0;   /* first window some later chunk needs */
        const bool serial = serial_min > 0 && (long long)U >= (long long)serial_min * (e_end - ew0);
        if (serial && skip_serial) continue;   /* k_sg_serial_lds takes it */
        int *const ocol = tcol + uoff[r];
        double *const oval = tval + uoff[r];
        if (threadIdx.x == 0) {
            ntouch = 0;
            cmin = INT_MAX;
            cmax = -1;
        }
        /* long right rows (distinct columns): one left entry at a time, its
           products added column-parallel - the same order per column */
        for (int e0 = ew0; serial && e0 < e_end; e0 += BS) {
            const int e = e0 + threadIdx.x;
            if (e < e_end) {
                const int k = Lci[e];
                wf0[threadIdx.x] = Rrp[k];
                ws0[threadIdx.x] = Rrp[k + 1];
                wlv[threadIdx.x] = Lv[e];
            }
            __syncthreads();
            const int ne = min(BS, e_end - e0);
            /* the lane's first product of the next entry is loaded before
               the current entry's accumulation, so its latency overlaps */
            int pf = wf0[0] + threadIdx.x;
            bool have = pf < ws0[0];
            int pc = have ? Rci[pf] : 0;
            double pv = have ? Rv[pf] : 0.0;
            for (int t = 0; t < ne; t++) {
                const int f1 = ws0[t];
                const double lv = wlv[t];
                const int nf = t + 1 < ne ? wf0[t + 1] + threadIdx.x : 0;
                const bool nhave = t + 1 < ne && nf < ws0[t + 1];
                const int nc = nhave ? Rci[nf] : 0;
                const double nv = nhave ? Rv[nf] : 0.0;
                for (int f = pf; have && f < f1; f += BS) {
                    const int col = f == pf ? pc : Rci[f];
                    const double prod = lv * (f == pf ? pv : Rv[f]);
                    const bool first = stamp[col] != id;
                    acc[col] = (first ? 0.0 : acc[col]) + prod;
                    if (first) {
                        stamp[col] = id;
                        ocol[atomicAdd(&ntouch, 1)] = col;
                        atomicMin(&cmin, col);
                        atomicMax(&cmax, col);
                    }
                }
                __threadfence_block();
                __syncthreads();
                pf = nf;
                have = nhave;
                pc = nc;
                pv = nv;
            }
        }
        for (int c0 = 0; !serial && c0 < U; c0 += SG_CH) {
            const int c1 = min(c0 + SG_CH, U);
            for (int t = threadIdx.x; t < SG_CH; t += BS) key[t] = SG_NONE;
            __syncthreads();
            int ew = ew0, pb = pb0;
            while (ew < e_end && pb < c1) {
                const int e = ew + threadIdx.x;
                int len = 0, f0 = 0;
                double lv = 0.0;
                if (e < e_end) {
                    const int k = Lci[e];
                    f0 = Rrp[k];
                    len = Rrp[k + 1] - f0;
                    lv = Lv[e];
                }
                int tot;
                const int s0 = pb + block_scan_excl(len, sh, &tot);
                ws0[threadIdx.x + 1] = s0 + len;
                wf0[threadIdx.x] = f0 - s0;
                wlv[threadIdx.x] = lv;
                if (threadIdx.x == 0) ws0[0] = pb;
                __syncthreads();
                /* the window's products in the chunk, spread over the lanes:
                   position p belongs to entry t with ws0[t] <= p < ws0[t + 1] */
                const int lo = max(pb, c0), hi = min(pb + tot, c1);
                for (int p = lo + threadIdx.x; p < hi; p += BS) {
                    int a = 0, b = BS;
                    while (b - a > 1) {
                        const int mid = (a + b) >> 1;
                        if (ws0[mid] <= p)
                            a = mid;
                        else
                            b = mid;
                    }
                    const int f = wf0[a] + p;
                    const int pos = p - c0;
                    key[pos] = ((unsigned long long)Rci[f] << POS_BITS) | (unsigned)pos;
                    val[pos] = wlv[a] * Rv[f];
                }
                __syncthreads();
                if (pb + tot <= c1) {
                    ew0 = ew + BS;
                    pb0 = pb + tot;
                }
                ew += BS;
                pb += tot;
            }
            __syncthreads();
            const int N = pow2_at_least(c1 - c0);
            bitonic_sort(key, N);
            for (int t = threadIdx.x; t < N; t += BS) {
                const unsigned long long k = key[t];
                if (k == SG_NONE || (t > 0 && (key[t - 1] >> POS_BITS) == (k >> POS_BITS)))
                    continue;
                const int col = (int)(k >> POS_BITS);
                const bool first = stamp[col] != id;
                double s = first ? 0.0 : acc[col];
                for (int t2 = t; t2 < N && (key[t2] >> POS_BITS) == (k >> POS_BITS); t2++)
                    s = s + val[key[t2] & ((1u << POS_BITS) - 1)];
                acc[col] = s;
                if (first) {
                    stamp[col] = id;
                    ocol[atomicAdd(&ntouch, 1)] = col;
                    atomicMin(&cmin, col);
                    atomicMax(&cmax, col);
                }
            }
            __threadfence_block();
            __syncthreads();
        }
        const int k = ntouch;
        if (k <= SG_SORT) {
            const int N = pow2_at_least(k);
            for (int t = threadIdx.x; t < N; t += BS) skey[t] = t < k ? ocol[t] : INT_MAX;
            __syncthreads();
            bitonic_sort(skey, N);
            for (int t = threadIdx.x; t < k; t += BS) {
                const int c = skey[t];
                ocol[t] = c;
                oval[t] = acc[c];
            }
        } else {
            const int j0 = cmin, j1 = cmax;
            int base = 0;
            for (int jb = j0; jb <= j1; jb += BS) {
                const int j = jb + threadIdx.x;
                const int fl = (j <= j1 && stamp[j] == id) ? 1 : 0;
                int tot;
                const int pos = base + block_scan_excl(fl, sh, &tot);
                if (fl) {
                    ocol[pos] = j;
                    oval[pos] = acc[j];
                }
                base += tot;
            }
        }
        if (threadIdx.x == 0) cnt[r] = k;
        __threadfence_block();
        __syncthreads();
    }
}

/* The serial-mode long rows when C has at most SG_LDS_COLS columns: the
   accumulator and the row marks live in LDS; the row's entries come out in
   column order from a scan of the marks (no sort). Same order per column as
   the serial mode of k_sg_dense. */
__global__ __launch_bounds__(BS) void k_sg_serial_lds(int nlarge, const int *__restrict__ large,
                                                      const int *__restrict__ Lrp,
                                                      const int *__restrict__ Lci,
                                                      const double *__restrict__ Lv,
                                                      const int *__restrict__ Rrp,
                                                      const int *__restrict__ Rci,
                                                      const double *__restrict__ Rv,
                                                      const int *__restrict__ u,
                                                      const int *__restrict__ uoff,
                                                      int *__restrict__ tcol,
                                                      double *__restrict__ tval, int m,
                                                      int serial_min, int *__restrict__ cnt)
{
    __shared__ double acc[SG_LDS_COLS];
    __shared__ int mark[SG_LDS_COLS];
    __shared__ int ws0[BS], wf0[BS];
    __shared__ double wlv[BS];
    __shared__ int sh[BS / 64];
    for (int j = threadIdx.x; j < m; j += BS) mark[j] = -1;
    __syncthreads();
    for (int q = blockIdx.x; q < nlarge; q += gridDim.x) {
        const int r = large[q];
        const int U = u[r];
        const int e_end = Lrp[r + 1];
        const int eb = Lrp[r];
        if (!((long long)U >= (long long)serial_min * (e_end - eb))) continue;
        for (int e0 = eb; e0 < e_end; e0 += BS) {
            const int e = e0 + threadIdx.x;
            if (e < e_end) {
                const int k = Lci[e];
                wf0[threadIdx.x] = Rrp[k];
                ws0[threadIdx.x] = Rrp[k + 1];
                wlv[threadIdx.x] = Lv[e];
            }
            __syncthreads();
            const int ne = min(BS, e_end - e0);
            int pf = wf0[0] + threadIdx.x;
            bool have = pf < ws0[0];
            int pc = have ? Rci[pf] : 0;
            double pv = have ? Rv[pf] : 0.0;
            for (int t = 0; t < ne; t++) {
                const int f1 = ws0[t];
                const double lv = wlv[t];
                const int nf = t + 1 < ne ? wf0[t + 1] + threadIdx.x : 0;
                const bool nhave = t + 1 < ne && nf < ws0[t + 1];
                const int nc = nhave ? Rci[nf] : 0;
                const double nv = nhave ? Rv[nf] : 0.0;
                for (int f = pf; have && f < f1; f += BS) {
                    const int col = f == pf ? pc : Rci[f];
                    const double prod = lv * (f == pf ? pv : Rv[f]);
                    const bool first = mark[col] != q;
                    acc[col] = (first ? 0.0 : acc[col]) + prod;
                    if (first) mark[col] = q;
                }
                __syncthreads();
                pf = nf;
                have = nhave;
                pc = nc;
                pv = nv;
            }
        }
        int *const ocol = tcol + uoff[r];
        double *const oval = tval + uoff[r];
        int base = 0;
        for (int jb = 0; jb < m; jb += BS) {
            const int j = jb + threadIdx.x;
            const int fl = (j < m && mark[j] == q) ? 1 : 0;
            int tot;
            const int pos = base + block_scan_excl(fl, sh, &tot);
            if (fl) {
                ocol[pos] = j;
                oval[pos] = acc[j];
            }
            base += tot;
        }
        if (threadIdx.x == 0) cnt[r] = base;
        __syncthreads();
    }
}

/* compact the rows from their product-count slots into C (G lanes a row) */
template <int G>
__global__ __launch_bounds__(BS) void k_sg_gather(int n, const int *__restrict__ uoff,
                                                  const int *__restrict__ Crp,
                                                  const int *__restrict__ tcol,
                                                  const double *__restrict__ tval,
                                                  int *__restrict__ ccol,
                                                  double *__restrict__ cval)
{
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int row = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= n) return;
    const int s = uoff[row], d = Crp[row], len = Crp[row + 1] - d;
    for (int q = lane; q < len; q += G) {
        ccol[d + q] = tcol[s + q];
        cval[d + q] = tval[s + q];
    }
}

/* ------------------------------------------------------------ transpose */

__global__ __launch_bounds__(BS) void k_col_count(int n, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  int *__restrict__ cnt)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; e++) atomicAdd(&cnt[ci[e]], 1);
}

__global__ __launch_bounds__(BS) void k_col_fill(int n, const int *__restrict__ rp,
                                                 const int *__restrict__ ci,
                                                 const double *__restrict__ v,
                                                 int *__restrict__ cursor, int *__restrict__ tcol,
                                                 double *__restrict__ tval)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
        const int pos = atomicAdd(&cursor[ci[e]], 1);
        tcol[pos] = i;
        tval[pos] = v[e];
    }
}

/* Each row of the transpose in ascending column order: an entry's place is
   its rank among the row's entries (column, then fill position for equal
   columns). Rows up to SHORT_ROW entries: one lane each; longer rows are
   listed for k_rank_long. */
__global__ __launch_bounds__(BS) void k_rank_short(int nr, const int *__restrict__ rp,
                                                   const int *__restrict__ cin,
                                                   const double *__restrict__ vin,
                                                   int *__restrict__ col, double *__restrict__ val,
                                                   int *__restrict__ longl, int *__restrict__ nlong)
{
    const int I = blockIdx.x * BS + threadIdx.x;
    if (I >= nr) return;
    const int a = rp[I], b = rp[I + 1];
    if (b - a > SHORT_ROW) {
        longl[atomicAdd(nlong, 1)] = I;
        return;
    }
    for (int e = a; e < b; e++) {
        const int c = cin[e];
        int rank = 0;
        for (int f = a; f < b; f++) {
            const int d = cin[f];
            rank += (d < c || (d == c && f < e)) ? 1 : 0;
        }
        col[a + rank] = c;
        val[a + rank] = vin[e];
    }
}

/* a long row per workgroup: ranks against the row's columns staged in LDS,
   RANK_CHUNK at a time */
__global__ __launch_bounds__(BS) void k_rank_long(const int *__restrict__ longl,
                                                  const int *__restrict__ rp,
                                                  const int *__restrict__ cin,
                                                  const double *__restrict__ vin,
                                                  int *__restrict__ col, double *__restrict__ val)
{
    __shared__ int buf[RANK_CHUNK];
    const int I = longl[blockIdx.x];
    const int a = rp[I], len = rp[I + 1] - a;
    if (len <= RANK_CHUNK) {
        for (int t = threadIdx.x; t < len; t += BS) buf[t] = cin[a + t];
        __syncthreads();
        for (int e = threadIdx.x; e < len; e += BS) {
            const int c = buf[e];
            int rank = 0;
            for (int f = 0; f < len; f++) {
                const int d = buf[f];
                rank += (d < c || (d == c && f < e)) ? 1 : 0;
            }
            col[a + rank] = c;
            val[a + rank] = vin[a + e];
        }
        return;
    }
    for (int e0 = 0; e0 < len; e0 += BS) {
        const int e = e0 + threadIdx.x;
        const int c = e < len ? cin[a + e] : 0;
        int rank = 0;
        for (int k0 = 0; k0 < len; k0 += RANK_CHUNK) {
            const int nk = min(RANK_CHUNK, len - k0);
            __syncthreads();
            for (int t = threadIdx.x; t < nk; t += BS) buf[t] = cin[a + k0 + t];
            __syncthreads();
            if (e < len)
                for (int t = 0; t < nk; t++) {
                    const int d = buf[t];
                    rank += (d < c || (d == c && k0 + t < e)) ? 1 : 0;
                }
        }
        if (e < len) {
            col[a + rank] = c;
            val[a + rank] = vin[a + e];
        }
    }
}

/* ------------------------------------------------ smoother S, omega */

/* s_ij = a_ij / d_i (CSRm::matrixVectorScaling, src/matrix/CSR.cu:520-545)
   and the largest 0 + sum_j |s_ij| in storage order (CSR.cu:386-416): row
   sums are >= 0, so their bit patterns order as the values do */
__global__ __launch_bounds__(BS) void k_sa_scale(int n, const int *__restrict__ rp,
                                                 const double *__restrict__ av,
                                                 const double *__restrict__ D,
                                                 double *__restrict__ sv,
                                                 unsigned long long *__restrict__ norm)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const double d = D[i];
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const double t = av[e] / d;
            sv[e] = t;
            s += fabs(t);
        }
    }
    unsigned long long b = (s > 0.0) ? (unsigned long long)__double_as_longlong(s) : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(b, o, 64);
        b = y > b ? y : b;
    }
    if ((threadIdx.x & 63) == 0 && b) atomicMax(norm, b);
}

/* s_ij = -omega s_ij, plus 1 on the diagonal (_applyOmega,
   src/AMG/matchingAggregation.cu:64-80) */
__global__ __launch_bounds__(BS) void k_sa_omega(int n, const int *__restrict__ rp,
                                                 const int *__restrict__ ci, double omega,
                                                 double *__restrict__ sv)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
        double t = -omega * sv[e];
        if (ci[e] == i) t += 1.;
        sv[e] = t;
    }
}

__global__ __launch_bounds__(BS) void k_iota(int n, int *__restrict__ a)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ __launch_bounds__(BS) void k_negate(long long n, double *__restrict__ v)
{
    const long long i = (long long)blockIdx.x * BS + threadIdx.x;
    if (i < n) v[i] = -v[i];
}

} // namespace

/* ================================================================ host */

namespace {

/* gather rows [0, n) from their product slots into (col, val) by rp */
void gather_rows(Ctx &c, int n, const int *uoff, const int *rp, const int *tcol,
                 const double *tval, long long nnz, int *col, double *val)
{
    if (n <= 0 || nnz <= 0) return;
    const double avg = (double)nnz / (double)n;
    const int G = avg <= 2 ? 1 : avg <= 8 ? 4 : avg <= 32 ? 16 : 64;
    const int gb = blocks_for((long long)n * G);
    switch (G) {
    case 1: k_sg_gather<1><<<gb, BS, 0, c.stream>>>(n, uoff, rp, tcol, tval, col, val); break;
    case 4: k_sg_gather<4><<<gb, BS, 0, c.stream>>>(n, uoff, rp, tcol, tval, col, val); break;
    case 16: k_sg_gather<16><<<gb, BS, 0, c.stream>>>(n, uoff, rp, tcol, tval, col, val); break;
    default: k_sg_gather<64><<<gb, BS, 0, c.stream>>>(n, uoff, rp, tcol, tval, col, val); break;
    }
    BCMG_CHECK_LAUNCH();
}

} // namespace

/* Rows are taken in batches of at most per_pass output slots plus packing
   weights (a row over it is a batch of its own), so that the temporaries
   stay bounded and their offsets fit 32 bits; one batch writes C directly. */
void spgemm(Ctx &c, const CsrRef &L, const CsrRef &R, DevCsr &C, long long per_pass,
            bool right_unique)
{
    if (L.m != R.n) throw Error(BCMG_EINVAL, "spgemm: inner dimensions differ");
    if (per_pass <= 0) per_pass = SG_PER_PASS;
    if (per_pass > INT_MAX / 2) per_pass = INT_MAX / 2;
    const int n = L.n;
    C.n = n;
    C.m = R.m;
    C.row.alloc((size_t)n + 1);
    if (n == 0) {
        BCMG_HIP(hipMemsetAsync(C.row.get(), 0, sizeof(int), c.stream));
        C.nnz = 0;
        C.col.alloc(1);
        C.val.alloc(1);
        launch_block_nnz(c, C);
        return;
    }
    const long long mcols = R.m > 0 ? R.m : 1;
    DBuf<int> u(n), sl(n), ws(n), cnt(n), large(n);
    DBuf<int> flags(2);   /* long-row count, row with too many products */
    DBuf<unsigned long long> total(1);
    BCMG_HIP(hipMemsetAsync(flags.get(), 0, 2 * sizeof(int), c.stream));
    BCMG_HIP(hipMemsetAsync(total.get(), 0, sizeof(unsigned long long), c.stream));
    k_sg_count<<<blocks_for(n), BS, 0, c.stream>>>(n, L.rp, L.ci, R.rp, mcols, u.get(), sl.get(),
                                                   ws.get(), cnt.get(), large.get(), flags.get(),
                                                   total.get(), flags.get() + 1);
    BCMG_CHECK_LAUNCH();
    unsigned long long h_total = 0;
    int h_flags[2] = {0, 0};
    BCMG_HIP(hipMemcpyAsync(&h_total, total.get(), sizeof(h_total), hipMemcpyDeviceToHost,
                            c.stream));
    BCMG_HIP(hipMemcpyAsync(h_flags, flags.get(), 2 * sizeof(int), hipMemcpyDeviceToHost,
                            c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    const int h_nlarge = h_flags[0];
    if (h_flags[1]) throw Error(BCMG_EUNSUP, "sparse product: a row with 2^30 products or more");
    /* batches of rows [lo, hi) by output slots and, per batch, its long rows
       (local ids) */
    std::vector<int> blo, bhi;
    std::vector<long long> bprod;
    std::vector<std::vector<int>> blarge;
    if ((long long)h_total <= per_pass) {
        blo.push_back(0);
        bhi.push_back(n);
        bprod.push_back((long long)h_total);   /* bounds the slots from above */
    } else {
        std::vector<int> hu(n), hw(n);
        BCMG_HIP(hipMemcpy(hu.data(), sl.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
        BCMG_HIP(hipMemcpy(hw.data(), ws.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
        long long acc = 0, slots = 0;
        int lo = 0;
        for (int i = 0; i < n; i++) {
            const long long cost = (long long)hu[i] + hw[i];
            if (i > lo && acc + cost > per_pass) {
                blo.push_back(lo);
                bhi.push_back(i);
                bprod.push_back(slots);
                lo = i;
                acc = 0;
                slots = 0;
            }
            acc += cost;
            slots += hu[i];
        }
        blo.push_back(lo);
        bhi.push_back(n);
        bprod.push_back(slots);
        std::vector<int> hl(h_nlarge);
        if (h_nlarge > 0)
            BCMG_HIP(hipMemcpy(hl.data(), large.get(), sizeof(int) * h_nlarge,
                               hipMemcpyDeviceToHost));
        std::sort(hl.begin(), hl.end());
        blarge.resize(blo.size());
        size_t q = 0;
        for (size_t b = 0; b < blo.size(); b++)
            while (q < hl.size() && hl[q] < bhi[b]) blarge[b].push_back(hl[q++] - blo[b]);
    }
    const size_t nbatch = blo.size();
    long long maxp = 1, maxrows = 1;
    for (size_t b = 0; b < nbatch; b++) {
        maxp = std::max(maxp, bprod[b]);
        maxrows = std::max<long long>(maxrows, bhi[b] - blo[b]);
    }
    DBuf<int> uoff((size_t)maxrows + 1), soff((size_t)maxrows + 1), brp;
    DBuf<int> tcol((size_t)maxp);
    DBuf<double> tval((size_t)maxp);
    std::vector<DBuf<int>> ocol(nbatch > 1 ? nbatch : 0);
    std::vector<DBuf<double>> oval(nbatch > 1 ? nbatch : 0);
    std::vector<long long> onnz(nbatch, 0);
    if (nbatch > 1) brp.alloc((size_t)maxrows + 1);
    DBuf<double> acc;
    DBuf<int> stamp;
    int dblocks = 0;
    for (size_t b = 0; b < nbatch; b++) {
        const int lo = blo[b], nr = bhi[b] - blo[b];
        const int *Lrp = L.rp + lo;
        exclusive_scan(c, sl.get() + lo, uoff.get(), nr, false);
        const long long S = exclusive_scan(c, ws.get() + lo, soff.get(), nr, true);
        const int nb = (int)((S + SG_CAP - 1) / SG_CAP);
        if (nb > 0)
            k_sg_small<<<nb, BS, 0, c.stream>>>(nr, Lrp, L.ci, L.v, R.rp, R.ci, R.v,
                                                ws.get() + lo, soff.get(), uoff.get(),
                                                tcol.get(), tval.get(), cnt.get() + lo);
        BCMG_CHECK_LAUNCH();
        /* long rows: dense accumulators */
        const int *lrows = large.get();
        int nl = h_nlarge;
        DBuf<int> lbuf;
        if (nbatch > 1) {
            nl = (int)blarge[b].size();
            if (nl > 0) {
                lbuf.alloc(nl);
                BCMG_HIP(hipMemcpyAsync(lbuf.get(), blarge[b].data(), sizeof(int) * nl,
                                        hipMemcpyHostToDevice, c.stream));
            }
            lrows = lbuf.get();
        }
        if (nl > 0) {
            if (!acc.get()) {
                /* dense accumulators: as many workgroups as the budget
                   allows, stamps cleared once for every batch */
                const long long per = 12 * mcols;
                dblocks = (int)std::min<long long>(
                    {(long long)h_nlarge, (long long)SG_DENSE_BLOCKS,
                     std::max<long long>(1, SG_DENSE_BYTES / per)});
                acc.alloc((size_t)dblocks * mcols);
                stamp.alloc((size_t)dblocks * mcols);
                BCMG_HIP(hipMemsetAsync(stamp.get(), 0xff, sizeof(int) * (size_t)dblocks * mcols,
                                        c.stream));
            }
            const int g = std::min(nl, dblocks);
            const int smin = right_unique ? SG_SERIAL_MIN : 0;
            const bool lds = smin > 0 && mcols <= SG_LDS_COLS;
            if (lds)
                k_sg_serial_lds<<<std::min(nl, SG_DENSE_BLOCKS), BS, 0, c.stream>>>(
                    nl, lrows, Lrp, L.ci, L.v, R.rp, R.ci, R.v, u.get() + lo, uoff.get(),
                    tcol.get(), tval.get(), (int)mcols, smin, cnt.get() + lo);
            k_sg_dense<<<g, BS, 0, c.stream>>>(nl, lrows, Lrp, L.ci, L.v, R.rp, R.ci, R.v,
                                               u.get() + lo, uoff.get(), tcol.get(), tval.get(),
                                               acc.get(), stamp.get(), mcols, lo, smin,
                                               lds ? 1 : 0, cnt.get() + lo);
            BCMG_CHECK_LAUNCH();
        }
        if (nbatch == 1) {
            const long long nnz = exclusive_scan(c, cnt.get(), C.row.get(), n, true);
            if (nnz > INT_MAX) throw Error(BCMG_EUNSUP, "sparse product: nnz above 2^31 - 1");
            C.nnz = (int)nnz;
            C.col.alloc(nnz > 0 ? (size_t)nnz : 1);
            C.val.alloc(nnz > 0 ? (size_t)nnz : 1);
            gather_rows(c, n, uoff.get(), C.row.get(), tcol.get(), tval.get(), nnz, C.col.get(),
                        C.val.get());
        } else {
            const long long nnz = exclusive_scan(c, cnt.get() + lo, brp.get(), nr, true);
            onnz[b] = nnz;
            ocol[b].alloc(nnz > 0 ? (size_t)nnz : 1);
            oval[b].alloc(nnz > 0 ? (size_t)nnz : 1);
            gather_rows(c, nr, uoff.get(), brp.get(), tcol.get(), tval.get(), nnz,
                        ocol[b].get(), oval[b].get());
        }
    }
    if (nbatch > 1) {
        const long long nnz = exclusive_scan(c, cnt.get(), C.row.get(), n, true);
        if (nnz > INT_MAX) throw Error(BCMG_EUNSUP, "sparse product: nnz above 2^31 - 1");
        C.nnz = (int)nnz;
        C.col.alloc(nnz > 0 ? (size_t)nnz : 1);
        C.val.alloc(nnz > 0 ? (size_t)nnz : 1);
        long long at = 0;
        for (size_t b = 0; b < nbatch; b++) {
            if (onnz[b] > 0) {
                BCMG_HIP(hipMemcpyAsync(C.col.get() + at, ocol[b].get(), sizeof(int) * onnz[b],
                                        hipMemcpyDeviceToDevice, c.stream));
                BCMG_HIP(hipMemcpyAsync(C.val.get() + at, oval[b].get(),
                                        sizeof(double) * onnz[b], hipMemcpyDeviceToDevice,
                                        c.stream));
            }
            at += onnz[b];
        }
    }
    C.group = lanes_for(n, C.nnz);
    launch_block_nnz(c, C);
    /* the temporaries are released on return, in the order of this stream */
}

void transpose_csr(Ctx &c, const CsrRef &A, DevCsr &T)
{
    const int nr = A.m;
    const size_t nz = A.nnz > 0 ? (size_t)A.nnz : 1;
    T.n = nr;
    T.m = A.n;
    T.nnz = A.nnz;
    T.row.alloc((size_t)nr + 1);
    T.col.alloc(nz);
    T.val.alloc(nz);
    DBuf<int> cnt(nr > 0 ? nr : 1), cursor((size_t)nr + 1), fcol(nz), longl(nr > 0 ? nr : 1);
    DBuf<double> fval(nz);
    DBuf<int> nlong(1);
    BCMG_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int) * (nr > 0 ? nr : 1), c.stream));
    BCMG_HIP(hipMemsetAsync(nlong.get(), 0, sizeof(int), c.stream));
    if (A.n > 0) k_col_count<<<blocks_for(A.n), BS, 0, c.stream>>>(A.n, A.rp, A.ci, cnt.get());
    BCMG_CHECK_LAUNCH();
    exclusive_scan(c, cnt.get(), T.row.get(), nr, false);
    BCMG_HIP(hipMemcpyAsync(cursor.get(), T.row.get(), sizeof(int) * ((size_t)nr + 1),
                            hipMemcpyDeviceToDevice, c.stream));
    if (A.n > 0)
        k_col_fill<<<blocks_for(A.n), BS, 0, c.stream>>>(A.n, A.rp, A.ci, A.v, cursor.get(),
                                                         fcol.get(), fval.get());
    if (nr > 0)
        k_rank_short<<<blocks_for(nr), BS, 0, c.stream>>>(nr, T.row.get(), fcol.get(),
                                                          fval.get(), T.col.get(), T.val.get(),
                                                          longl.get(), nlong.get());
    BCMG_CHECK_LAUNCH();
    int h_long = 0;
    BCMG_HIP(hipMemcpyAsync(&h_long, nlong.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    if (h_long > 0)
        k_rank_long<<<h_long, BS, 0, c.stream>>>(longl.get(), T.row.get(), fcol.get(), fval.get(),
                                                 T.col.get(), T.val.get());
    BCMG_CHECK_LAUNCH();
    T.group = lanes_for(nr, A.nnz);
    launch_block_nnz(c, T);
}

double sa_smoother(Ctx &c, const DevCsr &A, const double *D, DBuf<double> &S)
{
    const int n = A.n;
    S.alloc(A.nnz > 0 ? (size_t)A.nnz : 1);
    DBuf<unsigned long long> nb(1);
    BCMG_HIP(hipMemsetAsync(nb.get(), 0, sizeof(unsigned long long), c.stream));
    k_sa_scale<<<blocks_for(n), BS, 0, c.stream>>>(n, A.row.get(), A.val.get(), D, S.get(),
                                                   nb.get());
    BCMG_CHECK_LAUNCH();
    unsigned long long bits = 0;
    BCMG_HIP(hipMemcpyAsync(&bits, nb.get(), sizeof(bits), hipMemcpyDeviceToHost, c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    double norm;
    std::memcpy(&norm, &bits, sizeof(norm));
    const double omega = 4.0 / (3.0 * norm);
    k_sa_omega<<<blocks_for(n), BS, 0, c.stream>>>(n, A.row.get(), A.col.get(), omega, S.get());
    BCMG_CHECK_LAUNCH();
    return omega;
}

void launch_iota(hipStream_t s, int n, int *a)
{
    if (n <= 0) return;
    k_iota<<<blocks_for(n), BS, 0, s>>>(n, a);
    BCMG_CHECK_LAUNCH();
}

void launch_negate(hipStream_t s, long long n, double *v)
{
    if (n <= 0) return;
    k_negate<<<blocks_for(n), BS, 0, s>>>(n, v);
    BCMG_CHECK_LAUNCH();
}

} // namespace bcmg
