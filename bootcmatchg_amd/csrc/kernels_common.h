/*
 * kernels_common.h - device helpers shared by the single- and multi-column
 * solve kernels (one copy per translation unit).
 */
#pragma once
#include "bcmg_internal.h"

namespace bcmg {

namespace {

constexpr int BS = 256;          /* threads per block */
constexpr int RED_MAX_BLOCKS = 1024;
/* matrices whose value+column stream exceeds this are read non-temporally */
constexpr size_t NT_BYTES = (size_t)192 << 20;

/* Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
   Return the logical block index so that XCD x owns a contiguous range. */
__device__ __forceinline__ int xcd_block(int bid, int nb)
{
    const int q = nb >> 3, rm = nb & 7;
    const int x = bid & 7, l = bid >> 3;
    return x * q + (x < rm ? x : rm) + l;
}

/* stored matrix value -> compute type (bfloat16: upper half of a float) */
template <typename VT>
__device__ __forceinline__ VT mval(double v) { return (VT)v; }
template <typename VT>
__device__ __forceinline__ VT mval(float v) { return (VT)v; }
template <typename VT>
__device__ __forceinline__ VT mval(unsigned short h) { return (VT)__uint_as_float((unsigned)h << 16); }

template <bool NT, typename T>
__device__ __forceinline__ T ldm(const T *p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

/* Staged LDS layout: element e of a block's source range [e0, e1) sits at
   LDS index e - stage_base<T>(e0) (the range start rounded down to a 16-byte
   chunk), so that every whole 16-byte chunk of the source lands on a 16-byte
   LDS boundary. */
template <typename T>
__host__ __device__ __forceinline__ int stage_base(int e0)
{
    return e0 & ~(int)(16 / sizeof(T) - 1);
}

/* The part of [e0, e0 + cnt) made of whole 16-byte chunks, [a, b); the
   partial chunks [e0, a) and [b, e1) hold fewer than 16 / sizeof(T)
   elements each (thread t < V moves element t of each). */
template <typename T>
struct Chunks {
    static constexpr int V = 16 / sizeof(T);
    int a0, a, b, e0, e1;
    __device__ __forceinline__ Chunks(int s, int cnt) : e0(s), e1(s + cnt)
    {
        a0 = stage_base<T>(s);
        a = (s + V - 1) & ~(V - 1);
        b = e1 & ~(V - 1);
        if (a > b) a = b = e1;  /* no whole chunk: all of it element-wise */
    }
    __device__ __forceinline__ int count() const { return (b - a) / V; }
    __device__ __forceinline__ bool head() const { return (int)threadIdx.x < a - e0; }
    __device__ __forceinline__ bool tail() const { return (int)threadIdx.x < e1 - b; }
};

/* whether the staged kernels copy in 16-byte chunks, else element-wise: the
   chunks pay on streamed (non-temporal) operators and with 2-byte columns;
   a cache-resident operator with int columns is latency-bound and the
   element-wise copy issues less work there */
template <bool NT, typename IT>
constexpr bool stage_vec() { return NT || std::is_same<IT, short>::value; }

/* Phase 1 of the staged row-block kernels: the block's entries
   [e0, e0 + cnt) of col/val are copied into LDS (lcol, lval; layout above).
   The whole 16-byte chunks move as 16-byte loads and LDS stores, the column
   and value chunks of the same entries issued together (UV value chunks and
   the matching column chunks in flight per lane; non-temporal when NT); the
   partial chunks at either end go element by element, loaded before the
   chunks and stored after them so that the block waits on one round trip.
   Every access stays inside [e0, e0 + cnt). The sources must be 16-byte
   aligned. IT is the stored column type: int (the column) or short (column
   minus row, the narrow copy of a banded operator). */
template <typename MT, bool NT, typename IT = int>
__device__ __forceinline__ void stage_entries(const IT *__restrict__ ci,
                                              const MT *__restrict__ av, int e0, int cnt,
                                              IT *__restrict__ lcol, MT *__restrict__ lval)
{
    const Chunks<IT> sc(e0, cnt);
    const Chunks<MT> sv(e0, cnt);
    const int t = threadIdx.x;
    if constexpr (!stage_vec<NT, IT>()) {
        /* element-wise, consecutive lanes on consecutive entries, U in flight */
        constexpr int U = 8;
        IT *lc = lcol + (e0 - sc.a0);
        MT *lv = lval + (e0 - sv.a0);
        for (int k0 = t; k0 < cnt; k0 += BS * U) {
            IT c[U];
            MT v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = k0 + u * BS;
                if (k < cnt) {
                    c[u] = ldm<NT>(ci + e0 + k);
                    v[u] = ldm<NT>(av + e0 + k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = k0 + u * BS;
                if (k < cnt) {
                    lc[k] = c[u];
                    lv[k] = v[u];
                }
            }
        }
        return;
    }
    constexpr int UV = 4;
    constexpr int UC = UV * Chunks<MT>::V / Chunks<IT>::V;
    static_assert(UC >= 1, "column chunks per value chunk");
    IT ch = 0, ct = 0;
    MT vh = 0, vt = 0;
    if (sc.head()) ch = ldm<NT>(ci + e0 + t);
    if (sc.tail()) ct = ldm<NT>(ci + sc.b + t);
    if (sv.head()) vh = ldm<NT>(av + e0 + t);
    if (sv.tail()) vt = ldm<NT>(av + sv.b + t);
    const int nc = sc.count(), nv = sv.count();
    const u32x4 *cs = reinterpret_cast<const u32x4 *>(ci + sc.a);
    const u32x4 *vs = reinterpret_cast<const u32x4 *>(av + sv.a);
    u32x4 *cd = reinterpret_cast<u32x4 *>(lcol + (sc.a - sc.a0));
    u32x4 *vd = reinterpret_cast<u32x4 *>(lval + (sv.a - sv.a0));
    for (int j = 0; j * (BS * UV) < nv || j * (BS * UC) < nc; j++) {
        const int qv = j * (BS * UV) + t;
        const int qc = j * (BS * UC) + t;
        u32x4 c[UC], v[UV];
#pragma unroll
        for (int u = 0; u < UC; u++)
            if (qc + u * BS < nc) c[u] = ldm<NT>(cs + qc + u * BS);
#pragma unroll
        for (int u = 0; u < UV; u++)
            if (qv + u * BS < nv) v[u] = ldm<NT>(vs + qv + u * BS);
#pragma unroll
        for (int u = 0; u < UC; u++)
            if (qc + u * BS < nc) cd[qc + u * BS] = c[u];
#pragma unroll
        for (int u = 0; u < UV; u++)
            if (qv + u * BS < nv) vd[qv + u * BS] = v[u];
    }
    if (sc.head()) lcol[e0 - sc.a0 + t] = ch;
    if (sc.tail()) lcol[sc.b - sc.a0 + t] = ct;
    if (sv.head()) lval[e0 - sv.a0 + t] = vh;
    if (sv.tail()) lval[sv.b - sv.a0 + t] = vt;
}

/* Phase 2: one row per thread; the row's len staged entries (columns from
   LDS index kc, values from kv) gather x (consecutive rows gather
   neighbouring addresses, so the gathers coalesce across the wave), 8
   entries in flight, and p = a_ij x_j (rounded) is added left to right from
   0 - the serial order. With short columns the stored value is the offset
   from the thread's row r. FZ gathers f_j / d_j in place of x_j (the
   iterate of one sweep from zero, rounded as that sweep rounds it). */
template <typename MT, typename VT, typename IT = int, bool FZ = false>
__device__ __forceinline__ VT staged_row_sum(const IT *__restrict__ lcol,
                                             const MT *__restrict__ lval,
                                             const VT *__restrict__ x, int kc, int kv, int len,
                                             int r = 0, const VT *__restrict__ zf = nullptr,
                                             const VT *__restrict__ zd = nullptr)
{
    VT t = 0;
    constexpr int PF = 8;
    for (int k0 = 0; k0 < len; k0 += PF) {
        const int kn = len - k0;
        int c[PF];
        VT v[PF], g[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            if (u < kn) {
                if constexpr (std::is_same<IT, short>::value)
                    c[u] = r + (int)lcol[kc + k0 + u];
                else
                    c[u] = lcol[kc + k0 + u];
                v[u] = mval<VT>(lval[kv + k0 + u]);
            }
        }
        if constexpr (FZ) {
            VT gd[PF];
#pragma unroll
            for (int u = 0; u < PF; u++)
                if (u < kn) {
                    g[u] = zf[c[u]];
                    gd[u] = zd[c[u]];
                }
#pragma unroll
            for (int u = 0; u < PF; u++)
                if (u < kn) g[u] = g[u] / gd[u];
        } else {
#pragma unroll
            for (int u = 0; u < PF; u++)
                if (u < kn) g[u] = x[c[u]];
        }
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (u < kn) t = t + v[u] * g[u];
    }
    return t;
}

/* ------------------------------------------------ stencil-pattern rows */

/* the ascending column offsets of a pattern operator (kernel argument) */
struct PatOffsets {
    int d[PAT_MAX];
};

/* Phase 1 of the pattern kernels: the block's values [e0, e0 + cnt) into
   LDS at lval + (e - stage_base(e0)), the copy of stage_entries without the
   columns (16-byte chunks, UV in flight per lane, partial chunks element by
   element, loaded first and stored last). */
template <typename MT, bool NT>
__device__ __forceinline__ void stage_values(const MT *__restrict__ av, int e0, int cnt,
                                             MT *__restrict__ lval)
{
    const Chunks<MT> sv(e0, cnt);
    const int t = threadIdx.x;
    constexpr int UV = 4;
    MT vh = 0, vt = 0;
    if (sv.head()) vh = ldm<NT>(av + e0 + t);
    if (sv.tail()) vt = ldm<NT>(av + sv.b + t);
    const int nv = sv.count();
    const u32x4 *vs = reinterpret_cast<const u32x4 *>(av + sv.a);
    u32x4 *vd = reinterpret_cast<u32x4 *>(lval + (sv.a - sv.a0));
    for (int j = 0; j * (BS * UV) < nv; j++) {
        const int qv = j * (BS * UV) + t;
        u32x4 v[UV];
#pragma unroll
        for (int u = 0; u < UV; u++)
            if (qv + u * BS < nv) v[u] = ldm<NT>(vs + qv + u * BS);
#pragma unroll
        for (int u = 0; u < UV; u++)
            if (qv + u * BS < nv) vd[qv + u * BS] = v[u];
    }
    if (sv.head()) lval[e0 - sv.a0 + t] = vh;
    if (sv.tail()) lval[sv.b - sv.a0 + t] = vt;
}

/* The masks of a pattern block (BS = 256 rows) without a barrier: lane l of
   every wave holds word l of the block's masks (rows 4l .. 4l + 3, read
   before the value stream is issued, so that its latency overlaps it) and
   the waves scan the words' bit counts across their lanes; thread t takes
   its row's mask and the number of entries of the block's rows before it.
   The mask array is padded with zero masks to a whole number of blocks. */
__device__ __forceinline__ void pattern_block_rows(unsigned w, unsigned &m, int &pre)
{
    const int t = threadIdx.x, lane = t & 63;
    int s = __popc(w);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(s, o, 64);
        if (lane >= o) s += u;
    }
    const int g = t >> 2;
    const unsigned wg = __shfl(w, g, 64);
    const int sg = __shfl(s, g, 64);
    const int sh = 8 * (t & 3);
    m = (wg >> sh) & 0xFFu;
    pre = sg - __popc(wg) + __popc(wg & ((1u << sh) - 1u));
}

/* Phase 2 of the pattern kernels: row r's entries are the set bits k of its
   mask m in ascending order, column r + po.d[k], value lv[j] for the j-th
   set bit; all gathers go out at once and p = a_ij x_j (rounded) is added in
   that order - the CSR order, so the sum is staged_row_sum's bit for bit
   (FZ as there). */
template <typename MT, typename VT, bool FZ = false>
__device__ __forceinline__ VT pattern_row_sum(unsigned m, const PatOffsets &po,
                                              const MT *__restrict__ lv,
                                              const VT *__restrict__ x, int r,
                                              const VT *__restrict__ zf = nullptr,
                                              const VT *__restrict__ zd = nullptr)
{
    VT g[PAT_MAX], v[PAT_MAX];
#pragma unroll
    for (int k = 0; k < PAT_MAX; k++) {
        if ((m >> k) & 1u) {
            v[k] = mval<VT>(lv[__popc(m & ((1u << k) - 1u))]);
            const int c = r + po.d[k];
            if constexpr (FZ)
                g[k] = zf[c];
            else
                g[k] = x[c];
        }
    }
    if constexpr (FZ) {
        VT gd[PAT_MAX];
#pragma unroll
        for (int k = 0; k < PAT_MAX; k++)
            if ((m >> k) & 1u) gd[k] = zd[r + po.d[k]];
#pragma unroll
        for (int k = 0; k < PAT_MAX; k++)
            if ((m >> k) & 1u) g[k] = g[k] / gd[k];
    }
    VT t = 0;
#pragma unroll
    for (int k = 0; k < PAT_MAX; k++)
        if ((m >> k) & 1u) t = t + v[k] * g[k];
    return t;
}

/* LDS of the pattern kernels: the tile's values (room for the leading
   partial chunk) */
inline size_t pattern_lds_bytes(int tile, size_t mt_size)
{
    return ((size_t)tile + 16 / mt_size) * mt_size;
}

/* LDS of the staged row kernels: the tile's columns (room for the leading
   partial chunk, rounded to 16 bytes), then its values */
__host__ __device__ inline int staged_col_bytes(int tile, int it_size)
{
    return ((tile + 16 / it_size) * it_size + 15) & ~15;
}

inline size_t staged_lds_bytes(int tile, size_t mt_size, size_t it_size = 4)
{
    return (size_t)staged_col_bytes(tile, (int)it_size) +
           ((size_t)tile + 16 / mt_size) * mt_size;
}

template <typename MT, typename IT = int>
__device__ __forceinline__ void staged_lds(int tile, IT *&lcol, MT *&lval)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_staged[];
    lcol = reinterpret_cast<IT *>(smem_staged);
    lval = reinterpret_cast<MT *>(smem_staged + staged_col_bytes(tile, (int)sizeof(IT)));
}

/* the staged kernels read col/val in 16-byte chunks */
inline bool stage_aligned(const void *a, const void *b)
{
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

constexpr int RED_FINAL_BS = 1024; /* threads of the one-block final reductions */

/* K-cycle inner iteration scalars of one level (ks, KS_STRIDE doubles):
   [0] w.w  [1] w.d0  [2] d0.v  [3] alpha  [4] w'.w'  [5] d1.v  [6] d1.v1
   [7] d1.w'  [8] c0  [9] c1, [KS_FLAG] the level's breakdown flag */
constexpr int KS_FLAG = 15;

/* alpha = delta_old / tau; a non-positive delta_old (indefinite
   preconditioner) or tau (indefinite matrix) raises the flags - the
   reference exits there (src/SOLVER/FCG.cu:625-628, 640-643) */
__device__ __forceinline__ void kcycle_alpha_apply(double *ks, double *gflag)
{
    if (!(ks[1] > 0.0)) ks[KS_FLAG] = *gflag = 1.0;
    if (!(ks[2] > 0.0)) ks[KS_FLAG] = *gflag = 2.0;
    ks[3] = ks[1] / ks[2];
}

/* the combination coefficients of the inner iteration: one step when
   ||w'|| <= rtol ||w|| (c1 = 0), else the two-step update
   (src/SOLVER/FCG.cu:650-723) */
__device__ __forceinline__ void kcycle_coef_apply(double *ks, double rtol)
{
    const double delta0 = sqrt(ks[0]), l2 = sqrt(ks[4]);
    const double alpha = ks[3];
    if (l2 <= rtol * delta0) {
        ks[8] = alpha;
        ks[9] = 0.0;
        return;
    }
    const double tau = ks[2], tau1 = ks[5], tau2 = ks[6], tau3 = ks[7];
    const double tau4 = tau2 - (tau1 * tau1) / tau;
    ks[8] = alpha - (tau1 * tau3) / (tau * tau4);
    ks[9] = tau3 / tau4;
}

/* fixed-order block reduction of K partial sums; thread 0 writes out[0..K) */
template <int K>
__device__ __forceinline__ void block_reduce_store(double (&s)[K], double *out)
{
    __shared__ double sh[K][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_down(s[k], o, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) sh[k][wid] = s[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double t = 0.0;
            for (int w = 0; w < BS / 64; w++) t += sh[k][w];
            out[k] = t;
        }
    }
}

/* final pass of the plain reductions (one block of BS threads):
   out[k] = sum_b part[b*K + k], optionally square-rooted */
template <int K, bool SQRT>
__device__ __forceinline__ void reduce_final_body(int nb, const double *__restrict__ part,
                                                  double *__restrict__ out)
{
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += BS) {
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += part[(size_t)b * K + k];
    }
    __shared__ double res[K];
    block_reduce_store<K>(s, res);
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = SQRT ? sqrt(res[k]) : res[k];
    }
}

/* FCG v3 scalar recurrence (src/SOLVER/FCG.cu:531-534) on
   s = [alpha, beta, gamma, delta, theta, theta2, ...]:
   theta = gamma / delta; delta = beta - gamma^2 / delta; theta2 = alpha / delta */
__device__ __forceinline__ void fcg_recurrence(double *scal)
{
    const double alpha = scal[0], beta = scal[1], gamma = scal[2];
    double delta = scal[3];
    const double theta = gamma / delta;
    delta = beta - (gamma * gamma) / delta;
    scal[3] = delta;
    scal[4] = theta;
    scal[5] = alpha / delta;
}

/* first step (src/SOLVER/FCG.cu:475-478): delta = beta, theta = alpha / delta */
__device__ __forceinline__ void fcg_first_recurrence(double *scal)
{
    scal[3] = scal[1];
    scal[4] = scal[0] / scal[1];
}

/* Final pass of the FCG reductions (one block of RED_FINAL_BS threads): sums
   the per-block partials part[b*K + k] in a fixed order into res (valid in
   thread 0, which alone returns true). */
template <int K>
__device__ __forceinline__ bool fcg_final_sum(int nb, const double *__restrict__ part,
                                              double (&res)[K])
{
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_FINAL_BS) {
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += part[(size_t)b * K + k];
    }
    __shared__ double sh[K][RED_FINAL_BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_down(s[k], o, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) sh[k][wid] = s[k];
    }
    __syncthreads();
    if (threadIdx.x != 0) return false;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double t = 0.0;
        for (int w = 0; w < RED_FINAL_BS / 64; w++) t += sh[k][w];
        res[k] = t;
    }
    return true;
}

/* what the final pass stores:
   K = 1: scal[6] = sqrt(sum) (the residual norm);
   K = 2 (first step, src/SOLVER/FCG.cu:475-478): alpha, beta, delta = beta,
          theta = alpha / delta;
   K = 3 (src/SOLVER/FCG.cu:531-534): alpha, beta, gamma and the recurrence */
template <int K>
__device__ __forceinline__ void fcg_final_apply(const double (&res)[K], double *scal)
{
    if constexpr (K == 1) {
        scal[6] = sqrt(res[0]);
    } else if constexpr (K == 2) {
        scal[0] = res[0];
        scal[1] = res[1];
        scal[3] = res[1];
        scal[4] = res[0] / res[1];
    } else {
        scal[0] = res[0];
        scal[1] = res[1];
        scal[2] = res[2];
        fcg_recurrence(scal);
    }
}

inline int blocks_for(long long work) { return (int)((work + BS - 1) / BS); }

inline int red_blocks(int n)
{
    int nb = (int)(((long long)n + BS * 4 - 1) / (BS * 4));
    if (nb < 1) nb = 1;
    if (nb > RED_MAX_BLOCKS) nb = RED_MAX_BLOCKS;
    return nb;
}

/* rows that fit the LDS tile of the staged row kernel (48 KB for values of
   the vector type): these get the stream-order sum, others the group-order
   kernel; single- and multi-column paths decide alike */
template <typename VT>
inline bool stream_ok(const DevCsr &A)
{
    return A.tile > 0 && staged_lds_bytes(A.tile, sizeof(VT)) <= (size_t)48 * 1024;
}

/* All `sweeps` Jacobi sweeps of a small level in one workgroup of 1024
   threads: thread t owns rows t + 1024 q (q < RPT) and keeps their entries
   (at most MAXNZ per row) in registers for all sweeps; both iterates live in
   LDS and a barrier separates the sweeps. Row sums are sequential in column
   order (the order of k_rows_staged), so the result equals `sweeps` separate
   launches bit for bit. */
template <typename MT, typename VT, int RPT, int MAXNZ, bool PRO = false, bool RR = false>
__global__ __launch_bounds__(1024) void k_relax_block(int n, const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const MT *__restrict__ av,
                                                      const VT *__restrict__ d,
                                                      const VT *__restrict__ f,
                                                      const VT *__restrict__ xin,
                                                      VT *__restrict__ xout, int sweeps,
                                                      int zero, long cs, int rs,
                                                      const int *__restrict__ act,
                                                      BlockFuse<VT> fu)
{
    /* PRO: x starts as x_in + P x_c (k_prolong's fma; x_in = 0 when zero)
       and every sweep runs; RR: after the sweeps the residual rows go to the
       free LDS iterate and the restriction rows of R sum them (the residual
       and SpMV kernels' orders) into fu.fc. Single column only. */
    /* block j works on column j of (f, xin, xout): element i of column j at
       j * cs + i * rs; columns whose act flag is 0 are skipped (act ==
       nullptr: one column) */
    if (act && !act[blockIdx.x]) return;
    f += (size_t)blockIdx.x * cs;
    xin += (size_t)blockIdx.x * cs;
    xout += (size_t)blockIdx.x * cs;
    extern __shared__ double smem[];
    VT *xs = reinterpret_cast<VT *>(smem);
    VT *cur = xs, *nxt = xs + n;
    VT a[RPT][MAXNZ];
    int cj[RPT][MAXNZ];
    int len[RPT];
    VT fi[RPT], di[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + 1024 * q;
        len[q] = 0;
        fi[q] = 0;
        di[q] = 1;
        if (i < n) {
            const int j0 = rp[i];
            len[q] = rp[i + 1] - j0;
            fi[q] = f[(size_t)i * rs];
            di[q] = d[i];
#pragma unroll
            for (int k = 0; k < MAXNZ; k++) {
                a[q][k] = k < len[q] ? mval<VT>(av[j0 + k]) : VT(0);
                cj[q][k] = k < len[q] ? ci[j0 + k] : 0;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + 1024 * q;
        if (i >= n) continue;
        if constexpr (PRO)
            cur[i] = fma(fu.p[i], fu.xc[fu.agg[i]], zero ? VT(0) : xin[(size_t)i * rs]);
        else
            cur[i] = zero ? fi[q] / di[q] : xin[(size_t)i * rs];
    }
    __syncthreads();
    for (int s = (zero && !PRO) ? 1 : 0; s < sweeps; s++) {
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = threadIdx.x + 1024 * q;
            if (i < n) {
                VT t = 0;
#pragma unroll
                for (int k = 0; k < MAXNZ; k++)
                    if (k < len[q]) t = t + a[q][k] * cur[cj[q][k]];
                nxt[i] = (fi[q] - t) / di[q] + cur[i];
            }
        }
        __syncthreads();
        VT *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = threadIdx.x + 1024 * q;
        if (i < n) xout[(size_t)i * rs] = cur[i];
    }
    if constexpr (RR) {
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = threadIdx.x + 1024 * q;
            if (i < n) {
                VT t = 0;
#pragma unroll
                for (int k = 0; k < MAXNZ; k++)
                    if (k < len[q]) t = t + a[q][k] * cur[cj[q][k]];
                nxt[i] = fi[q] - t;
            }
        }
        __syncthreads();
        for (int I = threadIdx.x; I < fu.nc; I += 1024) {
            VT t = 0;
            const int e1 = fu.rrp[I + 1];
            for (int e = fu.rrp[I]; e < e1; e++) t = t + fu.rv[e] * nxt[fu.rci[e]];
            fu.fc[I] = t;
        }
    }
}

/* k_rows_staged's residual of each fine row of a coarse row's R entries,
   restricted as the SpMV with R restricts it: y_I = sum over R's row I of
   R_Ii (f_i - (A x)_i), every sum from 0 in storage order with products
   rounded. G lanes per coarse row: lane g forms the residual of the fine
   rows of R entries g, g + G, ... (8 entries' loads at once), and the
   group's first lane adds the products in R's entry order, fetched through
   shuffles. For small cache-resident levels, where a launch costs more than
   the walk. */
template <typename MT, typename VT, int G>
__global__ __launch_bounds__(BS) void k_resid_restrict(int nc, const int *__restrict__ rrp,
                                                       const int *__restrict__ rci,
                                                       const VT *__restrict__ rv,
                                                       const int *__restrict__ arp,
                                                       const int *__restrict__ aci,
                                                       const MT *__restrict__ av,
                                                       const VT *__restrict__ x,
                                                       const VT *__restrict__ f,
                                                       VT *__restrict__ y)
{
    const int t = blockIdx.x * BS + threadIdx.x;
    const int I = t / G, g = t % G;
    /* whole groups stay in the block (G divides BS), so the shuffles below
       see every lane of the group */
    int e0 = 0, e1 = 0;
    if (I < nc) {
        e0 = rrp[I];
        e1 = rrp[I + 1];
    }
    VT acc = 0;
    for (int b = e0; b < e1; b += G) {
        const int e = b + g;
        VT prod = 0;
        if (e < e1) {
            const int i = rci[e];
            const VT w = rv[e];
            const int j1 = arp[i + 1];
            VT s = 0;
            constexpr int PF = 8;
            for (int j0 = arp[i]; j0 < j1; j0 += PF) {
                int c[PF];
                VT v[PF], gx[PF];
#pragma unroll
                for (int u = 0; u < PF; u++)
                    if (j0 + u < j1) {
                        c[u] = aci[j0 + u];
                        v[u] = mval<VT>(av[j0 + u]);
                    }
#pragma unroll
                for (int u = 0; u < PF; u++)
                    if (j0 + u < j1) gx[u] = x[c[u]];
#pragma unroll
                for (int u = 0; u < PF; u++)
                    if (j0 + u < j1) s = s + v[u] * gx[u];
            }
            prod = w * (f[i] - s);
        }
        /* R's order: products of entries b, b + 1, ... added one by one */
#pragma unroll
        for (int q = 0; q < G; q++) {
            const VT pq = __shfl(prod, (int)(threadIdx.x & 63 & ~(G - 1)) + q, 64);
            if (b + q < e1) acc = acc + pq;
        }
    }
    if (I < nc && g == 0) y[I] = acc;
}

} // namespace

} // namespace bcmg
