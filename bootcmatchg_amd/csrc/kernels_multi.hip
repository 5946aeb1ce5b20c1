/*
 * kernels_multi.hip - multi-column ("multi-RHS") forms of the solve kernels.
 *
 * A column block holds nc vectors of one level, column j at base + j * ld.
 * Every kernel applies to column j exactly the arithmetic the single-column
 * kernel of kernels_solve.hip applies to one vector (same products, same
 * summation order, same reduction trees over the same block partition), so
 * a right-hand side solved inside a block gets the same bits as solved on its
 * own. What is shared is the traffic: a row's column indices and values are
 * read once for all columns of a chunk.
 *
 * Columns whose flag in `act` is 0 are skipped (act == nullptr: all nc).
 */
#include "kernels_common.h"

namespace bcmg {

namespace {

constexpr int CH = 8; /* columns per register chunk of the row kernels */

__device__ __forceinline__ bool col_on(int j, int nc, const int *act)
{
    return j < nc && (act == nullptr || act[j] != 0);
}

/* The flags of columns [0, nc) as a bit word (bit j: column j active). A
   flag array holds at least MULTI_MAX_COLS readable entries (its allocation
   is padded), so the loads are unconditional and wave-uniform: the scalar
   unit issues them. Kernels whose lanes own different columns test this
   word instead of loading their flags per lane, which added vector memory
   instructions to kernels bound by them (a fifth of the interleaved
   sweep's rate). */
static_assert(MULTI_MAX_COLS == 32, "column flags fit one 32-bit word");
__device__ __forceinline__ unsigned col_word(int nc, const int *__restrict__ act)
{
    const unsigned live = nc >= 32 ? ~0u : (1u << nc) - 1u;
    if (act == nullptr) return live;
    unsigned w = 0;
#pragma unroll
    for (int j = 0; j < 32; j++) w |= (act[j] != 0 ? 1u : 0u) << j;
    return w & live;
}

__device__ __forceinline__ bool word_on(unsigned w, int j)
{
    return j < 32 && ((w >> j) & 1u) != 0;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_nt(T *p, T v)
{
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

/* Stream-order row kernel (the arithmetic of k_rows_staged): one row per
   thread and CH columns per block row (chunk blockIdx.y); p = a_ij * x_j
   rounded, then added left to right from 0.
   OP_SPMV: y = A x; OP_RESID: y = f - A x; OP_JACOBI: y = (f - A x) / d + x.
   x has leading dimension ldx, f and y ldy (restriction maps a level of ldx
   rows onto one of ldy rows). When every column of the chunk is active the
   gathers are unconditional, so a row's loads are all in flight at once. */
template <typename MT, typename VT, int OP>
__global__ __launch_bounds__(BS) void k_rows_multi(int n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const MT *__restrict__ av,
                                                   const VT *__restrict__ x, long ldx,
                                                   const VT *__restrict__ f,
                                                   const VT *__restrict__ d,
                                                   VT *__restrict__ y, long ldy, int nc,
                                                   const int *__restrict__ act)
{
    const int j0 = blockIdx.y * CH;
    bool on[CH];
    bool any = false, all = true;
#pragma unroll
    for (int u = 0; u < CH; u++) {
        on[u] = col_on(j0 + u, nc, act);
        any |= on[u];
        all &= on[u];
    }
    if (!any) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r = b * BS + threadIdx.x;
    if (r >= n) return;
    const int ra = rp[r], re = rp[r + 1];
    VT t[CH];
#pragma unroll
    for (int u = 0; u < CH; u++) t[u] = 0;
    const VT *xj = x + (size_t)j0 * ldx;
    if (all) {
        for (int e = ra; e < re; e++) {
            const int c = ci[e];
            const VT v = mval<VT>(av[e]);
            VT g[CH];
#pragma unroll
            for (int u = 0; u < CH; u++) g[u] = xj[(size_t)u * ldx + c];
#pragma unroll
            for (int u = 0; u < CH; u++) t[u] = t[u] + v * g[u];
        }
    } else {
        for (int e = ra; e < re; e++) {
            const int c = ci[e];
            const VT v = mval<VT>(av[e]);
#pragma unroll
            for (int u = 0; u < CH; u++)
                if (on[u]) t[u] = t[u] + v * xj[(size_t)u * ldx + c];
        }
    }
    VT dr = 1;
    if (OP == OP_JACOBI) dr = d[r];
#pragma unroll
    for (int u = 0; u < CH; u++) {
        if (!on[u]) continue;
        const size_t o = (size_t)(j0 + u) * ldy + r;
        if (OP == OP_SPMV)
            y[o] = t[u];
        else if (OP == OP_RESID)
            y[o] = f[o] - t[u];
        else
            y[o] = (f[o] - t[u]) / dr + x[(size_t)(j0 + u) * ldx + r];
    }
}

/* Group-order row kernel (the arithmetic of k_rows<G>, used for matrices
   whose rows are too long for the stream kernel's LDS tile): G lanes per
   row, lane-strided fma sums, xor-shuffle reduction; one column at a time. */
template <typename MT, typename VT, int OP>
__global__ __launch_bounds__(BS) void k_rows_multi_g(int n, int G, const int *__restrict__ rp,
                                                     const int *__restrict__ ci,
                                                     const MT *__restrict__ av,
                                                     const VT *__restrict__ x, long ldx,
                                                     const VT *__restrict__ f,
                                                     const VT *__restrict__ d,
                                                     VT *__restrict__ y, long ldy, int nc,
                                                     const int *__restrict__ act)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tid = b * BS + threadIdx.x;
    const int row = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= n) return;
    const int j0 = rp[row], j1 = rp[row + 1];
    for (int j = 0; j < nc; j++) {
        if (!col_on(j, nc, act)) continue;
        const VT *xj = x + (size_t)j * ldx;
        VT t = 0;
        for (int e = j0 + lane; e < j1; e += G) t = fma(mval<VT>(av[e]), xj[ci[e]], t);
        for (int o = G >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, G);
        if (lane == 0) {
            const size_t oy = (size_t)j * ldy + row;
            if (OP == OP_SPMV)
                y[oy] = t;
            else if (OP == OP_RESID)
                y[oy] = f[oy] - t;
            else
                y[oy] = (f[oy] - t) / d[row] + xj[row];
        }
    }
}

/* column j = blockIdx.y of the element-wise kernels */
#define MCOL_BEGIN                                                      \
    const int j = blockIdx.y;                                           \
    if (act && !act[j]) return;                                         \
    const int i = blockIdx.x * BS + threadIdx.x;                        \
    if (i >= n) return;

template <typename VT>
__global__ __launch_bounds__(BS) void k_jz_multi(int n, const VT *__restrict__ f,
                                                 const VT *__restrict__ d, VT *__restrict__ y,
                                                 long ld, const int *__restrict__ act)
{
    MCOL_BEGIN
    const size_t o = (size_t)j * ld + i;
    y[o] = f[o] / d[i];
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_prolong_multi(int n, const int *__restrict__ agg,
                                                      const VT *__restrict__ p,
                                                      const VT *__restrict__ xc, long ldc,
                                                      VT *__restrict__ x, long ld,
                                                      const int *__restrict__ act)
{
    MCOL_BEGIN
    const size_t o = (size_t)j * ld + i;
    x[o] = fma(p[i], xc[(size_t)j * ldc + agg[i]], x[o]);
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_fill_multi(int n, VT v, VT *__restrict__ x, long ld,
                                                   const int *__restrict__ act)
{
    MCOL_BEGIN
    x[(size_t)j * ld + i] = v;
}

template <typename S, typename D>
__global__ __launch_bounds__(BS) void k_copy_multi(int n, const S *__restrict__ a, long lda,
                                                   D *__restrict__ b, long ldb,
                                                   const int *__restrict__ act)
{
    MCOL_BEGIN
    b[(size_t)j * ldb + i] = (D)a[(size_t)j * lda + i];
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_axpy_multi(int n, VT alpha, const VT *__restrict__ x,
                                                   VT *__restrict__ y, long ld,
                                                   const int *__restrict__ act)
{
    MCOL_BEGIN
    const size_t o = (size_t)j * ld + i;
    y[o] = fma(alpha, x[o], y[o]);
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_scale_multi(int n, VT alpha, VT *__restrict__ x, long ld,
                                                    const int *__restrict__ act)
{
    MCOL_BEGIN
    const size_t o = (size_t)j * ld + i;
    x[o] = alpha * x[o];
}

/* --------------------------------------------- interleaved column blocks
   The cycle's blocks keep a row's columns together: element (i, j) at
   i * kp + j, kp = 2^lkp >= the number of columns. A gathered row is then
   one contiguous run for all columns, whatever the gather pattern (stencil
   neighbours, an aggregate's scattered fine rows, a coarse level's long
   rows). Threads are (row, column) pairs, column fastest. */

__device__ __forceinline__ bool il_index(long t, int n, int nc, unsigned cw, int &i, int &j)
{
    /* threads cover the nc columns in use of each row (the row stride stays
       2^lkp) */
    i = (int)(t / nc);
    j = (int)(t % nc);
    return i < n && word_on(cw, j);
}

/* W adjacent columns of one row as a native vector (one 16-byte access) */
template <typename VT, int W>
struct VecW {
    typedef VT type __attribute__((ext_vector_type(W)));
};
template <typename VT>
struct VecW<VT, 1> {
    typedef VT type;
};
template <typename VT, int W>
using vec_t = typename VecW<VT, W>::type;

template <typename VT, int W>
__device__ __forceinline__ VT vget(const vec_t<VT, W> &v, int u)
{
    if constexpr (W == 1)
        return v;
    else
        return v[u];
}

/* the W columns [jg W, jg W + W) of a lane: any of them active? all? */
__device__ __forceinline__ void group_on(int jg, int W, unsigned cw, bool &any, bool &all)
{
    any = false;
    all = true;
    for (int u = 0; u < W; u++) {
        const bool on = word_on(cw, jg * W + u);
        any |= on;
        all &= on;
    }
}

/* stream-order row sums (k_rows_staged's arithmetic) on interleaved blocks:
   a lane owns W adjacent columns of one row and gathers them with one
   vector load per entry; per column p = a_ij x_j rounded, added left to
   right from 0. A partly active group computes all W columns (they share
   the gathered cache lines) and stores only the active ones; the columns of
   a group lie inside the block's 2^lkp stride, so the idle ones are
   allocated memory.
   il_row: the sums of lane (row r, column group jg) and the operator's
   epilogue. The row's own operands come first, so that their latency
   overlaps the gathers; then PF entries at a time: their indices and values,
   all their gathers in flight, the sums in entry order. NT: f is read
   non-temporally (a block beyond the last-level cache). */
template <typename MT, typename VT, int OP, int W, int PF, bool NT = false>
__device__ __forceinline__ vec_t<VT, W> il_row(int r, int jg, const int *__restrict__ rp,
                                               const int *__restrict__ ci,
                                               const MT *__restrict__ av,
                                               const VT *__restrict__ x,
                                               const VT *__restrict__ f,
                                               const VT *__restrict__ d, int lkp)
{
    using V = vec_t<VT, W>;
    const int ra = rp[r], re = rp[r + 1];
    V acc = V(0);
    const VT *xg = x + jg * W;
    const size_t o = ((size_t)r << lkp) + jg * W;
    V fo = V(0), xo = V(0);
    VT dr = VT(1);
    if (OP != OP_SPMV) fo = ldm<NT>(reinterpret_cast<const V *>(f + o));
    if (OP == OP_JACOBI) {
        xo = *reinterpret_cast<const V *>(x + o);
        dr = d[r];
    }
    for (int e0 = ra; e0 < re; e0 += PF) {
        int cc[PF];
        VT vv[PF];
        V gg[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            if (e0 + u < re) {
                cc[u] = ci[e0 + u];
                vv[u] = mval<VT>(av[e0 + u]);
            }
        }
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (e0 + u < re) gg[u] = *reinterpret_cast<const V *>(xg + ((size_t)cc[u] << lkp));
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (e0 + u < re) acc = acc + vv[u] * gg[u];
    }
    if (OP == OP_SPMV) return acc;
    if (OP == OP_RESID) return fo - acc;
    return (fo - acc) / dr + xo;
}

/* PF entries of a row are gathered at once. 7 waves per SIMD (at most 72
   VGPRs) hold without spills for every instance launched: PF 8 only with
   FP32 values and vectors, PF 6 otherwise. NT: f and y, each touched once,
   are read and written non-temporally (blocks beyond the last-level cache,
   where the next launch finds none of y there anyway). */
template <typename MT, typename VT, int OP, int W, int PF, bool NT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7))) void k_rows_il(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const MT *__restrict__ av,
    const VT *__restrict__ x, const VT *__restrict__ f, const VT *__restrict__ d,
    VT *__restrict__ y, int lkp, int nc, const int *__restrict__ act)
{
    using V = vec_t<VT, W>;
    const unsigned cw = col_word(nc, act);
    const int gpr = (nc + W - 1) / W; /* lanes per row: the columns in use */
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const long t = (long)b * BS + threadIdx.x;
    const int r = (int)(t / gpr);
    const int jg = (int)(t % gpr);
    if (r >= n) return;
    bool any, all;
    group_on(jg, W, cw, any, all);
    if (!any) return;
    const V res = il_row<MT, VT, OP, W, PF, NT>(r, jg, rp, ci, av, x, f, d, lkp);
    const size_t o = ((size_t)r << lkp) + jg * W;
    if (all) {
        if (NT)
            __builtin_nontemporal_store(res, reinterpret_cast<V *>(y + o));
        else
            *reinterpret_cast<V *>(y + o) = res;
    } else {
        for (int u = 0; u < W; u++)
            if (word_on(cw, jg * W + u)) y[o + u] = vget<VT, W>(res, u);
    }
}

/* A Jacobi/L1 sweep of k_rows_il whose result goes straight to a
   column-major FP64 block (column j at z + j ldz), widened exactly: the
   last sweep of a cycle hands its result to FCG this way instead of writing
   the interleaved block and converting it after. When the lanes per row
   divide the block (TILE), the block's rows are whole: the results pass
   through LDS so that each column's rows leave as one contiguous run;
   otherwise every lane stores its own columns. No thread leaves before the
   barrier. NT: f is read and z written non-temporally. */
template <typename MT, typename VT, int W, int PF, bool TILE, bool NT>
__global__ __launch_bounds__(BS) void k_jacobi_il_c64(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const MT *__restrict__ av,
    const VT *__restrict__ x, const VT *__restrict__ f, const VT *__restrict__ d,
    double *__restrict__ z, long ldz, int lkp, int nc, const int *__restrict__ act)
{
    const unsigned cw = col_word(nc, act);
    const int gpr = (nc + W - 1) / W;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const long t = (long)b * BS + threadIdx.x;
    const int r = (int)(t / gpr);
    const int jg = (int)(t % gpr);
    bool any = false, all = false;
    if (r < n) group_on(jg, W, cw, any, all);
    vec_t<VT, W> res{};
    if (any) res = il_row<MT, VT, OP_JACOBI, W, PF, NT>(r, jg, rp, ci, av, x, f, d, lkp);
    if constexpr (!TILE) {
        if (!any) return;
#pragma unroll
        for (int u = 0; u < W; u++) {
            const int j = jg * W + u;
            if (word_on(cw, j)) st_nt<NT>(z + (size_t)j * ldz + r, (double)vget<VT, W>(res, u));
        }
    } else {
        /* rows [r0, r0 + rb) of the block; column j of local row q at
           tile[j * (rb + 1) + q] (the pad spreads a row's columns over the
           banks) */
        __shared__ double tile[BS * W + MULTI_MAX_COLS];
        const int rb = BS / gpr, pitch = rb + 1;
        const int r0 = (int)((long)b * BS / gpr);
        const int q = threadIdx.x / gpr;
#pragma unroll
        for (int u = 0; u < W; u++) tile[(jg * W + u) * pitch + q] = (double)vget<VT, W>(res, u);
        __syncthreads();
        for (int e = threadIdx.x; e < gpr * W * rb; e += BS) {
            const int j = e / rb, qq = e - j * rb;
            if (r0 + qq < n && word_on(cw, j))
                st_nt<NT>(z + (size_t)j * ldz + r0 + qq, tile[j * pitch + qq]);
        }
    }
}

/* group-order row sums (k_rows<G>'s arithmetic) on interleaved blocks: G
   lanes per row, one column at a time */
template <typename MT, typename VT, int OP>
__global__ __launch_bounds__(BS) void k_rows_il_g(int n, int G, const int *__restrict__ rp,
                                                  const int *__restrict__ ci,
                                                  const MT *__restrict__ av,
                                                  const VT *__restrict__ x,
                                                  const VT *__restrict__ f,
                                                  const VT *__restrict__ d, VT *__restrict__ y,
                                                  int lkp, int nc, const int *__restrict__ act)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tid = b * BS + threadIdx.x;
    const int row = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= n) return;
    const int j0 = rp[row], j1 = rp[row + 1];
    for (int j = 0; j < nc; j++) {
        if (!col_on(j, nc, act)) continue;
        VT t = 0;
        for (int e = j0 + lane; e < j1; e += G)
            t = fma(mval<VT>(av[e]), x[((size_t)ci[e] << lkp) + j], t);
        for (int o = G >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, G);
        if (lane == 0) {
            const size_t o = ((size_t)row << lkp) + j;
            if (OP == OP_SPMV)
                y[o] = t;
            else if (OP == OP_RESID)
                y[o] = f[o] - t;
            else
                y[o] = (f[o] - t) / d[row] + x[o];
        }
    }
}

#define IL_BEGIN                                                                    \
    int i, j;                                                                       \
    const unsigned cw = col_word(nc, act);                                          \
    if (!il_index((long)blockIdx.x * BS + threadIdx.x, n, nc, cw, i, j)) return;    \
    const size_t o = ((size_t)i << lkp) + j;

/* lane = (row, W adjacent columns) for the vectorised element-wise kernels */
#define ILW_BEGIN                                                                   \
    using V = vec_t<VT, W>;                                                         \
    const unsigned cw = col_word(nc, act);                                          \
    const int gpr = (nc + W - 1) / W;                                               \
    const long t_ = (long)blockIdx.x * BS + threadIdx.x;                            \
    const int i = (int)(t_ / gpr), jg = (int)(t_ % gpr);                            \
    if (i >= n) return;                                                             \
    bool any, all;                                                                  \
    group_on(jg, W, cw, any, all);                                                  \
    if (!any) return;                                                               \
    const size_t o = ((size_t)i << lkp) + jg * W;

#define ILW_STORE(y, res)                                                           \
    if (all) {                                                                      \
        *reinterpret_cast<V *>((y) + o) = (res);                                    \
    } else {                                                                        \
        for (int u = 0; u < W; u++)                                                 \
            if (word_on(cw, jg * W + u)) (y)[o + u] = vget<VT, W>((res), u);        \
    }

template <typename VT, int W>
__global__ __launch_bounds__(BS) void k_jz_il(int n, const VT *__restrict__ f,
                                              const VT *__restrict__ d, VT *__restrict__ y,
                                              int lkp, int nc, const int *__restrict__ act)
{
    ILW_BEGIN
    const V res = *reinterpret_cast<const V *>(f + o) / d[i];
    ILW_STORE(y, res)
}

template <typename VT, int W>
__global__ __launch_bounds__(BS) void k_prolong_il(int n, const int *__restrict__ agg,
                                                   const VT *__restrict__ p,
                                                   const VT *__restrict__ xc,
                                                   VT *__restrict__ x, int lkp, int nc,
                                                   const int *__restrict__ act)
{
    ILW_BEGIN
    const VT pi = p[i];
    const V xcv = *reinterpret_cast<const V *>(xc + ((size_t)agg[i] << lkp) + jg * W);
    const V xo = *reinterpret_cast<const V *>(x + o);
    V res;
    for (int u = 0; u < W; u++) {
        if constexpr (W == 1)
            res = fma(pi, xcv, xo);
        else
            res[u] = fma(pi, xcv[u], xo[u]);
    }
    ILW_STORE(x, res)
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_fill_il(int n, VT v, VT *__restrict__ x, int lkp, int nc,
                                                const int *__restrict__ act)
{
    IL_BEGIN
    x[o] = v;
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_copy_il(int n, const VT *__restrict__ a,
                                                VT *__restrict__ b, int lkp, int nc,
                                                const int *__restrict__ act)
{
    IL_BEGIN
    b[o] = a[o];
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_axpy_il(int n, VT alpha, const VT *__restrict__ x,
                                                VT *__restrict__ y, int lkp, int nc,
                                                const int *__restrict__ act)
{
    IL_BEGIN
    y[o] = fma(alpha, x[o], y[o]);
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_scale_il(int n, VT alpha, VT *__restrict__ x, int lkp,
                                                 int nc, const int *__restrict__ act)
{
    IL_BEGIN
    x[o] = alpha * x[o];
}

/* column-major (column j at a + j * lda) <-> interleaved, with the type
   conversion of k_convert; 64 rows per block through an LDS tile so that
   both sides are read and written along their contiguous direction */
constexpr int TR_ROWS = 64;
template <typename S, typename D, bool TO_IL>
__global__ __launch_bounds__(BS) void k_transpose_il(int n, const S *__restrict__ a,
                                                     D *__restrict__ b, long ldc, int lkp, int nc,
                                                     const int *__restrict__ act)
{
    __shared__ D tile[TR_ROWS * (MULTI_MAX_COLS + 1)];
    const unsigned cw = col_word(nc, act);
    const int i0 = blockIdx.x * TR_ROWS;
    const int w = nc + 1;
    for (int q = threadIdx.x; q < TR_ROWS * nc; q += BS) {
        int ii, j;
        if (TO_IL) {
            /* read column-major along rows */
            j = q / TR_ROWS;
            ii = q % TR_ROWS;
        } else {
            ii = q / nc;
            j = q % nc;
        }
        const int i = i0 + ii;
        if (i >= n || !word_on(cw, j)) continue;
        if (TO_IL)
            tile[ii * w + j] = (D)a[(size_t)j * ldc + i];
        else
            tile[ii * w + j] = (D)a[((size_t)i << lkp) + j];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < TR_ROWS * nc; q += BS) {
        int ii, j;
        if (TO_IL) {
            ii = q / nc;
            j = q % nc;
        } else {
            j = q / TR_ROWS;
            ii = q % TR_ROWS;
        }
        const int i = i0 + ii;
        if (i >= n || !word_on(cw, j)) continue;
        if (TO_IL)
            b[((size_t)i << lkp) + j] = tile[ii * w + j];
        else
            b[(size_t)j * ldc + i] = tile[ii * w + j];
    }
}

/* column-major FP64 -> interleaved D with 16-byte accesses on both sides
   (W = 16 / sizeof(D) columns per store; needs 2^lkp >= W): 128 rows per
   block; a wave reads 128 rows of one column with one 16-byte load per lane
   (two rows), the tile holds the block's rows, and each lane then writes
   the W columns of one row with one 16-byte store (per column when the
   group is partly active or past nc). Inactive columns are neither read nor
   written, as in k_transpose_il. */
constexpr int TRV = 128;
template <typename D>
__global__ __launch_bounds__(BS) void k_to_il_v(int n, const double *__restrict__ a, long lda,
                                                D *__restrict__ b, int lkp, int nc,
                                                const int *__restrict__ act)
{
    constexpr int W = 16 / sizeof(D);
    using V = vec_t<D, W>;
    __shared__ D tile[TRV * (MULTI_MAX_COLS + 1)];
    constexpr int pitch = MULTI_MAX_COLS + 1;
    const unsigned cw = col_word(nc, act);
    const int i0 = blockIdx.x * TRV;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < nc; j += BS / 64) {
        if (!word_on(cw, j)) continue;
        const int i = i0 + 2 * lane;
        const double *col = a + (size_t)j * lda;
        double v0 = 0.0, v1 = 0.0;
        if (i + 1 < n) {
            const double2 v = *reinterpret_cast<const double2 *>(col + i);
            v0 = v.x;
            v1 = v.y;
        } else if (i < n) {
            v0 = col[i];
        }
        tile[(2 * lane) * pitch + j] = (D)v0;
        tile[(2 * lane + 1) * pitch + j] = (D)v1;
    }
    __syncthreads();
    const int lg = lkp - (W == 4 ? 2 : 1); /* log2 of the groups per row */
    for (int e = threadIdx.x; e < (TRV << lg); e += BS) {
        const int ii = e >> lg, g = e & ((1 << lg) - 1);
        const int i = i0 + ii;
        const int j0 = g * W;
        if (i >= n || j0 >= nc) continue;
        bool all = j0 + W <= nc;
        V v;
#pragma unroll
        for (int u = 0; u < W; u++) {
            v[u] = tile[ii * pitch + j0 + u];
            all = all && word_on(cw, j0 + u);
        }
        D *dst = b + ((size_t)i << lkp) + j0;
        if (all) {
            *reinterpret_cast<V *>(dst) = v;
        } else {
#pragma unroll
            for (int u = 0; u < W; u++)
                if (word_on(cw, j0 + u)) dst[u] = v[u];
        }
    }
}

/* --------------------------------------------------------- K-cycle
   The inner iteration of host_amg.hip's kcycle_inner on interleaved blocks:
   column j keeps its scalars at ks + j KS_STRIDE (the single-column layout)
   and gets the single path's arithmetic, reduction order included. */

/* k_kdots<K, VT> per column (column j = blockIdx.y): the same grid-stride
   partition and block tree, partials at part[(j * nb + b) * K + k] */
template <int K, typename VT>
__global__ __launch_bounds__(BS) void k_kdots_il(int n, const VT *__restrict__ a0,
                                                 const VT *__restrict__ a1,
                                                 const VT *__restrict__ a2,
                                                 const VT *__restrict__ b, int lkp, int nc,
                                                 const int *__restrict__ act,
                                                 double *__restrict__ part)
{
    const int j = blockIdx.y;
    if (!col_on(j, nc, act)) return;
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const size_t o = ((size_t)i << lkp) + j;
        const double bi = (double)b[o];
        s[0] = fma((double)a0[o], bi, s[0]);
        if constexpr (K > 1) s[1] = fma((double)a1[o], bi, s[1]);
        if constexpr (K > 2) s[2] = fma((double)a2[o], bi, s[2]);
    }
    block_reduce_store<K>(s, part + ((size_t)j * gridDim.x + blockIdx.x) * K);
}

/* k_reduce_final<K, false> per column (block j) into ks + j KS_STRIDE + at */
template <int K>
__global__ __launch_bounds__(BS) void k_kdots_final_il(int nb, const double *__restrict__ part,
                                                       double *__restrict__ ks, int at, int nc,
                                                       const int *__restrict__ act)
{
    const int j = blockIdx.x;
    if (!col_on(j, nc, act)) return;
    reduce_final_body<K, false>(nb, part + (size_t)j * nb * K, ks + (size_t)j * KS_STRIDE + at);
}

/* alpha per column; a breakdown raises the column's global flag */
__global__ void k_kcycle_alpha_il(double *ks, double *gflag, int nc, const int *__restrict__ act)
{
    const int j = threadIdx.x;
    if (!col_on(j, nc, act)) return;
    kcycle_alpha_apply(ks + (size_t)j * KS_STRIDE, gflag + (size_t)j * KS_STRIDE);
}

__global__ void k_kcycle_coef_il(double *ks, double rtol, int nc, const int *__restrict__ act)
{
    const int j = threadIdx.x;
    if (!col_on(j, nc, act)) return;
    kcycle_coef_apply(ks + (size_t)j * KS_STRIDE, rtol);
}

/* w = -alpha_j v + w */
template <typename VT>
__global__ __launch_bounds__(BS) void k_kcycle_axpy_il(int n, const double *__restrict__ ks,
                                                       const VT *__restrict__ v,
                                                       VT *__restrict__ w, int lkp, int nc,
                                                       const int *__restrict__ act)
{
    IL_BEGIN
    w[o] = fma(-(VT)ks[(size_t)j * KS_STRIDE + 3], v[o], w[o]);
}

/* x = c0_j d0 + c1_j d1 */
template <typename VT>
__global__ __launch_bounds__(BS) void k_kcycle_combine_il(int n, const double *__restrict__ ks,
                                                          const VT *__restrict__ d0,
                                                          const VT *__restrict__ d1,
                                                          VT *__restrict__ x, int lkp, int nc,
                                                          const int *__restrict__ act)
{
    IL_BEGIN
    const double *k = ks + (size_t)j * KS_STRIDE;
    x[o] = fma((VT)k[9], d1[o], fma((VT)k[8], d0[o], VT(0)));
}

/* ------------------------------------------------------------------ FCG */

/* Per-column FCG state. Directions live in two slot pairs (V[s], W[s]); the
   newest direction of column j is in pair par[j], an iteration writes the
   next one into pair par[j] ^ 1. The preconditioner output is the fixed
   block Z. scal holds 8 doubles per column (the single-column layout). */
struct FcgCols {
    double *V0, *V1, *W0, *W1;
    const double *Z;
    double *R, *X;
    const double *B;
    double *Y;
    long ld;
    double *scal;
    int *act, *par, *it;
    const double *tol;
    int maxit;
    /* K-cycle: column j's breakdown flag at kflag[j * KS_STRIDE] (or null) */
    const double *kflag;
};

__device__ __forceinline__ double *slot(double *s0, double *s1, int s, int j, long ld)
{
    return (s ? s1 : s0) + (size_t)j * ld;
}

/* the products of k_dots<K> (fma sums over a grid-stride partition),
   column j = blockIdx.y, written to part[(j * nb + b) * K + k].
   ROLE 0: ||r||^2. ROLE 1 (first step) / 2: (r.z, t.z[, to.z]) with t = A z
   taken from the scratch block Y and stored on the way into the pair the
   step writes (pair 0, or par ^ 1). */
template <int K, int ROLE>
__global__ __launch_bounds__(BS) void k_dots_multi(int n, FcgCols F,
                                                   const int *__restrict__ mask,
                                                   double *__restrict__ part)
{
    const int j = blockIdx.y;
    if (mask && !mask[j]) return;
    const long ld = F.ld;
    const double *a0 = F.R + (size_t)j * ld, *a1 = nullptr, *a2 = nullptr, *bv = a0;
    double *tdst = nullptr;
    if (ROLE != 0) {
        const int p = F.par[j];
        a1 = F.Y + (size_t)j * ld;
        tdst = slot(F.W0, F.W1, ROLE == 1 ? 0 : (p ^ 1), j, ld);
        a2 = slot(F.W0, F.W1, p, j, ld);
        bv = F.Z + (size_t)j * ld;
    }
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double bi = bv[i];
        s[0] = fma(a0[i], bi, s[0]);
        if constexpr (K > 1) {
            const double ti = a1[i];
            tdst[i] = ti;
            s[1] = fma(ti, bi, s[1]);
        }
        if constexpr (K > 2) s[2] = fma(a2[i], bi, s[2]);
    }
    block_reduce_store<K>(s, part + ((size_t)j * gridDim.x + blockIdx.x) * K);
}

/* k_reduce_final<K, false> per column (block j), into scal + 8 j (+ off) */
template <int K>
__global__ __launch_bounds__(BS) void k_reduce_final_multi(int nb, const double *__restrict__ part,
                                                           double *__restrict__ out, int off,
                                                           const int *__restrict__ mask)
{
    const int j = blockIdx.x;
    if (mask && !mask[j]) return;
    reduce_final_body<K, false>(nb, part + (size_t)j * nb * K, out + (size_t)j * 8 + off);
}

/* unfused FCG scalar recurrences per column (thread j) */
__global__ void k_fcg_scalars_multi(double *scal, int nc, int first, const int *__restrict__ mask)
{
    const int j = threadIdx.x;
    if (j >= nc || (mask && !mask[j])) return;
    if (first)
        fcg_first_recurrence(scal + (size_t)j * 8);
    else
        fcg_recurrence(scal + (size_t)j * 8);
}

/* SpMV fused with the FCG products (the arithmetic of k_spmv_dots): t = A z,
   per column the block partials of (r.z, t.z[, to.z]). K = 2: first step, t
   into pair 0; K = 3: t into pair par ^ 1, to from pair par. One row per
   thread, CH columns per block. The grid is one-dimensional: the nch column
   chunks of a row block are consecutive blocks of one XCD, so the matrix
   rows that the first chunk brings into that XCD's L2 serve the others.
   NT: r, the previous direction and t, touched once per launch, are read
   and written non-temporally (the columns' streams exceed the last-level
   cache). */
constexpr int SD_PF = 8; /* entries whose indices are loaded at once */
template <int K, bool NT>
__global__ __launch_bounds__(BS) void k_spmv_dots_multi(int n, const int *__restrict__ rp,
                                                        const int *__restrict__ ci,
                                                        const double *__restrict__ av, FcgCols F,
                                                        int nc, const int *__restrict__ mask,
                                                        double *__restrict__ part)
{
    __shared__ double sh[CH][K][BS / 64];
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int nch = (nc + CH - 1) / CH, nb = (int)(gridDim.x / nch);
    const int b = L / nch;
    const int j0 = (L - b * nch) * CH;
    bool on[CH];
    bool any = false, all = true;
#pragma unroll
    for (int u = 0; u < CH; u++) {
        on[u] = col_on(j0 + u, nc, mask);
        any |= on[u];
        all &= on[u];
    }
    if (!any) return;
    const int r = b * BS + threadIdx.x;
    const bool own = r < n;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const long ld = F.ld;
    int ra = 0, re = 0;
    if (own) {
        ra = rp[r];
        re = rp[r + 1];
    }
    double t[CH];
#pragma unroll
    for (int u = 0; u < CH; u++) t[u] = 0.0;
    const double *zj = F.Z + (size_t)j0 * ld;
    if (all) {
        /* SD_PF entries' indices and values at once, then their gathers */
        for (int e0 = ra; e0 < re; e0 += SD_PF) {
            int cc[SD_PF];
            double vv[SD_PF];
#pragma unroll
            for (int q = 0; q < SD_PF; q++) {
                if (e0 + q < re) {
                    cc[q] = ci[e0 + q];
                    vv[q] = av[e0 + q];
                }
            }
#pragma unroll
            for (int q = 0; q < SD_PF; q++) {
                if (e0 + q >= re) break;
                double g[CH];
#pragma unroll
                for (int u = 0; u < CH; u++) g[u] = zj[(size_t)u * ld + cc[q]];
#pragma unroll
                for (int u = 0; u < CH; u++) t[u] = t[u] + vv[q] * g[u];
            }
        }
    } else {
        for (int e = ra; e < re; e++) {
            const int c = ci[e];
            const double v = av[e];
#pragma unroll
            for (int u = 0; u < CH; u++)
                if (on[u]) t[u] = t[u] + v * zj[(size_t)u * ld + c];
        }
    }
#pragma unroll
    for (int u = 0; u < CH; u++) {
        if (!on[u]) continue;
        const int j = j0 + u;
        double sd[K];
#pragma unroll
        for (int k = 0; k < K; k++) sd[k] = 0.0;
        if (own) {
            const size_t o = (size_t)j * ld + r;
            const int p = F.par[j];
            const double zr = F.Z[o];
            const double rr = ldm<NT>(F.R + o);
            double *tv = slot(F.W0, F.W1, K == 2 ? 0 : (p ^ 1), j, ld);
            if (NT)
                __builtin_nontemporal_store(t[u], tv + r);
            else
                tv[r] = t[u];
            sd[0] = rr * zr;
            sd[1] = t[u] * zr;
            if constexpr (K > 2) sd[K - 1] = ldm<NT>(slot(F.W0, F.W1, p, j, ld) + r) * zr;
        }
        /* the in-wave part of block_reduce_store<K> */
#pragma unroll
        for (int k = 0; k < K; k++) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sd[k] += __shfl_down(sd[k], o, 64);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < K; k++) sh[u][k][wid] = sd[k];
        }
    }
    __syncthreads();
    /* the cross-wave part: thread u sums column j0 + u's wave results */
    const int u = threadIdx.x;
    if (u < CH && col_on(j0 + u, nc, mask)) {
        double *out = part + ((size_t)(j0 + u) * nb + b) * K;
#pragma unroll
        for (int k = 0; k < K; k++) {
            double tt = 0.0;
            for (int w = 0; w < BS / 64; w++) tt += sh[u][k][w];
            out[k] = tt;
        }
    }
}

/* Final pass per column (block j of RED_FINAL_BS threads), the arithmetic
   of k_fcg_final<K>. With COUNT (K = 1 at the end of an iteration) the
   column's stopping test follows: it += 1; the column stays active while
   ||r|| > tol and it < maxit (src/SOLVER/FCG.cu:558-565); its newest
   direction moves to the other pair. */
template <int K, bool COUNT>
__global__ __launch_bounds__(RED_FINAL_BS) void k_fcg_final_multi(int nb,
                                                                 const double *__restrict__ part,
                                                                 FcgCols F,
                                                                 const int *__restrict__ mask)
{
    const int j = blockIdx.x;
    if (mask && !mask[j]) return;
    double res[K];
    if (!fcg_final_sum<K>(nb, part + (size_t)j * nb * K, res)) return;
    double *scal = F.scal + (size_t)j * 8;
    fcg_final_apply<K>(res, scal);
    if constexpr (COUNT) {
        const int it = F.it[j] + 1;
        F.it[j] = it;
        F.par[j] ^= 1;
        if (!(scal[6] > F.tol[j] && it < F.maxit)) F.act[j] = 0;
        if (F.kflag && F.kflag[(size_t)j * KS_STRIDE] != 0.0) {
            /* an inner K-cycle iteration broke down: the column stops with a
               non-finite residual (the single-column solve's error) */
            scal[6] = __builtin_nan("");
            F.act[j] = 0;
        }
    }
}

/* FCG update (the arithmetic of k_fcg_update): z = -theta zo + Z into the
   new pair, x += theta2 z; t = -theta to + t; r += -theta2 t; ||r||^2.
   NT: the columns' ten streams exceed the last-level cache, so they are read
   and written non-temporally. */
template <bool NT>
__global__ __launch_bounds__(BS) void k_fcg_update_multi(int n, FcgCols F,
                                                         double *__restrict__ part)
{
    const int j = blockIdx.y;
    if (!F.act[j]) return;
    const long ld = F.ld;
    const double *scal = F.scal + (size_t)j * 8;
    const double theta = scal[4], theta2 = scal[5];
    const int p = F.par[j], q = p ^ 1;
    const double *zo = slot(F.V0, F.V1, p, j, ld);
    double *zn = slot(F.V0, F.V1, q, j, ld);
    const double *zc = F.Z + (size_t)j * ld;
    const double *to = slot(F.W0, F.W1, p, j, ld);
    double *t = slot(F.W0, F.W1, q, j, ld);
    double *x = F.X + (size_t)j * ld;
    double *r = F.R + (size_t)j * ld;
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double zi = fma(-theta, ldm<NT>(zo + i), ldm<NT>(zc + i));
        const double xi = fma(theta2, zi, ldm<NT>(x + i));
        const double ti = fma(-theta, ldm<NT>(to + i), ldm<NT>(t + i));
        const double ri = fma(-theta2, ti, ldm<NT>(r + i));
        if (NT) {
            __builtin_nontemporal_store(zi, zn + i);
            __builtin_nontemporal_store(xi, x + i);
            __builtin_nontemporal_store(ti, t + i);
            __builtin_nontemporal_store(ri, r + i);
        } else {
            zn[i] = zi;
            x[i] = xi;
            t[i] = ti;
            r[i] = ri;
        }
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + (size_t)j * gridDim.x + blockIdx.x);
}

/* first FCG step (the arithmetic of k_fcg_first): the direction Z moves to
   pair 0, x += theta v, r -= theta w, ||r||^2 */
__global__ __launch_bounds__(BS) void k_fcg_first_multi(int n, FcgCols F,
                                                        const int *__restrict__ mask,
                                                        double *__restrict__ part)
{
    const int j = blockIdx.y;
    if (!mask[j]) return;
    const long ld = F.ld;
    const double theta = F.scal[(size_t)j * 8 + 4];
    const double *v = F.Z + (size_t)j * ld;
    double *vs = F.V0 + (size_t)j * ld;
    const double *w = F.W0 + (size_t)j * ld;
    double *x = F.X + (size_t)j * ld;
    double *r = F.R + (size_t)j * ld;
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double vi = v[i];
        vs[i] = vi;
        x[i] = fma(theta, vi, x[i]);
        const double ri = fma(-theta, w[i], r[i]);
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + (size_t)j * gridDim.x + blockIdx.x);
}

/* r = b - y (y = A x in F.Y) and ||r||^2 partials (k_true_residual) */
__global__ __launch_bounds__(BS) void k_true_residual_multi(int n, FcgCols F,
                                                            const int *__restrict__ mask,
                                                            double *__restrict__ part)
{
    const int j = blockIdx.y;
    if (!mask[j]) return;
    const size_t o = (size_t)j * F.ld;
    const double *b = F.B + o;
    const double *y = F.Y + o;
    double *r = F.R + o;
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double ri = b[i] - y[i];
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + (size_t)j * gridDim.x + blockIdx.x);
}

/* ----------------------------------------------------------- dispatch */

template <typename MT, typename VT, int OP>
void rows_multi_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *x, long ldx,
                      const VT *f, const VT *d, VT *y, long ldy, int nc, const int *act)
{
    const int n = A.n;
    const int *rp = A.row.get(), *ci = A.col.get();
    if (stream_ok<VT>(A)) {
        const dim3 g((unsigned)blocks_for(n), (unsigned)((nc + CH - 1) / CH));
        k_rows_multi<MT, VT, OP><<<g, BS, 0, s>>>(n, rp, ci, av, x, ldx, f, d, y,
                                                              ldy, nc, act);
    } else {
        const int G = A.group;
        k_rows_multi_g<MT, VT, OP><<<blocks_for((long long)n * G), BS, 0, s>>>(
            n, G, rp, ci, av, x, ldx, f, d, y, ldy, nc, act);
    }
}

template <typename MT, typename VT>
void rows_multi_op(hipStream_t s, const DevCsr &A, const MT *av, RowOp op, const VT *x, long ldx,
                   const VT *f, const VT *d, VT *y, long ldy, int nc, const int *act)
{
    if (op == OP_SPMV)
        rows_multi_typed<MT, VT, OP_SPMV>(s, A, av, x, ldx, f, d, y, ldy, nc, act);
    else if (op == OP_RESID)
        rows_multi_typed<MT, VT, OP_RESID>(s, A, av, x, ldx, f, d, y, ldy, nc, act);
    else
        rows_multi_typed<MT, VT, OP_JACOBI>(s, A, av, x, ldx, f, d, y, ldy, nc, act);
}

inline dim3 grid2(int n, int nc) { return dim3((unsigned)blocks_for(n), (unsigned)nc); }

inline FcgCols fcg_cols(const MultiFcg &M)
{
    FcgCols F;
    F.V0 = M.V[0];
    F.V1 = M.V[1];
    F.W0 = M.W[0];
    F.W1 = M.W[1];
    F.Z = M.Z;
    F.R = M.R;
    F.X = M.X;
    F.B = M.B;
    F.Y = M.Y;
    F.ld = M.ld;
    F.scal = M.scal;
    F.act = M.act;
    F.par = M.par;
    F.it = M.it;
    F.tol = M.tol;
    F.maxit = M.maxit;
    F.kflag = M.kflag;
    return F;
}

} // namespace

template <typename MT, typename VT, int OP>
void rows_il_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *x, const VT *f,
                   const VT *d, VT *y, int lkp, int nc, const int *act)
{
    const int n = A.n;
    const int *rp = A.row.get(), *ci = A.col.get();
    if (stream_ok<VT>(A)) {
        /* W columns per lane: a lane's columns stay inside the 2^lkp stride
           (W <= nc <= 2^lkp, both powers of two, or W = 1). Entries in
           flight: 6 keeps 8 waves per SIMD and covers the short rows of 2D
           stencils in one pass; all-FP32 rows of 7 or 8 entries (3D
           stencils) take 8 at once at 7 waves rather than two passes (the
           other types would spill at 8). The sums are the same for any
           choice. */
        constexpr int WM = 16 / sizeof(VT);
        constexpr bool all32 = std::is_same<MT, float>::value && std::is_same<VT, float>::value;
        const bool pf8 = all32 && A.maxrow > 6 && A.maxrow <= 8;
        const bool nt = ((size_t)n << lkp) * sizeof(VT) > NT_BYTES;
#define ROWS_IL(W, PF)                                                                      \
    do {                                                                                    \
        const unsigned g = (unsigned)blocks_for((long long)n * ((nc + (W) - 1) / (W)));     \
        if (nt)                                                                             \
            k_rows_il<MT, VT, OP, W, PF, true><<<g, BS, 0, s>>>(n, rp, ci, av, x, f, d, y,  \
                                                                lkp, nc, act);              \
        else                                                                                \
            k_rows_il<MT, VT, OP, W, PF, false><<<g, BS, 0, s>>>(n, rp, ci, av, x, f, d, y, \
                                                                 lkp, nc, act);             \
    } while (0)
        if (nc >= WM && pf8)
            ROWS_IL(WM, (all32 ? 8 : 6));
        else if (nc >= WM)
            ROWS_IL(WM, 6);
        else if (nc >= 2)
            ROWS_IL(2, 6);
        else
            ROWS_IL(1, 6);
#undef ROWS_IL
    } else {
        const int G = A.group;
        k_rows_il_g<MT, VT, OP><<<blocks_for((long long)n * G), BS, 0, s>>>(
            n, G, rp, ci, av, x, f, d, y, lkp, nc, act);
    }
}

/* the c64 sweep where k_rows_il's vector path runs (no group fallback);
   false: not launched */
template <typename MT, typename VT>
bool jacobi_il_c64_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *x, const VT *f,
                         const VT *d, double *z, long ldz, int lkp, int nc, const int *act)
{
    constexpr int WM = 16 / sizeof(VT);
    if (!stream_ok<VT>(A) || nc < WM) return false;
    constexpr bool all32 = std::is_same<MT, float>::value && std::is_same<VT, float>::value;
    const bool pf8 = all32 && A.maxrow > 6 && A.maxrow <= 8;
    const int n = A.n;
    const unsigned nb = (unsigned)blocks_for((long long)n * ((nc + WM - 1) / WM));
    const int gpr = (nc + WM - 1) / WM;
    const bool tile = (BS % gpr) == 0;
    const bool nt = ((size_t)n << lkp) * sizeof(VT) > NT_BYTES;
#define C64(PF, TILE)                                                                        \
    do {                                                                                     \
        if (nt)                                                                              \
            k_jacobi_il_c64<MT, VT, WM, PF, TILE, true><<<nb, BS, 0, s>>>(                   \
                n, A.row.get(), A.col.get(), av, x, f, d, z, ldz, lkp, nc, act);             \
        else                                                                                 \
            k_jacobi_il_c64<MT, VT, WM, PF, TILE, false><<<nb, BS, 0, s>>>(                  \
                n, A.row.get(), A.col.get(), av, x, f, d, z, ldz, lkp, nc, act);             \
    } while (0)
    if (pf8 && tile)
        C64((all32 ? 8 : 6), true);
    else if (pf8)
        C64((all32 ? 8 : 6), false);
    else if (tile)
        C64(6, true);
    else
        C64(6, false);
#undef C64
    return true;
}

template <typename MT, typename VT>
void rows_il_op(hipStream_t s, const DevCsr &A, const MT *av, RowOp op, const VT *x, const VT *f,
                const VT *d, VT *y, int lkp, int nc, const int *act)
{
    if (op == OP_SPMV)
        rows_il_typed<MT, VT, OP_SPMV>(s, A, av, x, f, d, y, lkp, nc, act);
    else if (op == OP_RESID)
        rows_il_typed<MT, VT, OP_RESID>(s, A, av, x, f, d, y, lkp, nc, act);
    else
        rows_il_typed<MT, VT, OP_JACOBI>(s, A, av, x, f, d, y, lkp, nc, act);
}

inline int il_blocks(int n, int nc) { return blocks_for((long long)n * nc); }

/* ============================================================== host */

template <typename VT>
void launch_rows_multi(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x,
                       long ldx, const VT *f, const VT *d, VT *y, long ldy, int nc,
                       const int *act)
{
    if (A.n <= 0 || nc <= 0) return;
    if (nc > MULTI_MAX_COLS) throw Error(BCMG_EINVAL, "too many columns in one block");
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        rows_multi_op<double, double>(s, A, A.val.get(), op, x, ldx, f, d, y, ldy, nc, act);
    } else {
        if (mt == MT_F32)
            rows_multi_op<float, float>(s, A, A.val32.get(), op, x, ldx, f, d, y, ldy, nc, act);
        else if (mt == MT_BF16)
            rows_multi_op<unsigned short, float>(s, A, A.val16.get(), op, x, ldx, f, d, y, ldy,
                                                 nc, act);
        else
            rows_multi_op<double, float>(s, A, A.val.get(), op, x, ldx, f, d, y, ldy, nc, act);
    }
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_jacobi_zero_multi(hipStream_t s, int n, const VT *f, const VT *d, VT *y, long ld,
                              int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_jz_multi<VT><<<grid2(n, nc), BS, 0, s>>>(n, f, d, y, ld, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_prolong_multi(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc,
                          long ldc, VT *x, long ld, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_prolong_multi<VT><<<grid2(n, nc), BS, 0, s>>>(n, agg, p, xc, ldc, x, ld, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_fill_multi(hipStream_t s, int n, VT v, VT *x, long ld, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_fill_multi<VT><<<grid2(n, nc), BS, 0, s>>>(n, v, x, ld, act);
    BCMG_CHECK_LAUNCH();
}

template <typename S, typename D>
void launch_copy_multi(hipStream_t s, int n, const S *a, long lda, D *b, long ldb, int nc,
                       const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_copy_multi<S, D><<<grid2(n, nc), BS, 0, s>>>(n, a, lda, b, ldb, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_axpy_multi(hipStream_t s, int n, VT alpha, const VT *x, VT *y, long ld, int nc,
                       const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_axpy_multi<VT><<<grid2(n, nc), BS, 0, s>>>(n, alpha, x, y, ld, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_scale_multi(hipStream_t s, int n, VT alpha, VT *x, long ld, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_scale_multi<VT><<<grid2(n, nc), BS, 0, s>>>(n, alpha, x, ld, act);
    BCMG_CHECK_LAUNCH();
}

#define BCMG_INSTANTIATE(VT)                                                                 \
    template void launch_rows_multi<VT>(hipStream_t, const DevCsr &, MatType, RowOp,         \
                                        const VT *, long, const VT *, const VT *, VT *,      \
                                        long, int, const int *);                             \
    template void launch_jacobi_zero_multi<VT>(hipStream_t, int, const VT *, const VT *,     \
                                               VT *, long, int, const int *);                \
    template void launch_prolong_multi<VT>(hipStream_t, int, const int *, const VT *,        \
                                           const VT *, long, VT *, long, int, const int *);  \
    template void launch_fill_multi<VT>(hipStream_t, int, VT, VT *, long, int, const int *); \
    template void launch_axpy_multi<VT>(hipStream_t, int, VT, const VT *, VT *, long, int,   \
                                        const int *);                                        \
    template void launch_scale_multi<VT>(hipStream_t, int, VT, VT *, long, int, const int *);
BCMG_INSTANTIATE(double)
BCMG_INSTANTIATE(float)
#undef BCMG_INSTANTIATE
template void launch_copy_multi<double, double>(hipStream_t, int, const double *, long, double *,
                                                long, int, const int *);
template void launch_copy_multi<float, float>(hipStream_t, int, const float *, long, float *,
                                              long, int, const int *);
template void launch_copy_multi<double, float>(hipStream_t, int, const double *, long, float *,
                                               long, int, const int *);
template void launch_copy_multi<float, double>(hipStream_t, int, const float *, long, double *,
                                               long, int, const int *);

template <typename VT>
void launch_rows_il(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x,
                    const VT *f, const VT *d, VT *y, int lkp, int nc, const int *act)
{
    if (A.n <= 0 || nc <= 0) return;
    if (nc > MULTI_MAX_COLS || nc > (1 << lkp))
        throw Error(BCMG_EINVAL, "interleaved block narrower than its columns");
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        rows_il_op<double, double>(s, A, A.val.get(), op, x, f, d, y, lkp, nc, act);
    } else {
        if (mt == MT_F32)
            rows_il_op<float, float>(s, A, A.val32.get(), op, x, f, d, y, lkp, nc, act);
        else if (mt == MT_BF16)
            rows_il_op<unsigned short, float>(s, A, A.val16.get(), op, x, f, d, y, lkp, nc, act);
        else
            rows_il_op<double, float>(s, A, A.val.get(), op, x, f, d, y, lkp, nc, act);
    }
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
bool launch_jacobi_il_c64(hipStream_t s, const DevCsr &A, MatType mt, const VT *x, const VT *f,
                          const VT *d, double *z, long ldz, int lkp, int nc, const int *act)
{
    if (A.n <= 0 || nc <= 0) return true;
    if (nc > MULTI_MAX_COLS || nc > (1 << lkp))
        throw Error(BCMG_EINVAL, "interleaved block narrower than its columns");
    bool done;
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        done = jacobi_il_c64_typed<double, double>(s, A, A.val.get(), x, f, d, z, ldz, lkp, nc,
                                                   act);
    } else {
        if (mt == MT_F32)
            done = jacobi_il_c64_typed<float, float>(s, A, A.val32.get(), x, f, d, z, ldz, lkp,
                                                     nc, act);
        else if (mt == MT_BF16)
            done = jacobi_il_c64_typed<unsigned short, float>(s, A, A.val16.get(), x, f, d, z,
                                                              ldz, lkp, nc, act);
        else
            done = jacobi_il_c64_typed<double, float>(s, A, A.val.get(), x, f, d, z, ldz, lkp,
                                                      nc, act);
    }
    BCMG_CHECK_LAUNCH();
    return done;
}

/* lanes of W = 16 / sizeof(VT) columns when the block is that wide */
#define IL_VEC_LAUNCH(KERNEL, VT, ...)                                                   \
    do {                                                                                 \
        constexpr int WM = 16 / sizeof(VT);                                              \
        if (nc >= WM)                                                                    \
            KERNEL<VT, WM><<<blocks_for((long long)n * ((nc + WM - 1) / WM)), BS, 0, s>>>(  \
                __VA_ARGS__);                                                            \
        else if (nc >= 2)                                                                \
            KERNEL<VT, 2><<<blocks_for((long long)n * ((nc + 1) / 2)), BS, 0, s>>>(        \
                __VA_ARGS__);                                                            \
        else                                                                             \
            KERNEL<VT, 1><<<blocks_for((long long)n * nc), BS, 0, s>>>(__VA_ARGS__);     \
    } while (0)

template <typename VT>
void launch_jacobi_zero_il(hipStream_t s, int n, const VT *f, const VT *d, VT *y, int lkp,
                           int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    IL_VEC_LAUNCH(k_jz_il, VT, n, f, d, y, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_prolong_il(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc, VT *x,
                       int lkp, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    IL_VEC_LAUNCH(k_prolong_il, VT, n, agg, p, xc, x, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_fill_il(hipStream_t s, int n, VT v, VT *x, int lkp, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_fill_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, v, x, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_copy_il(hipStream_t s, int n, const VT *a, VT *b, int lkp, int nc, const int *act)
{
    if (n <= 0 || nc <= 0 || a == b) return;
    k_copy_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, a, b, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_axpy_il(hipStream_t s, int n, VT alpha, const VT *x, VT *y, int lkp, int nc,
                    const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_axpy_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, alpha, x, y, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_scale_il(hipStream_t s, int n, VT alpha, VT *x, int lkp, int nc, const int *act)
{
    if (n <= 0 || nc <= 0) return;
    k_scale_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, alpha, x, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename S, typename D>
void launch_to_il(hipStream_t s, int n, const S *a, long lda, D *b, int lkp, int nc,
                  const int *act)
{
    if (n <= 0 || nc <= 0) return;
    if (lkp > 5) throw Error(BCMG_EINVAL, "interleaved block wider than 32 columns");
    constexpr int W = 16 / sizeof(D);
    if (std::is_same<S, double>::value && (1 << lkp) >= W && lda % 2 == 0)
        k_to_il_v<D><<<(n + TRV - 1) / TRV, BS, 0, s>>>(n, reinterpret_cast<const double *>(a),
                                                         lda, b, lkp, nc, act);
    else
        k_transpose_il<S, D, true><<<(n + TR_ROWS - 1) / TR_ROWS, BS, 0, s>>>(n, a, b, lda, lkp,
                                                                              nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename S, typename D>
void launch_from_il(hipStream_t s, int n, const S *a, D *b, long ldb, int lkp, int nc,
                    const int *act)
{
    if (n <= 0 || nc <= 0) return;
    if (lkp > 5) throw Error(BCMG_EINVAL, "interleaved block wider than 32 columns");
    k_transpose_il<S, D, false><<<(n + TR_ROWS - 1) / TR_ROWS, BS, 0, s>>>(n, a, b, ldb, lkp,
                                                                           nc, act);
    BCMG_CHECK_LAUNCH();
}

#define BCMG_INSTANTIATE(VT)                                                                 \
    template bool launch_jacobi_il_c64<VT>(hipStream_t, const DevCsr &, MatType, const VT *, \
                                           const VT *, const VT *, double *, long, int, int,  \
                                           const int *);                                      \
    template void launch_rows_il<VT>(hipStream_t, const DevCsr &, MatType, RowOp, const VT *, \
                                     const VT *, const VT *, VT *, int, int, const int *);    \
    template void launch_jacobi_zero_il<VT>(hipStream_t, int, const VT *, const VT *, VT *,  \
                                            int, int, const int *);                          \
    template void launch_prolong_il<VT>(hipStream_t, int, const int *, const VT *, const VT *, \
                                        VT *, int, int, const int *);                        \
    template void launch_fill_il<VT>(hipStream_t, int, VT, VT *, int, int, const int *);     \
    template void launch_copy_il<VT>(hipStream_t, int, const VT *, VT *, int, int,           \
                                     const int *);                                           \
    template void launch_axpy_il<VT>(hipStream_t, int, VT, const VT *, VT *, int, int,       \
                                     const int *);                                           \
    template void launch_scale_il<VT>(hipStream_t, int, VT, VT *, int, int, const int *);
BCMG_INSTANTIATE(double)
BCMG_INSTANTIATE(float)
#undef BCMG_INSTANTIATE
template void launch_to_il<double, double>(hipStream_t, int, const double *, long, double *, int,
                                           int, const int *);
template void launch_to_il<double, float>(hipStream_t, int, const double *, long, float *, int,
                                          int, const int *);
template void launch_from_il<double, double>(hipStream_t, int, const double *, double *, long,
                                             int, int, const int *);
template void launch_from_il<float, double>(hipStream_t, int, const float *, double *, long,
                                            int, int, const int *);

/* ---------------------------------------------------- K-cycle launchers */

template <typename VT>
void launch_kdots_il(hipStream_t s, int n, int k, const VT *a0, const VT *a1, const VT *a2,
                     const VT *b, int lkp, int nc, const int *act, double *part, double *ks,
                     int at)
{
    const int nb = red_blocks(n);
    const dim3 g(nb, nc);
    if (k == 1) {
        k_kdots_il<1, VT><<<g, BS, 0, s>>>(n, a0, a1, a2, b, lkp, nc, act, part);
        k_kdots_final_il<1><<<nc, BS, 0, s>>>(nb, part, ks, at, nc, act);
    } else if (k == 2) {
        k_kdots_il<2, VT><<<g, BS, 0, s>>>(n, a0, a1, a2, b, lkp, nc, act, part);
        k_kdots_final_il<2><<<nc, BS, 0, s>>>(nb, part, ks, at, nc, act);
    } else {
        k_kdots_il<3, VT><<<g, BS, 0, s>>>(n, a0, a1, a2, b, lkp, nc, act, part);
        k_kdots_final_il<3><<<nc, BS, 0, s>>>(nb, part, ks, at, nc, act);
    }
    BCMG_CHECK_LAUNCH();
}

void launch_kcycle_alpha_il(hipStream_t s, double *ks, double *gflag, int nc, const int *act)
{
    k_kcycle_alpha_il<<<1, 64, 0, s>>>(ks, gflag, nc, act);
    BCMG_CHECK_LAUNCH();
}

void launch_kcycle_coef_il(hipStream_t s, double *ks, double rtol, int nc, const int *act)
{
    k_kcycle_coef_il<<<1, 64, 0, s>>>(ks, rtol, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_kcycle_axpy_il(hipStream_t s, int n, const double *ks, const VT *v, VT *w, int lkp,
                           int nc, const int *act)
{
    if (n <= 0) return;
    k_kcycle_axpy_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, ks, v, w, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_kcycle_combine_il(hipStream_t s, int n, const double *ks, const VT *d0, const VT *d1,
                              VT *x, int lkp, int nc, const int *act)
{
    if (n <= 0) return;
    k_kcycle_combine_il<VT><<<il_blocks(n, nc), BS, 0, s>>>(n, ks, d0, d1, x, lkp, nc, act);
    BCMG_CHECK_LAUNCH();
}

#define BCMG_INST_KCYCLE_IL(VT)                                                             \
    template void launch_kdots_il<VT>(hipStream_t, int, int, const VT *, const VT *,          \
                                      const VT *, const VT *, int, int, const int *, double *, \
                                      double *, int);                                         \
    template void launch_kcycle_axpy_il<VT>(hipStream_t, int, const double *, const VT *, VT *, \
                                            int, int, const int *);                           \
    template void launch_kcycle_combine_il<VT>(hipStream_t, int, const double *, const VT *,    \
                                               const VT *, VT *, int, int, const int *);
BCMG_INST_KCYCLE_IL(double)
BCMG_INST_KCYCLE_IL(float)
#undef BCMG_INST_KCYCLE_IL

size_t multi_part_size(int n, int nc)
{
    const size_t nbs = (size_t)((n + BS - 1) / BS);
    const size_t nbr = (size_t)red_blocks(n);
    return std::max(nbs, nbr) * 3 * (size_t)nc + 64;
}

void launch_norm2_multi(Ctx &c, int n, MultiFcg &M, const int *mask, double *part)
{
    /* ||r||^2 of every masked column into scal[8 j + 7] (launch_dots, k = 1) */
    const int nb = red_blocks(n);
    const FcgCols F = fcg_cols(M);
    k_dots_multi<1, 0><<<dim3(nb, M.nc), BS, 0, c.stream>>>(n, F, mask, part);
    k_reduce_final_multi<1><<<M.nc, BS, 0, c.stream>>>(nb, part, M.scal, 7, mask);
    BCMG_CHECK_LAUNCH();
}

void launch_spmv_dots_multi(Ctx &c, const DevCsr &A, int k, MultiFcg &M, const int *mask,
                            double *part)
{
    const int n = A.n;
    const FcgCols F = fcg_cols(M);
    if (!stream_ok<double>(A)) {
        /* long rows: the single-column path's separate SpMV (k_rows<G>
           arithmetic into Y), products (k_dots) and recurrence */
        launch_rows_multi<double>(c.stream, A, MT_F64, OP_SPMV, M.Z, M.ld, nullptr, nullptr,
                                  M.Y, M.ld, M.nc, mask);
        const int nbr = red_blocks(n);
        if (k == 2) {
            k_dots_multi<2, 1><<<dim3(nbr, M.nc), BS, 0, c.stream>>>(n, F, mask, part);
            k_reduce_final_multi<2><<<M.nc, BS, 0, c.stream>>>(nbr, part, M.scal, 0, mask);
        } else {
            k_dots_multi<3, 2><<<dim3(nbr, M.nc), BS, 0, c.stream>>>(n, F, mask, part);
            k_reduce_final_multi<3><<<M.nc, BS, 0, c.stream>>>(nbr, part, M.scal, 0, mask);
        }
        k_fcg_scalars_multi<<<1, 64, 0, c.stream>>>(M.scal, M.nc, k == 2 ? 1 : 0, mask);
        BCMG_CHECK_LAUNCH();
        return;
    }
    const int nb = (n + BS - 1) / BS;
    const int nch = (M.nc + CH - 1) / CH;
    const unsigned grid = (unsigned)nb * (unsigned)nch;
    /* r, the previous direction and t of every column */
    const bool nt = (size_t)n * M.nc * 3 * sizeof(double) > NT_BYTES;
    if (k == 2) {
        if (nt)
            k_spmv_dots_multi<2, true><<<grid, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(),
                                                                  A.val.get(), F, M.nc, mask,
                                                                  part);
        else
            k_spmv_dots_multi<2, false><<<grid, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(),
                                                                   A.val.get(), F, M.nc, mask,
                                                                   part);
        k_fcg_final_multi<2, false><<<M.nc, RED_FINAL_BS, 0, c.stream>>>(nb, part, F, mask);
    } else {
        if (nt)
            k_spmv_dots_multi<3, true><<<grid, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(),
                                                                  A.val.get(), F, M.nc, mask,
                                                                  part);
        else
            k_spmv_dots_multi<3, false><<<grid, BS, 0, c.stream>>>(n, A.row.get(), A.col.get(),
                                                                   A.val.get(), F, M.nc, mask,
                                                                   part);
        k_fcg_final_multi<3, false><<<M.nc, RED_FINAL_BS, 0, c.stream>>>(nb, part, F, mask);
    }
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_update_multi(Ctx &c, int n, MultiFcg &M, double *part)
{
    const int nb = red_blocks(n);
    const FcgCols F = fcg_cols(M);
    if ((size_t)n * M.nc * 10 * sizeof(double) > NT_BYTES)
        k_fcg_update_multi<true><<<dim3(nb, M.nc), BS, 0, c.stream>>>(n, F, part);
    else
        k_fcg_update_multi<false><<<dim3(nb, M.nc), BS, 0, c.stream>>>(n, F, part);
    k_fcg_final_multi<1, true><<<M.nc, RED_FINAL_BS, 0, c.stream>>>(nb, part, F, M.act);
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_first_multi(Ctx &c, int n, MultiFcg &M, const int *mask, double *part)
{
    const int nb = red_blocks(n);
    const FcgCols F = fcg_cols(M);
    k_fcg_first_multi<<<dim3(nb, M.nc), BS, 0, c.stream>>>(n, F, mask, part);
    k_fcg_final_multi<1, false><<<M.nc, RED_FINAL_BS, 0, c.stream>>>(nb, part, F, mask);
    BCMG_CHECK_LAUNCH();
}

void launch_true_residual_multi(Ctx &c, const DevCsr &A, MultiFcg &M, const int *mask,
                                double *part)
{
    const int n = A.n;
    launch_rows_multi<double>(c.stream, A, MT_F64, OP_SPMV, M.X, M.ld, nullptr, nullptr, M.Y,
                              M.ld, M.nc, mask);
    const int nb = red_blocks(n);
    const FcgCols F = fcg_cols(M);
    k_true_residual_multi<<<dim3(nb, M.nc), BS, 0, c.stream>>>(n, F, mask, part);
    k_fcg_final_multi<1, false><<<M.nc, RED_FINAL_BS, 0, c.stream>>>(nb, part, F, mask);
    BCMG_CHECK_LAUNCH();
}

} // namespace bcmg
