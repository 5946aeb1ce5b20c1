/*
 * kernels_solve.hip - solve-phase kernels for gfx950 (wave64, CDNA4).
 *
 * Everything here is HBM-bound CSR/vector streaming; there is no matrix-core
 * work. The row kernels are templated on the stored matrix value type MT
 * (double, float, or bfloat16 kept as its 16-bit pattern) and on the compute
 * and vector type VT (double or float), so the same code serves the FP64
 * drop-in and the reduced-precision preconditioner.
 *
 * Reductions (dot products, norms) are deterministic: a fixed grid that
 * depends only on n, a fixed in-block tree and a fixed final pass; no
 * floating-point atomics (the reference adds block results with atomicAdd,
 * src/SOLVER/FCG.cu:58-60, so its sums vary run to run).
 */
#include "kernels_common.h"

#include <algorithm>

namespace bcmg {

namespace {

/* layout of the context's reduction scratch (doubles) */
constexpr size_t RED_UPD = 3 * RED_MAX_BLOCKS + 16;   /* update/norm partials */
constexpr size_t RED_SPMV = 4 * RED_MAX_BLOCKS + 32;  /* SpMV+dots partials */

/* Row-block kernel: block b owns rows [256 b, 256 b + 256). Phase 1 streams
   the block's slice of col/val into LDS with consecutive lanes on
   consecutive entries (coalesced; non-temporal when the matrix is larger
   than the 256 MB last-level cache, so that the once-read stream does not
   push the gathered x out of the caches, while a smaller matrix stays cached
   for the next sweep); phase 2 gives one row to each thread, which gathers x
   at its columns (coalesced across the wave for banded and stencil
   matrices), sums the products in column order and applies the epilogue. The
   row's own operands (row pointers, f, d, x) are loaded before phase 1 so
   that their latency overlaps the stream. The sum order (products rounded,
   then added left to right) is the serial one. IT = short reads the narrow
   column copy (column minus row): same entries, 2 bytes fewer each. */
template <typename MT, typename VT, int OP, bool NT, typename IT>
__global__ __launch_bounds__(BS) void k_rows_staged(int n, const int *__restrict__ rp,
                                                    const IT *__restrict__ ci,
                                                    const MT *__restrict__ av,
                                                    const VT *__restrict__ x,
                                                    const VT *__restrict__ f,
                                                    const VT *__restrict__ d,
                                                    VT *__restrict__ y, int tile)
{
    IT *lcol;
    MT *lval;
    staged_lds<MT, IT>(tile, lcol, lval);
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * BS;
    const int r1 = min(r0 + BS, n);
    const int r = r0 + threadIdx.x;
    const bool own = r < r1;
    int ra = 0, re = 0;
    VT fr = 0, dr = 1, xr = 0;
    if (own) {
        ra = rp[r];
        re = rp[r + 1];
        if (OP != OP_SPMV) fr = f[r];
        if (OP == OP_JACOBI) {
            dr = d[r];
            xr = x[r];
        } else if (OP == OP_JACOBI_Z) {
            dr = d[r];
            xr = fr / dr;
        }
    }
    const int e0 = rp[r0];
    const int cnt = rp[r1] - e0;
    stage_entries<MT, NT, IT>(ci, av, e0, cnt, lcol, lval);
    __syncthreads();
    if (!own) return;
    const VT t = staged_row_sum<MT, VT, IT, OP == OP_JACOBI_Z>(
        lcol, lval, x, ra - stage_base<IT>(e0), ra - stage_base<MT>(e0), re - ra, r, f, d);
    if (OP == OP_SPMV)
        y[r] = t;
    else if (OP == OP_RESID)
        y[r] = fr - t;
    else
        y[r] = (fr - t) / dr + xr;
}

/* k_rows_staged on a stencil-pattern operator: the block reads its rows'
   one-byte masks, two row pointers (its first value and its end) and its
   values (into LDS); each row's value positions come from a block scan of
   the masks' bit counts and its columns from the offset table. Same
   products in the same order as k_rows_staged, so the same bits. */
template <typename MT, typename VT, int OP, bool NT>
__global__ __launch_bounds__(BS) void k_rows_pattern(int n, const int *__restrict__ rp,
                                                     const unsigned char *__restrict__ pm,
                                                     PatOffsets po, const MT *__restrict__ av,
                                                     const VT *__restrict__ x,
                                                     const VT *__restrict__ f,
                                                     const VT *__restrict__ d,
                                                     VT *__restrict__ y)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_pat[];
    MT *lval = reinterpret_cast<MT *>(smem_pat);
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * BS;
    const int r1 = min(r0 + BS, n);
    const int r = r0 + threadIdx.x;
    const bool own = r < r1;
    const unsigned w = ldm<NT>(reinterpret_cast<const unsigned *>(pm + r0) + (threadIdx.x & 63));
    VT fr = 0, dr = 1, xr = 0;
    if (own) {
        if (OP != OP_SPMV) fr = f[r];
        if (OP == OP_JACOBI) {
            dr = d[r];
            xr = x[r];
        } else if (OP == OP_JACOBI_Z) {
            dr = d[r];
            xr = fr / dr;
        }
    }
    const int e0 = rp[r0];
    const int cnt = rp[r1] - e0;
    stage_values<MT, NT>(av, e0, cnt, lval);
    unsigned m;
    int pre;
    pattern_block_rows(w, m, pre);
    __syncthreads();
    if (!own) return;
    const VT t = pattern_row_sum<MT, VT, OP == OP_JACOBI_Z>(
        m, po, lval + (e0 - stage_base<MT>(e0)) + pre, x, r, f, d);
    if (OP == OP_SPMV)
        y[r] = t;
    else if (OP == OP_RESID)
        y[r] = fr - t;
    else
        y[r] = (fr - t) / dr + xr;
}

/* The first Jacobi/L1 sweep after the prolongation, from x~ = x + P x_c
   formed where it is read (k_prolong's fma at every gather and for the row
   itself) and summed in column order from 0, so y equals a prolongation
   launch followed by a sweep launch bit for bit; x~ is not stored (the
   sweep's result replaces it). One thread per row: for small levels, where
   the launch saved costs more than the dependent loads. */
template <typename MT, typename VT>
__global__ __launch_bounds__(BS) void k_jacobi_prolong(int n, const int *__restrict__ rp,
                                                       const int *__restrict__ ci,
                                                       const MT *__restrict__ av,
                                                       const int *__restrict__ agg,
                                                       const VT *__restrict__ p,
                                                       const VT *__restrict__ xc,
                                                       const VT *__restrict__ x,
                                                       const VT *__restrict__ f,
                                                       const VT *__restrict__ d,
                                                       VT *__restrict__ y)
{
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r >= n) return;
    const VT xr = fma(p[r], xc[agg[r]], x[r]);
    const VT fr = f[r], dr = d[r];
    const int j1 = rp[r + 1];
    VT t = 0;
    constexpr int PF = 8;
    for (int j0 = rp[r]; j0 < j1; j0 += PF) {
        int c[PF];
        VT v[PF], g[PF];
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (j0 + u < j1) {
                c[u] = ci[j0 + u];
                v[u] = mval<VT>(av[j0 + u]);
            }
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (j0 + u < j1) g[u] = fma(p[c[u]], xc[agg[c[u]]], x[c[u]]);
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (j0 + u < j1) t = t + v[u] * g[u];
    }
    y[r] = (fr - t) / dr + xr;
}

/* Setup of the pattern, step 1: the distinct column offsets of each 256-row
   block (an LDS table of 2 PAT_MAX slots) are inserted into the operator's
   table gtab of as many slots (PAT_EMPTY where unused: a slot, once set,
   keeps its value, so no value is inserted twice); more distinct offsets
   than slots, in a block or in all, raise flag[0]. */
constexpr int PAT_SLOTS = 2 * PAT_MAX;
constexpr int PAT_EMPTY = (int)0x80000000; /* no offset col - row equals it */
__global__ __launch_bounds__(BS) void k_pattern_offsets(int n, const int *__restrict__ rp,
                                                        const int *__restrict__ ci,
                                                        int *__restrict__ gtab,
                                                        int *__restrict__ flag)
{
    __shared__ int tab[PAT_SLOTS];
    __shared__ int over;
    if (threadIdx.x < PAT_SLOTS) tab[threadIdx.x] = PAT_EMPTY;
    if (threadIdx.x == 0) over = 0;
    __syncthreads();
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r < n) {
        const int k1 = rp[r + 1];
        for (int k = rp[r]; k < k1; k++) {
            const int o = ci[k] - r;
            int s = 0;
            for (; s < PAT_SLOTS; s++) {
                const int v = __hip_atomic_load(&tab[s], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
                if (v == o) break;
                if (v == PAT_EMPTY) {
                    const int prev = atomicCAS(&tab[s], PAT_EMPTY, o);
                    if (prev == PAT_EMPTY || prev == o) break;
                }
            }
            if (s == PAT_SLOTS) over = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && over) atomicOr(flag, 1);
    if (threadIdx.x >= PAT_SLOTS) return;
    const int o = tab[threadIdx.x];
    if (o == PAT_EMPTY) return;
    int s = 0;
    for (; s < PAT_SLOTS; s++) {
        const int v = __hip_atomic_load(&gtab[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v == o) break;
        if (v == PAT_EMPTY) {
            const int prev = atomicCAS(&gtab[s], PAT_EMPTY, o);
            if (prev == PAT_EMPTY || prev == o) break;
        }
    }
    if (s == PAT_SLOTS) atomicOr(flag, 1);
}

/* step 3: the row masks over the sorted offset table (every offset of the
   operator is in it); the rows that pad the last 256-row block get 0 */
__global__ __launch_bounds__(BS) void k_pattern_masks(int n, const int *__restrict__ rp,
                                                      const int *__restrict__ ci, PatOffsets po,
                                                      int pk, unsigned char *__restrict__ pm)
{
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r >= n) {
        pm[r] = 0;
        return;
    }
    unsigned m = 0;
    const int k1 = rp[r + 1];
    for (int k = rp[r]; k < k1; k++) {
        const int o = ci[k] - r;
#pragma unroll
        for (int q = 0; q < PAT_MAX; q++)
            if (q < pk && po.d[q] == o) m |= 1u << q;
    }
    pm[r] = (unsigned char)m;
}

/* per 256-row block b: its nnz (the LDS tile of k_rows_staged) and its
   longest row, into part[2 b], part[2 b + 1] */
__global__ __launch_bounds__(BS) void k_block_nnz(int n, const int *__restrict__ rp,
                                                  int *__restrict__ part)
{
    __shared__ int sm[BS / 64];
    const int r0 = blockIdx.x * BS;
    const int r = r0 + threadIdx.x;
    int m = r < n ? rp[r + 1] - rp[r] : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < BS / 64; q++) m = max(m, sm[q]);
        part[2 * blockIdx.x] = rp[min(r0 + BS, n)] - rp[r0];
        part[2 * blockIdx.x + 1] = m;
    }
}

/* narrow column copy: c16[k] = col[k] - row for every entry of row r (one
   row per thread, setup only); wide[0] is raised when an offset needs more
   than 16 bits */
__global__ __launch_bounds__(BS) void k_narrow_cols(int n, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    short *__restrict__ c16,
                                                    int *__restrict__ wide)
{
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r >= n) return;
    int bad = 0;
    const int k1 = rp[r + 1];
    for (int k = rp[r]; k < k1; k++) {
        const int o = ci[k] - r;
        bad |= (o < -32768) | (o > 32767);
        c16[k] = (short)o;
    }
    if (bad) atomicOr(wide, 1);
}

/* maxima of the per-block pairs into out[0] (tile), out[1] (longest row) */
__global__ __launch_bounds__(BS) void k_block_nnz_final(int nb, const int *__restrict__ part,
                                                        int *__restrict__ out)
{
    __shared__ int sm[2][BS / 64];
    int a = 0, m = 0;
    for (int q = threadIdx.x; q < nb; q += BS) {
        a = max(a, part[2 * q]);
        m = max(m, part[2 * q + 1]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a = max(a, __shfl_xor(a, o, 64));
        m = max(m, __shfl_xor(m, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = a;
        sm[1][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < BS / 64; q++) {
            a = max(a, sm[0][q]);
            m = max(m, sm[1][q]);
        }
        out[0] = a;
        out[1] = m;
    }
}

/* Fallback for rows too long for the LDS tile: G lanes per row (a power of
   two from the average row length), xor-shuffle reduction. */
template <int G, typename MT, typename VT, int OP>
__global__ __launch_bounds__(BS) void k_rows(int n, const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const MT *__restrict__ av,
                                             const VT *__restrict__ x,
                                             const VT *__restrict__ f,
                                             const VT *__restrict__ d,
                                             VT *__restrict__ y)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tid = b * BS + threadIdx.x;
    const int row = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= n) return;
    const int j0 = rp[row], j1 = rp[row + 1];
    VT t = 0;
    for (int j = j0 + lane; j < j1; j += G) {
        const int c = ci[j];
        const VT g = (OP == OP_JACOBI_Z) ? f[c] / d[c] : x[c];
        t = fma(mval<VT>(av[j]), g, t);
    }
#pragma unroll
    for (int o = G >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, G);
    if (lane == 0) {
        if (OP == OP_SPMV)
            y[row] = t;
        else if (OP == OP_RESID)
            y[row] = f[row] - t;
        else if (OP == OP_JACOBI)
            y[row] = (f[row] - t) / d[row] + x[row];
        else
            y[row] = (f[row] - t) / d[row] + f[row] / d[row];
    }
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_jacobi_zero(int n, const VT *__restrict__ f,
                                                    const VT *__restrict__ d,
                                                    VT *__restrict__ y)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = f[i] / d[i];
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_prolong(int n, const int *__restrict__ agg,
                                                const VT *__restrict__ p,
                                                const VT *__restrict__ xc,
                                                VT *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = fma(p[i], xc[agg[i]], x[i]);
}

/* the prolongation into another iterate: y = x + P x_c (k_prolong's
   arithmetic) */
template <typename VT>
__global__ __launch_bounds__(BS) void k_prolong_to(int n, const int *__restrict__ agg,
                                                   const VT *__restrict__ p,
                                                   const VT *__restrict__ xc,
                                                   const VT *__restrict__ x,
                                                   VT *__restrict__ y)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = fma(p[i], xc[agg[i]], x[i]);
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_fill(int n, VT v, VT *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = v;
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_copy(int n, const VT *__restrict__ a,
                                             VT *__restrict__ b)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) b[i] = a[i];
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_axpy(int n, VT alpha, const VT *__restrict__ x,
                                             VT *__restrict__ y)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = fma(alpha, x[i], y[i]);
}

template <typename VT>
__global__ __launch_bounds__(BS) void k_scale(int n, VT alpha, VT *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = alpha * x[i];
}

template <typename S, typename D>
__global__ __launch_bounds__(BS) void k_convert(size_t n, const S *__restrict__ a,
                                                D *__restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    if (i < n) b[i] = (D)a[i];
}

/* double -> bfloat16 (through float, round to nearest even; NaN kept quiet) */
__global__ __launch_bounds__(BS) void k_to_bf16(size_t n, const double *__restrict__ a,
                                                unsigned short *__restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    const float f = (float)a[i];
    unsigned u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
        b[i] = (unsigned short)((u >> 16) | 0x40u);
        return;
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    b[i] = (unsigned short)(u >> 16);
}

template <int K>
__global__ __launch_bounds__(BS) void k_dots(int n, const double *__restrict__ a0,
                                             const double *__restrict__ a1,
                                             const double *__restrict__ a2,
                                             const double *__restrict__ b,
                                             double *__restrict__ part)
{
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double bi = b[i];
        s[0] = fma(a0[i], bi, s[0]);
        if constexpr (K > 1) s[1] = fma(a1[i], bi, s[1]);
        if constexpr (K > 2) s[2] = fma(a2[i], bi, s[2]);
    }
    block_reduce_store<K>(s, part + (size_t)blockIdx.x * K);
}

/* final pass of the reductions: out[k] = sum_b part[b*K + k] (+ optional sqrt) */
template <int K, bool SQRT>
__global__ __launch_bounds__(BS) void k_reduce_final(int nb, const double *__restrict__ part,
                                                     double *__restrict__ out)
{
    reduce_final_body<K, SQRT>(nb, part, out);
}

/* FCG v3 scalar recurrence (src/SOLVER/FCG.cu:531-534) for the unfused path */
__global__ void k_fcg_scalars(double *scal)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) fcg_recurrence(scal);
}

/* first step (src/SOLVER/FCG.cu:475-478): delta = beta, theta = alpha / delta */
__global__ void k_fcg_first_scalars(double *scal)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) fcg_first_recurrence(scal);
}

template <int K>
__global__ __launch_bounds__(RED_FINAL_BS) void k_fcg_final(int nb, const double *__restrict__ part,
                                                           double *__restrict__ scal)
{
    double res[K];
    if (fcg_final_sum<K>(nb, part, res)) fcg_final_apply<K>(res, scal);
}

/* FCG SpMV fused with its inner products (src/SOLVER/FCG.cu:508-528 and the
   triple inner product :6-83): t = A z, and per block the partial sums of
   alpha = r.z, beta = t.z (K = 2, first step) and gamma = to.z (K = 3),
   finished by k_fcg_final. The products use the z, r, to values the
   epilogue loads anyway, so the separate triple-product pass is saved. */
template <int K, bool NT, typename IT>
__global__ __launch_bounds__(BS) void k_spmv_dots(int n, const int *__restrict__ rp,
                                                  const IT *__restrict__ ci,
                                                  const double *__restrict__ av,
                                                  const double *__restrict__ z,
                                                  double *__restrict__ t,
                                                  const double *__restrict__ rv,
                                                  const double *__restrict__ to,
                                                  double *__restrict__ part, int tile)
{
    IT *lcol;
    double *lval;
    staged_lds<double, IT>(tile, lcol, lval);
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * BS;
    const int r1 = min(r0 + BS, n);
    const int r = r0 + threadIdx.x;
    const bool own = r < r1;
    int ra = 0, re = 0;
    double zr = 0.0, rr = 0.0, tor = 0.0;
    if (own) {
        ra = rp[r];
        re = rp[r + 1];
        zr = z[r];
        rr = rv[r];
        if (K > 2) tor = to[r];
    }
    const int e0 = rp[r0];
    const int cnt = rp[r1] - e0;
    stage_entries<double, NT, IT>(ci, av, e0, cnt, lcol, lval);
    __syncthreads();
    double sdot[K];
#pragma unroll
    for (int k = 0; k < K; k++) sdot[k] = 0.0;
    if (own) {
        const double tr = staged_row_sum<double, double, IT>(
            lcol, lval, z, ra - stage_base<IT>(e0), ra - stage_base<double>(e0), re - ra, r);
        t[r] = tr;
        sdot[0] = rr * zr;
        sdot[1] = tr * zr;
        if (K > 2) sdot[K - 1] = tor * zr;
    }
    block_reduce_store<K>(sdot, part + (size_t)b * K);
}

/* k_spmv_dots on a stencil-pattern operator (the reads of k_rows_pattern,
   the products and partials of k_spmv_dots) */
template <int K, bool NT>
__global__ __launch_bounds__(BS) void k_spmv_dots_pattern(int n, const int *__restrict__ rp,
                                                          const unsigned char *__restrict__ pm,
                                                          PatOffsets po,
                                                          const double *__restrict__ av,
                                                          const double *__restrict__ z,
                                                          double *__restrict__ t,
                                                          const double *__restrict__ rv,
                                                          const double *__restrict__ to,
                                                          double *__restrict__ part)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_pat[];
    double *lval = reinterpret_cast<double *>(smem_pat);
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * BS;
    const int r1 = min(r0 + BS, n);
    const int r = r0 + threadIdx.x;
    const bool own = r < r1;
    const unsigned w = ldm<NT>(reinterpret_cast<const unsigned *>(pm + r0) + (threadIdx.x & 63));
    double zr = 0.0, rr = 0.0, tor = 0.0;
    if (own) {
        zr = z[r];
        rr = rv[r];
        if (K > 2) tor = to[r];
    }
    const int e0 = rp[r0];
    const int cnt = rp[r1] - e0;
    stage_values<double, NT>(av, e0, cnt, lval);
    unsigned m;
    int pre;
    pattern_block_rows(w, m, pre);
    __syncthreads();
    double sdot[K];
#pragma unroll
    for (int k = 0; k < K; k++) sdot[k] = 0.0;
    if (own) {
        const double tr = pattern_row_sum<double, double>(
            m, po, lval + (e0 - stage_base<double>(e0)) + pre, z, r);
        t[r] = tr;
        sdot[0] = rr * zr;
        sdot[1] = tr * zr;
        if (K > 2) sdot[K - 1] = tor * zr;
    }
    block_reduce_store<K>(sdot, part + (size_t)b * K);
}

/* fused double_merged_axpy pair + ||r||^2 partials (src/SOLVER/FCG.cu:88-98,
   536-558): z = -theta zo + z; x += theta2 z; t = -theta to + t; r += -theta2 t */
__global__ __launch_bounds__(BS) void k_fcg_update(int n, const double *__restrict__ scal,
                                                   const double *__restrict__ zo,
                                                   double *__restrict__ z,
                                                   double *__restrict__ x,
                                                   const double *__restrict__ to,
                                                   double *__restrict__ t,
                                                   double *__restrict__ r,
                                                   double *__restrict__ part)
{
    const double theta = scal[4], theta2 = scal[5];
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double zi = fma(-theta, zo[i], z[i]);
        z[i] = zi;
        x[i] = fma(theta2, zi, x[i]);
        const double ti = fma(-theta, to[i], t[i]);
        t[i] = ti;
        const double ri = fma(-theta2, ti, r[i]);
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + blockIdx.x);
}

/* first FCG step (src/SOLVER/FCG.cu:475-486): x += theta v, r -= theta w, ||r||^2 */
__global__ __launch_bounds__(BS) void k_fcg_first(int n, const double *__restrict__ scal,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ x,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ r,
                                                  double *__restrict__ part)
{
    const double theta = scal[4];
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        x[i] = fma(theta, v[i], x[i]);
        const double ri = fma(-theta, w[i], r[i]);
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + blockIdx.x);
}

/* --------------------------------------------------------- K-cycle
   Device scalars of one level's inner iteration (src/SOLVER/FCG.cu:620-723),
   ks = [w.w (delta0^2), w.d0 (delta_old), d0.v (tau), alpha, w'.w' (l2^2),
         d1.v (tau1), d1.v1 (tau2), d1.w' (tau3), c0, c1, ..., flag]. */

/* up to 3 products of VT vectors with b, accumulated in double (fma over a
   grid-stride partition, then the fixed block tree) */
template <int K, typename VT>
__global__ __launch_bounds__(BS) void k_kdots(int n, const VT *__restrict__ a0,
                                              const VT *__restrict__ a1,
                                              const VT *__restrict__ a2,
                                              const VT *__restrict__ b,
                                              double *__restrict__ part)
{
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double bi = (double)b[i];
        s[0] = fma((double)a0[i], bi, s[0]);
        if constexpr (K > 1) s[1] = fma((double)a1[i], bi, s[1]);
        if constexpr (K > 2) s[2] = fma((double)a2[i], bi, s[2]);
    }
    block_reduce_store<K>(s, part + (size_t)blockIdx.x * K);
}

/* alpha = delta_old / tau; a non-positive delta_old (indefinite
   preconditioner) or tau (indefinite matrix) raises the flag - the
   reference exits there */
__global__ void k_kcycle_alpha(double *ks, double *gflag)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    kcycle_alpha_apply(ks, gflag);
}

/* w = -alpha v + w */
template <typename VT>
__global__ __launch_bounds__(BS) void k_kcycle_axpy(int n, const double *__restrict__ ks,
                                                    const VT *__restrict__ v,
                                                    VT *__restrict__ w)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) w[i] = fma(-(VT)ks[3], v[i], w[i]);
}

/* the combination coefficients of the inner iteration: one step when
   ||w'|| <= rtol ||w|| (c1 = 0), else the two-step update */
__global__ void k_kcycle_coef(double *ks, double rtol)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    kcycle_coef_apply(ks, rtol);
}

/* x = c0 d0 + c1 d1 (x starts at zero) */
template <typename VT>
__global__ __launch_bounds__(BS) void k_kcycle_combine(int n, const double *__restrict__ ks,
                                                       const VT *__restrict__ d0,
                                                       const VT *__restrict__ d1,
                                                       VT *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = fma((VT)ks[9], d1[i], fma((VT)ks[8], d0[i], VT(0)));
}

/* r = b - y (the true residual, y = A x) and its squared norm partials */
__global__ __launch_bounds__(BS) void k_true_residual(int n, const double *__restrict__ b,
                                                      const double *__restrict__ y,
                                                      double *__restrict__ r,
                                                      double *__restrict__ part)
{
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double ri = b[i] - y[i];
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + blockIdx.x);
}

PatOffsets pattern_offsets(const DevCsr &A)
{
    PatOffsets po;
    for (int k = 0; k < PAT_MAX; k++) po.d[k] = A.poff[k];
    return po;
}

template <typename MT, typename VT, int OP>
void rows_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *x, const VT *f,
                const VT *d, VT *y)
{
    const int n = A.n;
    const int *rp = A.row.get(), *ci = A.col.get();
    if (stream_ok<VT>(A)) {
        const int nb = (n + BS - 1) / BS;
        if (A.use_pattern()) {
            if (!stage_aligned(av, av))
                throw Error(BCMG_EINVAL, "row kernels: matrix values not 16-byte aligned");
            const size_t lds = pattern_lds_bytes(A.tile, sizeof(MT));
            if (lds > 65536) throw Error(BCMG_EUNSUP, "row tile too large for this value type");
            const PatOffsets po = pattern_offsets(A);
            if ((size_t)A.nnz * sizeof(MT) + (size_t)n > NT_BYTES)
                k_rows_pattern<MT, VT, OP, true><<<nb, BS, lds, s>>>(n, rp, A.pmask.get(), po, av,
                                                                     x, f, d, y);
            else
                k_rows_pattern<MT, VT, OP, false><<<nb, BS, lds, s>>>(n, rp, A.pmask.get(), po,
                                                                      av, x, f, d, y);
            return;
        }
        /* the narrow columns where the int32-column stream is non-temporal */
        const bool use16 = A.use_narrow() && (size_t)A.nnz * (4 + sizeof(MT)) > NT_BYTES;
        if (!stage_aligned(use16 ? (const void *)A.col16.get() : (const void *)ci, av))
            throw Error(BCMG_EINVAL, "row kernels: matrix arrays not 16-byte aligned");
        if (use16) {
            const short *c16 = A.col16.get();
            const size_t lds = staged_lds_bytes(A.tile, sizeof(MT), sizeof(short));
            if (lds > 65536) throw Error(BCMG_EUNSUP, "row tile too large for this value type");
            if ((size_t)A.nnz * (2 + sizeof(MT)) > NT_BYTES)
                k_rows_staged<MT, VT, OP, true, short><<<nb, BS, lds, s>>>(n, rp, c16, av, x, f,
                                                                           d, y, A.tile);
            else
                k_rows_staged<MT, VT, OP, false, short><<<nb, BS, lds, s>>>(n, rp, c16, av, x, f,
                                                                            d, y, A.tile);
            return;
        }
        const size_t lds = staged_lds_bytes(A.tile, sizeof(MT));
        if (lds > 65536) throw Error(BCMG_EUNSUP, "row tile too large for this value type");
        if ((size_t)A.nnz * (4 + sizeof(MT)) > NT_BYTES)
            k_rows_staged<MT, VT, OP, true, int><<<nb, BS, lds, s>>>(n, rp, ci, av, x, f, d, y,
                                                                     A.tile);
        else
            k_rows_staged<MT, VT, OP, false, int><<<nb, BS, lds, s>>>(n, rp, ci, av, x, f, d, y,
                                                                      A.tile);
        return;
    }
    const int G = A.group;
    const int nb = blocks_for((long long)n * G);
    switch (G) {
    case 1: k_rows<1, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 2: k_rows<2, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 4: k_rows<4, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 8: k_rows<8, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 16: k_rows<16, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 32: k_rows<32, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    default: k_rows<64, MT, VT, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    }
}

template <typename MT, typename VT>
void rows_op(hipStream_t s, const DevCsr &A, const MT *av, RowOp op, const VT *x, const VT *f,
             const VT *d, VT *y)
{
    if (op == OP_SPMV)
        rows_typed<MT, VT, OP_SPMV>(s, A, av, x, f, d, y);
    else if (op == OP_RESID)
        rows_typed<MT, VT, OP_RESID>(s, A, av, x, f, d, y);
    else if (op == OP_JACOBI)
        rows_typed<MT, VT, OP_JACOBI>(s, A, av, x, f, d, y);
    else
        rows_typed<MT, VT, OP_JACOBI_Z>(s, A, av, x, f, d, y);
}

template <typename MT, typename VT, bool PRO, bool RR>
void relax_block_shape(hipStream_t s, const DevCsr &A, const MT *av, const VT *d, const VT *f,
                       const VT *xin, VT *xout, int sweeps, int z, int ncols, long cs, int rs,
                       const int *act, const BlockFuse<VT> &fu)
{
    const size_t lds = (size_t)2 * A.n * sizeof(VT);
    const int *rp = A.row.get(), *ci = A.col.get();
    const int n = A.n;
    if (n <= 1024 && A.maxrow <= 8)
        k_relax_block<MT, VT, 1, 8, PRO, RR><<<ncols, 1024, lds, s>>>(
            n, rp, ci, av, d, f, xin, xout, sweeps, z, cs, rs, act, fu);
    else if (n <= 4096 && A.maxrow <= 8)
        k_relax_block<MT, VT, 4, 8, PRO, RR><<<ncols, 1024, lds, s>>>(
            n, rp, ci, av, d, f, xin, xout, sweeps, z, cs, rs, act, fu);
    else if (n <= 2048 && A.maxrow <= 16)
        k_relax_block<MT, VT, 2, 16, PRO, RR><<<ncols, 1024, lds, s>>>(
            n, rp, ci, av, d, f, xin, xout, sweeps, z, cs, rs, act, fu);
    else if (n <= 1024 && A.maxrow <= 32)
        k_relax_block<MT, VT, 1, 32, PRO, RR><<<ncols, 1024, lds, s>>>(
            n, rp, ci, av, d, f, xin, xout, sweeps, z, cs, rs, act, fu);
    else
        throw Error(BCMG_EINVAL, "relax_block: level too large");
}

template <typename MT, typename VT>
void relax_block_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *d, const VT *f,
                       const VT *xin, VT *xout, int sweeps, int z, int ncols, long cs,
                       int rs, const int *act, const BlockFuse<VT> &fu)
{
    const bool pro = fu.agg != nullptr, rr = fu.fc != nullptr;
    if ((pro || rr) && ncols != 1)
        throw Error(BCMG_EINVAL, "relax_block: fused work is single-column");
    if (pro && rr)
        throw Error(BCMG_EINVAL, "relax_block: prolongation and restriction in one launch");
    if (pro)
        relax_block_shape<MT, VT, true, false>(s, A, av, d, f, xin, xout, sweeps, z, ncols, cs,
                                               rs, act, fu);
    else if (rr)
        relax_block_shape<MT, VT, false, true>(s, A, av, d, f, xin, xout, sweeps, z, ncols, cs,
                                               rs, act, fu);
    else
        relax_block_shape<MT, VT, false, false>(s, A, av, d, f, xin, xout, sweeps, z, ncols,
                                                cs, rs, act, fu);
}

template <typename MT, typename VT>
void resid_restrict_typed(hipStream_t s, const DevCsr &A, const MT *av, const DevCsr &R,
                          const VT *rv, const VT *x, const VT *f, VT *y)
{
    const int nc = R.n;
    constexpr int G = 4;
    k_resid_restrict<MT, VT, G><<<blocks_for((long long)nc * G), BS, 0, s>>>(
        nc, R.row.get(), R.col.get(), rv, A.row.get(), A.col.get(), av, x, f, y);
}

} // namespace

/* ============================================================== host */

int lanes_for(int n, long long nnz)
{
    const double avg = n > 0 ? (double)nnz / (double)n : 1.0;
    int G = 1;
    while (G < 64 && G < avg) G <<= 1;
    return G;
}

void launch_block_nnz(Ctx &c, DevCsr &A)
{
    A.group = lanes_for(A.n, A.nnz);
    const int nblk = (A.n + BS - 1) / BS;
    A.tile_dev.ensure(2 + 2 * (size_t)nblk);
    if (A.n > 0) {
        k_block_nnz<<<nblk, BS, 0, c.stream>>>(A.n, A.row.get(), A.tile_dev.get() + 2);
        k_block_nnz_final<<<1, BS, 0, c.stream>>>(nblk, A.tile_dev.get() + 2, A.tile_dev.get());
    } else {
        BCMG_HIP(hipMemsetAsync(A.tile_dev.get(), 0, 2 * sizeof(int), c.stream));
    }
    BCMG_CHECK_LAUNCH();
    A.tile = -1; /* pending until resolve_tiles */
}

void resolve_tiles(Ctx &c, const std::vector<DevCsr *> &mats)
{
    for (size_t b = 0; b < mats.size(); b += 64) {
        const size_t e = std::min(mats.size(), b + 64);
        for (size_t k = b; k < e; k++)
            BCMG_HIP(hipMemcpyAsync(c.h_ints + 64 + 2 * (k - b), mats[k]->tile_dev.get(),
                                    2 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        BCMG_HIP(hipStreamSynchronize(c.stream));
        for (size_t k = b; k < e; k++) {
            const int t = c.h_ints[64 + 2 * (k - b)];
            mats[k]->tile = t > 0 ? t : 1;
            mats[k]->maxrow = c.h_ints[64 + 2 * (k - b) + 1];
        }
    }
}

void analyze_rows(Ctx &c, DevCsr &A)
{
    launch_block_nnz(c, A);
    resolve_tiles(c, {&A});
    make_narrow(c, {&A});
}


/* square operators whose FP64 stream (12 B per entry) exceeds this are
   checked for a stencil pattern; the 16-bit copy is kept for streamed ones
   (above NT_BYTES) only */
constexpr size_t PAT_MIN_BYTES = (size_t)128 << 20;

void make_narrow(Ctx &c, const std::vector<DevCsr *> &mats)
{
    std::vector<DevCsr *> cand;
    for (DevCsr *A : mats)
        if (A->n == A->m && (size_t)A->nnz * 12 > PAT_MIN_BYTES && !A->cols_analysed)
            cand.push_back(A);
    if (cand.empty()) return;
    const int nc = (int)cand.size();
    /* per candidate: the 16-bit overflow flag, the pattern overflow flag and
       the offset table */
    constexpr int W = 2 + PAT_SLOTS;
    DBuf<int> info((size_t)nc * W);
    std::vector<int> init((size_t)nc * W, PAT_EMPTY);
    for (int k = 0; k < nc; k++) init[(size_t)k * W] = init[(size_t)k * W + 1] = 0;
    BCMG_HIP(hipMemcpyAsync(info.get(), init.data(), init.size() * sizeof(int),
                            hipMemcpyHostToDevice, c.stream));
    for (int k = 0; k < nc; k++) {
        DevCsr &A = *cand[k];
        int *in = info.get() + (size_t)k * W;
        const int nblk = (A.n + BS - 1) / BS;
        if ((size_t)A.nnz * 12 > NT_BYTES) {
            A.col16.alloc(A.nnz);
            k_narrow_cols<<<nblk, BS, 0, c.stream>>>(A.n, A.row.get(), A.col.get(),
                                                     A.col16.get(), in);
        }
        k_pattern_offsets<<<nblk, BS, 0, c.stream>>>(A.n, A.row.get(), A.col.get(), in + 2,
                                                     in + 1);
    }
    BCMG_CHECK_LAUNCH();
    std::vector<int> h((size_t)nc * W);
    BCMG_HIP(hipMemcpyAsync(h.data(), info.get(), h.size() * sizeof(int), hipMemcpyDeviceToHost,
                            c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < nc; k++) {
        DevCsr &A = *cand[k];
        const int *hk = h.data() + (size_t)k * W;
        A.cols_analysed = true;
        A.narrow = A.col16.get() != nullptr && hk[0] == 0;
        if (!A.narrow) A.col16.release();
        std::vector<int> off;
        for (int q = 0; q < PAT_SLOTS; q++)
            if (hk[2 + q] != PAT_EMPTY) off.push_back(hk[2 + q]);
        const int cnt = (int)off.size();
        if (hk[1] != 0 || cnt < 1 || cnt > PAT_MAX) continue;
        std::sort(off.begin(), off.end());
        PatOffsets po;
        for (int q = 0; q < PAT_MAX; q++) po.d[q] = q < cnt ? off[q] : PAT_EMPTY;
        /* whole 256-row blocks: the kernels read every block's masks as 64
           words */
        const int nblk = (A.n + BS - 1) / BS;
        A.pmask.alloc((size_t)nblk * BS);
        k_pattern_masks<<<nblk, BS, 0, c.stream>>>(A.n, A.row.get(), A.col.get(), po, cnt,
                                                    A.pmask.get());
        BCMG_CHECK_LAUNCH();
        A.pk = cnt;
        for (int q = 0; q < PAT_MAX; q++) A.poff[q] = po.d[q];
    }
}

void DevCsr::make_lowp(hipStream_t s, bool bf16)
{
    if (nnz <= 0) return;
    const size_t nb = ((size_t)nnz + BS - 1) / BS;
    if (bf16) {
        if (!val16.get()) val16.alloc(nnz);
        k_to_bf16<<<nb, BS, 0, s>>>((size_t)nnz, val.get(), val16.get());
    } else {
        if (!val32.get()) val32.alloc(nnz);
        k_convert<double, float><<<nb, BS, 0, s>>>((size_t)nnz, val.get(), val32.get());
    }
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_rows(hipStream_t s, const DevCsr &A, MatType mt, RowOp op, const VT *x, const VT *f,
                 const VT *d, VT *y)
{
    if (A.n <= 0) return;
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        rows_op<double, double>(s, A, A.val.get(), op, x, f, d, y);
    } else {
        if (mt == MT_F32)
            rows_op<float, float>(s, A, A.val32.get(), op, x, f, d, y);
        else if (mt == MT_BF16)
            rows_op<unsigned short, float>(s, A, A.val16.get(), op, x, f, d, y);
        else
            rows_op<double, float>(s, A, A.val.get(), op, x, f, d, y);
    }
    BCMG_CHECK_LAUNCH();
}

void launch_rows(hipStream_t s, const DevCsr &A, RowOp op, const double *x, const double *f,
                 const double *d, double *y)
{
    launch_rows<double>(s, A, MT_F64, op, x, f, d, y);
}

template <typename VT>
bool fuse_zero_sweeps(const DevCsr &A, MatType mt)
{
    /* a streamed FP64 operator: the division at every gather costs more
       than the pass it saves */
    const size_t vb = mt == MT_F64 ? 8 : mt == MT_F32 ? 4 : 2;
    return !(std::is_same<VT, double>::value && (size_t)A.nnz * (4 + vb) > NT_BYTES);
}
template bool fuse_zero_sweeps<double>(const DevCsr &, MatType);
template bool fuse_zero_sweeps<float>(const DevCsr &, MatType);

template <typename VT>
bool relax_block_ok(const DevCsr &A)
{
    if (A.maxrow <= 0) return false;
    return (A.n <= 4096 && A.maxrow <= 8) || (A.n <= 2048 && A.maxrow <= 16) ||
           (A.n <= 1024 && A.maxrow <= 32);
}

template <typename VT>
void launch_relax_block(hipStream_t s, const DevCsr &A, MatType mt, const VT *d, const VT *f,
                        const VT *xin, VT *xout, int sweeps, bool zero, int ncols, long cs,
                        int rs, const int *act, const BlockFuse<VT> &fu)
{
    const int z = zero ? 1 : 0;
    if (ncols <= 0) return;
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        relax_block_typed<double, double>(s, A, A.val.get(), d, f, xin, xout, sweeps, z, ncols,
                                          cs, rs, act, fu);
    } else {
        if (mt == MT_F32)
            relax_block_typed<float, float>(s, A, A.val32.get(), d, f, xin, xout, sweeps, z,
                                            ncols, cs, rs, act, fu);
        else if (mt == MT_BF16)
            relax_block_typed<unsigned short, float>(s, A, A.val16.get(), d, f, xin, xout,
                                                     sweeps, z, ncols, cs, rs, act, fu);
        else
            relax_block_typed<double, float>(s, A, A.val.get(), d, f, xin, xout, sweeps, z,
                                             ncols, cs, rs, act, fu);
    }
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
bool rows_sequential(const DevCsr &A)
{
    return stream_ok<VT>(A);
}

template <typename VT>
bool launch_jacobi_prolong(hipStream_t s, const DevCsr &A, MatType mt, const int *agg,
                           const VT *p, const VT *xc, const VT *x, const VT *f, const VT *d,
                           VT *y)
{
    if (!stream_ok<VT>(A)) return false;
    if (x == y) throw Error(BCMG_EINVAL, "prolongation sweep into its own input");
    const int n = A.n;
    const int *rp = A.row.get(), *ci = A.col.get();
    if (n <= 0) return true;
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        k_jacobi_prolong<double, double><<<blocks_for(n), BS, 0, s>>>(n, rp, ci, A.val.get(), agg,
                                                                      p, xc, x, f, d, y);
    } else {
        if (mt == MT_F32)
            k_jacobi_prolong<float, float><<<blocks_for(n), BS, 0, s>>>(
                n, rp, ci, A.val32.get(), agg, p, xc, x, f, d, y);
        else if (mt == MT_BF16)
            k_jacobi_prolong<unsigned short, float><<<blocks_for(n), BS, 0, s>>>(
                n, rp, ci, A.val16.get(), agg, p, xc, x, f, d, y);
        else
            k_jacobi_prolong<double, float><<<blocks_for(n), BS, 0, s>>>(
                n, rp, ci, A.val.get(), agg, p, xc, x, f, d, y);
    }
    BCMG_CHECK_LAUNCH();
    return true;
}

template <typename VT>
bool launch_resid_restrict(hipStream_t s, const DevCsr &A, MatType mt, const DevCsr &R,
                           const VT *x, const VT *f, VT *y)
{
    /* the staged kernels' sequential sums are what this kernel repeats */
    if (!stream_ok<VT>(A) || !stream_ok<VT>(R) || R.n <= 0) return false;
    const VT *rv;
    if constexpr (std::is_same<VT, double>::value)
        rv = R.val.get();
    else
        rv = R.val32.get();
    if (!rv) return false;
    if constexpr (std::is_same<VT, double>::value) {
        if (mt != MT_F64) throw Error(BCMG_EINVAL, "FP64 vectors need FP64 matrix values");
        resid_restrict_typed<double, double>(s, A, A.val.get(), R, rv, x, f, y);
    } else {
        if (mt == MT_F32)
            resid_restrict_typed<float, float>(s, A, A.val32.get(), R, rv, x, f, y);
        else if (mt == MT_BF16)
            resid_restrict_typed<unsigned short, float>(s, A, A.val16.get(), R, rv, x, f, y);
        else
            resid_restrict_typed<double, float>(s, A, A.val.get(), R, rv, x, f, y);
    }
    BCMG_CHECK_LAUNCH();
    return true;
}

template <typename VT>
void launch_jacobi_zero(hipStream_t s, int n, const VT *f, const VT *d, VT *y)
{
    if (n <= 0) return;
    k_jacobi_zero<VT><<<blocks_for(n), BS, 0, s>>>(n, f, d, y);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_prolong(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc, VT *x)
{
    if (n <= 0) return;
    k_prolong<VT><<<blocks_for(n), BS, 0, s>>>(n, agg, p, xc, x);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_prolong_to(hipStream_t s, int n, const int *agg, const VT *p, const VT *xc,
                       const VT *x, VT *y)
{
    if (n <= 0) return;
    if (x == y) throw Error(BCMG_EINVAL, "prolongation into its own input");
    k_prolong_to<VT><<<blocks_for(n), BS, 0, s>>>(n, agg, p, xc, x, y);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_fill(hipStream_t s, int n, VT v, VT *x)
{
    if (n <= 0) return;
    k_fill<VT><<<blocks_for(n), BS, 0, s>>>(n, v, x);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_copy(hipStream_t s, int n, const VT *a, VT *b)
{
    if (n <= 0 || a == b) return;
    k_copy<VT><<<blocks_for(n), BS, 0, s>>>(n, a, b);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_axpy(hipStream_t s, int n, VT alpha, const VT *x, VT *y)
{
    if (n <= 0) return;
    k_axpy<VT><<<blocks_for(n), BS, 0, s>>>(n, alpha, x, y);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_scale(hipStream_t s, int n, VT alpha, VT *x)
{
    if (n <= 0) return;
    k_scale<VT><<<blocks_for(n), BS, 0, s>>>(n, alpha, x);
    BCMG_CHECK_LAUNCH();
}

void launch_d2f(hipStream_t s, size_t n, const double *a, float *b)
{
    if (!n) return;
    k_convert<double, float><<<(n + BS - 1) / BS, BS, 0, s>>>(n, a, b);
    BCMG_CHECK_LAUNCH();
}

void launch_f2d(hipStream_t s, size_t n, const float *a, double *b)
{
    if (!n) return;
    k_convert<float, double><<<(n + BS - 1) / BS, BS, 0, s>>>(n, a, b);
    BCMG_CHECK_LAUNCH();
}

#define BCMG_INSTANTIATE(VT)                                                                 \
    template void launch_rows<VT>(hipStream_t, const DevCsr &, MatType, RowOp, const VT *,   \
                                  const VT *, const VT *, VT *);                             \
    template bool relax_block_ok<VT>(const DevCsr &);                                        \
    template void launch_relax_block<VT>(hipStream_t, const DevCsr &, MatType, const VT *,   \
                                         const VT *, const VT *, VT *, int, bool, int, long, \
                                         int, const int *, const BlockFuse<VT> &);           \
    template bool rows_sequential<VT>(const DevCsr &);                                       \
    template bool launch_jacobi_prolong<VT>(hipStream_t, const DevCsr &, MatType, const int *,\
                                            const VT *, const VT *, const VT *, const VT *,  \
                                            const VT *, VT *);                               \
    template bool launch_resid_restrict<VT>(hipStream_t, const DevCsr &, MatType,            \
                                            const DevCsr &, const VT *, const VT *, VT *);   \
    template void launch_jacobi_zero<VT>(hipStream_t, int, const VT *, const VT *, VT *);    \
    template void launch_prolong<VT>(hipStream_t, int, const int *, const VT *, const VT *,  \
                                     VT *);                                                  \
    template void launch_prolong_to<VT>(hipStream_t, int, const int *, const VT *,           \
                                        const VT *, const VT *, VT *);                       \
    template void launch_fill<VT>(hipStream_t, int, VT, VT *);                               \
    template void launch_copy<VT>(hipStream_t, int, const VT *, VT *);                       \
    template void launch_axpy<VT>(hipStream_t, int, VT, const VT *, VT *);                   \
    template void launch_scale<VT>(hipStream_t, int, VT, VT *);
BCMG_INSTANTIATE(double)
BCMG_INSTANTIATE(float)
#undef BCMG_INSTANTIATE

void launch_dots(Ctx &c, int n, int k, const double *a0, const double *a1, const double *a2,
                 const double *b, double *out)
{
    const int nb = red_blocks(n);
    c.red.ensure((size_t)RED_MAX_BLOCKS * 3);
    double *part = c.red.get();
    if (k == 1) {
        k_dots<1><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<1, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    } else if (k == 2) {
        k_dots<2><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<2, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    } else {
        k_dots<3><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<3, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    }
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_scalars(hipStream_t s, double *scal)
{
    k_fcg_scalars<<<1, 64, 0, s>>>(scal);
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_update(Ctx &c, int n, double *scal, const double *zo, double *z, double *x,
                       const double *to, double *t, double *r)
{
    const int nb = red_blocks(n);
    k_fcg_update<<<nb, BS, 0, c.stream>>>(n, scal, zo, z, x, to, t, r, c.red.get() + RED_UPD);
    k_fcg_final<1><<<1, RED_FINAL_BS, 0, c.stream>>>(nb, c.red.get() + RED_UPD, scal);
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_first(Ctx &c, int n, double *scal, const double *v, double *x, const double *w,
                      double *r)
{
    const int nb = red_blocks(n);
    k_fcg_first<<<nb, BS, 0, c.stream>>>(n, scal, v, x, w, r, c.red.get() + RED_UPD);
    k_fcg_final<1><<<1, RED_FINAL_BS, 0, c.stream>>>(nb, c.red.get() + RED_UPD, scal);
    BCMG_CHECK_LAUNCH();
}

void launch_true_residual(Ctx &c, const DevCsr &A, const double *b, const double *x, double *y,
                          double *r, double *scal)
{
    const int n = A.n;
    launch_rows(c.stream, A, OP_SPMV, x, nullptr, nullptr, y);
    const int nb = red_blocks(n);
    k_true_residual<<<nb, BS, 0, c.stream>>>(n, b, y, r, c.red.get() + RED_UPD);
    k_fcg_final<1><<<1, RED_FINAL_BS, 0, c.stream>>>(nb, c.red.get() + RED_UPD, scal);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_kdots(Ctx &c, int n, int k, const VT *a0, const VT *a1, const VT *a2, const VT *b,
                  double *out)
{
    const int nb = red_blocks(n);
    c.red.ensure((size_t)RED_MAX_BLOCKS * 3);
    double *part = c.red.get();
    if (k == 1) {
        k_kdots<1, VT><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<1, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    } else if (k == 2) {
        k_kdots<2, VT><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<2, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    } else {
        k_kdots<3, VT><<<nb, BS, 0, c.stream>>>(n, a0, a1, a2, b, part);
        k_reduce_final<3, false><<<1, BS, 0, c.stream>>>(nb, part, out);
    }
    BCMG_CHECK_LAUNCH();
}

void launch_kcycle_alpha(hipStream_t s, double *ks, double *gflag)
{
    k_kcycle_alpha<<<1, 64, 0, s>>>(ks, gflag);
    BCMG_CHECK_LAUNCH();
}

void launch_kcycle_coef(hipStream_t s, double *ks, double rtol)
{
    k_kcycle_coef<<<1, 64, 0, s>>>(ks, rtol);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_kcycle_axpy(hipStream_t s, int n, const double *ks, const VT *v, VT *w)
{
    if (n <= 0) return;
    k_kcycle_axpy<VT><<<blocks_for(n), BS, 0, s>>>(n, ks, v, w);
    BCMG_CHECK_LAUNCH();
}

template <typename VT>
void launch_kcycle_combine(hipStream_t s, int n, const double *ks, const VT *d0, const VT *d1,
                           VT *x)
{
    if (n <= 0) return;
    k_kcycle_combine<VT><<<blocks_for(n), BS, 0, s>>>(n, ks, d0, d1, x);
    BCMG_CHECK_LAUNCH();
}

#define BCMG_INSTANTIATE_K(VT)                                                               \
    template void launch_kdots<VT>(Ctx &, int, int, const VT *, const VT *, const VT *,      \
                                   const VT *, double *);                                    \
    template void launch_kcycle_axpy<VT>(hipStream_t, int, const double *, const VT *, VT *); \
    template void launch_kcycle_combine<VT>(hipStream_t, int, const double *, const VT *,    \
                                            const VT *, VT *);
BCMG_INSTANTIATE_K(double)
BCMG_INSTANTIATE_K(float)
#undef BCMG_INSTANTIATE_K

/* the fused FCG SpMV + inner products over the stored column type IT */
template <typename IT>
static void spmv_dots_typed(hipStream_t s, int nb, int n, int k, const int *rp, const IT *ci,
                            const double *av, int nnz, int tile, const double *z, double *t,
                            const double *r, const double *to, double *part)
{
    const size_t lds = staged_lds_bytes(tile, sizeof(double), sizeof(IT));
    if (!stage_aligned(ci, av))
        throw Error(BCMG_EINVAL, "SpMV: matrix arrays not 16-byte aligned");
    const bool nt = (size_t)nnz * (sizeof(IT) + sizeof(double)) > NT_BYTES;
    if (k == 2) {
        if (nt)
            k_spmv_dots<2, true, IT><<<nb, BS, lds, s>>>(n, rp, ci, av, z, t, r, to, part, tile);
        else
            k_spmv_dots<2, false, IT><<<nb, BS, lds, s>>>(n, rp, ci, av, z, t, r, to, part, tile);
    } else {
        if (nt)
            k_spmv_dots<3, true, IT><<<nb, BS, lds, s>>>(n, rp, ci, av, z, t, r, to, part, tile);
        else
            k_spmv_dots<3, false, IT><<<nb, BS, lds, s>>>(n, rp, ci, av, z, t, r, to, part, tile);
    }
}

size_t red_size_for(int n) { return (size_t)RED_SPMV + 3 * (size_t)((n + BS - 1) / BS) + 64; }

void launch_spmv_dots(Ctx &c, const DevCsr &A, int k, const double *z, double *t,
                      const double *r, const double *to, double *scal)
{
    const int n = A.n;
    if (!stream_ok<double>(A)) {
        /* long rows: separate SpMV, products and recurrence */
        launch_rows(c.stream, A, OP_SPMV, z, nullptr, nullptr, t);
        launch_dots(c, n, k, r, t, to, z, scal);
        if (k == 2)
            k_fcg_first_scalars<<<1, 64, 0, c.stream>>>(scal);
        else
            k_fcg_scalars<<<1, 64, 0, c.stream>>>(scal);
        BCMG_CHECK_LAUNCH();
        return;
    }
    if (c.red.n < red_size_for(n))
        throw Error(BCMG_EINVAL, "reduction scratch not sized for this operator");
    const int nb = (n + BS - 1) / BS;
    double *part = c.red.get() + RED_SPMV;
    const int *rp = A.row.get();
    const double *av = A.val.get();
    if (A.use_pattern()) {
        if (!stage_aligned(av, av))
            throw Error(BCMG_EINVAL, "SpMV: matrix values not 16-byte aligned");
        const size_t lds = pattern_lds_bytes(A.tile, sizeof(double));
        const PatOffsets po = pattern_offsets(A);
        const bool nt = (size_t)A.nnz * sizeof(double) + (size_t)n > NT_BYTES;
        const unsigned char *pm = A.pmask.get();
        if (k == 2) {
            if (nt)
                k_spmv_dots_pattern<2, true><<<nb, BS, lds, c.stream>>>(n, rp, pm, po, av, z, t,
                                                                        r, to, part);
            else
                k_spmv_dots_pattern<2, false><<<nb, BS, lds, c.stream>>>(n, rp, pm, po, av, z, t,
                                                                         r, to, part);
        } else {
            if (nt)
                k_spmv_dots_pattern<3, true><<<nb, BS, lds, c.stream>>>(n, rp, pm, po, av, z, t,
                                                                        r, to, part);
            else
                k_spmv_dots_pattern<3, false><<<nb, BS, lds, c.stream>>>(n, rp, pm, po, av, z, t,
                                                                         r, to, part);
        }
    } else if (A.use_narrow())
        spmv_dots_typed<short>(c.stream, nb, n, k, rp, A.col16.get(), av, A.nnz, A.tile, z, t, r,
                               to, part);
    else
        spmv_dots_typed<int>(c.stream, nb, n, k, rp, A.col.get(), av, A.nnz, A.tile, z, t, r, to,
                             part);
    if (k == 2)
        k_fcg_final<2><<<1, RED_FINAL_BS, 0, c.stream>>>(nb, part, scal);
    else
        k_fcg_final<3><<<1, RED_FINAL_BS, 0, c.stream>>>(nb, part, scal);
    BCMG_CHECK_LAUNCH();
}

} // namespace bcmg
