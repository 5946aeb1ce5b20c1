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

template <typename MT, typename VT, int OP>
void rows_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *x, const VT *f,
                const VT *d, VT *y)
{
    const int n = A.n;
    const int *rp = A.row.get(), *ci = A.col.get();
    if (stream_ok<VT>(A)) {
        const int nb = (n + BS - 1) / BS;
        /* the narrow columns where the int32-column stream is non-temporal */
        const bool use16 = A.narrow && (size_t)A.nnz * (4 + sizeof(MT)) > NT_BYTES;
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

template <typename MT, typename VT>
void relax_block_typed(hipStream_t s, const DevCsr &A, const MT *av, const VT *d, const VT *f,
                       const VT *xin, VT *xout, int sweeps, int z, int ncols, long cs,
                       int rs, const int *act)
{
    const size_t lds = (size_t)2 * A.n * sizeof(VT);
    const int *rp = A.row.get(), *ci = A.col.get();
    const int n = A.n;
    if (n <= 1024 && A.maxrow <= 8)
        k_relax_block<MT, VT, 1, 8><<<ncols, 1024, lds, s>>>(n, rp, ci, av, d, f, xin, xout,
                                                              sweeps, z, cs, rs, act);
    else if (n <= 4096 && A.maxrow <= 8)
        k_relax_block<MT, VT, 4, 8><<<ncols, 1024, lds, s>>>(n, rp, ci, av, d, f, xin, xout,
                                                              sweeps, z, cs, rs, act);
    else if (n <= 2048 && A.maxrow <= 16)
        k_relax_block<MT, VT, 2, 16><<<ncols, 1024, lds, s>>>(n, rp, ci, av, d, f, xin, xout,
                                                               sweeps, z, cs, rs, act);
    else if (n <= 1024 && A.maxrow <= 32)
        k_relax_block<MT, VT, 1, 32><<<ncols, 1024, lds, s>>>(n, rp, ci, av, d, f, xin, xout,
                                                               sweeps, z, cs, rs, act);
    else
        throw Error(BCMG_EINVAL, "relax_block: level too large");
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


void make_narrow(Ctx &c, const std::vector<DevCsr *> &mats)
{
    std::vector<DevCsr *> cand;
    for (DevCsr *A : mats)
        if (A->n == A->m && (size_t)A->nnz * 12 > NT_BYTES && !A->narrow) cand.push_back(A);
    if (cand.empty()) return;
    const int nc = (int)cand.size();
    DBuf<int> wide(nc);
    k_fill<int><<<(nc + BS - 1) / BS, BS, 0, c.stream>>>(nc, 0, wide.get());
    for (int k = 0; k < nc; k++) {
        DevCsr &A = *cand[k];
        A.col16.alloc(A.nnz);
        k_narrow_cols<<<(A.n + BS - 1) / BS, BS, 0, c.stream>>>(A.n, A.row.get(), A.col.get(),
                                                                A.col16.get(), wide.get() + k);
    }
    BCMG_CHECK_LAUNCH();
    std::vector<int> h(nc);
    BCMG_HIP(hipMemcpyAsync(h.data(), wide.get(), nc * sizeof(int), hipMemcpyDeviceToHost,
                            c.stream));
    BCMG_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < nc; k++) {
        cand[k]->narrow = (h[k] == 0);
        if (!cand[k]->narrow) cand[k]->col16.release();
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
                        int rs, const int *act)
{
    const int z = zero ? 1 : 0;
    if (ncols <= 0) return;
    if constexpr (std::is_same<VT, double>::value) {
        relax_block_typed<double, double>(s, A, A.val.get(), d, f, xin, xout, sweeps, z, ncols,
                                          cs, rs, act);
    } else {
        if (mt == MT_F32)
            relax_block_typed<float, float>(s, A, A.val32.get(), d, f, xin, xout, sweeps, z,
                                            ncols, cs, rs, act);
        else if (mt == MT_BF16)
            relax_block_typed<unsigned short, float>(s, A, A.val16.get(), d, f, xin, xout,
                                                     sweeps, z, ncols, cs, rs, act);
        else
            relax_block_typed<double, float>(s, A, A.val.get(), d, f, xin, xout, sweeps, z,
                                             ncols, cs, rs, act);
    }
    BCMG_CHECK_LAUNCH();
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
                                         int, const int *);                                  \
    template void launch_jacobi_zero<VT>(hipStream_t, int, const VT *, const VT *, VT *);    \
    template void launch_prolong<VT>(hipStream_t, int, const int *, const VT *, const VT *,  \
                                     VT *);                                                  \
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
    if (A.narrow)
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
