/*
 * kernels_solve.hip - solve-phase kernels for gfx950 (wave64, CDNA4).
 *
 * Everything here is HBM-bound CSR/vector streaming; there is no matrix-core
 * work. Rows are processed by groups of G lanes (G a power of two chosen from
 * the average row length, so a 7-point row uses 8 lanes and eight rows share
 * one wave64); the group reduces with xor-shuffles. Blocks are remapped so
 * that each of the 8 XCDs streams one contiguous slice of the rows and keeps
 * the gathered x entries of its slice in its own L2.
 *
 * Reductions (dot products, norms) are deterministic: a fixed grid that
 * depends only on n, a fixed in-block tree and a fixed final pass; no
 * floating-point atomics (the reference adds block results with atomicAdd,
 * src/SOLVER/FCG.cu:58-60, so its sums vary run to run).
 */
#include "bcmg_internal.h"

namespace bcmg {

namespace {

constexpr int BS = 256;          /* threads per block */
constexpr int RED_MAX_BLOCKS = 1024;
constexpr int STREAM_TILE_MAX = 4096;  /* doubles of LDS per block (32 KB) */

/* Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
   Return the logical block index so that XCD x owns a contiguous range. */
__device__ __forceinline__ int xcd_block(int bid, int nb)
{
    const int q = nb >> 3, rm = nb & 7;
    const int x = bid & 7, l = bid >> 3;
    return x * q + (x < rm ? x : rm) + l;
}

/* Row-block kernel: block b owns rows [256 b, 256 b + 256). Phase 1 streams
   the block's slice of val/col with consecutive lanes on consecutive entries
   (fully coalesced) and stores the products a_ij x_j in LDS; phase 2 gives one
   row to each thread, which sums its products in column order and applies the
   epilogue with coalesced loads of f, d, x and a coalesced store. The sum
   order (products rounded, then added left to right) is the serial one. */
template <int OP>
__global__ __launch_bounds__(BS) void k_rows_stream(int n, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ av,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ f,
                                                    const double *__restrict__ d,
                                                    double *__restrict__ y)
{
    extern __shared__ double prod[];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * BS;
    const int r1 = min(r0 + BS, n);
    const int e0 = rp[r0];
    const int cnt = rp[r1] - e0;
    constexpr int U = 4;
    for (int k0 = threadIdx.x; k0 < cnt; k0 += BS * U) {
        int c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u * BS;
            if (k < cnt) {
                c[u] = ci[e0 + k];
                v[u] = av[e0 + k];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + u * BS;
            if (k < cnt) prod[k] = v[u] * x[c[u]];
        }
    }
    __syncthreads();
    const int r = r0 + threadIdx.x;
    if (r >= r1) return;
    const int a = rp[r] - e0, e = rp[r + 1] - e0;
    double t = 0.0;
    for (int k = a; k < e; k++) t = t + prod[k];
    if (OP == OP_SPMV)
        y[r] = t;
    else if (OP == OP_RESID)
        y[r] = f[r] - t;
    else
        y[r] = (f[r] - t) / d[r] + x[r];
}

/* largest nnz of any 256-row block (the LDS tile of k_rows_stream) */
__global__ __launch_bounds__(BS) void k_block_nnz(int n, const int *__restrict__ rp,
                                                  int *__restrict__ maxv)
{
    const int b = blockIdx.x * BS + threadIdx.x;
    const long long r0 = (long long)b * BS;
    if (r0 >= n) return;
    const int r1 = (int)min(r0 + BS, (long long)n);
    atomicMax(maxv, rp[r1] - rp[r0]);
}

template <int G, int OP>
__global__ __launch_bounds__(BS) void k_rows(int n, const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const double *__restrict__ av,
                                             const double *__restrict__ x,
                                             const double *__restrict__ f,
                                             const double *__restrict__ d,
                                             double *__restrict__ y)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tid = b * BS + threadIdx.x;
    const int row = tid / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= n) return;
    const int j0 = rp[row], j1 = rp[row + 1];
    double t = 0.0;
    for (int j = j0 + lane; j < j1; j += G) t = fma(av[j], x[ci[j]], t);
#pragma unroll
    for (int o = G >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, G);
    if (lane == 0) {
        if (OP == OP_SPMV)
            y[row] = t;
        else if (OP == OP_RESID)
            y[row] = f[row] - t;
        else
            y[row] = (f[row] - t) / d[row] + x[row];
    }
}

__global__ __launch_bounds__(BS) void k_jacobi_zero(int n, const double *__restrict__ f,
                                                    const double *__restrict__ d,
                                                    double *__restrict__ y)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = f[i] / d[i];
}

__global__ __launch_bounds__(BS) void k_prolong(int n, const int *__restrict__ agg,
                                                const double *__restrict__ p,
                                                const double *__restrict__ xc,
                                                double *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = fma(p[i], xc[agg[i]], x[i]);
}

__global__ __launch_bounds__(BS) void k_fill(int n, double v, double *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ __launch_bounds__(BS) void k_copy(int n, const double *__restrict__ a,
                                             double *__restrict__ b)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(BS) void k_axpy(int n, double alpha, const double *__restrict__ x,
                                             double *__restrict__ y)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = fma(alpha, x[i], y[i]);
}

__global__ __launch_bounds__(BS) void k_scale(int n, double alpha, double *__restrict__ x)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = alpha * x[i];
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

/* FCG v3 scalar recurrence (src/SOLVER/FCG.cu:531-534):
   theta = gamma / delta; delta = beta - gamma^2 / delta; theta2 = alpha / delta */
__global__ void k_fcg_scalars(double *scal)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double alpha = scal[0], beta = scal[1], gamma = scal[2];
        double delta = scal[3];
        const double theta = gamma / delta;
        delta = beta - (gamma * gamma) / delta;
        scal[3] = delta;
        scal[4] = theta;
        scal[5] = alpha / delta;
    }
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

/* first FCG step (src/SOLVER/FCG.cu:475-486): delta = beta, theta = alpha/delta,
   x += theta v, r -= theta w */
__global__ __launch_bounds__(BS) void k_fcg_first(int n, double *__restrict__ scal,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ x,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ r,
                                                  double *__restrict__ part)
{
    const double alpha = scal[0], beta = scal[1];
    const double theta = alpha / beta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[3] = beta;
        scal[4] = theta;
    }
    double s[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        x[i] = fma(theta, v[i], x[i]);
        const double ri = fma(-theta, w[i], r[i]);
        r[i] = ri;
        s[0] = fma(ri, ri, s[0]);
    }
    block_reduce_store<1>(s, part + blockIdx.x);
}

inline int blocks_for(long long work) { return (int)((work + BS - 1) / BS); }

inline int red_blocks(int n)
{
    int nb = (int)(((long long)n + BS * 4 - 1) / (BS * 4));
    if (nb < 1) nb = 1;
    if (nb > RED_MAX_BLOCKS) nb = RED_MAX_BLOCKS;
    return nb;
}

template <int OP>
void rows_dispatch(hipStream_t s, const DevCsr &A, const double *x, const double *f,
                   const double *d, double *y)
{
    const int n = A.n;
    if (A.tile > 0 && A.tile <= STREAM_TILE_MAX) {
        const int nb = (n + BS - 1) / BS;
        k_rows_stream<OP><<<nb, BS, (size_t)A.tile * sizeof(double), s>>>(
            n, A.row.get(), A.col.get(), A.val.get(), x, f, d, y);
        return;
    }
    const int G = A.group;
    const int nb = blocks_for((long long)n * G);
    const int *rp = A.row.get(), *ci = A.col.get();
    const double *av = A.val.get();
    switch (G) {
    case 1: k_rows<1, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 2: k_rows<2, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 4: k_rows<4, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 8: k_rows<8, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 16: k_rows<16, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    case 32: k_rows<32, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    default: k_rows<64, OP><<<nb, BS, 0, s>>>(n, rp, ci, av, x, f, d, y); break;
    }
}

} // namespace

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
    A.tile_dev.ensure(1);
    BCMG_HIP(hipMemsetAsync(A.tile_dev.get(), 0, sizeof(int), c.stream));
    const int nblk = (A.n + BS - 1) / BS;
    if (A.n > 0)
        k_block_nnz<<<blocks_for(nblk), BS, 0, c.stream>>>(A.n, A.row.get(), A.tile_dev.get());
    BCMG_CHECK_LAUNCH();
    A.tile = -1; /* pending until resolve_tiles */
}

void resolve_tiles(Ctx &c, const std::vector<DevCsr *> &mats)
{
    for (size_t b = 0; b < mats.size(); b += 64) {
        const size_t e = std::min(mats.size(), b + 64);
        for (size_t k = b; k < e; k++)
            BCMG_HIP(hipMemcpyAsync(c.h_ints + 64 + (k - b), mats[k]->tile_dev.get(), sizeof(int),
                                    hipMemcpyDeviceToHost, c.stream));
        BCMG_HIP(hipStreamSynchronize(c.stream));
        for (size_t k = b; k < e; k++) {
            const int t = c.h_ints[64 + (k - b)];
            mats[k]->tile = t > 0 ? t : 1;
        }
    }
}

void analyze_rows(Ctx &c, DevCsr &A)
{
    launch_block_nnz(c, A);
    resolve_tiles(c, {&A});
}

void launch_rows(hipStream_t s, const DevCsr &A, RowOp op, const double *x, const double *f,
                 const double *d, double *y)
{
    if (A.n <= 0) return;
    if (op == OP_SPMV)
        rows_dispatch<OP_SPMV>(s, A, x, f, d, y);
    else if (op == OP_RESID)
        rows_dispatch<OP_RESID>(s, A, x, f, d, y);
    else
        rows_dispatch<OP_JACOBI>(s, A, x, f, d, y);
    BCMG_CHECK_LAUNCH();
}

void launch_jacobi_zero(hipStream_t s, int n, const double *f, const double *d, double *y)
{
    if (n <= 0) return;
    k_jacobi_zero<<<blocks_for(n), BS, 0, s>>>(n, f, d, y);
    BCMG_CHECK_LAUNCH();
}

void launch_prolong(hipStream_t s, int n, const int *agg, const double *p, const double *xc,
                    double *x)
{
    if (n <= 0) return;
    k_prolong<<<blocks_for(n), BS, 0, s>>>(n, agg, p, xc, x);
    BCMG_CHECK_LAUNCH();
}

void launch_fill(hipStream_t s, int n, double v, double *x)
{
    if (n <= 0) return;
    k_fill<<<blocks_for(n), BS, 0, s>>>(n, v, x);
    BCMG_CHECK_LAUNCH();
}

void launch_copy(hipStream_t s, int n, const double *a, double *b)
{
    if (n <= 0 || a == b) return;
    k_copy<<<blocks_for(n), BS, 0, s>>>(n, a, b);
    BCMG_CHECK_LAUNCH();
}

void launch_axpy(hipStream_t s, int n, double alpha, const double *x, double *y)
{
    if (n <= 0) return;
    k_axpy<<<blocks_for(n), BS, 0, s>>>(n, alpha, x, y);
    BCMG_CHECK_LAUNCH();
}

void launch_scale(hipStream_t s, int n, double alpha, double *x)
{
    if (n <= 0) return;
    k_scale<<<blocks_for(n), BS, 0, s>>>(n, alpha, x);
    BCMG_CHECK_LAUNCH();
}

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
    c.red.ensure((size_t)RED_MAX_BLOCKS * 3);
    k_fcg_update<<<nb, BS, 0, c.stream>>>(n, scal, zo, z, x, to, t, r, c.red.get());
    k_reduce_final<1, true><<<1, BS, 0, c.stream>>>(nb, c.red.get(), scal + 6);
    BCMG_CHECK_LAUNCH();
}

void launch_fcg_first(Ctx &c, int n, double *scal, const double *v, double *x, const double *w,
                      double *r)
{
    const int nb = red_blocks(n);
    c.red.ensure((size_t)RED_MAX_BLOCKS * 3);
    k_fcg_first<<<nb, BS, 0, c.stream>>>(n, scal, v, x, w, r, c.red.get());
    k_reduce_final<1, true><<<1, BS, 0, c.stream>>>(nb, c.red.get(), scal + 6);
    BCMG_CHECK_LAUNCH();
}

} // namespace bcmg
