"""Synthetic sparse operators and seeded right-hand sides.

All operators are returned as CSR triples (row_ptr int32[n+1], col int32[nnz],
val float64[nnz]) with 0-based indices, columns ascending and an explicit
diagonal in every row, the input format the solver requires (the reference
reads the same layout: src/matrix/CSR.h:13-28, and needs sorted columns with a
diagonal: src/AMG/matching.cu:216-218).

Dirichlet boundary rows are eliminated, so a boundary cell keeps its full
diagonal and loses the couplings that leave the grid.

Random numbers come from SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), vectorised: draw number k
(k = 0, 1, ...) of stream `seed` is mix(seed + (k + 1) * 0x9E3779B97F4A7C15),
u = (z >> 11) * 2^-53 in [0, 1). Uniform[-1, 1] = 2u - 1. Normal variates use
Box-Muller on the draw pairs (2c, 2c + 1).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """Return `count` raw 64-bit SplitMix64 outputs of stream `seed`."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count, start=0):
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_pm1(seed, n):
    """n draws of Uniform[-1, 1] from stream `seed`."""
    return 2.0 * uniform01(seed, n) - 1.0


def normal(seed, n, sigma=1.0):
    """n draws of N(0, sigma^2) by Box-Muller from stream `seed`."""
    u = uniform01(seed, 2 * n)
    u1 = 1.0 - u[0::2]                       # (0, 1]
    u2 = u[1::2]
    return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def _assemble(n, offsets, vals, masks):
    """Stack per-offset diagonals (each length n, offsets ascending) into CSR."""
    rows = np.arange(n, dtype=np.int64)
    cols = np.stack([rows + o for o in offsets], axis=1)
    v = np.stack(vals, axis=1)
    m = np.stack(masks, axis=1)
    counts = m.sum(axis=1)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    if row_ptr[-1] >= 2 ** 31:
        raise ValueError("nnz does not fit int32 indices")
    col = cols[m].astype(np.int32)
    val = np.ascontiguousarray(v[m], dtype=np.float64)
    return row_ptr.astype(np.int32), col, val


def poisson3d(N):
    """7-point Laplacian on an N^3 grid: diagonal 6, off-diagonal -1."""
    n = N ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = (idx // N) % N
    z = idx // (N * N)
    one = np.ones(n, dtype=bool)
    offs = [-N * N, -N, -1, 0, 1, N, N * N]
    masks = [z > 0, y > 0, x > 0, one, x < N - 1, y < N - 1, z < N - 1]
    neg = -np.ones(n)
    vals = [neg, neg, neg, np.full(n, 6.0), neg, neg, neg]
    return _assemble(n, offs, vals, masks)


def aniso2d(N, eps=0.01, cx=1.0):
    """5-point Laplacian on an N^2 grid, coupling cx in x and eps in y.
    Diagonal 2 cx + 2 eps (Dirichlet, eliminated)."""
    n = N * N
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = idx // N
    one = np.ones(n, dtype=bool)
    offs = [-N, -1, 0, 1, N]
    masks = [y > 0, x > 0, one, x < N - 1, y < N - 1]
    vals = [np.full(n, -eps), np.full(n, -cx), np.full(n, 2 * cx + 2 * eps),
            np.full(n, -cx), np.full(n, -eps)]
    return _assemble(n, offs, vals, masks)


def poisson2d(N):
    return aniso2d(N, eps=1.0, cx=1.0)


def lognormal3d(N, sigma2=4.0, seed=0):
    """Cell-centred 7-point diffusion on an N^3 grid with cell coefficient
    k = exp(g), g ~ N(0, sigma2) i.i.d. per cell (stream `seed`). Interior
    faces use the harmonic mean 2 k1 k2 / (k1 + k2); a boundary face uses the
    cell's own k. Row i: diagonal = sum of its 6 face coefficients,
    off-diagonal = -face coefficient."""
    n = N ** 3
    k = np.exp(normal(seed, n, sigma=np.sqrt(sigma2)))
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = (idx // N) % N
    z = idx // (N * N)

    def face(mask, off):
        nb = np.where(mask, idx + off, idx)
        k2 = k[nb]
        h = 2.0 * k * k2 / (k + k2)
        return np.where(mask, h, k)          # boundary face: the cell's k

    specs = [(z > 0, -N * N), (y > 0, -N), (x > 0, -1),
             (x < N - 1, 1), (y < N - 1, N), (z < N - 1, N * N)]
    faces = [face(m, o) for m, o in specs]
    diagv = faces[0] + faces[1] + faces[2] + faces[3] + faces[4] + faces[5]
    one = np.ones(n, dtype=bool)
    offs = [-N * N, -N, -1, 0, 1, N, N * N]
    masks = [specs[0][0], specs[1][0], specs[2][0], one, specs[3][0], specs[4][0], specs[5][0]]
    vals = [-faces[0], -faces[1], -faces[2], diagv, -faces[3], -faces[4], -faces[5]]
    # interior faces are symmetric by construction: the harmonic mean is
    # symmetric in (k1, k2) bit for bit because + and * commute in IEEE-754.
    return _assemble(n, offs, vals, masks)


def to_scipy(A):
    import scipy.sparse as sp
    row, col, val = A
    n = len(row) - 1
    return sp.csr_matrix((val, col, row), shape=(n, n))
