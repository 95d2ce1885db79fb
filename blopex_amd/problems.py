"""Test-problem builders for the LOBPCG path (numpy only, host side).

The reference has no CSR operator and no variable-coefficient generator; its
7-point convention (diagonal = sum of the face coefficients, off-diagonal =
-face coefficient) is the one GenerateLaplacian is driven with in
blopex_hypre/test/ij.c:3853-3873.  The builders here fix the remaining
conventions:

* unknown (x, y, z) of an nx*ny*nz grid has index x + nx*(y + ny*z);
* every cell carries a coefficient c; the face between two cells gets the
  harmonic mean 2*c_a*c_b/(c_a + c_b); a face on the Dirichlet boundary gets the
  cell's own c (so with c == 1 the matrix is the constant 7-point stencil with
  diagonal 6 and off-diagonals -1);
* each CSR row lists its columns in ascending order;
* per-cell c = exp(sigma * z) with z standard normal from a counter-based
  generator (splitmix64 of (seed, cell), Box-Muller), so any cell's value is
  reproducible without generating the others.
"""
from __future__ import annotations

import math

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _uniform01(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits for counter idx."""
    with np.errstate(over="ignore"):
        key = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0x8CB92BA72F3D8DD7) & 0xFFFFFFFFFFFFFFFF)
        u = _splitmix64(idx.astype(np.uint64) * np.uint64(2) + key)
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def standard_normal(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """count standard normal values, Box-Muller over two counter-based uniforms."""
    idx = np.arange(count, dtype=np.uint64)
    u1 = _uniform01(seed, 2 * stream, idx)
    u2 = _uniform01(seed, 2 * stream + 1, idx)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def lognormal_coefficients(ncells: int, sigma: float, seed: int, stream: int = 0) -> np.ndarray:
    """Per-cell c ~ lognormal(0, sigma)."""
    if sigma == 0.0:
        return np.ones(ncells)
    return np.exp(sigma * standard_normal(seed, ncells, stream))


def uniform(seed: int, count: int, lo: float, hi: float, stream: int = 0) -> np.ndarray:
    idx = np.arange(count, dtype=np.uint64)
    return lo + (hi - lo) * _uniform01(seed, 1000 + stream, idx)


class CSR:
    """Square CSR matrix: int32 row pointers / column indices, float64 values."""

    def __init__(self, rowptr: np.ndarray, colind: np.ndarray, val: np.ndarray):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def diagonal(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rowptr))
        d = np.zeros(self.n)
        m = rows == self.colind
        d[rows[m]] = self.val[m]
        return d

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.colind, self.rowptr), shape=(self.n, self.n))

    def matvec(self, x: np.ndarray) -> np.ndarray:
        return self.to_scipy() @ x


def _face(ca: np.ndarray, cb: np.ndarray) -> np.ndarray:
    return 2.0 * ca * cb / (ca + cb)


def diffusion_stencil(dims: tuple[int, ...], coef: np.ndarray | None = None) -> CSR:
    """Cell-centred diffusion operator on a 2D or 3D grid, Dirichlet boundary."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    c = np.ones(n) if coef is None else np.asarray(coef, dtype=np.float64).reshape(n)
    # grid array with x fastest: shape (..., ny, nx)
    shape = dims[::-1]
    cg = c.reshape(shape)
    diag = np.zeros(shape)
    strides = [int(np.prod(shape[ax + 1:])) for ax in range(len(shape))]
    entries = []  # (flattened column offset, value over all cells, row has this entry)
    for ax in range(len(shape)):
        for step in (-1, 1):
            nb = np.roll(cg, -step, axis=ax)  # nb[i] = c[i + step] along ax
            valid = np.ones(shape, dtype=bool)
            sl = [slice(None)] * len(shape)
            sl[ax] = 0 if step == -1 else -1
            valid[tuple(sl)] = False
            f = np.where(valid, _face(cg, nb), cg)
            diag += f
            entries.append((step * strides[ax], -f, valid))
    entries.append((0, diag, np.ones(shape, dtype=bool)))
    entries.sort(key=lambda e: e[0])
    cnt = np.zeros(n, dtype=np.int64)
    for _, _, valid in entries:
        cnt += valid.reshape(n)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt, out=rowptr[1:])
    nnz = int(rowptr[-1])
    colind = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz)
    pos = rowptr[:-1].copy()
    rows = np.arange(n, dtype=np.int64)
    for off, v, valid in entries:
        m = valid.reshape(n)
        r = rows[m]
        colind[pos[m]] = (r + off).astype(np.int32)
        val[pos[m]] = v.reshape(n)[m]
        pos[m] += 1
    return CSR(rowptr.astype(np.int32), colind, val)


def laplacian_7pt(nx: int, ny: int, nz: int, sigma: float = 0.0, seed: int = 0) -> CSR:
    c = None if sigma == 0.0 else lognormal_coefficients(nx * ny * nz, sigma, seed)
    return diffusion_stencil((nx, ny, nz), c)


def laplacian_5pt(nx: int, ny: int, sigma: float = 0.0, seed: int = 0, stream: int = 0) -> CSR:
    c = None if sigma == 0.0 else lognormal_coefficients(nx * ny, sigma, seed, stream)
    return diffusion_stencil((nx, ny), c)


def batched_problems(count: int, smin: int = 32, smax: int = 256, ks=(4, 8, 16),
                     sigma: float = 0.5, seed: int = 4) -> list[tuple[CSR, int]]:
    """The batched workload: count independent 2D 5-point diffusion problems,
    side s_i uniform in [smin, smax], cell coefficients lognormal(0, sigma)
    (problem i draws from stream i), block size k_i uniform over ks.  Sides
    and block sizes come from the counter-based generator (streams 500, 501)."""
    idx = np.arange(count, dtype=np.uint64)
    us = _uniform01(seed, 500, idx)
    uk = _uniform01(seed, 501, idx)
    out = []
    for i in range(count):
        sd = smin + int(us[i] * (smax - smin + 1))
        k = ks[int(uk[i] * len(ks))]
        out.append((laplacian_5pt(sd, sd, sigma=sigma, seed=seed, stream=1 + i), k))
    return out


def dirichlet_eigenvalues(dims: tuple[int, ...], m: int) -> np.ndarray:
    """Smallest m eigenvalues of the constant-coefficient Dirichlet stencil:
    sum over axes of 4 sin^2(pi j / (2 (N + 1))), j = 1..N
    (blopex_tools/matlab/laplacian/laplacian.m:253-327)."""
    lam = np.zeros(1)
    for N in dims:
        l1 = 4.0 * np.sin(np.pi * np.arange(1, N + 1) / (2.0 * (N + 1))) ** 2
        lam = (lam[:, None] + l1[None, :]).reshape(-1)
        lam = np.sort(lam)[: max(m, 1) * 4 + 8] if lam.size > 4 * m + 8 else lam
    return np.sort(lam)[:m]


def dense_to_csr(a: np.ndarray) -> CSR:
    import scipy.sparse as sp

    s = sp.csr_matrix(a)
    s.sort_indices()
    return CSR(s.indptr, s.indices, s.data)


# --------------------------------------------------------------------------
# drand48 stream (POSIX: x' = (0x5DEECE66D x + 0xB) mod 2^48, seed -> (seed<<16)|0x330E,
# value x'/2^48), reproduced in bulk by jumping ahead in blocks.
# --------------------------------------------------------------------------
_A48 = 0x5DEECE66D
_C48 = 0xB
_M48 = (1 << 48) - 1


def drand48_stream(seed: int, count: int, block: int = 1 << 14) -> np.ndarray:
    """First `count` values of drand48() after srand48(seed)."""
    # affine maps x -> A_t x + C_t for t = 1..block
    A = np.empty(block, dtype=np.uint64)
    C = np.empty(block, dtype=np.uint64)
    a, c = 1, 0
    for t in range(block):
        a, c = (_A48 * a) & _M48, (_A48 * c + _C48) & _M48
        A[t], C[t] = a, c
    a_blk, c_blk = int(A[-1]), int(C[-1])
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    out = np.empty(count)
    m48 = np.uint64(_M48)
    with np.errstate(over="ignore"):
        for s in range(0, count, block):
            e = min(count, s + block)
            st = (A[: e - s] * np.uint64(x) + C[: e - s]) & m48
            out[s:e] = st.astype(np.float64) * (1.0 / 281474976710656.0)
            x = (a_blk * x + c_blk) & _M48
    return out


def drand48_block(n: int, k: int, seed: int) -> np.ndarray:
    """X0 of serial_Multi_VectorSetRandomValues (multi_vector.c:166-184):
    column-major n x k, 2*drand48()-1, column by column."""
    v = drand48_stream(seed, n * k)
    return np.asfortranarray((2.0 * v - 1.0).reshape(k, n).T)
