"""Test harness of the eigenpair adjoint: per-cell coefficients -> CSR values
(in torch, so autograd carries dL/dval back to the coefficients), a
torch.autograd.Function around the device solve and hip_LobpcgAdjoint, and the
dense FP64 references (torch.linalg.eigh, finite differences) they are checked
against.  Used by the tests only."""
from __future__ import annotations

import numpy as np
import torch

from blopex_amd import lobpcg as bl
from blopex_amd import problems as pb


def stencil_structure(dims):
    """CSR pattern of problems.diffusion_stencil on dims, and for each of its
    7 entry groups (offset, valid cells, CSR positions of those entries)."""
    A1 = pb.diffusion_stencil(dims)
    n = A1.n
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(A1.rowptr))
    keys = rows * n + A1.colind.astype(np.int64)
    shape = tuple(dims)[::-1]
    strides = [int(np.prod(shape[ax + 1:])) for ax in range(len(shape))]
    groups = []
    cells = np.arange(n, dtype=np.int64)
    for ax in range(len(shape)):
        for step in (-1, 1):
            valid = np.ones(shape, dtype=bool)
            sl = [slice(None)] * len(shape)
            sl[ax] = 0 if step == -1 else -1
            valid[tuple(sl)] = False
            v = valid.reshape(n)
            r = cells[v]
            pos = np.searchsorted(keys, r * n + r + step * strides[ax])
            groups.append((ax, step, valid, pos))
    dpos = np.searchsorted(keys, cells * n + cells)
    return A1, groups, dpos


def stencil_values(c: torch.Tensor, dims, structure=None) -> torch.Tensor:
    """CSR values of problems.diffusion_stencil(dims, c) as a torch function of c."""
    A1, groups, dpos = structure if structure is not None else stencil_structure(dims)
    n = A1.n
    shape = tuple(dims)[::-1]
    cg = c.reshape(shape)
    val = torch.zeros(A1.nnz, dtype=c.dtype)
    diag = torch.zeros(shape, dtype=c.dtype)
    for ax, step, valid, pos in groups:
        nb = torch.roll(cg, -step, dims=ax)
        vt = torch.from_numpy(valid)
        f = torch.where(vt, 2.0 * cg * nb / (cg + nb), cg)
        diag = diag + f
        val = val.index_put((torch.from_numpy(pos),), -f.reshape(n)[torch.from_numpy(valid.reshape(n))])
    val = val.index_put((torch.from_numpy(dpos),), diag.reshape(n))
    return val


def box_mask(dims, lo, hi) -> np.ndarray:
    """Cells with lo <= coordinate < hi along every axis (x fastest)."""
    shape = tuple(dims)[::-1]
    m = np.zeros(shape, dtype=bool)
    m[tuple(slice(lo, hi) for _ in shape)] = True
    return m.reshape(-1)


class DeviceEig(torch.autograd.Function):
    """(lambda, X) of the device solve as a function of the CSR values of A;
    backward runs hip_LobpcgAdjoint."""

    @staticmethod
    def forward(ctx, val, pattern, opts):
        A = pb.CSR(pattern.rowptr, pattern.colind, val.detach().numpy())
        r = bl.lobpcg(opts.get("X0"), A, T=opts.get("T"), k=opts["k"], tol=opts["tol"],
                      maxit=opts["maxit"], mode=opts["mode"], final_rr=True)
        if r.flag != 0:
            raise RuntimeError(f"forward solve flag {r.flag}")
        ctx.A, ctx.opts = A, opts
        ctx.lam, ctx.X = r.lam.copy(), r.X.copy()
        return torch.from_numpy(r.lam.copy()), torch.from_numpy(np.ascontiguousarray(r.X))

    @staticmethod
    def backward(ctx, glam, gX):
        o = ctx.opts
        ad = bl.adjoint(ctx.A, ctx.X, ctx.lam, glam.numpy(), gX.numpy(), T=o.get("T"),
                        mode=o["mode"], tol=o.get("adj_tol", 1e-10), maxit=o.get("adj_maxit", 5000))
        if ad.flag != 0:
            raise RuntimeError(f"adjoint flag {ad.flag}: {ad.stats}")
        ctx.adjoint_stats = ad.stats
        o["adjoint_stats"] = ad.stats
        return torch.from_numpy(ad.gval), None, None


def loss_terms(lam, X, box, m=8):
    """sum of the smallest m eigenvalues + ||x_1 restricted to box||^2"""
    return lam[:m].sum() + (X[box, 0] ** 2).sum()


def dense_gradient(c: np.ndarray, dims, box, m=8):
    """dL/dc by torch.linalg.eigh on the dense matrix (FP64, CPU)."""
    st = stencil_structure(dims)
    A1 = st[0]
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    val = stencil_values(ct, dims, st)
    rows = np.repeat(np.arange(A1.n, dtype=np.int64), np.diff(A1.rowptr))
    M = torch.zeros((A1.n, A1.n), dtype=torch.float64).index_put(
        (torch.from_numpy(rows), torch.from_numpy(A1.colind.astype(np.int64))), val)
    lam, V = torch.linalg.eigh(M)
    L = loss_terms(lam, V, torch.from_numpy(box), m)
    L.backward()
    return float(L.detach()), ct.grad.numpy().copy()


def dense_loss(c: np.ndarray, dims, box, m=8) -> float:
    A = pb.diffusion_stencil(dims, c).to_scipy().toarray()
    lam, V = np.linalg.eigh(A)
    return float(lam[:m].sum() + np.sum(V[box, 0] ** 2))


def adjoint_dense(A: np.ndarray, Bd: np.ndarray | None, X: np.ndarray, lam: np.ndarray,
                  glam: np.ndarray, GX: np.ndarray, rows: np.ndarray, cols: np.ndarray):
    """The adjoint formulas of include/lobpcg_adjoint.h with dense linear
    algebra: u_i from the full generalized eigendecomposition (the spectral
    sum over the eigenvectors outside X) instead of conjugate gradients.
    Returns gval at the entries (rows, cols), gb and W."""
    import scipy.linalg as sl

    n, k = X.shape
    Bm = np.eye(n) if Bd is None else np.diag(Bd)
    mu, V = sl.eigh(A, Bm)
    C = X.T @ GX
    W = np.zeros((n, k))
    for i in range(k):
        for j in range(k):
            if j != i:
                W[:, i] += X[:, j] * C[j, i] / (lam[j] - lam[i])
        rhs = GX[:, i] - Bm @ X @ (X.T @ GX[:, i])
        # rhs is orthogonal to X; expand in the eigenvectors outside X (X
        # spans the first k of them)
        Vo, muo = V[:, k:], mu[k:]
        W[:, i] += Vo @ ((Vo.T @ rhs) / (muo - lam[i]))
    gval = (X[rows] * X[cols]) @ glam - 0.5 * np.sum(W[rows] * X[cols] + X[rows] * W[cols], axis=1)
    gb = (X ** 2) @ (-glam * lam - 0.5 * np.diag(C)) + np.sum(X * W * lam, axis=1)
    return gval, gb, W
