"""CPU checks of what the eigenpair adjoint tests compare against (no GPU):
the torch restatement of the coefficient -> CSR value map, the dense torch
eigh gradient against central finite differences, and the adjoint formulas of
include/lobpcg_adjoint.h (restated densely in tests/adjoint_harness.py)
against finite differences of the eigenpairs -- for A and for a diagonal B --
and against the torch gradient of the differentiation loss."""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

from blopex_amd import problems as pb
from tests import adjoint_harness as ah


@pytest.mark.parametrize("dims", [(5, 4, 3), (6, 5)])
def test_stencil_values_match_numpy_builder(dims):
    n = int(np.prod(dims))
    c = pb.lognormal_coefficients(n, 1.0, 3)
    ref = pb.diffusion_stencil(dims, c)
    got = ah.stencil_values(torch.from_numpy(c), dims).numpy()
    np.testing.assert_array_equal(got, ref.val)


def test_dense_torch_gradient_against_finite_differences():
    """torch.linalg.eigh's gradient of sum(lambda_1..8) + ||x_1 on a box||^2
    against central differences with step 1e-6 (numpy eigh) along three
    random directions, 5^3 grid."""
    dims = (5, 5, 5)
    n = 125
    c = pb.lognormal_coefficients(n, 1.0, 11)
    box = ah.box_mask(dims, 1, 4)
    _, g = ah.dense_gradient(c, dims, box)
    rng = np.random.default_rng(0)
    h = 1e-6
    for _ in range(3):
        v = rng.standard_normal(n) * c
        fd = (ah.dense_loss(c + h * v, dims, box) - ah.dense_loss(c - h * v, dims, box)) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * abs(fd) + 1e-10, (fd, g @ v)


def _pencil(seed):
    dims = (5, 5, 4)
    A = pb.diffusion_stencil(dims, pb.lognormal_coefficients(100, 1.0, seed))
    Bd = pb.uniform(seed, 100, 0.5, 1.5)
    return A, Bd


def test_adjoint_formulas_against_finite_differences():
    """gval and gb of the header's formulas give the first-order change of
    L = glam . lambda + sum_i g_i . x_i under a symmetric perturbation of the
    stored values of A and a perturbation of the diagonal of B (generalized
    pencil, k = 5, eigenvector signs aligned with the unperturbed ones)."""
    A, Bd = _pencil(4)
    S = A.to_scipy().toarray()
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    cols = A.colind.astype(np.int64)
    k = 5
    mu, V = sl.eigh(S, np.diag(Bd))
    X, lam = V[:, :k], mu[:k]
    rng = np.random.default_rng(5)
    glam = rng.standard_normal(k)
    GX = rng.standard_normal((A.n, k))
    gval, gb, _ = ah.adjoint_dense(S, Bd, X, lam, glam, GX, rows, cols)

    def L(Sp, Bp):
        m, U = sl.eigh(Sp, np.diag(Bp))
        U = U[:, :k] * np.sign(np.sum(U[:, :k] * (Bp[:, None] * X), axis=0))
        return glam @ m[:k] + np.sum(GX * U)

    h = 1e-6
    E = np.zeros_like(S)
    E[rows, cols] = rng.standard_normal(len(rows))
    E = E + E.T
    fd = (L(S + h * E, Bd) - L(S - h * E, Bd)) / (2 * h)
    assert abs(fd - gval @ E[rows, cols]) <= 1e-6 * abs(fd), (fd, gval @ E[rows, cols])
    dB = rng.standard_normal(A.n)
    fd = (L(S, Bd + h * dB) - L(S, Bd - h * dB)) / (2 * h)
    assert abs(fd - gb @ dB) <= 1e-6 * abs(fd), (fd, gb @ dB)


def test_adjoint_formulas_against_torch_eigh_gradient():
    """Chained through the coefficient -> value map, the header's formulas
    (k = 10 computed pairs, loss on the first 8 and on x_1) give torch eigh's
    coefficient gradient, 6^3 grid."""
    dims = (6, 6, 6)
    n = 216
    c = pb.lognormal_coefficients(n, 1.0, 21)
    box = ah.box_mask(dims, 2, 4)
    _, gref = ah.dense_gradient(c, dims, box)
    A = pb.diffusion_stencil(dims, c)
    S = A.to_scipy().toarray()
    lam, V = np.linalg.eigh(S)
    k = 10
    X = V[:, :k]
    glam = np.r_[np.ones(8), np.zeros(k - 8)]
    GX = np.zeros((n, k))
    GX[:, 0] = 2.0 * X[:, 0] * box
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    gval, _, _ = ah.adjoint_dense(S, None, X, lam[:k], glam, GX, rows, A.colind.astype(np.int64))
    ct = torch.tensor(c, requires_grad=True)
    val = ah.stencil_values(ct, dims)
    (g,) = torch.autograd.grad(val, ct, torch.from_numpy(gval))
    np.testing.assert_allclose(g.numpy(), gref, rtol=1e-8, atol=1e-10 * np.abs(gref).max())
