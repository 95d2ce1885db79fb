"""GPU checks of the eigenpair adjoint (hip_LobpcgAdjoint) through the C ABI.

* Against the header's formulas restated densely (tests/adjoint_harness.py,
  themselves checked against finite differences on the CPU): the same X,
  lambda, dL/dlambda and dL/dX give the same gradients of the stored values
  of A, of the diagonal of B and the same adjoint vectors -- to 1e-9 relative
  in FP64, 1e-7 in MIXED (CG tolerance 1e-12 / 1e-10 relative).
* The differentiation workload: loss = sum of the 8 smallest eigenvalues +
  ||x_1||^2 on a central sub-box of a 7-point lognormal(0, 1) diffusion
  operator, gradients with respect to every per-cell coefficient (device
  solve, adjoint, torch autograd through the coefficient -> value map):
  - 12^3 against torch.linalg.eigh's dense FP64 gradient (1e-6 relative) and
    central differences with step 1e-6 (dense eigh);
  - 64^3 (central 16^3 box) against central differences of device solves
    (relative step 1e-4, tolerance 1e-11: lambda_8 and lambda_9 are 6e-5
    apart, which bounds the step from above, and the eigenvector error over
    the step bounds the quotient's accuracy, so the bar is 1e-3), and FP64
    against MIXED.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl
import torch

from blopex_amd import lobpcg as bl
from blopex_amd import problems as pb
from tests import adjoint_harness as ah

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("B", [None, "diag"])
def test_adjoint_matches_dense_formulas(mode, B):
    say(f"adjoint vs dense formulas, 8^3, k=6, mode {mode}, B {B}")
    dims = (8, 8, 8)
    n = 512
    A = pb.diffusion_stencil(dims, pb.lognormal_coefficients(n, 1.0, 8))
    Bd = None if B is None else pb.uniform(3, n, 0.5, 1.5)
    S = A.to_scipy().toarray()
    mu, V = sl.eigh(S, None if Bd is None else np.diag(Bd))
    k = 6
    X, lam = V[:, :k], mu[:k]
    rng = np.random.default_rng(1)
    glam = rng.standard_normal(k)
    GX = rng.standard_normal((n, k))
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    gval, gb, W = ah.adjoint_dense(S, Bd, X, lam, glam, GX, rows, A.colind.astype(np.int64))
    tol = 1e-12 if mode == "fp64" else 1e-10
    r = bl.adjoint(A, X, lam, glam, GX, B=Bd, T="jacobi", mode=mode, tol=tol, maxit=5000,
                   want_gb=True, want_W=True)
    say(f"  stats {r.stats}")
    assert r.flag == 0
    bar = 1e-9 if mode == "fp64" else 1e-7
    assert rel(r.W, W) < bar
    assert rel(r.gval, gval) < bar
    assert rel(r.gb, gb) < bar
    if mode == "mixed":
        assert r.stats["corrections"] >= 2


def test_adjoint_eigenvalue_terms_only():
    """GX absent: no solve, gval = sum_i glam_i x_i[r] x_i[c] exactly."""
    say("adjoint, eigenvalue terms only")
    A = pb.laplacian_7pt(6, 5, 4, sigma=1.0, seed=2)
    S = A.to_scipy().toarray()
    lam, V = np.linalg.eigh(S)
    X = V[:, :4]
    glam = np.array([1.0, -2.0, 0.5, 3.0])
    r = bl.adjoint(A, X, lam[:4], glam, None, want_gb=True)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    ref = (X[rows] * X[A.colind]) @ glam
    assert r.stats["iterations"] == 0
    np.testing.assert_allclose(r.gval, ref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(r.gb, (X ** 2) @ (-glam * lam[:4]), rtol=1e-13, atol=1e-15)


def _device_gradient(c, dims, box, k, mode, tol, maxit=3000, X0=None):
    st = ah.stencil_structure(dims)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    val = ah.stencil_values(ct, dims, st)
    opts = dict(k=k, T="jacobi", tol=tol, maxit=maxit, mode=mode, X0=X0)
    lam, X = ah.DeviceEig.apply(val, st[0], opts)
    L = ah.loss_terms(lam, X, torch.from_numpy(box))
    L.backward()
    return (float(L.detach()), ct.grad.numpy().copy(), opts["adjoint_stats"],
            X.detach().numpy())


@pytest.fixture(scope="module")
def dense12():
    dims = (12, 12, 12)
    c = pb.lognormal_coefficients(1728, 1.0, 5)
    box = ah.box_mask(dims, 4, 8)
    L, g = ah.dense_gradient(c, dims, box)
    return dims, c, box, L, g


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_adjoint_12cube_against_dense_eigh_and_fd(dense12, mode):
    dims, c, box, Lref, gref = dense12
    say(f"adjoint 12^3 (k=12 computed, loss on 8 + x_1 box), mode {mode}")
    L, g, stats, _ = _device_gradient(c, dims, box, 12, mode, 1e-12)
    say(f"  loss {L:.15g} (dense {Lref:.15g}); adjoint {stats}")
    assert abs(L - Lref) < 1e-10 * abs(Lref)
    assert rel(g, gref) < 1e-6, rel(g, gref)
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(2):
        v = rng.standard_normal(c.size) * c
        fd = (ah.dense_loss(c + h * v, dims, box) - ah.dense_loss(c - h * v, dims, box)) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * abs(fd), (fd, g @ v)


def _smooth_directions(N):
    x = (np.arange(N) + 0.5) / N
    s1, s2, c1 = np.sin(np.pi * x), np.sin(2 * np.pi * x), np.cos(np.pi * x)
    v1 = np.einsum("z,y,x->zyx", s1, s1, s1).reshape(-1)
    v2 = np.einsum("z,y,x->zyx", c1, s1, s2).reshape(-1)
    return [v1, v2]


@pytest.fixture(scope="module")
def cube64():
    N = 64
    dims = (N, N, N)
    c = pb.lognormal_coefficients(N ** 3, 1.0, 64)
    box = ah.box_mask(dims, 24, 40)
    return dims, c, box, {}


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_adjoint_64cube_against_finite_differences(cube64, mode):
    dims, c, box, memo = cube64
    say(f"adjoint 64^3 (k=11, loss on 8 + x_1 on the central 16^3), mode {mode}")
    # 11 columns: lambda_11 = 0.0272 and lambda_12 = 0.0316 are well apart,
    # so the block edge converges quickly
    L, g, stats, X = _device_gradient(c, dims, box, 11, mode, 1e-11)
    say(f"  loss {L:.15g}; adjoint {stats}")
    memo[mode] = g
    if "fd" not in memo:
        h = 1e-4
        fds = []
        for v in _smooth_directions(dims[0]):
            lv = []
            for sgn in (1, -1):
                A = pb.diffusion_stencil(dims, c * (1 + sgn * h * v))
                r = bl.lobpcg(np.ascontiguousarray(X), A, T="jacobi", tol=1e-11, maxit=3000,
                              final_rr=True)
                assert r.flag == 0
                lv.append(r.lam[:8].sum() + np.sum(r.X[box, 0] ** 2))
            fds.append((lv[0] - lv[1]) / (2 * h))
        memo["fd"] = fds
    for v, fd in zip(_smooth_directions(dims[0]), memo["fd"]):
        dd = g @ (c * v)
        say(f"  directional derivative {dd:.10g}, central difference {fd:.10g}")
        assert abs(fd - dd) <= 1e-3 * abs(fd)
    if "fp64" in memo and "mixed" in memo:
        assert rel(memo["mixed"], memo["fp64"]) < 1e-6


def test_adjoint_breakdown_is_reported():
    """X skips the lowest eigenvector: on the complement of X the shifted
    operator A - lambda_i B is indefinite (lambda_1 - lambda_i < 0), the CG
    meets p'(A - lambda B)p <= 0 and the call returns -1 with a message."""
    say("adjoint breakdown: X = eigenvectors 2..4 of an 8^3 operator")
    dims = (8, 8, 8)
    A = pb.diffusion_stencil(dims, pb.lognormal_coefficients(512, 1.0, 8))
    mu, V = np.linalg.eigh(A.to_scipy().toarray())
    X, lam = V[:, 1:4], mu[1:4]
    GX = np.random.default_rng(3).standard_normal((512, 3))
    with pytest.raises(RuntimeError, match="breakdown in column"):
        bl.adjoint(A, X, lam, np.ones(3), GX, mode="fp64", tol=1e-12, maxit=2000)


@pytest.mark.parametrize("maxit", [5, 11])
def test_adjoint_maxit_exit(maxit):
    """maxit runs out before the tolerance (at iterations that are not
    multiples of the host's check interval): flag 1 and exactly maxit CG
    iterations reported."""
    say(f"adjoint maxit exit, maxit {maxit}")
    dims = (10, 10, 10)
    A = pb.diffusion_stencil(dims, pb.lognormal_coefficients(1000, 1.0, 9))
    mu, V = np.linalg.eigh(A.to_scipy().toarray())
    k = 4
    GX = np.random.default_rng(4).standard_normal((1000, k))
    r = bl.adjoint(A, V[:, :k], mu[:k], np.ones(k), GX, mode="fp64", tol=1e-14, maxit=maxit)
    say(f"  stats {r.stats}")
    assert r.flag == 1
    assert r.stats["iterations"] == maxit


def test_adjoint_counts_only_running_iterations():
    """The iteration count is that of the iterations in which some column
    still ran: a converged solve reports the same count whatever the host's
    check interval, and fewer than the check interval past the last column's
    finish (the count of a run with a cap just below it fails to converge)."""
    say("adjoint iteration count, 10^3 FP64")
    dims = (10, 10, 10)
    A = pb.diffusion_stencil(dims, pb.lognormal_coefficients(1000, 1.0, 9))
    mu, V = np.linalg.eigh(A.to_scipy().toarray())
    k = 4
    GX = np.random.default_rng(4).standard_normal((1000, k))
    r = bl.adjoint(A, V[:, :k], mu[:k], np.ones(k), GX, mode="fp64", tol=1e-10, maxit=2000)
    assert r.flag == 0
    it = r.stats["iterations"]
    r2 = bl.adjoint(A, V[:, :k], mu[:k], np.ones(k), GX, mode="fp64", tol=1e-10, maxit=it - 1)
    say(f"  converged in {it}; capped at {it - 1}: flag {r2.flag}")
    assert r2.flag == 1
