"""GPU checks of the device Rayleigh-Ritz step (blopex_amd_kernel_rr) against
scipy on the same pencil.

The Gram buffers are built in numpy the way the solver fills them: raw
products of [X R P] with A[X R P] and B[X R P], X a B-orthonormal Ritz basis
(so X'AX = diag(lambda), X'BX = I).  The reference pencil is the one
lobpcg_solve assembles after B-orthonormalising R and P explicitly
(krylov/lobpcg.c:1027-1165), solved with scipy.linalg.eigh (LAPACK
dsygvd).  Bars: Ritz values within 1e-12 relative, Ritz block
B-orthonormal to 1e-12 and spanning the same subspace (projector
difference 1e-9; the eigenvalue gaps of these pencils are >= 1e-3), for both
the tridiagonal and the Jacobi eigensolvers.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl

from blopex_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def make_case(k, a, p, seed, degenerate=False, bdiag=False):
    rng = np.random.default_rng(seed)
    N = 400
    M = rng.standard_normal((N, N))
    A = M @ M.T / N + np.diag(np.linspace(0.5, 3.0, N))
    if degenerate:  # a triple eigenvalue at the bottom
        w, V = np.linalg.eigh(A)
        w[:3] = w[0]
        A = (V * w) @ V.T
    Bd = rng.uniform(0.5, 1.5, N) if bdiag else np.ones(N)
    B = np.diag(Bd)
    Q = rng.standard_normal((N, k))
    lam, V = sl.eigh(Q.T @ A @ Q, Q.T @ B @ Q)
    X = Q @ V
    R = rng.standard_normal((N, a)) * np.logspace(-4, 0, a)
    P = rng.standard_normal((N, p))
    return A, B, X, R, P, lam


def reference(A, B, X, R, P, lam):
    k = X.shape[1]

    def orth(V):
        U = np.linalg.cholesky(V.T @ B @ V).T
        return V @ np.linalg.inv(U)

    Rn = orth(R)
    Pn = orth(P) if P.shape[1] else P
    Bas = np.hstack([X, Rn, Pn])
    GA, GB = Bas.T @ A @ Bas, Bas.T @ B @ Bas
    GA[:k, :k] = np.diag(lam)
    GB[:k, :k] = np.eye(k)
    a, p = R.shape[1], P.shape[1]
    GB[k:k + a, k:k + a] = np.eye(a)
    if p:
        GB[k + a:, k + a:] = np.eye(p)
    ev, Y = sl.eigh(GA, GB)
    return ev[:k], Bas @ Y[:, :k]


def run_kernel(A, B, X, R, P, lam, jacobi, explicit_x=0, retry=0):
    k, a, p = X.shape[1], R.shape[1], P.shape[1]
    Raw = np.hstack([X, R, P])
    n = k + a + p
    lda = n + 3
    ga = np.zeros((lda, n), order="F")
    gb = np.zeros((lda, n), order="F")
    ga[:n] = Raw.T @ A @ Raw
    gb[:n] = Raw.T @ B @ Raw
    lam_out = np.zeros(k)
    coef = np.zeros((n, k), order="F")
    flags = np.zeros(4, dtype=np.int32)
    lam_in = np.ascontiguousarray(lam)
    L.load().blopex_amd_kernel_rr(k, a, p, ga.ctypes.data, gb.ctypes.data, lda, lam_in.ctypes.data,
                                  explicit_x, retry, jacobi, lam_out.ctypes.data, coef.ctypes.data,
                                  flags.ctypes.data)
    return lam_out, Raw @ coef, flags, coef


@pytest.mark.parametrize("jacobi", [0, 1])
@pytest.mark.parametrize("k,a,p,degenerate,bdiag", [
    (32, 32, 32, False, False), (10, 10, 10, True, False), (8, 5, 5, False, True),
    (16, 16, 0, False, False), (5, 3, 2, True, True), (32, 20, 20, False, True),
    # few active columns: the inverse-iteration workspace (2 n k) exceeds n^2
    (10, 1, 1, False, False), (32, 2, 2, False, True), (32, 1, 0, False, False)])
def test_rr_matches_scipy(jacobi, k, a, p, degenerate, bdiag):
    say(f"rr jacobi={jacobi} k={k} a={a} p={p} degenerate={degenerate} bdiag={bdiag}")
    A, B, X, R, P, lam = make_case(k, a, p, seed=k * 100 + a + p, degenerate=degenerate,
                                   bdiag=bdiag)
    ev, Xr = reference(A, B, X, R, P, lam)
    lg, Xg, flags, _ = run_kernel(A, B, X, R, P, lam, jacobi)
    assert flags[0] == 0 and flags[2] == 0
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-12
    np.testing.assert_allclose(Xg.T @ B @ Xg, np.eye(k), atol=1e-12)
    # same Ritz subspace (B-projectors)
    Pg, Pr = Xg @ Xg.T @ B, Xr @ Xr.T @ B
    assert np.linalg.norm(Pg - Pr) < 1e-9


@pytest.mark.parametrize("k,a,p,degenerate,bdiag", [
    # the large variant: pencils up to 128 (B_n in LDS), up to 192 (B_n in LDS
    # up to the LDS size, else in global memory), block widths above 32
    (40, 40, 40, False, False), (40, 40, 40, True, True), (33, 33, 0, False, False),
    (40, 8, 8, False, True), (44, 42, 42, False, False), (64, 64, 64, False, True),
    (48, 48, 48, True, False), (64, 3, 3, False, False), (36, 36, 36, False, True),
    (64, 40, 0, False, False)])
def test_rr_large_pencils_match_scipy(k, a, p, degenerate, bdiag):
    say(f"rr large k={k} a={a} p={p} degenerate={degenerate} bdiag={bdiag}")
    A, B, X, R, P, lam = make_case(k, a, p, seed=k * 100 + a + p + 1, degenerate=degenerate,
                                   bdiag=bdiag)
    ev, Xr = reference(A, B, X, R, P, lam)
    lg, Xg, flags, _ = run_kernel(A, B, X, R, P, lam, 0)
    assert flags[0] == 0 and flags[2] == 0
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-12
    np.testing.assert_allclose(Xg.T @ B @ Xg, np.eye(k), atol=1e-12)
    Pg, Pr = Xg @ Xg.T @ B, Xr @ Xr.T @ B
    assert np.linalg.norm(Pg - Pr) < 1e-9


def test_rr_large_explicit_x_and_retry():
    """The large variant with computed X blocks (explicit_x), then a pencil
    that is not positive definite solved again without P."""
    say("rr large explicit_x and retry without P (k=40)")
    A, B, X, R, P, lam = make_case(40, 40, 40, seed=17, bdiag=True)
    k = 40
    Xp = X @ (np.eye(k) + 1e-3 * np.random.default_rng(2).standard_normal((k, k)))
    Bas = np.hstack([Xp, R, P])
    ev = sl.eigh(Bas.T @ A @ Bas, Bas.T @ B @ Bas, eigvals_only=True)[:k]
    lg, Xg, flags, _ = run_kernel(A, B, Xp, R, P, np.zeros(k), 0, explicit_x=1)
    assert flags[0] == 0
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-11
    np.testing.assert_allclose(Xg.T @ B @ Xg, np.eye(k), atol=1e-11)
    # R B-orthogonal to X, so that [X R] stays positive definite with X
    # doubled (at k = 40 the cross block 2 X'B R of random R alone has norm
    # above one); a P column along X makes the full pencil indefinite
    R = R - X @ (X.T @ B @ R)
    Xs = 2.0 * X
    P[:, 0] = X[:, 0]
    _, _, flags, _ = run_kernel(A, B, Xs, R, P, lam, 0)
    assert flags[0] == 1 and flags[3] > 0
    lg, Xg, flags, coef = run_kernel(A, B, Xs, R, P, lam, 0, retry=1)
    assert flags[0] == 0 and flags[2] == -1 and flags[3] == 0
    assert np.all(coef[80:] == 0.0)
    ev, _ = reference(A, B, Xs, R, P[:, :0], lam)
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-12


@pytest.mark.parametrize("jacobi", [0, 1])
def test_rr_explicit_x_blocks(jacobi):
    """explicit_x uses the computed X'AX / X'BX: X need not be B-orthonormal."""
    say(f"rr explicit_x jacobi={jacobi}")
    A, B, X, R, P, lam = make_case(12, 12, 12, seed=7, bdiag=True)
    Xp = X @ (np.eye(12) + 1e-3 * np.random.default_rng(1).standard_normal((12, 12)))
    k = 12
    Bas = np.hstack([Xp, R, P])
    ev = sl.eigh(Bas.T @ A @ Bas, Bas.T @ B @ Bas, eigvals_only=True)[:k]
    lg, Xg, flags, _ = run_kernel(A, B, Xp, R, P, np.zeros(k), jacobi, explicit_x=1)
    assert flags[0] == 0
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-11
    np.testing.assert_allclose(Xg.T @ B @ Xg, np.eye(k), atol=1e-11)


def test_rr_explicit_x_not_positive_definite():
    """explicit_x with a zero X column: X'BX fails at pivot 3, reported as
    dsygv does for the pencil (n + 3); with retry_without_p, after P is
    dropped (p_info = -1, n = k + a)."""
    say("rr explicit_x with X'BX not positive definite")
    A, B, X, R, P, lam = make_case(6, 6, 6, seed=5)
    X[:, 2] = 0.0
    _, _, flags, _ = run_kernel(A, B, X, R, P, lam, 0, explicit_x=1)
    assert flags[0] == 1 and flags[1] == 0 and flags[3] == 18 + 3
    _, _, flags, _ = run_kernel(A, B, X, R, P, lam, 0, explicit_x=1, retry=1)
    assert flags[0] == 1 and flags[2] == -1 and flags[3] == 12 + 3


@pytest.mark.parametrize("jacobi", [0, 1])
def test_rr_drops_rank_deficient_p(jacobi):
    say(f"rr drop P jacobi={jacobi}")
    A, B, X, R, P, lam = make_case(6, 6, 6, seed=3)
    P[:, 3] = 0.0  # zero pivot: dpotrf fails at column 4
    lg, Xg, flags, coef = run_kernel(A, B, X, R, P, lam, jacobi)
    assert flags[0] == 0 and flags[2] == 4
    assert np.all(coef[12:] == 0.0)
    ev, _ = reference(A, B, X, R, P[:, :0], lam)
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-12


@pytest.mark.parametrize("jacobi", [0, 1])
def test_rr_reports_rank_deficient_r(jacobi):
    say(f"rr R failure jacobi={jacobi}")
    A, B, X, R, P, lam = make_case(6, 6, 6, seed=4)
    R[:, 2] = 0.0
    _, _, flags, _ = run_kernel(A, B, X, R, P, lam, jacobi)
    assert flags[0] == 1 and flags[1] == 3


def test_rr_retries_without_p():
    """X taken as B-orthonormal (implicit X block) while the data hold 2 X, and
    a P column along X: the pencil is indefinite.  Plain, the step fails like
    dsygv; with retry_without_p it is solved on [X R] alone (p_info = -1, P
    coefficients zero)."""
    say("rr retry without P")
    A, B, X, R, P, lam = make_case(8, 8, 4, seed=11)
    Xs = 2.0 * X
    P[:, 0] = X[:, 0]
    _, _, flags, _ = run_kernel(A, B, Xs, R, P, lam, 0)
    assert flags[0] == 1 and flags[3] > 0
    lg, Xg, flags, coef = run_kernel(A, B, Xs, R, P, lam, 0, retry=1)
    assert flags[0] == 0 and flags[2] == -1 and flags[3] == 0
    assert np.all(coef[16:] == 0.0)
    ev, _ = reference(A, B, Xs, R, P[:, :0], lam)
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-12


@pytest.mark.parametrize("hint", ["exact", "above", "below", "far_below", "tiny", "nan", "huge"])
def test_rr_eigenvalue_hints(hint):
    """The previous Ritz values (lam_in) only seed the eigenvalue brackets in
    the explicit-X mode: wrong, NaN or huge hints must give the same Ritz
    values as exact ones."""
    say(f"rr hint {hint}")
    k = 32
    A, B, X, R, P, lam = make_case(k, 32, 32, seed=21, bdiag=True)
    Bas = np.hstack([X, R, P])
    ev = sl.eigh(Bas.T @ A @ Bas, Bas.T @ B @ Bas, eigvals_only=True)[:k]
    rng = np.random.default_rng(5)
    hints = {"exact": ev, "above": ev * (1 + 1e-6), "below": ev * (1 - 1e-9),
             "far_below": ev * (1 - 0.3 * rng.random(k)), "tiny": np.full(k, 1e-300),
             "nan": np.full(k, np.nan), "huge": np.full(k, 1e300)}
    lg, Xg, flags, _ = run_kernel(A, B, X, R, P, hints[hint], 0, explicit_x=1)
    assert flags[0] == 0
    assert np.max(np.abs(lg - ev) / np.abs(ev)) < 1e-11
    np.testing.assert_allclose(Xg.T @ B @ Xg, np.eye(k), atol=1e-11)
