"""Pins the FP64 oracle (oracle/lobpcg_oracle.c) before anything is checked against it.

* bit-for-bit against the reference's own lobpcg_solve_double on the golden
  fixtures (tests/golden/ref_fixtures.npz, written by make_golden.py from
  the compiled reference) and, where oracle/_ref is built, live;
* against closed forms: the Dirichlet Laplacian spectrum (the 20^3 case),
  diag(1..40) -> 1..5 (serial_driver.c), diag(1..n) with 3 constraints ->
  4, 5, 6 (test_diag.m), and a generalized pencil against scipy's eigh;
* a planted mistake in the oracle's inputs must fail the checks.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from blopex_amd import problems as pb
from oracle import pyoracle as po


def _case(name):
    from tests.golden.make_golden import cases

    for nm, kw in cases(which="oracle"):  # the oracle's drand48: the reference's stream
        if nm == name:
            return kw
    raise KeyError(name)


@pytest.mark.parametrize("name", ["cfg1", "diag40", "testdiag", "gen12", "cfg2s"])
def test_oracle_matches_reference_fixtures(golden, name):
    r = po.solve(which="oracle", **_case(name))
    assert r.flag == int(golden[name + "_flag"])
    assert r.iterations == int(golden[name + "_iters"])
    np.testing.assert_array_equal(r.lam_hist, golden[name + "_lam_hist"])
    np.testing.assert_array_equal(r.res_hist, golden[name + "_res_hist"])
    if name + "_X" in golden:
        np.testing.assert_array_equal(r.X, golden[name + "_X"])


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_oracle_matches_reference_live():
    A = pb.laplacian_7pt(9, 8, 7, sigma=0.7, seed=11)
    kw = dict(A=A, X0=po.drand48_block(A.n, 6, 7), tdiag=1.0 / A.diagonal(),
              bdiag=pb.uniform(2, A.n, 0.5, 1.5), maxit=150)
    o = po.solve(which="oracle", **kw)
    r = po.solve(which="ref", **kw)
    assert (o.flag, o.iterations) == (r.flag, r.iterations)
    np.testing.assert_array_equal(o.lam_hist, r.lam_hist)
    np.testing.assert_array_equal(o.X, r.X)


def test_drand48_stream(golden):
    x = po.drand48_block(64, 1, 1)[:, 0]
    np.testing.assert_array_equal(x, golden["drand48_seed1_first64"])
    # the numpy jump-ahead restatement reproduces glibc's stream across block edges
    a = pb.drand48_block(5000, 3, 7)
    b = po.drand48_block(5000, 3, 7)
    np.testing.assert_array_equal(a, b)
    c = pb.drand48_stream(1, 40000, block=1024)
    np.testing.assert_array_equal(2 * c - 1, po.drand48_block(40000, 1, 1)[:, 0])


def test_laplacian_20cube_analytic():
    kw = _case("cfg1")
    r = po.solve(which="oracle", **kw)
    exact = pb.dirichlet_eigenvalues((20, 20, 20), 10)
    assert r.flag == 0
    assert np.max(np.abs(r.lam - exact) / exact) < 1e-8
    # the eigenvectors are B-orthonormal
    np.testing.assert_allclose(r.X.T @ r.X, np.eye(10), atol=1e-7)


def test_known_answers():
    r = po.solve(which="oracle", **_case("diag40"))
    np.testing.assert_allclose(r.lam, [1, 2, 3, 4, 5], rtol=1e-10)
    r = po.solve(which="oracle", **_case("testdiag"))
    np.testing.assert_allclose(r.lam, [4, 5, 6], rtol=1e-8)


def test_generalized_against_eigh():
    kw = _case("gen12")
    r = po.solve(which="oracle", **kw)
    A = kw["A"].to_scipy().toarray()
    ev = sl.eigh(A, np.diag(kw["bdiag"]), eigvals_only=True)[:8]
    assert r.flag == 0
    assert np.max(np.abs(r.lam - ev) / ev) < 1e-8
    # B-orthonormal eigenvectors
    np.testing.assert_allclose(r.X.T @ (kw["bdiag"][:, None] * r.X), np.eye(8), atol=1e-7)


def test_exit_codes():
    A = pb.laplacian_7pt(6, 6, 6)
    # duplicated initial columns: orthonormalisation of X fails -> -1, 0 iterations
    X = po.drand48_block(A.n, 3, 1)
    X[:, 1] = X[:, 0]
    r = po.solve(A=A, X0=X, maxit=10)
    assert (r.flag, r.iterations) == (-1, 0)
    # dependent constraints -> 3
    Y = np.ones((A.n, 2))
    r = po.solve(A=A, X0=po.drand48_block(A.n, 3, 1), Y=Y, maxit=10)
    assert r.flag == 3
    # max iterations reached -> -1 with maxit iterations
    r = po.solve(A=A, X0=po.drand48_block(A.n, 3, 1), maxit=2, atol=1e-14)
    assert (r.flag, r.iterations) == (-1, 2)


def test_planted_mistake_is_caught():
    """Flip one off-diagonal sign of A: the spectrum check must fail."""
    A = pb.laplacian_7pt(20, 20, 20)
    bad = pb.CSR(A.rowptr, A.colind, A.val.copy())
    off = np.nonzero(bad.val < 0)[0][5]
    bad.val[off] = -bad.val[off]
    r = po.solve(A=bad, X0=po.drand48_block(A.n, 10, 1), maxit=200)
    exact = pb.dirichlet_eigenvalues((20, 20, 20), 10)
    assert np.max(np.abs(r.lam - exact) / exact) > 1e-6


def test_spmm_oracle_matches_scipy():
    A = pb.laplacian_7pt(7, 5, 6, sigma=1.0, seed=2)
    X = np.random.default_rng(0).standard_normal((A.n, 4))
    np.testing.assert_allclose(po.spmm(A, X), A.to_scipy() @ X, rtol=1e-13, atol=1e-13)


def test_stencil_conventions():
    # sigma=0 reproduces the 6/-1 stencil of ij.c:3853-3873; rows sorted; symmetric
    A = pb.laplacian_7pt(4, 3, 5)
    S = A.to_scipy()
    assert np.all(A.diagonal() == 6.0)
    assert set(np.unique(A.val)) == {6.0, -1.0}
    assert (abs(S - S.T)).max() == 0
    for r in range(A.n):
        cols = A.colind[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0)
    # variable coefficients: symmetric, diagonally dominant, harmonic faces
    A = pb.laplacian_7pt(6, 5, 4, sigma=1.0, seed=9)
    S = A.to_scipy()
    assert abs(S - S.T).max() < 1e-15
    rowsum = np.asarray(S.sum(axis=1)).ravel()
    assert np.all(rowsum > -1e-12 * A.diagonal())  # zero inside, positive next to the boundary
    c = pb.lognormal_coefficients(A.n, 1.0, 9)
    # face between cells 0 and 1 (x-neighbours)
    assert np.isclose(-S[0, 1], 2 * c[0] * c[1] / (c[0] + c[1]))
    # 2D five-point: nnz = 5 s^2 - 4 s
    A2 = pb.laplacian_5pt(32, 32)
    assert A2.nnz == 5 * 32 * 32 - 4 * 32
    # 128^3 nnz formula 7n - 6*128^2 at a small size: 7n - 2(nx ny + ny nz + nx nz)
    A3 = pb.laplacian_7pt(8, 8, 8)
    assert A3.nnz == 7 * 512 - 6 * 64


@pytest.mark.parametrize("n,ld", [(1, 1), (7, 9), (24, 24)])
def test_oracle_upper_inv_against_numpy(n, ld):
    """oracle_upper_inv (the restatement of utilities_FortranMatrixUpperInv,
    fortran_matrix.c:1368-1416) inverts the upper triangle in place: equal to
    numpy's inverse of the triangular matrix to 1e-12, entries below the
    diagonal and padding rows of the leading dimension untouched."""
    rng = np.random.default_rng(n)
    U = np.triu(rng.standard_normal((n, n))) + np.diag(2.0 + rng.random(n))
    buf = np.full((ld, n), 7.0, order="F")
    buf[:n, :n] = U + np.tril(np.full((n, n), 5.0), -1)
    po.oracle_lib().oracle_upper_inv(n, ld, buf.ctypes.data)
    inv = np.linalg.inv(U)
    got = np.triu(buf[:n, :n])
    np.testing.assert_allclose(got, inv, rtol=1e-12, atol=1e-13 * np.abs(inv).max())
    assert np.all(np.tril(buf[:n, :n], -1)[np.tril_indices(n, -1)] == 5.0)
    assert np.all(buf[n:, :] == 7.0)
