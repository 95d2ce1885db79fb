"""GPU checks of complex Hermitian solves (lobpcg_solve_complex on HIPMV_C64
blocks) through the C ABI.

* Every multivector slot the generic algorithm uses, on masked complex
  columns, against numpy: X^H Y into a sub-block, its diagonal, X C, Y += X C,
  X diag(d), real axpy, copy, clear; the drand48 fill bit-exact against the
  stream (real and imaginary parts consecutive, as the reference's complex
  SetRandomValues draws them); the dense complex operator.
* Solves on the reference's complex driver data (fixtures from the compiled
  reference): same flag, iteration count within max(2, 5%), eigenvalues to
  1e-10 relative (numpy eigvalsh agrees with the fixture to 1e-12), the same
  initial Rayleigh-Ritz values, residuals at the tolerance, orthonormal X.
* The driver's random mode (A = Hermitian part of a drand48 fill, X0 the
  same stream) against the reference run in the test on the same inputs.
* The device path (the real form of the pencil on the real device solver)
  beyond the driver's cases, against built-in spectra: a generalized pencil
  with constraints, doubly degenerate eigenvalues (the returned vectors are
  B-orthonormal and span the eigenspaces), a preconditioner, a 4000-row
  problem, MIXED mode;
  and a block wider than the device solver takes, which runs the generic
  algorithm with the caller's zpotrf/zhegv.
"""
import ctypes
import sys

import numpy as np
import pytest

from blopex_amd import _lib as L
from blopex_amd import lobpcg as bl
from blopex_amd import problems as pb

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _wrap(ii, blk):
    return L.load().mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(blk.ptr, ctypes.c_void_p), 0)


def _crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def test_complex_interpreter_slots_with_masks():
    say("complex interpreter slots")
    lib = L.load()
    rng = np.random.default_rng(3)
    n, k = 2001, 6
    Xh, Yh = _crand(rng, n, k), _crand(rng, n, k)
    X, Y = bl.DeviceBlock.from_numpy(Xh), bl.DeviceBlock.from_numpy(Yh)
    assert X.ptr.contents.precision == L.HIPMV_C64
    np.testing.assert_array_equal(X.to_numpy(), Xh)
    ii = bl.interpreter()
    xx, yy = _wrap(ii, X), _wrap(ii, Y)
    mx = np.array([1, 0, 1, 1, 0, 1], dtype=np.int32)
    my = np.array([0, 1, 1, 0, 1, 1], dtype=np.int32)
    ax, ay = np.nonzero(mx)[0], np.nonzero(my)[0]
    lib.mv_MultiVectorSetMask(xx, mx.ctypes.data)
    lib.mv_MultiVectorSetMask(yy, my.ctypes.data)
    tol = 1e-12
    gh = 7
    G = np.full((gh, 4), np.nan + 0j, order="F")
    lib.mv_MultiVectorByMultiVector(xx, yy, gh, 4, 4, G.ctypes.data)
    ref = Xh[:, ax].conj().T @ Yh[:, ay]
    np.testing.assert_allclose(G[:4, :], ref, rtol=tol, atol=tol * np.abs(ref).max())
    assert np.all(np.isnan(G[4:, :]))
    d = np.full(6, -7.0 + 0j)
    lib.mv_MultiVectorByMultiVectorDiag(xx, yy, mx.ctypes.data, 6, d.ctypes.data)
    np.testing.assert_allclose(d[ax], np.sum(Xh[:, ax].conj() * Yh[:, ay], axis=0), rtol=tol,
                               atol=tol * n)
    assert np.all(d[mx == 0] == -7.0)
    C = _crand(rng, 5, 4)
    lib.mv_MultiVectorByMatrix(xx, 5, 4, 4, C.ctypes.data, yy)
    Yn = Yh.copy()
    Yn[:, ay] = Xh[:, ax] @ C[:4, :]
    np.testing.assert_allclose(Y.to_numpy(), Yn, rtol=tol, atol=tol * 10)
    lib.mv_MultiVectorXapy(xx, 5, 4, 4, C.ctypes.data, yy)
    Yn[:, ay] += Xh[:, ax] @ C[:4, :]
    np.testing.assert_allclose(Y.to_numpy(), Yn, rtol=tol, atol=tol * 20)
    dd = np.arange(1.0, 7.0) + 1j * np.arange(6.0, 0.0, -1.0)
    lib.mv_MultiVectorByDiagonal(xx, mx.ctypes.data, 6, dd.ctypes.data, yy)
    Yn[:, ay] = Xh[:, ax] * dd[ax]
    np.testing.assert_allclose(Y.to_numpy(), Yn, rtol=tol, atol=tol)
    lib.mv_MultiVectorAxpy(-2.5, xx, yy)
    Yn[:, ay] += -2.5 * Xh[:, ax]
    np.testing.assert_allclose(Y.to_numpy(), Yn, rtol=tol, atol=tol * 4)
    lib.mv_MultiVectorCopy(xx, yy)
    Yn[:, ay] = Xh[:, ax]
    np.testing.assert_array_equal(Y.to_numpy(), Yn)
    lib.mv_MultiVectorClear(yy)
    Yn[:, ay] = 0.0
    np.testing.assert_array_equal(Y.to_numpy(), Yn)
    zz = lib.mv_MultiVectorCreateCopy(xx, 1)
    Z = L.HipMultiVector.from_address(lib.mv_MultiVectorGetData(zz))
    assert Z.precision == L.HIPMV_C64 and Z.num_vectors == 6
    lib.mv_MultiVectorDestroy(zz)
    # drand48 fill of the active columns: consecutive values, real then imaginary
    lib.mv_MultiVectorSetRandom(yy, 5)
    u = 2.0 * pb.drand48_stream(5, 2 * n * len(ay)) - 1.0
    Yn[:, ay] = (u[0::2] + 1j * u[1::2]).reshape(len(ay), n).T
    np.testing.assert_array_equal(Y.to_numpy(), Yn)
    for h in (xx, yy):
        lib.mv_MultiVectorDestroy(h)
    X.close()
    Y.close()


def test_dense_complex_operator():
    say("dense complex operator")
    lib = L.load()
    rng = np.random.default_rng(4)
    n = 300
    Ah, Xh = _crand(rng, n, n), _crand(rng, n, 5)
    A, X, Y = (bl.DeviceBlock.from_numpy(Ah), bl.DeviceBlock.from_numpy(Xh),
               bl.DeviceBlock.from_numpy(np.zeros((n, 5), dtype=np.complex128)))
    ii = bl.interpreter()
    xx, yy = _wrap(ii, X), _wrap(ii, Y)
    m = np.array([1, 1, 0, 1, 0], dtype=np.int32)
    lib.mv_MultiVectorSetMask(xx, m.ctypes.data)
    lib.mv_MultiVectorSetMask(yy, m.ctypes.data)
    op = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(
        L.symbol_address("HipDenseMatMultiVec"))
    op(ctypes.cast(A.ptr, ctypes.c_void_p), ctypes.cast(X.ptr, ctypes.c_void_p),
       ctypes.cast(Y.ptr, ctypes.c_void_p))
    ref = np.zeros((n, 5), dtype=np.complex128)
    ref[:, m == 1] = Ah @ Xh[:, m == 1]
    np.testing.assert_allclose(Y.to_numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    for h in (xx, yy):
        lib.mv_MultiVectorDestroy(h)
    for b in (A, X, Y):
        b.close()


@pytest.mark.parametrize("name", ["ztest1", "ztest2"])
def test_complex_solve_matches_reference(golden, name):
    say(f"complex solve {name}")
    A, X0 = golden[name + "_A"], golden[name + "_X0"]
    k = X0.shape[1]
    r = bl.lobpcg_complex(X0, A, tol=1e-6, maxit=1000)
    ref_it = int(golden[name + "_iters"])
    say(f"  flag {r.flag} iterations {r.iterations} (reference {ref_it})")
    assert r.flag == int(golden[name + "_flag"]) == 0
    assert r.stats["iterations_fp64"] == r.iterations  # the device solver ran
    assert abs(r.iterations - ref_it) <= max(2, ref_it // 20)
    assert np.all(r.lam.imag == 0.0)
    ref_lam = golden[name + "_lam"]
    assert np.max(np.abs(r.lam.real - ref_lam) / np.abs(ref_lam)) < 1e-10
    np.testing.assert_allclose(r.lam_hist[:, 0].real, golden[name + "_lam_hist"][:, 0], rtol=1e-12)
    assert np.all(r.res <= 1e-6 + 1e-15)
    X = r.X
    assert np.max(np.abs(X.conj().T @ X - np.eye(k))) < 1e-12
    # the residuals of the returned pairs meet the tolerance (the reported
    # norms of columns locked early are the ones measured when they locked,
    # as in lobpcg.c)
    res = np.linalg.norm(A @ X - X * r.lam.real, axis=0)
    assert np.all(res <= 1e-6 + 1e-12)


def test_complex_solve_random_hermitian_mode():
    """The complex driver's random mode: A = (R + R^H)/2 with R a drand48
    fill (seed 1), X0 a fill of the same seed; n = 100, k = 10."""
    from oracle import pyoracle as po

    say("complex solve, random Hermitian (driver mode R)")
    n, k = 100, 10
    u = 2.0 * pb.drand48_stream(1, 2 * n * n) - 1.0
    R = (u[0::2] + 1j * u[1::2]).reshape(n, n).T
    A = np.asfortranarray((R + R.conj().T) / 2)
    v = 2.0 * pb.drand48_stream(1, 2 * n * k) - 1.0
    X0 = np.asfortranarray((v[0::2] + 1j * v[1::2]).reshape(k, n).T)
    o = po.ref_solve_complex(A, X0, maxit=1000)
    r = bl.lobpcg_complex(X0, A, tol=1e-6, maxit=1000)
    say(f"  flag {r.flag} iterations {r.iterations} (reference {o.flag}, {o.iterations})")
    assert r.flag == o.flag == 0
    assert abs(r.iterations - o.iterations) <= max(2, o.iterations // 20)
    ev = np.linalg.eigvalsh(A)[:k]
    assert np.max(np.abs(r.lam.real - ev) / np.abs(ev)) < 1e-10
    assert np.max(np.abs(r.lam.real - o.lam.real) / np.abs(ev)) < 1e-10


def _hermitian_with_spectrum(rng, d):
    n = len(d)
    Q, _ = np.linalg.qr(_crand(rng, n, n))
    return np.asfortranarray((Q * d) @ Q.conj().T), Q


def test_complex_generalized_with_constraints():
    """(A, B) with B dense Hermitian positive definite and, as constraints,
    its two lowest eigenvectors (the residual lobpcg.c measures is that of
    the unprojected pencil, so the constraints are eigenvectors, as in the
    reference's test_diag.m): the next k eigenpairs, built in."""
    say("complex generalized pencil with constraints")
    rng = np.random.default_rng(21)
    n, k, m = 400, 6, 2
    d = np.concatenate([np.arange(1.0, 13.0), np.linspace(40.0, 400.0, n - 12)])
    Cb = _crand(rng, n, n) * (0.2 / n)
    B = np.asfortranarray(np.eye(n) + (Cb + Cb.conj().T) / 2)
    s, U = np.linalg.eigh(B)
    Q, _ = np.linalg.qr(_crand(rng, n, n))
    Z = (U / np.sqrt(s)) @ U.conj().T @ Q  # Z^H B Z = I
    A = np.asfortranarray(B @ (Z * d) @ Z.conj().T @ B)
    A = (A + A.conj().T) / 2
    Y = np.asfortranarray(Z[:, :m])
    r = bl.lobpcg_complex(_crand(rng, n, k), A, tol=1e-9, maxit=500, B=B, Y=Y)
    say(f"  flag {r.flag} iterations {r.iterations}")
    assert r.flag == 0 and r.stats["iterations_fp64"] == r.iterations
    np.testing.assert_allclose(r.lam.real, d[m:m + k], rtol=1e-9)
    X = r.X
    assert np.max(np.abs(X.conj().T @ B @ X - np.eye(k))) < 1e-10
    assert np.max(np.abs(Y.conj().T @ B @ X)) < 1e-10
    res = np.linalg.norm(A @ X - B @ X * r.lam.real, axis=0)
    assert np.all(res <= 1.5e-9 * np.abs(A).max())


def test_complex_degenerate_eigenvalues():
    """Eigenvalues 1, 1, 2, 2, 3, 3: each pair's real form has four Ritz
    vectors; the returned complex vectors are orthonormal and span each
    two-dimensional eigenspace."""
    say("complex doubly degenerate eigenvalues")
    rng = np.random.default_rng(22)
    n, k = 200, 6
    d = np.concatenate([[1.0, 1.0, 2.0, 2.0, 3.0, 3.0], np.linspace(20.0, 200.0, n - 6)])
    A, Q = _hermitian_with_spectrum(rng, d)
    r = bl.lobpcg_complex(_crand(rng, n, k), A, tol=1e-10, maxit=400)
    say(f"  flag {r.flag} iterations {r.iterations}")
    assert r.flag == 0
    np.testing.assert_allclose(r.lam.real, d[:k], rtol=1e-12)
    X = r.X
    assert np.max(np.abs(X.conj().T @ X - np.eye(k))) < 1e-12
    for j in range(0, k, 2):  # the projector onto each eigenspace from X and from Q agree
        Px, Pq = X[:, j:j + 2] @ X[:, j:j + 2].conj().T, Q[:, j:j + 2] @ Q[:, j:j + 2].conj().T
        assert np.max(np.abs(Px - Pq)) < 1e-8


def test_complex_larger_problem():
    """n = 4000, k = 16: eigenvalues against the known spectrum, residuals
    recomputed on the host."""
    say("complex n=4000 k=16")
    rng = np.random.default_rng(23)
    n, k = 4000, 16
    d = np.concatenate([np.linspace(1.0, 4.0, k), np.linspace(30.0, 3000.0, n - k)])
    A, _ = _hermitian_with_spectrum(rng, d)
    r = bl.lobpcg_complex(_crand(rng, n, k), A, tol=1e-8, maxit=300)
    say(f"  flag {r.flag} iterations {r.iterations} in {r.stats['seconds']:.3f} s")
    assert r.flag == 0 and r.stats["iterations_fp64"] == r.iterations
    np.testing.assert_allclose(r.lam.real, d[:k], rtol=1e-10)
    res = np.linalg.norm(A @ r.X - r.X * r.lam.real, axis=0)
    assert np.all(res <= 1.5e-8)


def test_complex_preconditioned():
    """A preconditioner T (here the inverse diagonal of A as a dense complex
    operator) through the adapter: the known spectrum, in fewer iterations
    than without T on a badly scaled A."""
    say("complex preconditioned (Jacobi as a dense operator)")
    rng = np.random.default_rng(25)
    n, k = 600, 8
    A0, _ = _hermitian_with_spectrum(rng, np.concatenate([np.linspace(1.0, 2.0, k),
                                                          np.linspace(10.0, 100.0, n - k)]))
    s = np.exp(rng.uniform(-2.0, 2.0, n))  # diagonal scaling: A = S A0 S
    A = np.asfortranarray((s[:, None] * A0) * s[None, :])
    T = np.asfortranarray(np.diag(1.0 / np.diag(A).real).astype(np.complex128))
    X0 = _crand(rng, n, k)
    r = bl.lobpcg_complex(X0, A, tol=1e-8, maxit=1000, T=T)
    r0 = bl.lobpcg_complex(X0, A, tol=1e-8, maxit=1000)
    say(f"  with T: flag {r.flag} iterations {r.iterations}; without: {r0.flag} {r0.iterations}")
    assert r.flag == 0 and r.stats["iterations_fp64"] == r.iterations
    ev = np.linalg.eigvalsh(A)[:k]
    assert np.max(np.abs(r.lam.real - ev) / np.abs(ev)) < 1e-10
    assert r.iterations < r0.iterations
    res = np.linalg.norm(A @ r.X - r.X * r.lam.real, axis=0)
    assert np.all(res <= 1.5e-8)


def test_complex_mixed_mode(golden):
    """MIXED mode on the driver's second case: FP32 iterations first, then
    the reference's eigenvalues to 1e-8 (the bar of the real MIXED parity
    tests at the same 1e-6 residual tolerance) after the FP64 finish."""
    say("complex MIXED mode (ztest2)")
    A, X0 = golden["ztest2_A"], golden["ztest2_X0"]
    r = bl.lobpcg_complex(X0, A, tol=1e-6, maxit=1000, mode="mixed")
    say(f"  flag {r.flag} iterations {r.iterations} (fp32 {r.stats['iterations_fp32']})")
    assert r.flag == 0 and r.stats["iterations_fp32"] > 0
    ref_lam = golden["ztest2_lam"]
    assert np.max(np.abs(r.lam.real - ref_lam) / np.abs(ref_lam)) < 1e-8
    assert np.all(r.res <= 1e-6 + 1e-15)


def test_complex_wide_block_runs_generic_path():
    """33 columns (66 in real form, above the device solver's 64): the
    generic algorithm with scipy's zpotrf/zhegv, against the spectrum."""
    say("complex k=33 (generic path)")
    rng = np.random.default_rng(24)
    n, k = 300, 33
    d = np.concatenate([np.linspace(1.0, 3.0, k + 3), np.linspace(20.0, 300.0, n - k - 3)])
    A, _ = _hermitian_with_spectrum(rng, d)
    r = bl.lobpcg_complex(_crand(rng, n, k), A, tol=1e-8, maxit=300)
    say(f"  flag {r.flag} iterations {r.iterations}")
    assert r.flag == 0 and r.stats["iterations_fp64"] == 0
    np.testing.assert_allclose(r.lam.real, d[:k], rtol=1e-10)
