"""CPU checks of the complex Hermitian path's host side (no GPU):

* the complex small-dense routines (zutilities_FortranMatrix*) against
  numpy: conjugate-transpose products, Hermitian symmetrisation, in-place
  conjugate transpose, triangular inverse, copies, moduli;
* the library's lobpcg_solve_complex (the generic algorithm with these
  routines and scipy's zpotrf / zhegv) driven over the reference's own complex
  serial backend (oracle/_ref) against the reference's lobpcg_solve_complex on
  the same backend: same flag and iteration count, eigenvalues and residual
  histories to rounding, on the complex driver's own data (fixtures);
* the reference's results themselves against numpy's eigvalsh."""
import ctypes
import os

import numpy as np
import pytest

from blopex_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_COMPLEX = os.path.join(ROOT, "oracle", "_ref", "libblopex_ref_complex.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    return L.load()


def _fm(a: np.ndarray) -> L.FortranMatrix:
    return L.FortranMatrix(a.shape[0], a.shape[0], a.shape[1], a.ctypes.data, 0)


def _rand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_complex_multiply(lib, ta, tb):
    rng = np.random.default_rng(ta * 2 + tb)
    A = _rand(rng, 5, 4) if not ta else _rand(rng, 4, 5)
    B = _rand(rng, 4, 3) if not tb else _rand(rng, 3, 4)
    C = np.zeros((5, 3), dtype=np.complex128, order="F")
    fa, fb, fc = _fm(A), _fm(B), _fm(C)
    lib.zutilities_FortranMatrixMultiply(ctypes.byref(fa), ta, ctypes.byref(fb), tb,
                                         ctypes.byref(fc))
    opa = A.conj().T if ta else A
    opb = B.conj().T if tb else B
    np.testing.assert_allclose(C, opa @ opb, rtol=1e-13, atol=1e-13)


def test_complex_symmetrize_transpose_upperinv_copy_abs(lib):
    rng = np.random.default_rng(9)
    M = _rand(rng, 6, 6)
    S = M.copy(order="F")
    lib.zutilities_FortranMatrixSymmetrize(ctypes.byref(_fm(S)))
    H = (M + M.conj().T) / 2
    # lower triangle from (lower + conj(upper)) / 2, upper its conjugate, real diagonal
    np.testing.assert_allclose(S, H, rtol=1e-15, atol=1e-15)
    assert np.all(S.diagonal().imag == 0.0)
    T = M.copy(order="F")
    lib.zutilities_FortranMatrixTransposeSquare(ctypes.byref(_fm(T)))
    np.testing.assert_array_equal(T, M.conj().T)
    U = np.asfortranarray(np.triu(M) + 4 * np.eye(6))
    Ui = U.copy(order="F")
    lib.zutilities_FortranMatrixUpperInv(ctypes.byref(_fm(Ui)))
    np.testing.assert_allclose(Ui @ U, np.eye(6), atol=1e-13)
    D = np.zeros((6, 6), dtype=np.complex128, order="F")
    lib.zutilities_FortranMatrixCopy(ctypes.byref(_fm(M)), 1, ctypes.byref(_fm(D)))
    np.testing.assert_array_equal(D, M.conj().T)
    idx = np.array([3, 1], dtype=np.int32)
    E = np.zeros((6, 2), dtype=np.complex128, order="F")
    lib.zutilities_FortranMatrixIndexCopy(idx.ctypes.data, ctypes.byref(_fm(M)), 1,
                                          ctypes.byref(_fm(E)))
    np.testing.assert_array_equal(E, M[[2, 0], :].conj().T)
    assert lib.zutilities_FortranMatrixAbs(ctypes.byref(_fm(M)), 2, 3) == pytest.approx(
        abs(M[1, 2]), rel=1e-15)


def _ref():
    if not os.path.exists(REF_COMPLEX):
        pytest.skip("oracle/_ref/libblopex_ref_complex.so not built")
    r = ctypes.CDLL(REF_COMPLEX)
    r.ref_complex_block.restype = ctypes.c_void_p
    r.ref_complex_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    r.ref_complex_block_get.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    r.ref_complex_block_destroy.argtypes = [ctypes.c_void_p]
    r.ref_complex_setup_interpreter.argtypes = [ctypes.c_void_p]
    r.ref_complex_matmultivec.restype = ctypes.c_void_p
    return r


@pytest.mark.parametrize("name", ["ztest1", "ztest2"])
def test_generic_complex_solver_matches_reference(lib, golden, name):
    from oracle import pyoracle as po

    ref = _ref()
    A, X0 = golden[name + "_A"], golden[name + "_X0"]
    n, k = X0.shape
    # the fixture is the reference's own run; its eigenvalues are A's (numpy)
    ev = np.linalg.eigvalsh(A)[:k]
    assert np.max(np.abs(golden[name + "_lam"] - ev) / np.abs(ev)) < 1e-12
    a = ref.ref_complex_block(n, n, np.asfortranarray(A).ctypes.data)
    x = ref.ref_complex_block(n, k, np.asfortranarray(X0).ctypes.data)
    ii = L.Interpreter()
    ref.ref_complex_setup_interpreter(ctypes.byref(ii))
    xx = lib.mv_MultiVectorWrap(ctypes.byref(ii), x, 0)
    zp, zh = po.zlapack_pointers()
    maxit = 1000
    lam = np.zeros(k, dtype=np.complex128)
    lh = np.zeros((k, maxit + 1), dtype=np.complex128, order="F")
    res = np.zeros(k)
    rh = np.zeros((k, maxit + 1), order="F")
    it = ctypes.c_int(0)
    flag = lib.lobpcg_solve_complex(xx, a, ref.ref_complex_matmultivec(), None, None, None, None,
                                    None, L.BlasLapack(None, None, zp, zh),
                                    L.Tolerance(1e-6, 1e-50), maxit, 0, ctypes.byref(it),
                                    lam.ctypes.data, lh.ctypes.data, k, res.ctypes.data,
                                    rh.ctypes.data, k)
    lib.mv_MultiVectorDestroy(xx)
    X = np.zeros((n, k), dtype=np.complex128, order="F")
    ref.ref_complex_block_get(x, X.ctypes.data)
    ref.ref_complex_block_destroy(x)
    ref.ref_complex_block_destroy(a)
    m = it.value
    assert flag == int(golden[name + "_flag"]) == 0
    assert m == int(golden[name + "_iters"])
    assert np.all(lam.imag == 0.0)
    np.testing.assert_allclose(lam.real, golden[name + "_lam"], rtol=1e-12)
    np.testing.assert_allclose(lh[:, : m + 1].real, golden[name + "_lam_hist"], rtol=1e-9)
    # residual norms near 1e-7 carry the rounding of the products that form
    # them (the symmetrisation and triangular inverse sum in their own order)
    np.testing.assert_allclose(rh[:, : m + 1], golden[name + "_res_hist"], rtol=1e-4, atol=1e-12)
    # X: orthonormal eigenvectors of A
    assert np.max(np.abs(X.conj().T @ X - np.eye(k))) < 1e-12
    r = np.linalg.norm(A @ X - X * lam.real, axis=0)
    assert np.all(r <= 1e-6 + 1e-12)
