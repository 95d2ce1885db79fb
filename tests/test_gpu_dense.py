"""GPU: the MatMultiVec operator slot of the serial driver
(blopex_serial_double/pcg_multi/pcg_multi.c:167-193) through
HipDenseMatMultiVec -- a dense real A on FP64 MFMA, or, with no matrix, the
1D Laplacian of matmultivec.c:11-48.

Bars: products in FP64 agree with numpy to 1e-13 of |A| |X| (summation order
differs); FP32 blocks (FP64 arithmetic, one rounding at the store) to 2^-23 of
|A| |X|; the stencil to one rounding (exact on FP32 blocks).  Operand
mismatches abort the process, as the reference's assert does, so they are not
exercised here.  The stock serial driver (40 x 40 diag(1..40) stored dense,
k = 5, drand48 seed 1, tol 1e-6, maxit 100) reproduces the compiled
reference's flag, eigenvalues (1e-8 relative), first Ritz values (1e-12) and
iteration count (within max(2, 5%)).
"""
import ctypes
import sys

import numpy as np
import pytest

from blopex_amd import _lib as L
from blopex_amd import lobpcg as bl

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _op():
    return ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(
        L.symbol_address("HipDenseMatMultiVec"))


def _vp(p):
    return None if p is None else ctypes.cast(p, ctypes.c_void_p)


def _masked_product(n, k, mask, prec, seed, A=None):
    """Y(active) = op(A) X(active) on blocks of precision prec; returns the
    host X (as the device holds it), the device Y and the untouched Y0."""
    lib = L.load()
    rng = np.random.default_rng(seed)
    Xh = rng.standard_normal((n, k))
    Y0 = rng.standard_normal((n, k))
    if prec == L.HIPMV_FP32:
        Xh = Xh.astype(np.float32).astype(np.float64)
        Y0 = Y0.astype(np.float32).astype(np.float64)
    X, Y = bl.DeviceBlock.from_numpy(Xh, prec), bl.DeviceBlock.from_numpy(Y0, prec)
    dA = bl.DeviceDense(A, n)
    ii = bl.interpreter()
    xx = lib.mv_MultiVectorWrap(ctypes.byref(ii), _vp(X.ptr), 0)
    yy = lib.mv_MultiVectorWrap(ctypes.byref(ii), _vp(Y.ptr), 0)
    m = np.asarray(mask, dtype=np.int32)
    lib.mv_MultiVectorSetMask(xx, m.ctypes.data)
    lib.mv_MultiVectorSetMask(yy, m.ctypes.data)
    _op()(_vp(dA.ptr), _vp(X.ptr), _vp(Y.ptr))
    got = Y.to_numpy()
    for h in (xx, yy):
        lib.mv_MultiVectorDestroy(h)
    for b in (X, Y, dA):
        b.close()
    return Xh, got, Y0


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
@pytest.mark.parametrize("n,k", [(37, 3), (300, 5), (4000, 32), (1000, 70)])
def test_dense_real_operator(prec, n, k):
    """One inner-index slice (n = 37, 300) and several with the ordered sum
    (n = 4000); more than 64 active columns run as two column groups."""
    say(f"dense real operator n={n} k={k} prec={prec}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    mask = np.ones(k, dtype=np.int32)
    mask[1::3] = 0
    Xh, got, Y0 = _masked_product(n, k, mask, prec, seed=k, A=A)
    act = mask == 1
    ref = A @ Xh[:, act]
    scale = np.abs(A) @ np.abs(Xh[:, act])
    tol = 1e-13 if prec == L.HIPMV_FP64 else 2.0 ** -23
    assert np.max(np.abs(got[:, act] - ref) / scale) < tol
    np.testing.assert_array_equal(got[:, ~act], Y0[:, ~act])


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
def test_null_matrix_applies_1d_laplacian(prec):
    say(f"1D Laplacian n=50 k=4 prec={prec}")
    n = 50
    Xh, got, Y0 = _masked_product(n, 4, [1, 0, 1, 1], prec, seed=3)
    L1 = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    act = np.array([True, False, True, True])
    ref = L1 @ Xh[:, act]
    if prec == L.HIPMV_FP32:
        # FP32 inputs: the three-term sum is exact in FP64, one rounding at the store
        np.testing.assert_array_equal(got[:, act], ref.astype(np.float32).astype(np.float64))
    else:
        scale = np.abs(L1) @ np.abs(Xh[:, act])
        assert np.max(np.abs(got[:, act] - ref) / scale) < 2.0 ** -52
    np.testing.assert_array_equal(got[:, ~act], Y0[:, ~act])


def test_stock_serial_driver_dense_diag40(golden):
    """serial_driver.c:44-175 on the device: the same dense operator slot."""
    say("stock serial driver: dense diag(1..40), k=5")
    dA = bl.DeviceDense(np.diag(np.arange(1.0, 41.0)))
    X = bl.DeviceBlock(40, 5)
    X.set_random(1)
    r = bl.solve_device(dA, X, tol=1e-6, maxit=100)
    say(f"  flag {r.flag} iterations {r.iterations} (reference {int(golden['diag40_iters'])})")
    assert r.flag == int(golden["diag40_flag"]) == 0
    ref_it = int(golden["diag40_iters"])
    assert abs(r.iterations - ref_it) <= max(2, ref_it // 20)
    ref_lam = golden["diag40_lam"]
    assert np.max(np.abs(r.lam - ref_lam) / ref_lam) < 1e-8
    np.testing.assert_allclose(r.lam_hist[:, 0], golden["diag40_lam_hist"][:, 0], rtol=1e-12)
    np.testing.assert_allclose(r.lam, np.arange(1.0, 6.0), rtol=1e-8)
    X.close()
    dA.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_1d_laplacian_solve_closed_form(mode):
    """A = NULL: the 1D Laplacian, eigenvalues 2 - 2 cos(j pi / (n + 1))."""
    n, k = 100, 5
    say(f"1D Laplacian solve n={n} k={k} {mode}")
    dA = bl.DeviceDense(None, n)
    X = bl.DeviceBlock(n, k)
    X.set_random(1)
    r = bl.solve_device(dA, X, tol=1e-8, maxit=2000, mode=mode)
    say(f"  flag {r.flag} iterations {r.iterations}")
    exact = 2.0 - 2.0 * np.cos(np.arange(1, k + 1) * np.pi / (n + 1))
    assert r.flag == 0
    assert np.max(np.abs(r.lam - exact) / exact) < 1e-8
    X.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_dense_random_symmetric_solve(mode):
    """A dense symmetric A = Q diag(d) Q^T, n = 3000, k = 8; the eigenvalues
    are d's smallest, the residuals meet the tolerance."""
    n, k = 3000, 8
    say(f"dense symmetric solve n={n} k={k} {mode}")
    rng = np.random.default_rng(11)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.concatenate([np.linspace(1.0, 2.0, k), np.linspace(4.0, 400.0, n - k)])
    A = (Q * d) @ Q.T
    A = (A + A.T) / 2
    dA = bl.DeviceDense(A)
    X = bl.DeviceBlock(n, k)
    X.set_random(1)
    r = bl.solve_device(dA, X, tol=1e-6, maxit=500, mode=mode)
    say(f"  flag {r.flag} iterations {r.iterations}")
    assert r.flag == 0
    assert np.max(np.abs(r.lam - d[:k]) / d[:k]) < 1e-8
    Xh = X.to_numpy()
    res = np.linalg.norm(A @ Xh - Xh * r.lam, axis=0)
    assert np.all(res <= 1e-6 * 1.01)
    X.close()
    dA.close()
