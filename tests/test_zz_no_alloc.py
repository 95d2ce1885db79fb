"""hip_Multi_VectorInnerProdNoAlloc (the allocation-free MultiInnerProd) against
the interpreter's MultiInnerProd slot (mv_MultiVectorByMultiVector on HIP
blocks), both through the C ABI: two calls in a row, the second smaller and
masked, FP64 and FP32, a padded leading dimension; the same kernels with the
same partition, so the results are bitwise equal.  Sizes above the maximum
return an error without computing."""
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


def _both(X, Y, xmask, ymask, prec, gh_pad):
    lib = L.load()
    dX = bl.DeviceBlock.from_numpy(X, prec)
    dY = bl.DeviceBlock.from_numpy(Y, prec)
    ii = bl.interpreter()
    xx = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(dX.ptr, ctypes.c_void_p), 0)
    yy = lib.mv_MultiVectorWrap(ctypes.byref(ii), ctypes.cast(dY.ptr, ctypes.c_void_p), 0)
    for mv, m in ((xx, xmask), (yy, ymask)):
        if m is not None:
            mm = np.ascontiguousarray(m, dtype=np.int32)
            lib.mv_MultiVectorSetMask(mv, mm.ctypes.data)
    h = int(np.sum(xmask)) if xmask is not None else X.shape[1]
    w = int(np.sum(ymask)) if ymask is not None else Y.shape[1]
    gh = h + gh_pad
    ref = np.full((gh, w), -7.0, order="F")
    got = np.full((gh, w), -7.0, order="F")
    lib.mv_MultiVectorByMultiVector(xx, yy, gh, h, w, ref.ctypes.data)
    rc = lib.hip_Multi_VectorInnerProdNoAlloc(dX.ptr, dY.ptr, gh, h, w, got.ctypes.data)
    assert rc == 0, L.last_error()
    lib.mv_MultiVectorDestroy(xx)
    lib.mv_MultiVectorDestroy(yy)
    dX.close()
    dY.close()
    return ref, got, h, w


@pytest.mark.parametrize("prec", [L.HIPMV_FP64, L.HIPMV_FP32])
def test_no_alloc_inner_product_matches_the_slot(prec):
    rng = np.random.default_rng(11)
    cases = [((300001, 40, 33), None, None, 0),  # many partial sets, 3 x 3 tiles
             ((777, 9, 4), [1, 0, 1, 0, 0, 1, 1, 0, 1], [0, 1, 1, 1], 3)]  # smaller, masked
    for (n, kx, ky), xm, ym, pad in cases:
        say(f"no-alloc inner product n={n} kx={kx} ky={ky} prec={prec} masks={xm is not None}")
        X = rng.standard_normal((n, kx))
        Y = rng.standard_normal((n, ky))
        ref, got, h, w = _both(X, Y, xm, ym, prec, pad)
        np.testing.assert_array_equal(got, ref)
        Xa = X if xm is None else X[:, np.flatnonzero(xm)]
        Ya = Y if ym is None else Y[:, np.flatnonzero(ym)]
        if prec == L.HIPMV_FP32:
            Xa, Ya = Xa.astype(np.float32).astype(float), Ya.astype(np.float32).astype(float)
        want = Xa.T @ Ya
        scale = np.abs(Xa).T @ np.abs(Ya)
        assert np.max(np.abs(got[:h] - want) / scale) < (1e-13 if prec == L.HIPMV_FP64 else 1e-6)
        assert np.all(got[h:] == -7.0)  # rows past h of the leading dimension untouched


def test_no_alloc_rejects_results_above_the_maximum():
    say("no-alloc inner product: 65 columns rejected")
    lib = L.load()
    X = bl.DeviceBlock(100, 65)
    v = np.zeros((65, 65), order="F")
    assert lib.hip_Multi_VectorInnerProdNoAlloc(X.ptr, X.ptr, 65, 65, 65, v.ctypes.data) == 1
    assert "maximum" in L.last_error()
    X.close()
